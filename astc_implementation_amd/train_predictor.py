"""Trains the candidate predictor of the LEARNED policy on the GPU.

Labels come from the encoder's own plain (exhaustive compat) search on
synthetic images that are not part of any benchmark (seeds >= 10000): half
gradient + noise + rectangles images, a quarter normal maps, a quarter flat
colour + noise.  Inputs are the blocks' normalised RGB texels.  Every
forward / backward / Adam step runs in the library's HIP kernels
(``predictor.Predictor``).

    python -m astc_implementation_amd.train_predictor --block 4x4 \\
        --blocks 2000000 --out astc_implementation_amd/predictor_4x4.npz
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import astc_implementation_amd as A  # noqa: E402
from astc_implementation_amd import synth  # noqa: E402
from astc_implementation_amd.predictor import Predictor  # noqa: E402


def training_image(i, size):
    kind = i % 4
    seed = 10000 + i
    if kind < 2:
        return synth.gradient_image(size, size, seed)
    if kind == 2:
        return synth.normal_map_image(size, size, seed)
    return synth.flat_noise_image(size, size, seed)


def collect(P, enc, n_blocks, size=1024, log=print):
    """Encodes training images with the plain search until n_blocks blocks
    are labelled; returns device tensors X (bf16 features) and Y (int32
    head labels: pc - 1, best seed for 2 / 3 / 4 partitions, trial mode)."""
    import torch
    xs, ys, total, i = [], [], 0, 0
    while total < n_blocks:
        img = training_image(i, size)
        h, w = img.shape[:2]
        enc.configure(w, h, P.bx, P.by, labels=True)
        d_img = torch.from_numpy(img.reshape(-1).copy()).cuda()
        d_out = torch.empty(enc.num_blocks * 16, dtype=torch.uint8, device="cuda")
        enc.encode_device(d_img.data_ptr(), d_out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        lab = enc.labels()
        y = np.stack([lab[:, 0] - 1, lab[:, 3], lab[:, 4], lab[:, 5], lab[:, 2]], axis=1).astype(np.int32)
        xs.append(P.features(d_img, w, h, 0, enc.num_blocks))
        ys.append(torch.from_numpy(y).cuda())
        total += enc.num_blocks
        i += 1
        log("image %d: %d blocks labelled" % (i, total))
    return torch.cat(xs)[:n_blocks], torch.cat(ys)[:n_blocks]


def recall(P, X, Y, ks=(1, 2, 4, 16), batch=8192):
    """Fraction of rows whose label is within the top-k of its head."""
    import torch
    hits = {k: np.zeros(5) for k in ks}
    n = 0
    kmax = max(ks)
    for b0 in range(0, X.shape[0], batch):
        x, y = X[b0:b0 + batch], Y[b0:b0 + batch]
        idx, _ = P.topk(P.forward(x), kmax)
        idx, y = idx.cpu().numpy(), y.cpu().numpy()
        for k in ks:
            hits[k] += (idx[:, :, :k] == y[:, :, None]).any(axis=2).sum(axis=0)
        n += x.shape[0]
    return {k: (hits[k] / n).round(4).tolist() for k in ks}


def train(P, X, Y, epochs=4, batch=4096, lr=2e-3, log=print):
    import torch
    n = X.shape[0]
    steps = 0
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda")
        tot, cnt = 0.0, 0
        lr_ep = lr * (0.5 ** ep)
        for b0 in range(0, n - batch + 1, batch):
            sel = perm[b0:b0 + batch]
            loss = P.train_step(X[sel].contiguous(), Y[sel].contiguous(), lr=lr_ep)
            steps += 1
            if steps % 50 == 0:
                tot += float(loss.mean())
                cnt += 1
        log("epoch %d: mean loss %.4f (lr %.2e)" % (ep, tot / max(cnt, 1), lr_ep))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--block", default="4x4")
    ap.add_argument("--blocks", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    bx, by = (int(v) for v in a.block.split("x"))
    import torch
    torch.manual_seed(0)
    P = Predictor((bx, by), seed=0)
    enc = A.Encoder(0)
    t0 = time.time()
    X, Y = collect(P, enc, a.blocks)
    torch.cuda.synchronize()
    t1 = time.time()
    n_val = min(65536, X.shape[0] // 10)
    Xv, Yv, Xt, Yt = X[:n_val], Y[:n_val], X[n_val:], Y[n_val:]
    train(P, Xt, Yt, epochs=a.epochs)
    torch.cuda.synchronize()
    t2 = time.time()
    rec = recall(P, Xv, Yv)
    print(json.dumps({"block": a.block, "train_blocks": int(Xt.shape[0]), "label_s": round(t1 - t0, 1),
                      "train_s": round(t2 - t1, 1), "val_recall_topk": rec}))
    out = a.out or os.path.join(HERE, "predictor_%dx%d.npz" % (bx, by))
    st = P.state()
    np.savez_compressed(out, **{"p%d" % i: w.astype(np.float16) for i, w in enumerate(st["params"])},
                        hidden=np.int32(P.hidden), block=np.array([bx, by], np.int32))
    enc.close()


def load_params(path):
    """The six fp32 parameter arrays of a saved predictor."""
    z = np.load(path, allow_pickle=False)
    return [z["p%d" % i].astype(np.float32) for i in range(6)], int(z["hidden"])


if __name__ == "__main__":
    main()
