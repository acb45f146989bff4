"""Trains the candidate predictor of the LEARNED policy on the GPU.

Labels come from the encoder's own plain search on synthetic images that
are not part of any benchmark (seeds 10000..19999): by default the
EXHAUSTIVE partition search (every selected partitioning ranked by its line
fits, code/best_partitionings.cpp:498-737), or the COMPAT search; half
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


def rank_targets(seeds, errors):
    """Logit columns of the 4 compat positions of each partition count
    ((n, 12) int32) and the winning position ((n, 3) int32: the first
    position with the lowest error, as the search picks it; -1 when every
    position reached the same error, which says nothing about the ranking)."""
    cols = (4 + np.arange(3)[None, :, None] * 1024 + seeds.astype(np.int32)).reshape(-1, 12)
    cols[seeds.reshape(-1, 12) < 0] = -1
    tgt = np.argmin(errors, axis=2).astype(np.int32)
    tgt[(errors == errors[:, :, :1]).all(axis=2)] = -1
    return cols.astype(np.int32), tgt


def collect(P, enc, n_blocks, size=1024, log=print, policy=None):
    """Encodes training images with the plain search (`policy`, default
    EXHAUSTIVE) until n_blocks blocks are labelled; returns device tensors X (bf16 features), Y (int32 head
    labels: pc - 1, best seed for 2 / 3 / 4 partitions, trial mode) and the
    position-ranking labels R (logit columns, (n, 12)) and T (winners, (n, 3))."""
    import torch
    policy = A.POLICY_EXHAUSTIVE if policy is None else policy
    xs, ys, rs, ts, total, i = [], [], [], [], 0, 0
    while total < n_blocks:
        img = training_image(i, size)
        h, w = img.shape[:2]
        enc.configure(w, h, P.bx, P.by, labels=True, policy=policy)
        d_img = torch.from_numpy(img.reshape(-1).copy()).cuda()
        d_out = torch.empty(enc.num_blocks * 16, dtype=torch.uint8, device="cuda")
        enc.encode_device(d_img.data_ptr(), d_out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        lab = enc.labels()
        y = np.stack([lab[:, 0] - 1, lab[:, 3], lab[:, 4], lab[:, 5], lab[:, 2]], axis=1).astype(np.int32)
        cols, tgt = rank_targets(*enc.candidates())
        xs.append(P.features(d_img, w, h, 0, enc.num_blocks))
        ys.append(torch.from_numpy(y).cuda())
        rs.append(torch.from_numpy(cols).cuda())
        ts.append(torch.from_numpy(tgt).cuda())
        total += enc.num_blocks
        i += 1
        log("image %d: %d blocks labelled" % (i, total))
    return (torch.cat(xs)[:n_blocks], torch.cat(ys)[:n_blocks], torch.cat(rs)[:n_blocks],
            torch.cat(ts)[:n_blocks])


def recall(P, X, Y, R=None, T=None, ks=(1, 2, 4, 16), batch=8192):
    """Fraction of rows whose label is within the top-k of its head, and
    (with R, T) the fraction of partition counts whose winning compat
    position is ranked first / within the first two by its seed logit."""
    hits = {k: np.zeros(5) for k in ks}
    rank_hits, rank_n = np.zeros(2), 0
    n = 0
    kmax = max(ks)
    for b0 in range(0, X.shape[0], batch):
        x, y = X[b0:b0 + batch], Y[b0:b0 + batch]
        logits = P.forward(x)
        idx, _ = P.topk(logits, kmax)
        idx, y = idx.cpu().numpy(), y.cpu().numpy()
        for k in ks:
            hits[k] += (idx[:, :, :k] == y[:, :, None]).any(axis=2).sum(axis=0)
        if R is not None:
            lg = logits.cpu().numpy()
            cols = R[b0:b0 + batch].cpu().numpy().reshape(-1, 3, 4)
            tgt = T[b0:b0 + batch].cpu().numpy()
            sc = np.take_along_axis(lg[:, None, :].repeat(3, 1), np.maximum(cols, 0), axis=2)
            sc[cols < 0] = -np.inf
            # rank of the winner: positions scoring higher, or equal and earlier
            win = np.take_along_axis(sc, np.maximum(tgt, 0)[:, :, None], axis=2)
            pos = np.arange(4)[None, None, :]
            ahead = ((sc > win) | ((sc == win) & (pos < np.maximum(tgt, 0)[:, :, None]))).sum(axis=2)
            ok = tgt >= 0
            rank_hits += [((ahead == 0) & ok).sum(), ((ahead <= 1) & ok).sum()]
            rank_n += int(ok.sum())
        n += x.shape[0]
    out = {k: (hits[k] / n).round(4).tolist() for k in ks}
    if rank_n:
        out["rank_top1_top2"] = (rank_hits / rank_n).round(4).tolist()
    return out


def train(P, X, Y, R=None, T=None, epochs=4, batch=4096, lr=2e-3, log=print):
    import torch
    n = X.shape[0]
    steps = 0
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda")
        tot, cnt = 0.0, 0
        lr_ep = lr * (0.5 ** ep)
        for b0 in range(0, n - batch + 1, batch):
            sel = perm[b0:b0 + batch]
            rank = (R[sel].contiguous(), T[sel].contiguous()) if R is not None else None
            loss = P.train_step(X[sel].contiguous(), Y[sel].contiguous(), lr=lr_ep, rank=rank)
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
    ap.add_argument("--rank-loss", type=int, default=0,
                    help="also train the seed heads to rank the 4 compat positions of each partition count")
    ap.add_argument("--out", default=None)
    ap.add_argument("--policy", choices=["exhaustive", "compat"], default="exhaustive",
                    help="the plain search whose outcome labels the blocks")
    # pruning saved with the weights (tuned on a held-out image)
    ap.add_argument("--seeds", type=int, default=0)
    ap.add_argument("--modes", type=int, default=16)
    ap.add_argument("--slots", type=int, default=5)
    ap.add_argument("--fallback-below", type=float, default=0.5)
    ap.add_argument("--count-below", type=float, default=0.0)
    ap.add_argument("--partitionings", type=int, default=0)
    ap.add_argument("--refine-partitionings", type=int, default=0)
    ap.add_argument("--early-out", type=float, default=0.0)
    ap.add_argument("--full-search-above", type=float, default=0.0)
    ap.add_argument("--rounds", type=int, default=0)
    ap.add_argument("--partitioned-slots", type=int, default=0)
    ap.add_argument("--candidates", type=int, default=1)
    ap.add_argument("--count-above", type=float, default=0.0)
    ap.add_argument("--full-search-modes", type=int, default=0)
    ap.add_argument("--mismatch-candidates", type=int, default=0)
    ap.add_argument("--fallback-above", type=float, default=0.0)
    ap.add_argument("--retune", default=None, metavar="NPZ",
                    help="no training: write NPZ's weights again (to --out, default NPZ itself) with the pruning "
                         "settings given here")
    a = ap.parse_args()
    bx, by = (int(v) for v in a.block.split("x"))
    prune = dict(seeds=a.seeds, modes=a.modes, slots=a.slots, fallback_below=a.fallback_below,
                 count_below=a.count_below, partitionings=a.partitionings,
                 refine_partitionings=a.refine_partitionings, early_out=a.early_out,
                 full_search_above=a.full_search_above, rounds=a.rounds,
                 partitioned_slots=a.partitioned_slots, candidates=a.candidates,
                 count_above=a.count_above, full_search_modes=a.full_search_modes,
                 mismatch_candidates=a.mismatch_candidates, fallback_above=a.fallback_above)
    if a.retune:
        params, hidden = load_params(a.retune)
        save_arrays(a.out or a.retune, params, hidden, (bx, by), prune)
        print(json.dumps({"block": a.block, "retuned": a.out or a.retune, "pruning": prune}))
        return
    import torch
    torch.manual_seed(0)
    P = Predictor((bx, by), seed=0)
    enc = A.Encoder(0)
    t0 = time.time()
    X, Y, R, T = collect(P, enc, a.blocks,
                         policy=A.POLICY_EXHAUSTIVE if a.policy == "exhaustive" else A.POLICY_COMPAT)
    torch.cuda.synchronize()
    t1 = time.time()
    n_val = min(65536, X.shape[0] // 10)
    if not a.rank_loss:
        R = T = None
    sl_v, sl_t = slice(0, n_val), slice(n_val, None)
    train(P, X[sl_t], Y[sl_t], R[sl_t] if R is not None else None, T[sl_t] if T is not None else None,
          epochs=a.epochs)
    torch.cuda.synchronize()
    t2 = time.time()
    rec = recall(P, X[sl_v], Y[sl_v], R[sl_v] if R is not None else None, T[sl_v] if T is not None else None)
    print(json.dumps({"block": a.block, "train_blocks": int(X.shape[0] - n_val), "label_s": round(t1 - t0, 1),
                      "train_s": round(t2 - t1, 1), "val_recall_topk": rec}))
    out = a.out or os.path.join(HERE, "predictor_%dx%d.npz" % (bx, by))
    save(out, P, prune)
    enc.close()


def save(path, P, prune):
    """Weights as fp16 plus the pruning settings (LearnedParams fields) to
    install with them."""
    st = P.state()
    save_arrays(path, st["params"], P.hidden, (P.bx, P.by), prune)


def save_arrays(path, params, hidden, block, prune):
    """The six parameter arrays (stored as fp16: a saved file read back and
    written again keeps its weights bit for bit) and the pruning settings."""
    np.savez_compressed(path, **{"p%d" % i: np.asarray(w).astype(np.float16) for i, w in enumerate(params)},
                        hidden=np.int32(hidden), block=np.array(block, np.int32),
                        prune_i=np.array([prune["seeds"], prune["modes"], prune["slots"]], np.int32),
                        prune_f=np.array([prune["fallback_below"], prune["count_below"]], np.float32),
                        prune_p=np.int32(prune.get("partitionings", 0)),
                        prune_r=np.int32(prune.get("refine_partitionings", 0)),
                        prune_e=np.float32(prune.get("early_out", 0.0)),
                        prune_h=np.float32(prune.get("full_search_above", 0.0)),
                        prune_n=np.int32(prune.get("rounds", 0)),
                        prune_s=np.int32(prune.get("partitioned_slots", 0)),
                        prune_c=np.int32(prune.get("candidates", 0)),
                        prune_a=np.float32(prune.get("count_above", 0.0)),
                        prune_m=np.int32(prune.get("full_search_modes", 0)),
                        prune_k=np.int32(prune.get("mismatch_candidates", 0)),
                        prune_l=np.float32(prune.get("fallback_above", 0.0)))


def load_params(path):
    """The six fp32 parameter arrays of a saved predictor and its hidden width."""
    params, hidden, _ = A.load_predictor_file(path)
    return params, hidden


if __name__ == "__main__":
    main()
