"""Speed / quality sweep of the LEARNED policy against the plain search.

    python -m astc_implementation_amd.eval_predictor --block 4x4 --size 4096 --seed 3

Encodes one synthetic gradient + noise + rectangles image (device-resident,
one warm-up and `--steps` timed encodes per setting) with the plain search
and with the learned policy under several pruning settings; prints one JSON
line per setting with Mtexel/s, decoded PSNR, the PSNR loss against the plain
search and the fallback rate.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import astc_implementation_amd as A  # noqa: E402
from astc_implementation_amd import synth  # noqa: E402
from astc_implementation_amd.train_predictor import load_params  # noqa: E402

SETTINGS = [
    dict(seeds=0, modes=16, slots=5, fallback_below=0.5),
    dict(seeds=0, modes=16, slots=8, fallback_below=0.5, early_out=1.0),
    dict(seeds=0, modes=16, slots=8, fallback_below=0.5, early_out=1.2, full_search_above=8.0),
    dict(seeds=0, modes=16, slots=8, fallback_below=0.5, early_out=1.25, full_search_above=8.0),
    dict(seeds=0, modes=12, slots=6, fallback_below=0.5, early_out=1.3, full_search_above=10.0),
]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--block", default="4x4")
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--weights", default=None)
    ap.add_argument("--settings", default=None,
                    help="JSON list of astcg_learned_params dicts to sweep (default: the built-in list)")
    a = ap.parse_args()
    settings = json.loads(a.settings) if a.settings else SETTINGS
    bx, by = (int(v) for v in a.block.split("x"))
    import torch
    params, hidden = load_params(a.weights or os.path.join(HERE, "predictor_%dx%d.npz" % (bx, by)))
    img = synth.gradient_image(a.size, a.size, a.seed)
    enc = A.Encoder(0)
    d_in = torch.from_numpy(img.reshape(-1).copy()).cuda()

    def run(policy, **lp):
        if policy == A.POLICY_LEARNED:
            enc.set_predictor(bx, by, params, hidden=hidden, **lp)
        enc.configure(a.size, a.size, bx, by, policy=policy)
        d_out = torch.empty(enc.num_blocks * 16, dtype=torch.uint8, device="cuda")
        s = torch.cuda.current_stream().cuda_stream
        enc.encode_device(d_in.data_ptr(), d_out.data_ptr(), s)
        torch.cuda.synchronize()
        t0 = time.time()
        for _ in range(a.steps):
            enc.encode_device(d_in.data_ptr(), d_out.data_ptr(), s)
        torch.cuda.synchronize()
        dt = (time.time() - t0) / a.steps
        dec = enc.decode(bx, by, a.size, a.size, d_out.cpu().numpy())
        fb = enc.fallback_rate() if policy == A.POLICY_LEARNED else 0.0
        return a.size * a.size / dt / 1e6, A.psnr(dec, img), fb

    mt, ps, _ = run(A.POLICY_COMPAT)
    print(json.dumps({"policy": "plain", "mtexel_s": round(mt, 2), "psnr": round(ps, 4)}), flush=True)
    for lp in settings:
        mt2, ps2, fb = run(A.POLICY_LEARNED, **lp)
        print(json.dumps({"policy": "learned", **lp, "mtexel_s": round(mt2, 2), "psnr": round(ps2, 4),
                          "dpsnr": round(ps - ps2, 4), "speedup": round(mt2 / mt, 2), "fallback": round(fb, 4)}),
              flush=True)
    enc.close()


if __name__ == "__main__":
    main()
