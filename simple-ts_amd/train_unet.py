"""Trains the warm-start U-Net on the GPU and writes its parameters.

    python simple-ts_amd/train_unet.py --out simple-ts_amd/unet_b32.npz
    python simple-ts_amd/train_unet.py --corrector --out simple-ts_amd/unet_corr_b32.npz

Generates the training set with the solver on random confined-square cases
on a 512 x 128 grid and trains with the network's own HIP forward/backward
kernels and Adam; prints one JSON line with the loss history.  Default: the
p/T predictor (cin 5, cout 2; one sample per plain FP64 time step).
--corrector: the corrector of the extrapolated start (cin 9, cout 4; one
sample per predicted step of the accelerated FP64 method, 40 cases x 100
samples = 4000, a diverging case replaced by the next seeded one)."""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", type=int, default=40)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", required=True)
    ap.add_argument("--corrector", action="store_true")
    a = ap.parse_args()
    pkg = importlib.import_module("simple-ts_amd")
    ws = importlib.import_module("simple-ts_amd.warmstart")
    cin, cout = (ws.CORR_CIN, ws.CORR_COUT) if a.corrector else (5, 2)
    net = pkg.UNet(cin, cout, a.base, a.batch, a.ny, a.nx)
    net.set_params(ws.he_init(cin, cout, a.base, a.seed))
    t0 = time.time()
    if a.corrector:
        used = ws.generate_corrections(net, a.cases, a.steps, seed=a.seed, log=lambda m: print(m, flush=True))
        n = a.cases * a.steps
    else:
        n, used = ws.generate(net, a.cases, a.steps, seed=a.seed, log=lambda m: print(m, flush=True))
    t1 = time.time()
    hist, hold = ws.train(net, n, a.epochs, a.batch, a.lr, seed=a.seed + 1, log=lambda m: print(m, flush=True))
    t2 = time.time()
    np.savez(a.out, params=net.params(), base=a.base, cin=cin, cout=cout, train_loss=np.array(hist),
             holdout_loss=np.array(hold), samples=n, grid=np.array([a.nx, a.ny]),
             corr_fraction=ws.CORR_FRACTION if a.corrector else 0.0,
             cases=np.array([[p["Ma"], p["Kn"], p["side"], p["xc"]] for p in used]))
    print(json.dumps({"samples": n, "cases_used": len(used), "generate_s": t1 - t0, "train_s": t2 - t1,
                      "train_loss": hist, "holdout_loss": hold, "out": a.out}), flush=True)


if __name__ == "__main__":
    main()
