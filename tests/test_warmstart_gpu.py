"""Warm start by the U-Net: the Python set-up of a case matches the program's,
the fallback keeps the plain path bit for bit when the prediction does not
help, and a small trained network runs through the whole pipeline (record,
train, predict, choose) with results converged to the same tolerance."""
import importlib
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "simple-ts_amd", "sts_gpu")
NX, NY = 128, 32  # network grid: W = Nx (multiple of 64), H = Ny (multiple of 8)


def mods():
    return importlib.import_module("simple-ts_amd"), importlib.import_module("simple-ts_amd.warmstart")


def fields(ctx):
    return {k: ctx.get(k) for k in ("u", "v", "p", "Temper")}


PRM = dict(Ma=0.8, Kn=0.01, Re=0.8 / (0.01 * 0.6067627457812106), side=0.25, xc=4.0)


def test_python_setup_matches_program(tmp_path):
    pkg, ws = mods()
    steps = 5
    d1 = ws.write_case(str(tmp_path / "py"), NX, NY, PRM, steps)
    d2 = ws.write_case(str(tmp_path / "prog"), NX, NY, PRM, steps)
    print("sts_gpu", d2, flush=True)
    subprocess.run([DRIVER, d2], check=True, timeout=300, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    run = ws.Run(d1)
    for _ in range(steps):
        run.step()
    f = fields(run.ctx)
    m = run.case.vol == 0
    run.close()
    for k, fn in (("u", "u.bin"), ("v", "v.bin"), ("p", "p.bin"), ("Temper", "Temper.bin")):
        ref = np.fromfile(os.path.join(d2, fn), dtype=np.float64)
        a = f[k] if k != "Temper" else np.where(m, f[k], ref)  # body-cell T is overwritten at saves
        assert np.array_equal(a, ref), k


def test_fallback_keeps_plain_path(tmp_path):
    """A useless prediction (huge output bias) is never taken, and the steps
    are then bit-identical to sts_time_step."""
    pkg, ws = mods()
    net = pkg.UNet(5, 2, 8, 1, NY, NX)
    p = np.zeros(net.n_params, np.float32)
    p[-2:] = 1e3  # head bias
    net.set_params(p)
    a = ws.Run(ws.write_case(str(tmp_path / "a"), NX, NY, PRM, 4))
    b = ws.Run(ws.write_case(str(tmp_path / "b"), NX, NY, PRM, 4))
    for _ in range(4):
        used, n32, n64 = a.step_ws(net)
        it = b.step()
        assert used == 0 and n32 == 0 and n64 == it
    fa, fb = fields(a.ctx), fields(b.ctx)
    a.close()
    b.close()
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k


def test_trained_network_pipeline(tmp_path):
    """Record -> train -> warm-started steps.  The warm-started run ends on the
    same loop-2 test, so its fields agree with the plain run to well under
    1e-6 relative L2; the outer-iteration counts of both are printed."""
    pkg, ws = mods()
    net = pkg.UNet(5, 2, 8, 4, NY, NX)
    rng = np.random.default_rng(0)
    # He-initialised parameters
    from oracle import unet as U
    P = U.init_params(5, 2, 8, 0)
    net.set_params(np.concatenate([np.concatenate([P[n + ".w"].ravel(), P[n + ".b"]])
                                   for n, _, _, _ in U.conv_specs(5, 2, 8)]))
    n, used_prm = ws.generate(net, 3, 10, seed=1)
    assert n >= 10
    hist, hold = ws.train(net, n, epochs=4, batch=4, lr=1e-3, seed=2)
    assert all(np.isfinite(hist)) and hist[-1] < hist[0]
    steps = 6
    a = ws.Run(ws.write_case(str(tmp_path / "a"), NX, NY, PRM, steps))
    b = ws.Run(ws.write_case(str(tmp_path / "b"), NX, NY, PRM, steps))
    ia, ib, taken = 0, 0, 0
    for _ in range(steps):
        used, n32, n64 = a.step_ws(net)
        taken += used
        ia += n64
        ib += b.step()
    print(f"outer iterations: warm-started {ia}, plain {ib}; prediction taken in {taken}/{steps} steps", flush=True)
    fa, fb = fields(a.ctx), fields(b.ctx)
    a.close()
    b.close()
    for k in fa:
        e = np.linalg.norm(fa[k] - fb[k]) / np.linalg.norm(fb[k])
        assert e < 1e-6, (k, e)
