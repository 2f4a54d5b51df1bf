"""Parity at the benchmark's own size, the error path of a non-finite residual,
the predicted start under a time-dependent inflow, and the continuity norms.

* The supersonic confined square on the full 1920 x 384 grid: the GPU program
  (sts_gpu) and the reference solver (oracle/_ref/sts_ref, built from the
  reference's sources) run the same 2 time steps of a fixed 10 outer
  iterations each (tolerances 0, so both run exactly N_I iterations: the
  reference's loop-2 test S:4433-4437 never stops early), and write the same
  bytes: u/v/p/Temper.bin, Tmp.txt (every iteration's residual maxima) and
  Iterations_in_loops.txt.
* The subsonic square on 240 x 48 at dt = 0.01 H/u_in, where the reference's
  own solve goes non-finite in step 1.  Its maxima skip NaN (fabs(x) > max is
  false for a NaN cell), so it keeps logging finite maxima while NaN spreads
  and ends loop 2 as if converged (zero residuals) at iteration 752.  The
  library returns STS_ERR_NONFINITE at the first outer iteration that leaves a
  non-finite value in u, v, p or T -- the iteration the CPU restatement of the
  reference (oracle/flow.py) names -- after logging the same maxima as the
  reference, digit for digit, for every iteration before it.
"""
import ctypes
import importlib
import os
import subprocess

import numpy as np
import pytest

from casefiles import have_reference, run_reference
from golden_cases import fresh_input, ref_dir

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "simple-ts_amd", "sts_gpu")
STS_ERR_NONFINITE = -3


def run_driver(d, *args, timeout=300):
    print("sts_gpu", d, *args, flush=True)
    subprocess.run([DRIVER, d, *args], check=True, timeout=timeout, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a - b))


@pytest.mark.skipif(not have_reference(), reason="oracle/_ref/sts_ref not built")
def test_headline_grid_matches_live_reference(tmp_path):
    cases = importlib.import_module("simple-ts_amd.cases")
    kw = dict(tol=0.0, N_I=10, extra=dict(Nt_save_temp_data=1))
    dt = 0.01 * 240 / 1920
    mine = cases.confined_square(str(tmp_path / "gpu"), 1920, 384, 1415, 0.00283, 2, dt_over_tin=dt, **kw)
    ref = cases.confined_square(str(tmp_path / "ref"), 1920, 384, 1415, 0.00283, 2, dt_over_tin=dt, **kw)
    print("reference solver, 1920x384, 2 steps x 10 outer iterations", flush=True)
    run_reference(ref, timeout=600)
    run_driver(mine)
    for f in ("u.bin", "v.bin", "p.bin", "Temper.bin"):
        a = np.fromfile(os.path.join(mine, f), dtype=np.float64)
        b = np.fromfile(os.path.join(ref, f), dtype=np.float64)
        assert a.size == 1920 * 384 and np.all(np.isfinite(a)), f
        print(f"{f}: relative L2 {rel_l2(a, b):.3e}", flush=True)
        assert a.tobytes() == b.tobytes(), f"{f} differs from the reference's"
    for f in ("Tmp.txt", "Iterations_in_loops.txt"):
        assert open(os.path.join(mine, f), "rb").read() == open(os.path.join(ref, f), "rb").read(), f
    rows = open(os.path.join(ref, "Iterations_in_loops.txt")).read().split("\n")[1:]
    assert [int(r.split()[2]) for r in rows if r.strip()] == [10, 10]


def _tmp_residuals(case_dir):
    """(iteration, the four p/T maxima as printed) of step 1 in Tmp.txt."""
    out = []
    for line in open(os.path.join(case_dir, "Tmp.txt")).read().split("\n")[1:]:
        if line.startswith("it ="):
            break
        w = line.split()
        if len(w) == 6:
            out.append((int(w[0]), w[2:]))
    return out


@pytest.mark.skipif(not have_reference(), reason="oracle/_ref/sts_ref not built")
def test_nonfinite_residual_is_an_error(tmp_path, sts):
    cases = importlib.import_module("simple-ts_amd.cases")
    ws = importlib.import_module("simple-ts_amd.warmstart")
    kw = dict(extra=dict(Nt_save_temp_data=1))
    ref = cases.confined_square(str(tmp_path / "ref"), 240, 48, 85, 0.00194, 2, dt_over_tin=0.01, **kw)
    mine = cases.confined_square(str(tmp_path / "gpu"), 240, 48, 85, 0.00194, 2, dt_over_tin=0.01, **kw)
    run_reference(ref)
    ref_rows = _tmp_residuals(ref)
    last = ref_rows[-1]
    # the reference's own run: its final iteration logs zeros (NaN skipped)
    assert all(float(x) == 0.0 for x in last[1])
    n_ref = last[0] + 1
    run = ws.Run(mine)
    N_I, n_pc, tu, tp, tT = run.loop
    lc = sts.LoopControl(N_I, n_pc, tu, tp, tT, 0)
    it = ctypes.c_int()
    log = np.zeros(6 * N_I)
    rc = sts.lib().sts_time_step(run.ctx.h, ctypes.byref(lc), ctypes.byref(it),
                                 log.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None)
    # the first outer iteration whose fields hold a non-finite value, from the
    # CPU restatement (whose maxima, like the reference's, skip NaN)
    flow = importlib.import_module("oracle.flow")
    first = []

    def on_iteration(step, k, o):
        if not first and not all(np.all(np.isfinite(o.get(f))) for f in ("u", "v", "p", "Temper")):
            first.append(k)
    ora_case = cases.confined_square(str(tmp_path / "ora"), 240, 48, 85, 0.00194, 2, dt_over_tin=0.01)
    flow.run_case(ora_case, max_steps=1, on_iteration=on_iteration)
    print(f"rc {rc}, iterations {it.value}; restatement: first non-finite state after iteration {first}; "
          f"reference loop ended at {n_ref}, N_I {N_I}", flush=True)
    assert rc == STS_ERR_NONFINITE
    assert first and it.value == first[0] <= n_ref < N_I
    log = log[:6 * it.value].reshape(-1, 6)
    assert not np.all(np.isfinite(log[-1]))
    # Tmp.txt rows: first-sweep and last-sweep max|dp|, then the same for T
    for k, vals in ref_rows[:it.value - 1]:
        assert ["%.15e" % log[k, q] for q in (4, 2, 5, 3)] == vals, k
    run.close()
    # the predicted-start entry point (guard on; step 1 runs from x^n) and the
    # batched step report the same error
    for path in ("mixed", "batch"):
        r = ws.Run(cases.confined_square(str(tmp_path / path), 240, 48, 85, 0.00194, 2, dt_over_tin=0.01))
        with pytest.raises(sts.StsError, match="STS_ERR_NONFINITE"):
            if path == "mixed":
                r.ctx.time_step_mixed(*r.loop, 0, (1e-7,) * 3, extrapolate=3, guard=64)
            else:
                b = sts.Batch([r.ctx])
                try:
                    b.time_step(*r.loop)
                finally:
                    b.close()
        r.close()


def test_predicted_start_follows_the_inflow_ramp(tmp_path):
    """Golden case 'square' ramps the inflow velocity in time
    (use_interpolation_for_u_BC_by_time): the caller sets u on the inflow
    columns after each step, as the program does.  A predicted start must
    keep that value through the step (no kernel of loop 2 rewrites it)
    instead of extrapolating it one more ramp increment ahead."""
    ws = importlib.import_module("simple-ts_amd.warmstart")
    d = fresh_input("square", tmp_path)
    r = ws.Run(d)
    c = r.case
    tb, te, ub, ue = (c.value(k) for k in ("t_BC_b", "t_BC_e", "u_BC_xb_t_BC_b", "u_BC_xb_t_BC_e"))
    ht = c.value("ht")
    nxu0 = c.boxes["Nx_u"][0]
    fluid = (c.problem().mask == 1).reshape(c.Ny, c.Nx)[:, :nxu0]
    assert nxu0 > 0 and fluid.any()
    u_set = r.ctx.get("u").reshape(c.Ny, c.Nx)[:, :nxu0][fluid]
    t = 0.0
    for s in range(8):
        r.ctx.time_step_mixed(*r.loop, 0, (1e-7,) * 3, extrapolate=2)
        u = r.ctx.get("u").reshape(c.Ny, c.Nx)[:, :nxu0][fluid]
        assert np.array_equal(u, u_set), f"step {s}: inflow velocity changed inside the step"
        t += ht
        ug = ub if t <= tb else (ue if t > te else (ue - ub) / (te - tb) * (t - tb) + ub)
        r.ctx.set_inflow_u(ug)
        u_set = np.full(u_set.shape, ug)
    r.close()


def test_continuity_norms_match_the_reference_field(tmp_path):
    """sts_continuity's max / L2 / L1 over the fluid cells the reference fills
    (i < Nx-1, j < Ny-1) against numpy's norms of the reference's own
    ContinuityEquation.bin of golden case 'square' (the program's last
    evaluation is the one the file holds)."""
    d = fresh_input("square", tmp_path)
    norms = os.path.join(str(tmp_path), "norms.txt")
    run_driver(d, "--continuity-norms", norms)
    last = open(norms).read().split("\n")[-2].split()
    got = np.array([float(x) for x in last[1:]])
    ref = ref_dir("square")
    cf = np.fromfile(os.path.join(ref, "ContinuityEquation.bin"), dtype=np.float64)
    vol = np.fromfile(os.path.join(ref, "vol_inf_fb.bin"), dtype=np.uint32)
    Nx, Ny = 60, 14
    assert cf.size == Nx * Ny
    sel = np.zeros((Ny, Nx), bool)
    sel[:Ny - 1, :Nx - 1] = True
    sel &= (vol == 0).reshape(Ny, Nx)
    c = cf.reshape(Ny, Nx)[sel]
    want = np.array([np.max(np.abs(c)), np.sqrt(np.sum(c * c)), np.sum(np.abs(c))])
    print("norms", got, "numpy", want, flush=True)
    assert got[0] == want[0]
    assert np.allclose(got[1:], want[1:], rtol=1e-12, atol=0.0)
