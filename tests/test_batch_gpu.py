"""Batched time steps: cases of different grid sizes stepped together in
shared launches give, case by case, exactly the fields and iteration counts
of the same cases stepped one by one -- for every tile shape, after the
autotuner has run, and for the mixed-precision step.  The set holds
confined squares, a channel periodic about OX and two walled
Rayleigh-Benard boxes (gravity term c_Fr_v in the v equation,
SIMPLE_TS_MPI.cpp:3404, :3585, c_Fr_v set at :2066)."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (Nx, Ny, Ma, Kn, side, xc)
CASES = [(60, 14, 0.1, 0.00194, 0.25, 4.0), (96, 20, 1.2, 0.01, 0.3, 3.0), (128, 32, 2.0, 0.003, 0.2, 5.0)]
# (Nx, Ny, Ra) of the Rayleigh-Benard boxes, stepped at dt = 2e-3 (the oracle
# needs 14 and 28-34 outer iterations per step over the first 10 steps; at
# dt = 5e-3 the 100 x 51 box goes non-finite in its third step, on the oracle
# as on the GPU)
RB = [(64, 32, 3000.0), (100, 51, 20000.0)]


def runs(tmp_path, tag, precision=64):
    ws = importlib.import_module("simple-ts_amd.warmstart")
    out = []
    for k, (Nx, Ny, Ma, Kn, side, xc) in enumerate(CASES):
        prm = dict(Ma=Ma, Kn=Kn, Re=Ma / (Kn * ws.SQ), side=side, xc=xc)
        d = ws.write_case(str(tmp_path / f"{tag}{k}"), Nx, Ny, prm, 10, dt_over_tin=0.005)
        out.append(ws.Run(d, precision))
    # a channel periodic about OX, with bodies across the seam
    cs = importlib.import_module("simple-ts_amd.cases")
    d = cs.periodic_channel(str(tmp_path / f"{tag}p"), 100, 22, 0.3, 0.02, 10, dt=0.005)
    out.append(ws.Run(d, precision))
    # walled Rayleigh-Benard boxes (2:1, hot bottom, cold top) at Ra 3000 and 20000
    for k, (Nx, Ny, Ra) in enumerate(RB):
        d = cs.rayleigh_benard(str(tmp_path / f"{tag}rb{k}"), Nx, Ny, Kn=0.05, Fr=cs.rayleigh_fr(Ra, 0.05), steps=10,
                               dt=2e-3)
        out.append(ws.Run(d, precision))
    return out


def state(r):
    return [r.ctx.get(k) for k in ("u", "v", "p", "Temper", "rho")]


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_batch_matches_single_cases(variant, tmp_path):
    pkg = importlib.import_module("simple-ts_amd")
    single, batched = runs(tmp_path, "s"), runs(tmp_path, "b")
    B = pkg.Batch([r.ctx for r in batched])
    B.set_tiles(variant, variant)
    loop = single[0].loop
    for step in range(3):
        it_single = [r.ctx.time_step(*loop)[0] for r in single]
        _, it_batch = B.time_step(*loop)
        print(f"variant {variant} step {step}: single {it_single}, batched {it_batch}", flush=True)
        assert it_batch == it_single
    for a, b in zip(single, batched):
        for x, y in zip(state(a), state(b)):
            assert np.array_equal(x, y)
    B.close()
    for r in single + batched:
        r.close()


def test_autotune_then_step_matches(tmp_path):
    pkg = importlib.import_module("simple-ts_amd")
    single, batched = runs(tmp_path, "s"), runs(tmp_path, "b")
    B = pkg.Batch([r.ctx for r in batched])
    cv, sv, ms = B.autotune(3)
    print(f"autotune: cell variant {cv}, sweep variant {sv}, ms {np.round(ms, 4).tolist()}", flush=True)
    assert 0 <= cv < 4 and 0 <= sv < 4 and np.all(ms > 0)
    loop = single[0].loop
    for step in range(2):
        assert B.time_step(*loop)[1] == [r.ctx.time_step(*loop)[0] for r in single]
    for a, b in zip(single, batched):
        for x, y in zip(state(a), state(b)):
            assert np.array_equal(x, y)
    B.close()
    for r in single + batched:
        r.close()


def test_mixed_batch_matches_single_mixed(tmp_path):
    pkg = importlib.import_module("simple-ts_amd")
    single, batched = runs(tmp_path, "s", 32), runs(tmp_path, "b", 32)
    B = pkg.Batch([r.ctx for r in batched])
    loop = single[0].loop
    mixed = (loop[0], (1e-6, 1e-6, 1e-6), 4, 0.7)
    for step in range(2):
        ref = [r.ctx.time_step_mixed(*loop, mixed[0], mixed[1], mixed[2], mixed[3])[:2] for r in single]
        n32, n64 = B.time_step(*loop, mixed=mixed)
        assert list(zip(n32, n64)) == ref
    for a, b in zip(single, batched):
        for x, y in zip(state(a), state(b)):
            assert np.array_equal(x, y)
    B.close()
    for r in single + batched:
        r.close()


@pytest.mark.parametrize("fp32_above", [0, 18])
def test_predicted_batch_matches_single(fp32_above, tmp_path):
    """Predicted starts (degree 3) and the FP32 gate in shared launches: case
    by case the single-case counts and fields, bit for bit, including the
    steps where some cases skip the FP32 phase and others run it."""
    pkg = importlib.import_module("simple-ts_amd")
    single, batched = runs(tmp_path, "s", 32), runs(tmp_path, "b", 32)
    B = pkg.Batch([r.ctx for r in batched])
    loop = single[0].loop
    mixed = (loop[0], (1e-6, 1e-6, 1e-6), 4, 0.7, 3, fp32_above)
    for step in range(6):
        ref = [r.ctx.time_step_mixed(*loop, *mixed[:4], extrapolate=mixed[4], fp32_above=mixed[5])[:2]
               for r in single]
        n32, n64 = B.time_step(*loop, mixed=mixed)
        print(f"step {step}: single {ref}, batched {list(zip(n32, n64))}", flush=True)
        assert list(zip(n32, n64)) == ref
    for a, b in zip(single, batched):
        for x, y in zip(state(a), state(b)):
            assert np.array_equal(x, y)
    B.close()
    for r in single + batched:
        r.close()


def test_guarded_batch_stays_on_the_plain_fields(tmp_path):
    """Batched predicted starts with a guard (pauses per case, no restart):
    the fields stay within 1e-6 relative L2 of the same cases stepped from
    x^n, and every step is counted as predicted or from x^n."""
    pkg = importlib.import_module("simple-ts_amd")
    plain, batched = runs(tmp_path, "p", 32), runs(tmp_path, "b", 32)
    B = pkg.Batch([r.ctx for r in batched])
    loop = plain[0].loop
    steps = 8
    for _ in range(steps):
        for r in plain:
            r.ctx.time_step(*loop)
        B.time_step(*loop, mixed=(200, (1e-7,) * 3, 6, 0.95, 2, 80, 16))
    for a, b in zip(plain, batched):
        st = b.ctx.predict_stats()
        assert st[0] + st[4] == steps, st
        for x, y in zip(state(a), state(b)):
            assert float(np.linalg.norm(x - y) / np.linalg.norm(x)) <= 1e-6
    B.close()
    for r in plain + batched:
        r.close()


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_anderson_batch_follows_single(variant, tmp_path):
    """Anderson acceleration (depth 2) from the degree-3 predicted start in
    shared launches, per tile shape: the same acceleration as each case
    stepped alone, up to the order of the dot-product sums (workgroups of
    tiles instead of runs of the flattened grid), so the outer-iteration
    counts agree to one and the fields to 1e-9 relative L2 after 8 steps, and
    both meet the reference's test on every step with fewer iterations than
    the plain batched step."""
    pkg = importlib.import_module("simple-ts_amd")
    single, batched, plain = runs(tmp_path, "s"), runs(tmp_path, "b"), runs(tmp_path, "p")
    B = pkg.Batch([r.ctx for r in batched])
    P = pkg.Batch([r.ctx for r in plain])
    B.set_tiles(variant, variant)
    P.set_tiles(variant, variant)
    loop = single[0].loop
    acc = (0, (1e-7,) * 3, 6, 0.95, 3, 0, 64, 2)
    tot_b = tot_p = 0
    for step in range(8):
        it_single = [r.ctx.time_step_mixed(*loop, *acc[:4], extrapolate=3, guard=64, anderson=2)[1] for r in single]
        _, it_batch = B.time_step(*loop, mixed=acc)
        _, it_plain = P.time_step(*loop)
        print(f"variant {variant} step {step}: single {it_single}, batched {it_batch}, plain {it_plain}", flush=True)
        assert all(abs(a - b) <= 1 for a, b in zip(it_batch, it_single))
        tot_b += sum(it_batch)
        tot_p += sum(it_plain)
    for a, b in zip(single, batched):
        for x, y in zip(state(a), state(b)):
            assert np.all(np.isfinite(y))
            assert np.linalg.norm(x - y) <= 1e-9 * np.linalg.norm(x)
    assert tot_b < tot_p
    for X in (B, P):
        X.close()
    for r in single + batched + plain:
        r.close()
