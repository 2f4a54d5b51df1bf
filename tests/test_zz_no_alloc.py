"""The allocation-free time step (sts_time_step_noalloc) against sts_time_step,
through the C ABI: two calls in a row, the second on a smaller case than the
first (the fixed per-GPU buffer then holds the larger call's contents), each
bit-identical to the allocating entry point on the same state (fields,
outer-iteration count, residual log, probe).  Sizes beyond the buffer are
refused with STS_ERR_ARG."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "p", "Temper", "rho", "Fx", "DTy")


def _pair(tmp_path, tag, Nx, Ny):
    cases = importlib.import_module("simple-ts_amd.cases")
    ws = importlib.import_module("simple-ts_amd.warmstart")
    runs = []
    for side in ("alloc", "noalloc"):
        d = cases.confined_square(str(tmp_path / f"{tag}_{side}"), Nx, Ny, 1415, 0.00283, 4,
                                  dt_over_tin=0.01 * 240 / Nx, N_I=2000)
        runs.append(ws.Run(d))
    return runs


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("shape", [((240, 48), (120, 24))])
def test_noalloc_step_matches_time_step(shape, tmp_path, sts):
    for call, (Nx, Ny) in enumerate(shape):
        ref, mine = _pair(tmp_path, f"c{call}", Nx, Ny)
        for step in range(2):
            print(f"call {call} ({Nx}x{Ny}) step {step}", flush=True)
            n_a, log_a, pr_a = ref.ctx.time_step(*ref.loop)
            n_b, log_b, pr_b = mine.ctx.time_step_noalloc(*mine.loop)
            assert n_a == n_b > 1
            assert _same(log_a, log_b)
            assert _same(pr_a, pr_b)
            for f in FIELDS:
                assert _same(ref.ctx.get(f), mine.ctx.get(f)), (call, step, f)
        ref.close()
        mine.close()


def test_noalloc_rejects_a_log_beyond_its_buffer(tmp_path, sts):
    _, mine = _pair(tmp_path, "big", 60, 14)
    N_I, n_pc, tu, tp, tT = mine.loop
    with pytest.raises(sts.StsError, match="STS_NOALLOC_MAX_N_I"):
        mine.ctx.time_step_noalloc(4097, n_pc, tu, tp, tT)
    n, _, _ = mine.ctx.time_step_noalloc(4096, n_pc, tu, tp, tT)
    assert n > 0
    mine.close()
