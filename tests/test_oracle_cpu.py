"""Pins the CPU restatement (oracle/) before it is trusted as the checker:

* bit-exact against the reference solver's own outputs for every golden case
  (final u, v, p, Temper after N steps, per-step outer-iteration counts,
  continuity fields), and against the intermediate arrays the reference's
  DEBUG_PROGRAM build dumps;
* against facts the mathematics fixes: a gas at rest with uniform p and T is
  a discrete fixed point to the last bit; the loop-3 sweep count composes
  (two single sweeps == one double sweep); the mirror image of a symmetric
  channel stays symmetric.
"""
import os

import numpy as np
import pytest

from golden_cases import RUN_CASES, fresh_input, ref_dir, ref_field, ref_iterations, same_bits


@pytest.mark.parametrize("name", RUN_CASES)
def test_oracle_matches_reference_outputs(name, tmp_path):
    from oracle import flow

    d = fresh_input(name, tmp_path)
    run = flow.run_case(d)
    assert run.iters == ref_iterations(name), "outer-iteration counts per step differ"
    for f in ("u", "v", "p", "Temper"):
        ref = ref_field(name, f)
        assert same_bits(run.fields[f], ref), f"{name}: {f} differs (max {np.max(np.abs(run.fields[f] - ref))})"
    cpath = os.path.join(ref_dir(name), "ContinuityEquation.bin")
    if os.path.exists(cpath):
        cs, cf = run.continuity
        assert same_bits(cf, np.fromfile(cpath)), "ContinuityEquation differs"
        assert same_bits(cs, np.fromfile(os.path.join(ref_dir(name), "ContinuityEquation_stationary.bin")))


DBG_LAST_ITER = ["u_pseudo", "v_pseudo", "d_p_c_12", "d_p_c_34", "apx", "apy", "bpx", "bpy", "Fx", "Fy",
                 "Gamma_xf", "Gamma_yf", "D_ux", "D_uy", "D_vx", "D_vy", "DTx", "DTy", "rho_u", "rho_v", "u", "v", "p",
                 "Temper", "Temper_of_fluid_on_the_wall_u", "Temper_of_fluid_on_the_wall_v"]


def test_oracle_matches_reference_debug_dumps(tmp_path):
    """square_dbg: one step of two outer iterations.  Each dump holds its
    array's last value; rho is dumped inside loop 3, i.e. it is the value left
    by the previous outer iteration."""
    from oracle import flow

    run = flow.run_case(fresh_input("square_dbg", tmp_path))
    o = run.oracle
    for f in DBG_LAST_ITER:
        ref = np.fromfile(os.path.join(ref_dir("square_dbg"), f"dbg_{f}.bin"))
        assert same_bits(o.get(f), ref), f"{f} differs from the reference's dump"
    # rho after the first outer iteration
    import shutil

    d1 = fresh_input("square_dbg", tmp_path / "one")
    text = open(os.path.join(d1, "EnteredData.txt")).read().split("\n")
    text[16] = "1" + text[16][text[16].index("\t"):]  # N_I = 1
    open(os.path.join(d1, "EnteredData.txt"), "w").write("\n".join(text))
    r1 = flow.run_case(d1)
    assert same_bits(r1.oracle.get("rho"), np.fromfile(os.path.join(ref_dir("square_dbg"), "dbg_rho.bin")))
    shutil.rmtree(d1)


def test_gas_at_rest_is_a_fixed_point(tmp_path):
    """u = v = 0, uniform p = 1 and T = 1, no gravity: every flux and source
    vanishes and the T and p updates reduce to x/x, so the state is
    reproduced exactly, step after step."""
    import importlib

    from oracle import flow

    cases = importlib.import_module("simple-ts_amd.cases")
    d = cases.rayleigh_benard(str(tmp_path / "rest"), 20, 12, Fr=1.0, steps=3, N_I=3, tol=0.0, T_hot=1.0,
                              T_cold=1.0, extra=dict(SolveGravityField=0))
    run = flow.run_case(d)
    m = run.oracle.problem.mask.astype(bool)
    assert np.all(run.fields["u"][m] == 0.0) and np.all(run.fields["v"][m] == 0.0)
    assert np.all(run.fields["p"][m] == 1.0) and np.all(run.fields["Temper"][m] == 1.0)


def test_sweeps_compose(oracle):
    """Loop 3 restarts from the current p and T, so one call of two sweeps
    equals two calls of one sweep (catches a lost or stale buffer)."""
    from cases import channel_case, random_state

    prob = channel_case(Nx=33, Ny=14, seed=2)
    st = random_state(prob, seed=4)
    a, b = oracle.Oracle(prob), oracle.Oracle(prob)
    for k, v in st.items():
        a.set(k, v)
        b.set(k, v)
    ra = a.pt_jacobi(2)
    rb1 = b.pt_jacobi(1)
    rb2 = b.pt_jacobi(1)
    assert np.array_equal(ra[0], rb1[0]) and np.array_equal(ra[1], rb2[0])
    for f in ("p", "Temper"):
        assert same_bits(a.get(f), b.get(f))
    # three sweeps equal one then two
    c3, c12 = oracle.Oracle(prob), oracle.Oracle(prob)
    for k, v in st.items():
        c3.set(k, v)
        c12.set(k, v)
    c3.pt_jacobi(3)
    c12.pt_jacobi(1)
    c12.pt_jacobi(2)
    for f in ("p", "Temper"):
        assert same_bits(c3.get(f), c12.get(f))


def test_mirror_symmetry_of_a_symmetric_channel(tmp_path):
    """Square centred on the channel axis: u, p, T even and v odd about the
    axis.  The scheme is symmetric up to the order of floating-point sums in
    mirrored stencils, so the mirror agrees to round-off."""
    import importlib

    from oracle import flow

    cases = importlib.import_module("simple-ts_amd.cases")
    Nx, Ny = 50, 17
    d = cases.confined_square(str(tmp_path / "sym"), Nx, Ny, 85, 0.01, 3, N_I=6, tol=0.0)
    run = flow.run_case(d)
    F = {k: v.reshape(Ny, Nx) for k, v in run.fields.items()}
    rows = np.arange(1, Ny - 2)           # fluid rows 1..Ny-3, mirror j <-> Ny-2-j
    mir = Ny - 2 - rows
    for k in ("u", "p", "Temper"):
        a, b = F[k][rows], F[k][mir]
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a)), k
    # v faces: face j (between rows j-1 and j) mirrors face Ny-1-j
    faces = np.arange(2, Ny - 2)
    a, b = F["v"][faces], -F["v"][Ny - 1 - faces]
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(F["u"]))
    assert np.max(np.abs(F["v"][faces])) > 1e-6  # the flow is not trivially at rest
