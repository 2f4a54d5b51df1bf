"""End-to-end parity of the GPU program (sts_gpu) with the reference solver:
the same case directory in, the same files out, byte for byte (FP64)."""
import importlib
import os
import shutil
import subprocess

import numpy as np
import pytest

from casefiles import REF_BIN, have_reference, run_reference
from golden_cases import RUN_CASES, fresh_input, ref_dir

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "simple-ts_amd", "sts_gpu")
SAME_BYTES = ["u.bin", "v.bin", "p.bin", "Temper.bin", "vol_inf_fb.bin", "x_f.txt", "y_f.txt", "hx.txt", "hy.txt",
              "x_v.txt", "y_v.txt", "Tmp.txt", "Iterations_in_loops.txt", "it_begin.txt",
              "ContinuityEquation.bin", "ContinuityEquation_stationary.bin"]


def run_driver(d, timeout=300, env=None):
    print(f"sts_gpu {d}", flush=True)
    e = dict(os.environ)
    e.update(env or {})
    subprocess.run([DRIVER, d], check=True, timeout=timeout, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, env=e)


def compare_dirs(mine, ref):
    bad = []
    drag = sorted(f for f in os.listdir(ref) if f.startswith("CD."))  # WriteDragCoefficient files
    for f in SAME_BYTES + drag:
        rp = os.path.join(ref, f)
        if not os.path.exists(rp):
            continue
        a = open(os.path.join(mine, f), "rb").read()
        b = open(rp, "rb").read()
        if a != b:
            bad.append(f)
    # EnteredData.txt: line 44 holds the run's elapsed hours
    a = open(os.path.join(mine, "EnteredData.txt")).read().split("\n")
    b = open(os.path.join(ref, "EnteredData.txt")).read().split("\n")
    if [x for k, x in enumerate(a) if k != 43] != [x for k, x in enumerate(b) if k != 43]:
        bad.append("EnteredData.txt")
    return bad


@pytest.mark.parametrize("fused", ["1", "0"])
@pytest.mark.parametrize("name", RUN_CASES)
def test_driver_matches_reference_files(name, fused, tmp_path):
    """fused: loop 3 as one two-sweep pass (k_sweep2) or as single sweeps."""
    d = fresh_input(name, tmp_path)
    run_driver(d, env={"STS_FUSED_SWEEPS": fused})
    bad = compare_dirs(d, ref_dir(name))
    assert not bad, f"{name}: files differing from the reference's: {bad}"


@pytest.mark.skipif(not have_reference(), reason="oracle/_ref/sts_ref not built")
def test_driver_matches_live_reference_subsonic_square(tmp_path):
    """The subsonic confined square on the 240x48 grid for 12 steps, each
    converged to 1e-8, against the reference run in the same job (dt =
    0.005 H/u_in: at 0.01 H/u_in the reference itself diverges on this grid)."""
    cases = importlib.import_module("simple-ts_amd.cases")
    mine = cases.confined_square(str(tmp_path / "gpu"), 240, 48, 85, 0.00194, 12, dt_over_tin=0.005)
    ref = cases.confined_square(str(tmp_path / "ref"), 240, 48, 85, 0.00194, 12, dt_over_tin=0.005)
    run_reference(ref)
    run_driver(mine)
    bad = compare_dirs(mine, ref)
    assert not bad, f"files differing from the reference's: {bad}"


@pytest.mark.skipif(not have_reference(), reason="oracle/_ref/sts_ref not built")
@pytest.mark.parametrize("Nx", [256, 200])
def test_driver_matches_live_reference_periodic_layer(Nx, tmp_path):
    """A Rayleigh-Benard layer periodic about OX with a heated cylinder, on
    a grid whose width is (256) and is not (200) a whole number of sweep
    tiles, 6 steps each converged to 1e-8, against the reference."""
    cases = importlib.import_module("simple-ts_amd.cases")

    def make(d):
        return cases.rayleigh_benard(str(d), Nx, 40, W=4.0, Fr=cases.rayleigh_fr(3000.0, 0.05), steps=6, dt=2e-3,
                                     periodic=True, jump=True,
                                     circles=[dict(x_c=3.9, y_c=0.35, r_c=0.2, T_body=2.0, jump=1)])
    mine, ref = make(tmp_path / "gpu"), make(tmp_path / "ref")
    run_reference(ref)
    run_driver(mine)
    bad = compare_dirs(mine, ref)
    assert not bad, f"files differing from the reference's: {bad}"


@pytest.mark.skipif(not have_reference(), reason="oracle/_ref/sts_ref not built")
@pytest.mark.parametrize("n_pc", [3, 5])
def test_driver_odd_sweep_counts_match_live_reference(n_pc, tmp_path):
    """Odd N_I_p_c leaves each outer iteration's p/T in the partner buffers;
    the device-side loop must follow that parity.  (N_I_p_c = 1 diverges in
    the reference, as its input notes warn.)"""
    cases = importlib.import_module("simple-ts_amd.cases")
    kw = dict(N_I=400, N_I_p_c=n_pc, extra=dict(Nt_save_temp_data=1))
    mine = cases.confined_square(str(tmp_path / "gpu"), 60, 14, 85, 0.00194, 4, **kw)
    ref = cases.confined_square(str(tmp_path / "ref"), 60, 14, 85, 0.00194, 4, **kw)
    run_reference(ref)
    run_driver(mine)
    bad = compare_dirs(mine, ref)
    assert not bad, f"files differing from the reference's: {bad}"


def _set_entries(case_dir, **values):
    """Rewrites entries of a case's EnteredData.txt (by the reference's line
    order) in place."""
    cases = importlib.import_module("simple-ts_amd.cases")
    names = cases.entry_names()
    path = os.path.join(case_dir, "EnteredData.txt")
    lines = open(path).read().split("\n")
    for k, name in enumerate(names):
        if name in values:
            val, rest = lines[k].split("\t", 1)
            lines[k] = cases.fmt(values[name]) + "\t" + rest
    open(path, "w").write("\n".join(lines))


@pytest.mark.skipif(not have_reference(), reason="oracle/_ref/sts_ref not built")
def test_driver_restart_matches_live_reference(tmp_path):
    """Restart (ToReadSolvedDataFromFile, S:5734-5835, S:6762-6787): both
    programs run 3 steps, then continue 2 more from the files they wrote;
    every file of the continued runs is the reference's, byte for byte."""
    cases = importlib.import_module("simple-ts_amd.cases")
    kw = dict(extra=dict(Nt_save_solved_data=3, Nt_save_temp_data=1))
    mine = cases.confined_square(str(tmp_path / "gpu"), 60, 14, 85, 0.00194, 3, **kw)
    ref = cases.confined_square(str(tmp_path / "ref"), 60, 14, 85, 0.00194, 3, **kw)
    run_reference(ref)
    run_driver(mine)
    assert not compare_dirs(mine, ref)
    # both programs write their elapsed hours over max_time_for_solve at the
    # end (S:4702); a restart would inherit that as its time limit, an input
    # that depends on the wall clock, so the continued runs get a fixed one
    for d in (mine, ref):
        _set_entries(d, ToReadSolvedDataFromFile=1, Nt=2, max_time_for_solve=1000.0)
    run_reference(ref)
    run_driver(mine)
    bad = compare_dirs(mine, ref)
    assert not bad, f"files differing from the reference's after the restart: {bad}"


@pytest.mark.skipif(not have_reference(), reason="oracle/_ref/sts_ref not built")
@pytest.mark.parametrize("name", ["shipped_coarse", "square_cd"])
def test_short_format_and_inflow_files_match_live_reference(name, tmp_path):
    """DataInShortFormat.txt (Re, Kn, CD, CD_p, u_gas_nd, the mass flows at
    the inlet and at column Nx_u[1], the box-end pressures; S:8438-8500) and
    p_and_u_xb_it.txt (inflow pressure and velocities when Pressure_BC is
    set; S:8504-8565), written at every drag-coefficient write: the
    pressure-driven microchannel (Pressure_BC, no drag bodies) and the
    confined square with drag bodies, 3 steps, byte for byte against the
    reference run in the same job; then 2 more steps continued from the saved
    files (p_and_u_xb_it.txt appended to)."""
    mine, ref = fresh_input(name, tmp_path / "gpu"), fresh_input(name, tmp_path / "ref")
    for d in (mine, ref):
        _set_entries(d, WriteDataInShortFormat_kind=1, Nt_save_DragCoefficient=1, Nt=3, Nt_save_solved_data=3)
    run_reference(ref)
    run_driver(mine)
    files = ["DataInShortFormat.txt"] + (["p_and_u_xb_it.txt"] if name == "shipped_coarse" else [])
    for f in files:
        assert os.path.exists(os.path.join(ref, f)), f
        assert open(os.path.join(mine, f), "rb").read() == open(os.path.join(ref, f), "rb").read(), f
    assert not compare_dirs(mine, ref)
    for d in (mine, ref):
        _set_entries(d, ToReadSolvedDataFromFile=1, Nt=2, max_time_for_solve=1000.0)
    run_reference(ref)
    run_driver(mine)
    for f in files:
        assert open(os.path.join(mine, f), "rb").read() == open(os.path.join(ref, f), "rb").read(), f + " (restart)"
    assert not compare_dirs(mine, ref)


@pytest.mark.skipif(not have_reference(), reason="oracle/_ref/sts_ref not built")
@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("nd_type", [3, 1])
def test_inflow_velocity_correction_methods_match_live_reference(method, nd_type, tmp_path):
    """p_BC_xb_correction_method 1 / 2 (S:7496-7521): the step's inflow
    velocity is the largest / the hy-weighted mean u of the inflow column,
    which enters the inflow trace and, with the nondimensionalisation of type
    1, cT of every outer iteration (S:4347).  The pressure-driven
    microchannel, 3 steps, every file byte for byte against the reference."""
    mine, ref = fresh_input("shipped_coarse", tmp_path / "gpu"), fresh_input("shipped_coarse", tmp_path / "ref")
    for d in (mine, ref):
        _set_entries(d, p_BC_xb_correction_method=method, TypeOfNonDimensiolization=nd_type, Nt=3,
                     Nt_save_DragCoefficient=1)
    run_reference(ref)
    run_driver(mine)
    f = "p_and_u_xb_it.txt"
    assert open(os.path.join(mine, f), "rb").read() == open(os.path.join(ref, f), "rb").read()
    bad = compare_dirs(mine, ref)
    assert not bad, f"files differing from the reference's: {bad}"


@pytest.mark.skipif(not have_reference(), reason="oracle/_ref/sts_ref not built")
@pytest.mark.parametrize("name", ["shipped_coarse", "square"])
def test_exact_solution_start_matches_live_reference(name, tmp_path):
    """ToStartFromExactSolutionForChannelFlowWithPressureBC (S:2566-2830).
    The pressure-driven microchannel starts from the exact slip-flow solution
    between its box-end pressures (p on the solve columns, u on every fluid
    cell; polyhedron 0's slip coefficient); the confined square, whose inflow
    velocity is given, takes the parabolic profile on its inflow columns and
    the outflow pressure that gives u_given_xb at the channel's centre.
    3 steps, every file byte for byte against the reference run in the same
    job (EnteredData.txt carries the box-end pressures it rewrote)."""
    mine, ref = fresh_input(name, tmp_path / "gpu"), fresh_input(name, tmp_path / "ref")
    for d in (mine, ref):
        _set_entries(d, ToStartFromExactSolutionForChannelFlowWithPressureBC=1, Nt=3, Nt_save_solved_data=3,
                     Nt_save_DragCoefficient=1, WriteDataInShortFormat_kind=1)
    run_reference(ref)
    run_driver(mine)
    for f in ("DataInShortFormat.txt",):
        assert open(os.path.join(mine, f), "rb").read() == open(os.path.join(ref, f), "rb").read(), f
    bad = compare_dirs(mine, ref)
    assert not bad, f"files differing from the reference's: {bad}"


@pytest.mark.skipif(not have_reference(), reason="oracle/_ref/sts_ref not built")
def test_continue_from_interpolation_matches_live_reference(tmp_path):
    """ToContinueFromInterpolation (S:2982-2991): a run continued from saved
    fields first applies the body conditions, the velocity-slip and
    temperature-jump wall conditions and the body-cell temperature overwrite
    to them.  The microchannel (slip walls), 3 steps, then 2 more continued
    that way; every file byte for byte against the reference."""
    mine, ref = fresh_input("shipped_coarse", tmp_path / "gpu"), fresh_input("shipped_coarse", tmp_path / "ref")
    for d in (mine, ref):
        _set_entries(d, Nt=3, Nt_save_solved_data=3)
    run_reference(ref)
    run_driver(mine)
    assert not compare_dirs(mine, ref)
    for d in (mine, ref):
        _set_entries(d, ToReadSolvedDataFromFile=1, ToContinueFromInterpolation=1, Nt=2, max_time_for_solve=1000.0)
    run_reference(ref)
    run_driver(mine)
    bad = compare_dirs(mine, ref)
    assert not bad, f"files differing from the reference's after the continued run: {bad}"


@pytest.mark.skipif(not have_reference(), reason="oracle/_ref/sts_ref not built")
@pytest.mark.parametrize("method", [0, 2])
@pytest.mark.parametrize("at_xb", [1, 0])
def test_pressure_ratio_correction_matches_live_reference(method, at_xb, tmp_path):
    """Pressure_ratio_correct (S:7525-7640), inside every outer iteration:
    p_correction adapts to the correction's velocity residuals, and the
    inflow (correct_p_BC_xb = 1) or outflow pressure moves by p_correction
    times the squared miss of u_given_xb when the step's inflow velocity (the
    given one, or with p_BC_xb_correction_method 2 the measured mean) lies
    outside its band; that end's columns take the pressure and its density.
    The microchannel from rest with a target velocity near its developing
    flow (u_given_xb 0.05; the shipped 2.2147 drives the pressure up
    tenfold per iteration and the reference itself overflows), p_correction
    at most 0.1, 3 steps, every file byte for byte against the reference
    (p_and_u_xb_it.txt traces p_BC_xb and p_correction)."""
    mine, ref = fresh_input("shipped_coarse", tmp_path / "gpu"), fresh_input("shipped_coarse", tmp_path / "ref")
    for d in (mine, ref):
        _set_entries(d, Pressure_ratio_correct=1, correct_p_BC_xb=at_xb, p_BC_xb_correction_method=method,
                     p_correction_auto=1, p_correction_max=0.1, u_given_xb=0.05, Nt=3, Nt_save_DragCoefficient=1,
                     WriteDataInShortFormat_kind=1)
    run_reference(ref)
    run_driver(mine)
    for f in ("p_and_u_xb_it.txt", "DataInShortFormat.txt"):
        assert open(os.path.join(mine, f), "rb").read() == open(os.path.join(ref, f), "rb").read(), f
    bad = compare_dirs(mine, ref)
    assert not bad, f"files differing from the reference's: {bad}"
