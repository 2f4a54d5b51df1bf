"""Writes the golden fixtures under tests/golden/ from the reference solver.

The reference (kshterev/SIMPLE-TS) ships no tests, fixtures or outputs; these
are produced by running its own serial build (oracle/_ref/sts_ref, built from
the reference sources by oracle/Makefile.ref) on small cases in the
reference's input format:

  <case>/input/  EnteredData.txt and body files given to the reference
  <case>/ref/    the files the reference wrote: final u, v, p, Temper and
                 vol_inf_fb (.bin), the mesh text files, Tmp.txt,
                 Iterations_in_loops.txt, it_begin.txt, the rewritten
                 EnteredData.txt, ContinuityEquation*.bin when enabled
  square_dbg/ref/dbg_*.bin  the intermediate arrays the reference's
                 DEBUG_PROGRAM build dumps (after 1 step of 2 outer
                 iterations)

Run from the repository root:  python tests/golden/make_golden.py [case ...]
(no names: every case).
"""
import importlib
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
cases = importlib.import_module("simple-ts_amd.cases")

REF = os.path.join(ROOT, "oracle", "_ref", "sts_ref")
REF_DBG = os.path.join(ROOT, "oracle", "_ref", "sts_ref_dbg")

KEEP = ["u.bin", "v.bin", "p.bin", "Temper.bin", "vol_inf_fb.bin", "x_f.txt", "y_f.txt", "hx.txt", "hy.txt",
        "x_v.txt", "y_v.txt", "Tmp.txt", "Iterations_in_loops.txt", "it_begin.txt", "EnteredData.txt",
        "ContinuityEquation.bin", "ContinuityEquation_stationary.bin", "CD.0.txt", "CD.1.txt", "CD.2.txt", "CD.3.txt"]

MESH_KEEP = ["vol_inf_fb.bin", "x_f.txt", "y_f.txt", "hx.txt", "hy.txt", "x_v.txt", "y_v.txt", "EnteredData.txt"]

DBG_ARRAYS = ["DTx", "DTy", "D_ux", "D_uy", "D_vx", "D_vy", "Fx", "Fy", "Gamma_xf", "Gamma_yf", "Temper",
              "Temper_of_fluid_on_the_wall_u", "Temper_of_fluid_on_the_wall_v", "apx", "apy", "bpx", "bpy",
              "d_p_c_12", "d_p_c_34", "p", "rho", "rho_u", "rho_v", "u", "u_pseudo", "v", "v_pseudo"]

CONV = dict(N_I=200, MaxError_Velocities=1e-8, MaxError_p_c=1e-8, MaxError_T=1e-8)


def case_specs():
    """name -> function(dir) writing the case input."""
    return {
        # the shipped microchannel on a coarser mesh: pressure-driven, slip walls
        # (y_e keeps the top row inside the upper body: a fluid top row would make
        # the reference's drhodt_... edge condition read past its arrays, S:6906)
        "shipped_coarse": lambda d: cases.write_case(d, dict(hxmin=0.1, hxmax=0.1, hymin=0.1, hymax=0.1, y_e=2.15,
                                                             Nt=6, Nt_save_temp_data=1, **CONV)),
        # subsonic confined square with time-dependent inflow and continuity output
        "square": lambda d: cases.confined_square(
            d, 60, 14, 85, 0.00194, 4, extra=dict(ToSolveContinuityEquation=1, use_interpolation_for_u_BC_by_time=1,
                                                  t_BC_b=0.0, t_BC_e=100.0, u_BC_xb_t_BC_b=0.1,
                                                  u_BC_xb_t_BC_e=0.2, Nt_save_solved_data=2)),
        # supersonic confined square
        "square_super": lambda d: cases.confined_square(d, 60, 14, 1415, 0.00283, 4),
        # drag coefficients of a circle and the square (WriteDragCoefficient, S:8147-8430)
        "square_cd": lambda d: cases.confined_square(
            d, 60, 14, 85, 0.00194, 5, drag=1, extra=dict(Nt_save_DragCoefficient=2),
            circles=[dict(x_c=8.0, y_c=0.5, r_c=0.2, drag=1)]),
        # heated box with gravity and temperature jump
        "rb_jump": lambda d: cases.rayleigh_benard(d, 24, 12, Fr=1.0, steps=3, N_I=40, tol=0.0, jump=True),
        # circle body in a slip channel, local Kn on the walls
        "circle": lambda d: _circle_case(d),
        # periodic about OX: channel with slip/jump walls, staggered blocks
        # across the seam and a heated cylinder, gas started at u = 0.3
        "periodic_channel": lambda d: cases.periodic_channel(
            d, 40, 14, 0.3, 0.02, 4, circles=[dict(x_c=1.0, y_c=0.45, r_c=0.18, T_body=1.2, slip=1, jump=1)]),
        # periodic about OX with an open top whose zero-gradient conditions
        # the reference switches off for periodic cases
        "periodic_open": lambda d: cases.periodic_channel(d, 40, 14, 0.3, 0.02, 3, top_wall=False),
        # periodic Rayleigh-Benard layer with a hot cylinder
        "periodic_rb": lambda d: cases.rayleigh_benard(
            d, 32, 12, Fr=0.5, steps=3, N_I=60, jump=True, periodic=True,
            circles=[dict(x_c=0.6, y_c=0.3, r_c=0.15, T_body=2.0, jump=1)]),
        # mesh kinds 1-4 (setup checks)
        "mesh1": lambda d: _mesh_case(d, 1),
        "mesh2": lambda d: _mesh_case(d, 2),
        "mesh3": lambda d: _mesh_case(d, 3),
        "mesh4": lambda d: _mesh_case(d, 4),
    }


def _circle_case(d):
    cases.confined_square(d, 48, 16, 100, 0.01, 3, side=0.0, wall_slip=1,
                          circles=[dict(x_c=4.0, y_c=0.45, r_c=0.2, T_body=1.2, slip=1, jump=1)],
                          extra=dict(To_Use_Kn_local_in_wall_BC=1))


def _mesh_case(d, kind):
    cases.write_case(d, dict(kind_of_mesh=kind, hxmin=0.05, hxmax=0.2, hymin=0.03, hymax=0.1, xfmin=0.5, xfmax=-1.0,
                             xbmin=9.5, xbmax=11.0, xmid=5.0, ybmin=0.3, ybmax=-0.5, ytmin=0.7, ytmax=1.5, ymid=0.5,
                             Nx=60, Ny=20, N_I_T=3, Nt=1, N_I=1, x_e=11.0, y_e=1.0, x_b=-1.0, y_b=-0.2))


def run(binary, d):
    subprocess.run([binary], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.STDOUT, check=True, timeout=600)


def main():
    for path in (REF, REF_DBG):
        if not os.access(path, os.X_OK):
            sys.exit(f"{path} missing: make -f oracle/Makefile.ref")
    only = set(sys.argv[1:])
    for name, make in case_specs().items():
        if only and name not in only:
            continue
        base = os.path.join(HERE, name)
        shutil.rmtree(base, ignore_errors=True)
        inp, ref = os.path.join(base, "input"), os.path.join(base, "ref")
        make(inp)
        shutil.copytree(inp, ref)
        run(REF, ref)
        keep = MESH_KEEP if name.startswith("mesh") else KEEP
        for f in os.listdir(ref):
            if f not in keep:
                os.remove(os.path.join(ref, f))
        print(name, sorted(os.listdir(ref)))
    if only and "square_dbg" not in only:
        return
    # DEBUG_PROGRAM dumps of one step with two outer iterations
    base = os.path.join(HERE, "square_dbg")
    shutil.rmtree(base, ignore_errors=True)
    inp, ref = os.path.join(base, "input"), os.path.join(base, "ref")
    cases.confined_square(inp, 40, 12, 85, 0.00194, 1, N_I=2, tol=0.0)
    shutil.copytree(inp, ref)
    run(REF_DBG, ref)
    for f in os.listdir(ref):
        stem = f[:-4] if f.endswith(".bin") else None
        if stem in DBG_ARRAYS:
            os.rename(os.path.join(ref, f), os.path.join(ref, "dbg_" + f))
        elif f not in KEEP:
            os.remove(os.path.join(ref, f))
    print("square_dbg", sorted(os.listdir(ref)))


if __name__ == "__main__":
    main()
