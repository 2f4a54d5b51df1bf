"""Synthetic SIMPLE-TS cases for stage-level parity tests.

A channel with wall bodies along the bottom and top rows, a square body and a
one-cell notch (so that body corners, thin bodies and both slip branches
occur), on a slightly non-uniform mesh, with seeded random fields.  The solve
boxes follow define_area_for_solving (SIMPLE_TS_MPI.cpp:6009-6267).
"""
import importlib

import numpy as np

sts = importlib.import_module("simple-ts_amd")


def boxes_for(Nx, Ny, pressure_bc=False, given_xb=False, given_xe=False, periodic=False):
    """Solve boxes of the serial build (S:6009-6267); periodic about OX the
    x boxes span every column (S:6055-6102)."""
    if periodic:
        xu = xv = xo = (0, Nx)
    elif pressure_bc:
        xu, xv, xo = (1, Nx - 1), (1, Nx - 2), (1, Nx - 2)
    else:
        xu = (2 if given_xb else 1, Nx - 1)
        xv = (1, Nx - 2)
        xo = (1, Nx - 1 if given_xe else Nx - 2)
    return {"Nx_u": xu, "Ny_u": (1, Ny - 2), "Nx_v": xv, "Ny_v": (1, Ny - 1), "Nx_o": xo, "Ny_o": (1, Ny - 2)}


def channel_case(Nx=40, Ny=16, seed=0, slip=(1, 0, 1, 1), jump=(0, 1, 1, 0), kn_local=1, given_xb=1,
                 gravity=0.0, bc=None, uniform=False, periodic=False):
    rng = np.random.default_rng(seed)
    if uniform:
        hx = np.full(Nx, 0.05)
        hy = np.full(Ny, 0.04)
    else:
        hx = 0.05 * (1.0 + 0.2 * rng.random(Nx))
        hy = 0.04 * (1.0 + 0.2 * rng.random(Ny))
    vol = np.zeros((Ny, Nx), dtype=np.uint32)
    vol[0, :] = 1                     # bottom wall
    vol[Ny - 1, :] = 2                # top wall
    a, b = Nx // 3, Nx // 3 + max(2, Nx // 8)
    c, d = Ny // 3, Ny // 3 + max(2, Ny // 4)
    vol[c:d, a:b] = 3                 # square
    vol[1, Nx // 2 + 3] = 4           # one-cell notch on the bottom wall
    vol[Ny - 2, Nx // 2 - 2: Nx // 2] = 4
    if periodic:
        # bodies against the periodic seam: a block across it, one that stops
        # a column short of it, and single cells in the first and last columns
        r = Ny // 2
        vol[r - 2:r + 1, Nx - 3:] = 3
        vol[r - 2:r + 1, :2] = 3
        vol[r + 2:r + 4, Nx - 4:Nx - 1] = 4
        vol[2, 0] = 4
        vol[Ny - 3, Nx - 1] = 4
    bodies = []
    for k in range(4):
        bodies.append(dict(u_body=0.0, v_body=0.0, p_body=0.0, T_body=1.0 + 0.25 * k, is_slip=int(slip[k]),
                           F_slip=1.1466, is_jump=int(jump[k]), F_jump=2.1904, dTdn_on_wall_0=0.0))
    bodies[2]["dTdn_on_wall_0"] = 1.0 if not jump[2] else 0.0
    Kn = 0.05
    gamma1 = 5.0 / 3.0
    Ma, Pr = np.sqrt(6.0 / 5.0), 2.0 / 3.0
    c_mu = (5.0 / 16.0) * np.sqrt(2.0 * np.pi / gamma1)
    scal = dict(ht=0.01, gamma1=gamma1, Kn=Kn, cT=Kn * 5.0 * np.sqrt(np.pi) / 16.0,
                cT1=gamma1 * c_mu * Kn / (Ma * Pr), cT2=gamma1 * (gamma1 - 1.0) * c_mu * Kn * Ma, cPch=0.5,
                c_Fr_v=gravity)
    if bc is None:
        bc = dict(dvdx_0_xb=1, dvdx_0_xe=1, dudx_0_xe=1, dTdx_0_xe=1, dpdx_0_xb=1, dTdy_0_yb=1, dudy_0_ye=1)
    return sts.Problem(hx, hy, vol.reshape(-1), bodies,
                       boxes_for(Nx, Ny, given_xb=bool(given_xb), periodic=periodic), scal,
                       given_xb=given_xb, kn_local=kn_local, bc=bc, periodic_x=int(periodic))


def random_state(problem, seed=1, zero_frac=0.05):
    """Physically plausible random fields for every array of the time step."""
    rng = np.random.default_rng(seed)
    n = problem.Na
    m = problem.mask.astype(bool)

    def U(lo, hi):
        return lo + (hi - lo) * rng.random(n)

    st = {}
    st["u"] = U(-0.3, 0.8)
    st["v"] = U(-0.4, 0.4)
    for k in ("u", "v"):
        z = rng.random(n) < zero_frac
        st[k][z] = 0.0
    st["p"] = U(0.6, 1.4)
    st["Temper"] = U(0.7, 1.4)
    st["rho"] = np.where(m, st["p"] / st["Temper"], 0.0)
    st["sqrt_T"] = np.sqrt(st["Temper"])
    st["u_pr"] = st["u"] + U(-0.01, 0.01)
    st["v_pr"] = st["v"] + U(-0.01, 0.01)
    st["p_pr"] = st["p"] * U(0.99, 1.01)
    st["Temper_pr"] = st["Temper"] * U(0.99, 1.01)
    st["rho_pr"] = np.where(m, st["p_pr"] / st["Temper_pr"], 0.0)
    st["rho_u"] = U(0.4, 1.6)
    st["rho_v"] = U(0.4, 1.6)
    st["Fx"] = U(-0.05, 0.08)
    st["Fy"] = U(-0.05, 0.05)
    for k in ("Gamma_xf", "Gamma_yf"):
        st[k] = U(0.8, 1.2)
    for k in ("D_ux", "D_uy", "D_vx", "D_vy", "DTx", "DTy"):
        st[k] = U(0.001, 0.02)
    for k in ("u_pseudo", "v_pseudo"):
        st[k] = U(-0.3, 0.8)
    for k in ("d_p_c_12", "d_p_c_34"):
        st[k] = U(0.01, 0.05)
    for k in ("apx", "apy"):
        st[k] = U(0.0, 0.05)
    for k in ("bpx", "bpy"):
        st[k] = U(-0.02, 0.04)
    st["ScT1"] = U(0.0, 0.003)
    st["Temper_of_fluid_on_the_wall_u"] = U(0.8, 1.3)
    st["Temper_of_fluid_on_the_wall_v"] = U(0.8, 1.3)
    return st


def bit_mismatch(a, b):
    """Number of entries whose IEEE-754 bits differ (signed zeros count)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return int(np.count_nonzero(a.view(np.uint64) != b.view(np.uint64)))
