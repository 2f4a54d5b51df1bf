"""CPU restatement of the reference's main() control flow over the oracle's
stages (test infrastructure only).

Setup sequence S:2476-3226 and loops 1-3 S:3230-4691 of SIMPLE_TS_MPI.cpp for
the options the time-step library provides (no exact-solution start, p_BC_xb_correction_method 0, no pressure-ratio correction).  Mesh, mask
and boxes come from the host setup library (simple_ts_host.h), which is
checked separately against the reference's own mesh and mask files.
"""
import importlib
import math
import os
import sys

import numpy as np

from . import oracle as ora

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _ROOT not in sys.path:
    sys.path.insert(0, _ROOT)
sts = importlib.import_module("simple-ts_amd")


class Run:
    """Result of a run: final fields, per-step outer-iteration counts and the
    residual log (the numbers Tmp.txt carries)."""

    def __init__(self):
        self.iters = []
        self.residuals = []
        self.fields = {}


def run_case(case_dir, max_steps=None, on_step=None, on_iteration=None):
    c = sts.Case(case_dir)
    prob = c.problem()
    o = ora.Oracle(prob)
    Na = prob.Na
    m = prob.mask.astype(bool)
    V = c.value
    type_nd = int(V("TypeOfNonDimensiolization"))
    for flag in ("ToContinueFromInterpolation", "ToStartFromExactSolutionForChannelFlowWithPressureBC",
                 "Pressure_ratio_correct", "ToReadSolvedDataFromFile"):
        if V(flag):
            raise NotImplementedError(flag)
    N_I, N_I_p_c = max(1, int(V("N_I"))), int(V("N_I_p_c"))
    top = (prob.Ny - 1) * prob.Nx  # S:6906-6912: rho_v one row up, outside the arrays on the top row
    if (V("drhodt_durhodx_dvrhody_0_BC_xb") and m[top + prob.boxes["Nx_u"][0] - 1]) or (
            V("drhodt_durhodx_dvrhody_0_BC_xe") and m[top + prob.boxes["Nx_u"][1] - 1]):
        raise NotImplementedError("drhodt_durhodx_dvrhody edge condition on a fluid cell of the top row")
    Nt = V("Nt") if max_steps is None else min(V("Nt"), max_steps)
    p_BC_xb = V("p_BC_xb")
    u_given_xb = V("u_given_xb")
    u_in_it = u_given_xb if type_nd in (1, 2) else 0.0

    def scalars():
        if type_nd in (1, 2):
            cT = (1.0 / V("Re")) * (u_in_it * p_BC_xb / V("T_BC_xb"))
            cPch = 0.5 if type_nd == 1 else u_in_it * u_in_it / (V("gamma1") * V("Ma") * V("Ma"))
        else:
            cT, cPch = V("Kn") * 5.0 * math.sqrt(math.pi) / 16.0, 0.5
        o.set_scalars(cT, cPch)

    def pressure_begin():
        nonlocal u_given_xb, u_in_it, p_BC_xb
        if prob.periodic_x:  # S:7436
            return
        if V("GivenReKn"):
            u_given_xb = V("Re") * V("Kn") * math.sqrt(15.0 * math.pi / 128.0)
        u_in_it = u_given_xb
        if V("Given_velocity_on_xb"):
            jj = int((prob.boxes["Ny_o"][1] - 1.0) - prob.boxes["Ny_o"][0])
            p_BC_xb = o.f["p"][prob.boxes["Nx_o"][0] + jj * prob.Nx]

    def columns(with_rho):
        Nx = prob.Nx
        for j in range(prob.Ny):
            for i in list(range(0, prob.boxes["Nx_o"][0])) + list(range(prob.boxes["Nx_o"][1], Nx)):
                n = i + j * Nx
                if not m[n]:
                    continue
                lo = i < prob.boxes["Nx_o"][0]
                pb, Tb = (V("p_BC_xb"), V("T_BC_xb")) if lo else (V("p_BC_xe"), V("T_BC_xe"))
                o.f["p"][n], o.f["Temper"][n] = pb, Tb
                if with_rho:
                    o.f["rho"][n] = pb / Tb

    # S:2476-2483 uniform fields
    o.f["u"][:] = V("u_gas_nd")
    o.f["v"][:] = V("v_gas_nd")
    o.f["p"][:] = V("p_gas_nd")
    o.f["rho"][:] = c.value("rho_gas_nd")
    o.f["Temper"][:] = V("T_gas_nd")
    # S:2502-2864
    o.apply_bodies()
    if V("Pressure_BC"):
        columns(True)
    o.save_previous()
    o.rho_in_middle_points()
    o.apply_bodies()
    o.boundary_conditions(4)
    o.save_previous()
    o.rho_in_middle_points()
    o.jump_overwrite()  # WriteSolvedDataToFile
    for k in ("d_p_c_12", "d_p_c_34", "apx", "apy", "bpx", "bpy", "Gamma_yf", "Gamma_xf", "Fx", "Fy", "u_pseudo",
              "v_pseudo", "D_ux", "D_uy", "D_vx", "D_vy", "DTx", "DTy", "ScT1", "rho_u", "rho_v"):
        o.f[k][:] = 0.0
    if V("Pressure_BC"):
        columns(False)
    o.properties()
    o.rho_in_middle_points()
    o.save_previous()
    scalars()
    o.rho_in_middle_points()
    o.boundary_conditions(1 | 2 | 4)
    pressure_begin()
    o.boundary_conditions(8 | 16 | 32)
    o.fluxes_diffusion()

    run = Run()
    save_every, save_ctr = int(V("Nt_save_solved_data")), 0
    it, it_counter = 0.0, 0
    ht = V("ht")
    eV, ep, eT, es = V("MaxError_Velocities"), V("MaxError_p_c"), V("MaxError_T"), V("MaxError_to_stop_program")
    while True:  # loop 1
        o.save_previous()
        it += ht
        it_counter += 1
        Iter = 0
        step_res = []
        while True:  # loop 2
            o.assemble()
            rpt = o.pt_jacobi(N_I_p_c)
            ruv = o.correct()
            scalars()
            o.boundary_conditions(1 | 2 | 4)
            pressure_begin()
            o.boundary_conditions(8 | 16 | 32)
            o.fluxes_diffusion()
            Iter += 1
            if on_iteration:
                on_iteration(it_counter, Iter, o)
            rp, rT = rpt[-1]
            step_res.append((ruv[0], ruv[1], rp, rT, rpt[0][0], rpt[0][1]))
            if V("ToSolveContinuityEquation"):  # S:4586-4589 (fluxes recomputed first)
                o.fluxes()
                run.continuity = o.continuity()
            cont_iter = Iter < N_I and (eV < ruv[0] or eV < ruv[1] or ep < rp or eT < rT)
            cont_time = it_counter < Nt and (
                (not (Iter <= 1 and N_I_p_c <= 2)) or es < ruv[0] or es < ruv[1] or es < rpt[0][0]
                or es < rpt[0][1])
            if not cont_iter:
                break
        if V("use_interpolation_for_u_BC_by_time") and not prob.periodic_x:  # S:8102-8144
            tb, te, ub, ue = V("t_BC_b"), V("t_BC_e"), V("u_BC_xb_t_BC_b"), V("u_BC_xb_t_BC_e")
            if it <= tb:
                ug = ub
            elif tb < it <= te:
                ug = ((ue - ub) / (te - tb)) * (it - tb) + ub
            else:
                ug = ue
            u_given_xb = ug
            for j in range(prob.Ny):
                for i in range(prob.boxes["Nx_u"][0]):
                    if m[i + j * prob.Nx]:
                        o.f["u"][i + j * prob.Nx] = ug
        run.iters.append(Iter)
        run.residuals.append(step_res)
        save_ctr += 1
        if save_ctr >= save_every:  # WriteSolvedDataToFile at the save cadence (S:4625-4650)
            o.jump_overwrite()
            if V("ToSolveContinuityEquation"):
                o.fluxes()
                run.continuity = o.continuity()
            save_ctr = 0
        if on_step:
            on_step(it_counter, o)
        if not cont_time:
            break
    o.jump_overwrite()
    run.fields = {k: o.get(k) for k in ("u", "v", "p", "Temper")}
    run.oracle = o
    run.it = it
    return run
