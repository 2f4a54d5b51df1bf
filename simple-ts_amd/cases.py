"""Case directories in the reference's input format (EnteredData.txt and
body files, Objects_for_Solving.h:306-412), and the benchmark workloads.

The reference ships one input (a pressure-driven microchannel); the other
workloads -- subsonic and supersonic flow past a confined square, and
Rayleigh-Benard boxes -- are written here in the same format.  The shipped
EnteredData.txt (tests/golden/shipped) is the template that provides every
line's comment text; entries are overridden by their variable names.
"""
import math
import os
import shutil

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SHIPPED = os.path.join(ROOT, "tests", "golden", "shipped")

NAMES = None


def entry_names():
    """Variable names of the 126 EnteredData.txt lines, from their comments."""
    global NAMES
    if NAMES is None:
        names = []
        with open(os.path.join(SHIPPED, "EnteredData.txt")) as f:
            for line in f:
                com = line.split("//", 1)[1].strip()
                names.append(com.split()[0].split("=")[0].strip())
        NAMES = names
    return NAMES


def fmt(v):
    if isinstance(v, bool):
        return str(int(v))
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(v))


def write_circles(dst, circles):
    with open(os.path.join(dst, "solve_circles_b.txt"), "w") as f:
        f.write(f"{len(circles)}\t//N_circles\n")
        for b in circles:
            f.write(f"{fmt(b['x_c'])}\t{fmt(b['y_c'])}\t{fmt(b['r_c'])}\t//x_c y_c r_c\n")
            f.write(f"{fmt(b.get('u_body', 0.0))}\t{fmt(b.get('v_body', 0.0))}\t//u_body v_body\n")
            f.write("0\t//p_body\n")
            f.write(f"{fmt(b.get('T_body', 1.0))}\t//T_body\n")
            f.write(f"{int(b.get('slip', 0))}\t//is_VelocitySlipBC_body\n")
            f.write(f"{fmt(b.get('F_slip', 1.1466))}\t//F_VelocitySlip\n")
            f.write("0\t//w_VelocitySlipBC\n")
            f.write(f"{int(b.get('jump', 0))}\t//is_TemperatureJumpBC_body\n")
            f.write(f"{fmt(b.get('F_jump', 2.1904))}\t//F_TemperatureJump\n")
            f.write(f"{fmt(b.get('dTdn', 0.0))}\t//dTdn_on_wall_0\n")
            f.write(f"{int(b.get('drag', 0))}\t//ToSolveDragCoefficientForThisBody\n")


def write_case(dst, overrides=None, polyhedrons=None, template=SHIPPED, circles=None):
    """Writes dst/EnteredData.txt (template with overrides) and
    dst/solve_polyhedrons_b.txt (template's, or the given polyhedron list:
    dicts with vertices [(x, y), ...], inside points p1, p2 and body
    parameters)."""
    os.makedirs(dst, exist_ok=True)
    names = entry_names()
    lines = open(os.path.join(template, "EnteredData.txt")).read().splitlines()
    overrides = dict(overrides or {})
    out = []
    for k, line in enumerate(lines):
        val, rest = line.split("\t", 1)
        if names[k] in overrides:
            val = fmt(overrides.pop(names[k]))
        out.append(val + "\t" + rest)
    if overrides:
        raise KeyError(f"unknown entries: {sorted(overrides)}")
    with open(os.path.join(dst, "EnteredData.txt"), "w") as f:
        f.write("\n".join(out) + "\n")
    if polyhedrons is None:
        shutil.copy(os.path.join(template, "solve_polyhedrons_b.txt"), dst)
    else:
        with open(os.path.join(dst, "solve_polyhedrons_b.txt"), "w") as f:
            f.write(f"{len(polyhedrons)}\t//N_polyhedrons\n")
            for b in polyhedrons:
                v = b["vertices"]
                f.write(f"{len(v)}\t//N_vertex_body\n")
                f.write(f"{fmt(b['p1'][0])}\t{fmt(b['p1'][1])}\t//point1_inside_x point1_inside_y\n")
                f.write(f"{fmt(b['p2'][0])}\t{fmt(b['p2'][1])}\t//point2_inside_x point2_inside_y\n")
                f.write(f"{fmt(b.get('u_body', 0.0))}\t{fmt(b.get('v_body', 0.0))}\t//u_body v_body\n")
                f.write("0\t//pressure in polyhedron - must be 0\n")
                f.write(f"{fmt(b.get('T_body', 1.0))}\t//Temperature in polyhedron\n")
                f.write(f"{int(b.get('slip', 0))}\t//is_VelocitySlipBC_body\n")
                f.write(f"{fmt(b.get('F_slip', 1.1466))}\t//F_VelocitySlip\n")
                f.write("0\t//w_VelocitySlipBC\n")
                f.write(f"{int(b.get('jump', 0))}\t//is_TemperatureJumpBC_body\n")
                f.write(f"{fmt(b.get('F_jump', 2.1904))}\t//F_TemperatureJump\n")
                f.write(f"{fmt(b.get('dTdn', 0.0))}\t//dTdn_on_wall_0\n")
                f.write(f"{int(b.get('drag', 0))}\t//ToSolveDragCoefficientForThisBody\n")
                for x, y in v:
                    f.write(f"{fmt(x)}\t{fmt(y)}\n")
    if circles:
        write_circles(dst, circles)
    return dst


def u_inlet(Re, Kn):
    """Inlet velocity of a (Re, Kn) pair, u = Re*Kn*sqrt(15*pi/128), the
    relation the reference applies when GivenReKn is set (S:2022, S:7491)."""
    return Re * Kn * math.sqrt(15.0 * math.pi / 128.0)


def confined_square(dst, Nx, Ny, Re, Kn, steps, dt_over_tin=0.01, L=15.0, H=1.0, side=0.25, xc=4.0, yc=0.5,
                    tol=1e-8, N_I=5000, N_I_p_c=2, extra=None, wall_slip=0, circles=None, drag=0):
    """Flow past a confined square in a channel of length L and height H:
    square of side `side` centred at (xc, yc), isothermal no-slip walls (T=1),
    uniform inlet velocity u_in = u_inlet(Re, Kn) held on the first two u
    columns, outlet pressure fixed (the last two p columns are never updated,
    S:6120-6122), uniform initial field, dt = dt_over_tin * H / u_in.

    Nx*Ny is the reference's array shape, boundary cells included (the
    update boxes of S:6009-6267 leave out row 0 and the last two rows), so
    the channel holds Ny-3 fluid rows between one bottom and two top wall rows,
    and the cell centres run from x = 0 (inlet column) to x = L."""
    hx, hy = L / (Nx - 1), H / (Ny - 3)
    u_in = u_inlet(Re, Kn)
    s2 = 0.5 * side
    big = 1e5
    walls = [
        dict(vertices=[(-big, -big), (big, -big), (big, 0.0), (-big, 0.0)], p1=(0.0, -5e4), p2=(1.0, -6e4),
             slip=wall_slip),
        dict(vertices=[(-big, H), (big, H), (big, big), (-big, big)], p1=(0.0, 5e4), p2=(1.0, 6e4), slip=wall_slip),
    ]
    if side > 0:
        walls.append(dict(vertices=[(xc - s2, yc - s2), (xc + s2, yc - s2), (xc + s2, yc + s2), (xc - s2, yc + s2)],
                          p1=(xc - 0.3 * s2, yc - 0.2 * s2), p2=(xc + 0.3 * s2, yc + 0.4 * s2), drag=drag))
    ov = dict(u_gas_nd=u_in, v_gas_nd=0.0, T_gas_nd=1.0, p_gas_nd=1.0, Kn=Kn, Re=Re, TypeOfNonDimensiolization=3,
              Nt=steps, Nx=Nx, Ny=Ny, N_I=N_I, N_I_p_c=N_I_p_c, N_I_T=0, ht=dt_over_tin * H / u_in,
              x_b=0.0, x_e=L, y_b=-0.5 * hy, y_e=H + 1.5 * hy, kind_of_mesh=5, hxmin=hx, hxmax=hx, hymin=hy,
              hymax=hy, MaxError_Velocities=tol, MaxError_p_c=tol, MaxError_T=tol, MaxError_to_stop_program=1e-14,
              max_time_for_solve=1000.0, Nt_save_solved_data=10 ** 9, Nt_save_DragCoefficient=10 ** 9,
              Nt_save_temp_data=1, Given_velocity_on_xb=1, Given_velocity_on_xe=0, Pressure_BC=0, GivenReKn=1,
              u_given_xb=u_in, drhodt_durhodx_dvrhody_0_BC_xb=0, drhodt_durhodx_dvrhody_0_BC_xe=0, dudx_0_BC_xe=1,
              dvdx_0_BC_xb=0, dvdx_0_BC_xe=1, dTdx_0_BC_xe=1, dpdx_0_BC_xb=1, To_Use_Kn_local_in_wall_BC=0,
              N_SubDomains_x=1)
    if circles:
        ov["ToImport_data_for_circles_from_file_b"] = 1
    ov.update(extra or {})
    return write_case(dst, ov, walls, circles=circles)


def rayleigh_benard(dst, Nx, Ny, Kn=0.05, Fr=1.0, steps=10, dt=0.01, W=2.0, H=1.0, T_hot=2.0, T_cold=1.0,
                    tol=1e-8, N_I=5000, N_I_p_c=2, jump=False, extra=None, periodic=False, circles=None):
    """Closed box W*H of gas at rest between a hot bottom wall (T_hot) and a
    cold top wall (T_cold), side walls at the mean temperature, gravity along
    -y with c_Fr_v = 2/(gamma*Fr*Ma^2) (S:2066).  As in confined_square, the
    fluid occupies columns 1..Nx-3 and rows 1..Ny-3 of the Nx*Ny array.

    periodic: no side walls; the layer is periodic about OX with period W
    over all Nx columns (cell centres at (i + 1/2) W/Nx).  circles: extra
    bodies (write_circles dicts), e.g. a heated cylinder."""
    hx, hy = (W / Nx if periodic else W / (Nx - 3)), H / (Ny - 3)
    Tm = 0.5 * (T_hot + T_cold)
    big = 1e5
    walls = [
        dict(vertices=[(-big, -big), (big, -big), (big, 0.0), (-big, 0.0)], p1=(0.0, -5e4), p2=(1.0, -6e4),
             T_body=T_hot, jump=int(jump)),
        dict(vertices=[(-big, H), (big, H), (big, big), (-big, big)], p1=(0.0, 5e4), p2=(1.0, 6e4), T_body=T_cold,
             jump=int(jump)),
    ]
    if not periodic:
        walls += [
            dict(vertices=[(-big, 0.0), (0.0, 0.0), (0.0, H), (-big, H)], p1=(-5e4, 0.3 * H), p2=(-6e4, 0.6 * H),
                 T_body=Tm, jump=int(jump)),
            dict(vertices=[(W, 0.0), (big, 0.0), (big, H), (W, H)], p1=(5e4, 0.3 * H), p2=(6e4, 0.6 * H),
                 T_body=Tm, jump=int(jump)),
        ]
    x_b, x_e = (0.0, W) if periodic else (-0.5 * hx, W + 1.5 * hx)
    ov = dict(u_gas_nd=0.0, v_gas_nd=0.0, T_gas_nd=Tm, p_gas_nd=1.0, Kn=Kn, Fr=Fr, SolveGravityField=1,
              TypeOfNonDimensiolization=3, Nt=steps, Nx=Nx, Ny=Ny, N_I=N_I, N_I_p_c=N_I_p_c, N_I_T=0, ht=dt,
              x_b=x_b, x_e=x_e, y_b=-0.5 * hy, y_e=H + 1.5 * hy, kind_of_mesh=5, hxmin=hx, hxmax=hx,
              hymin=hy, hymax=hy, MaxError_Velocities=tol, MaxError_p_c=tol, MaxError_T=tol,
              MaxError_to_stop_program=1e-14, max_time_for_solve=1000.0, Nt_save_solved_data=10 ** 9,
              Nt_save_DragCoefficient=10 ** 9, Nt_save_temp_data=1, Given_velocity_on_xb=0, Given_velocity_on_xe=0,
              Pressure_BC=0, GivenReKn=0, drhodt_durhodx_dvrhody_0_BC_xb=0, drhodt_durhodx_dvrhody_0_BC_xe=0,
              dvdx_0_BC_xb=0, dvdx_0_BC_xe=0, dTdx_0_BC_xe=0, To_Use_Kn_local_in_wall_BC=1, N_SubDomains_x=1,
              Periodic_boundary_conditions_about_OX=int(periodic))
    if circles:
        ov["ToImport_data_for_circles_from_file_b"] = 1
    ov.update(extra or {})
    return write_case(dst, ov, walls, circles=circles)


def periodic_channel(dst, Nx, Ny, u0, Kn, steps, dt=0.01, W=2.0, H=1.0, tol=1e-8, N_I=5000, N_I_p_c=2,
                     slip=1, jump=1, seam_block=True, circles=None, top_wall=True, extra=None):
    """Channel periodic about OX (period W over all Nx columns) between
    walls at y = 0 and y = H (T = 1, velocity slip and temperature jump
    switchable), gas starting at a uniform velocity u0 that the walls and
    bodies slow down.  seam_block: staggered blocks at the periodic seam
    x = 0 = W: one from the seam to x = 0.15 W over 0.4-0.8 H, one over
    0.2-0.6 H that ends one column short of the seam, and one over 0.08-0.2 H
    that reaches the seam from the left, so that across the seam every
    fluid/body combination of the last two and first two columns occurs.
    top_wall=False leaves the top rows fluid and asks for zero-gradient
    conditions there, which the reference switches off for periodic cases
    (S:6795, S:7340, S:7655): the top rows then keep their initial state."""
    hx, hy = W / Nx, H / (Ny - 3)
    big = 1e5
    walls = [
        dict(vertices=[(-big, -big), (big, -big), (big, 0.0), (-big, 0.0)], p1=(0.0, -5e4), p2=(1.0, -6e4),
             slip=slip, jump=jump),
    ]
    if top_wall:
        walls.append(dict(vertices=[(-big, H), (big, H), (big, big), (-big, big)], p1=(0.0, 5e4), p2=(1.0, 6e4),
                          slip=slip, jump=jump, T_body=1.0))
    if seam_block:
        w, xr = 0.15 * W, W - 1.2 * hx
        walls += [
            dict(vertices=[(-1.0, 0.4 * H), (w, 0.4 * H), (w, 0.8 * H), (-1.0, 0.8 * H)], p1=(0.2 * w, 0.5 * H),
                 p2=(0.3 * w, 0.7 * H), slip=slip, jump=jump, T_body=1.3),
            dict(vertices=[(xr - w, 0.2 * H), (xr, 0.2 * H), (xr, 0.6 * H), (xr - w, 0.6 * H)],
                 p1=(xr - 0.5 * w, 0.3 * H), p2=(xr - 0.4 * w, 0.5 * H), slip=slip, jump=jump, T_body=0.8),
            dict(vertices=[(W - 3.5 * hx, 0.08 * H), (W + 1.0, 0.08 * H), (W + 1.0, 0.2 * H), (W - 3.5 * hx, 0.2 * H)],
                 p1=(W - 2.0 * hx, 0.1 * H), p2=(W - 1.0 * hx, 0.15 * H), slip=slip, jump=jump, T_body=1.1),
        ]
    ov = dict(u_gas_nd=u0, v_gas_nd=0.0, T_gas_nd=1.0, p_gas_nd=1.0, Kn=Kn, TypeOfNonDimensiolization=3,
              Nt=steps, Nx=Nx, Ny=Ny, N_I=N_I, N_I_p_c=N_I_p_c, N_I_T=0, ht=dt, x_b=0.0, x_e=W, y_b=-0.5 * hy,
              y_e=H + 1.5 * hy, kind_of_mesh=5, hxmin=hx, hxmax=hx, hymin=hy, hymax=hy, MaxError_Velocities=tol,
              MaxError_p_c=tol, MaxError_T=tol, MaxError_to_stop_program=1e-14, max_time_for_solve=1000.0,
              Nt_save_solved_data=10 ** 9, Nt_save_DragCoefficient=10 ** 9, Nt_save_temp_data=1,
              Given_velocity_on_xb=0, Given_velocity_on_xe=0, Pressure_BC=0, GivenReKn=0, SolveGravityField=0,
              drhodt_durhodx_dvrhody_0_BC_xb=0, drhodt_durhodx_dvrhody_0_BC_xe=0, dudx_0_BC_xe=0, dvdx_0_BC_xb=0,
              dvdx_0_BC_xe=0, dTdx_0_BC_xe=0, dpdx_0_BC_xb=0, To_Use_Kn_local_in_wall_BC=1, N_SubDomains_x=1,
              Periodic_boundary_conditions_about_OX=1)
    if not top_wall:
        ov.update(dudy_0_BC_ye=1, dvdy_0_BC_ye=1, dpdy_0_BC_ye=1, dTdy_0_BC_ye=1, dudx_0_BC_xe=1, dTdx_0_BC_xe=1)
    if circles:
        ov["ToImport_data_for_circles_from_file_b"] = 1
    ov.update(extra or {})
    return write_case(dst, ov, walls, circles=circles)


# Mode-3 constants (S:2013-2071): gamma = 5/3, Ma = sqrt(6/5), Pr = 2/3,
# c_mu = (5/16) sqrt(2 pi / gamma); cT = Kn 5 sqrt(pi)/16 multiplies sqrt(T)
# in the momentum diffusion, cT1 = gamma c_mu Kn / (Ma Pr) in the heat
# diffusion, and c_Fr_v = 2 / (gamma Fr Ma^2) = 1/Fr is the gravity term.
def rayleigh_fr(Ra, Kn, T_hot=2.0, T_cold=1.0, p=1.0):
    """Fr that gives a box of height H = 1 the Rayleigh number Ra.

    The reference's momentum equation is rho Du/Dt = ... + div(cT sqrt(T)
    grad u) - c_Fr_v rho e_y and its temperature equation rho DT/Dt = ... +
    div(cT1 sqrt(T) grad T) (the ht*sum(aT) + hx*hy*rho denominators of
    S:3973-3983), so at the mean state (T_m, rho_m = p/T_m) the kinematic
    viscosity is nu = cT sqrt(T_m)/rho_m, the thermal diffusivity
    alpha = cT1 sqrt(T_m)/rho_m, the gravity g = c_Fr_v, and an ideal gas has
    beta = 1/T_m.  Ra = g beta dT H^3 / (nu alpha) then gives
    Fr = dT rho_m^2 / (Ra T_m^2 cT cT1).  This relation is not in the
    reference; it is the Boussinesq reading of its nondimensional equations."""
    gamma, Ma, Pr = 5.0 / 3.0, math.sqrt(6.0 / 5.0), 2.0 / 3.0
    c_mu = (5.0 / 16.0) * math.sqrt(2.0 * math.pi / gamma)
    cT = Kn * 5.0 * math.sqrt(math.pi) / 16.0
    cT1 = gamma * c_mu * Kn / (Ma * Pr)
    Tm = 0.5 * (T_hot + T_cold)
    rho_m = p / Tm
    return (T_hot - T_cold) * rho_m ** 2 / (Ra * Tm ** 2 * cT * cT1)


def batch_cases(root, n=64, steps=300, seed=0, nx_range=(256, 1024), Ra=3000.0):
    """The mixed set of independent cases stepped together: seeded grids
    Nx ~ U{256..1024} (multiples of 64) with Nx:Ny = 4:1; the first half are
    confined squares with Ma ~ U[0.1, 2.5], Kn ~ log-U[1e-3, 5e-2] and
    dt = 0.005 (240/Nx) H/u_in; the second half Rayleigh-Benard boxes (2:1,
    T = 2 below, 1 above, Kn = 0.05, Fr from rayleigh_fr(Ra)) with
    dt = 8e-4 (256/Nx)^2.  Returns [(case_dir, description)]."""
    rng = np.random.default_rng(seed)
    out = []
    sq = math.sqrt(15 * math.pi / 128)
    for k in range(n):
        Nx = int(rng.integers(nx_range[0] // 64, nx_range[1] // 64 + 1)) * 64
        Ny = Nx // 4
        d = os.path.join(root, f"case{k:02d}")
        if k < n // 2:
            Ma = rng.uniform(0.1, 2.5)
            Kn = 10 ** rng.uniform(math.log10(1e-3), math.log10(5e-2))
            confined_square(d, Nx, Ny, Ma / (Kn * sq), Kn, steps, dt_over_tin=0.005 * 240 / Nx)
            out.append((d, f"square {Nx}x{Ny} Ma {Ma:.2f} Kn {Kn:.4f}"))
        else:
            Fr = rayleigh_fr(Ra, 0.05)
            # the buoyancy-driven start needs dt ~ 1/Nx^2 (1e-3 is the limit at Nx = 256)
            rayleigh_benard(d, Nx, Ny, Kn=0.05, Fr=Fr, steps=steps, dt=8e-4 * (256.0 / Nx) ** 2)
            out.append((d, f"rayleigh-benard {Nx}x{Ny} Ra {Ra:g} Fr {Fr:.4f}"))
    return out
