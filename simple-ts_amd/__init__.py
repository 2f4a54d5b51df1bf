"""simple-ts_amd: MI355X-native SIMPLE-TS time step (host-side Python mirror).

The product is the C ABI in include/simple_ts.h, implemented by the HIP
library libsimple_ts_hip.so in this directory, and the host driver
``sts_gpu`` (C++) that mirrors the reference's main() loops.  This module
binds the same C ABI with ctypes for tests and benchmarks; it never falls
back to a CPU path: loading fails loudly if the HIP library is missing.

The reference (kshterev/SIMPLE-TS) exposes no API of its own; the methods of
:class:`Context` carry the names of the reference's call sites in
SIMPLE_TS_MPI.cpp that they replace (cited as S:line).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_HIP = os.path.join(HERE, "libsimple_ts_hip.so")
LIB_HOST = os.path.join(HERE, "libsimple_ts_host.so")
DRIVER = os.path.join(HERE, "sts_gpu")

STS_OK = 0
ERRORS = {-1: "STS_ERR_ARG", -2: "STS_ERR_HIP", -3: "STS_ERR_NONFINITE", -4: "STS_ERR_UNSUPPORTED",
          -5: "STS_ERR_IO"}

# enum sts_field (include/simple_ts.h), the reference's array names
FIELDS = ["u", "v", "p", "Temper", "rho", "sqrt_T", "u_pr", "v_pr", "p_pr", "Temper_pr", "rho_pr",
          "rho_u", "rho_v", "Fx", "Fy", "Gamma_xf", "Gamma_yf", "D_ux", "D_uy", "D_vx", "D_vy", "DTx", "DTy",
          "u_pseudo", "v_pseudo", "d_p_c_12", "d_p_c_34", "apx", "apy", "bpx", "bpy", "ScT1",
          "Temper_of_fluid_on_the_wall_u", "Temper_of_fluid_on_the_wall_v"]
FIELD_ID = {n: k for k, n in enumerate(FIELDS)}

BC_SLIP_U, BC_SLIP_V, BC_VELOCITIES, BC_JUMP, BC_TEMPERATURE, BC_PRESSURE = 1, 2, 4, 8, 16, 32
BC_ALL = 63

# edge boundary-condition switches in sts_grid order (EnteredData.txt:81-109)
BC_FLAGS = ["dudx_0_xb", "durhodx_0_xb", "drhodt_xb", "dudx_0_xe", "durhodx_0_xe", "drhodt_xe",
            "u_Min_Mout_xe", "dudy_0_yb", "dudy_0_ye", "dvdx_0_xb", "dvdx_0_xe", "dvdy_0_yb", "dvdy_0_ye",
            "dpdx_0_xb", "dpdx_0_xe", "dpdy_0_yb", "dpdy_0_ye", "dTdx_0_xb", "dTdx_0_xe", "dTdy_0_yb",
            "dTdy_0_ye"]


class Body(ctypes.Structure):
    _fields_ = [("u_body", ctypes.c_double), ("v_body", ctypes.c_double), ("p_body", ctypes.c_double),
                ("T_body", ctypes.c_double), ("is_slip", ctypes.c_int), ("F_slip", ctypes.c_double),
                ("is_jump", ctypes.c_int), ("F_jump", ctypes.c_double), ("dTdn_on_wall_0", ctypes.c_double)]


_I2 = ctypes.c_int * 2
_DP = ctypes.POINTER(ctypes.c_double)
_UP = ctypes.POINTER(ctypes.c_uint)
_FP = ctypes.POINTER(ctypes.c_float)
_BP = ctypes.POINTER(ctypes.c_ubyte)
_IP = ctypes.POINTER(ctypes.c_int)


def grid_fields(with_mask):
    """ctypes field list of sts_grid (and of the oracle's ora_grid, which
    carries the boolean mask pointer after vol_inf_fb)."""
    f = [("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("hx", _DP), ("hy", _DP), ("vol_inf_fb", _UP)]
    if with_mask:
        f.append(("m", ctypes.POINTER(ctypes.c_ubyte)))
    f += [("nbodies", ctypes.c_int), ("bodies", ctypes.POINTER(Body))]
    f += [(n, _I2) for n in ("Nx_u", "Ny_u", "Nx_v", "Ny_v", "Nx_o", "Ny_o")]
    f += [(n, ctypes.c_double) for n in ("ht", "gamma1", "Kn", "cT", "cT1", "cT2", "cPch", "c_Fr_v")]
    f += [(n, ctypes.c_int) for n in ("given_xb", "given_xe", "kn_local")]
    f += [(n, ctypes.c_int) for n in BC_FLAGS]
    f.append(("periodic_x", ctypes.c_int))
    return f


class Grid(ctypes.Structure):
    _fields_ = grid_fields(False)


class LoopControl(ctypes.Structure):
    _fields_ = [("N_I", ctypes.c_int), ("N_I_p_c", ctypes.c_int), ("tol_uv", ctypes.c_double),
                ("tol_p", ctypes.c_double), ("tol_T", ctypes.c_double), ("batch", ctypes.c_int)]


class MixedControl(ctypes.Structure):
    _fields_ = [("N_I32", ctypes.c_int), ("tol32_uv", ctypes.c_double), ("tol32_p", ctypes.c_double),
                ("tol32_T", ctypes.c_double), ("stall_w", ctypes.c_int), ("stall_ratio", ctypes.c_double),
                ("extrapolate", ctypes.c_int), ("fp32_above", ctypes.c_int), ("guard", ctypes.c_int),
                ("anderson", ctypes.c_int), ("anderson_every", ctypes.c_int), ("anderson32", ctypes.c_int)]


class UNetNorm(ctypes.Structure):
    _fields_ = [("u_ref", ctypes.c_double), ("p_ref", ctypes.c_double), ("dp_scale", ctypes.c_double),
                ("dT_scale", ctypes.c_double)]


class CorrNorm(ctypes.Structure):
    """sts_unet_corr_norm: the corrector's normalisation."""
    _fields_ = [("u_ref", ctypes.c_double), ("p_ref", ctypes.c_double), ("step", ctypes.c_double),
                ("corr", ctypes.c_double)]


class Problem:
    """A SIMPLE-TS case as the time step sees it: mesh, body mask and table,
    solve boxes, scalars and boundary-condition switches.  Keeps the numpy
    arrays alive while ctypes structures point into them."""

    SCALARS = ("ht", "gamma1", "Kn", "cT", "cT1", "cT2", "cPch", "c_Fr_v")

    def __init__(self, hx, hy, vol_inf_fb, bodies, boxes, scalars, given_xb=0, given_xe=0, kn_local=0,
                 bc=None, periodic_x=0):
        self.hx = np.ascontiguousarray(hx, dtype=np.float64)
        self.hy = np.ascontiguousarray(hy, dtype=np.float64)
        self.Nx, self.Ny = self.hx.size, self.hy.size
        self.vol = np.ascontiguousarray(vol_inf_fb, dtype=np.uint32).reshape(-1)
        if self.vol.size != self.Nx * self.Ny:
            raise ValueError("vol_inf_fb must have Nx*Ny entries")
        self.mask = (self.vol == 0).astype(np.uint8)
        self.bodies = list(bodies)
        self.boxes = {k: tuple(int(x) for x in v) for k, v in boxes.items()}
        self.scalars = dict(scalars)
        self.given_xb, self.given_xe, self.kn_local = int(given_xb), int(given_xe), int(kn_local)
        self.periodic_x = int(periodic_x)
        self.bc = {k: 0 for k in BC_FLAGS}
        if bc:
            for k, v in bc.items():
                if k not in self.bc:
                    raise KeyError(k)
                self.bc[k] = int(v)
        self._body_arr = (Body * max(1, len(self.bodies)))()
        for k, b in enumerate(self.bodies):
            self._body_arr[k] = Body(**b) if isinstance(b, dict) else b

    @property
    def Na(self):
        return self.Nx * self.Ny

    def fill(self, st):
        """Fill a ctypes sts_grid / ora_grid structure."""
        st.Nx, st.Ny = self.Nx, self.Ny
        st.hx = self.hx.ctypes.data_as(_DP)
        st.hy = self.hy.ctypes.data_as(_DP)
        st.vol_inf_fb = self.vol.ctypes.data_as(_UP)
        if hasattr(st, "m"):
            st.m = self.mask.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))
        st.nbodies = len(self.bodies)
        st.bodies = ctypes.cast(self._body_arr, ctypes.POINTER(Body))
        for k in ("Nx_u", "Ny_u", "Nx_v", "Ny_v", "Nx_o", "Ny_o"):
            setattr(st, k, _I2(*self.boxes[k]))
        for k in self.SCALARS:
            setattr(st, k, float(self.scalars[k]))
        st.given_xb, st.given_xe, st.kn_local = self.given_xb, self.given_xe, self.kn_local
        st.periodic_x = self.periodic_x
        for k, v in self.bc.items():
            setattr(st, k, v)
        return st


_lib = None


def lib():
    """The HIP library.  Raises if it is not built: there is no CPU path."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_HIP):
            raise RuntimeError(f"{LIB_HIP} is not built; run __graft_entry__.build() (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_HIP)
        c_int, c_double, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        L.sts_last_error.restype = ctypes.c_char_p
        L.sts_create.argtypes = [ctypes.POINTER(Grid), c_int, ctypes.POINTER(vp)]
        for name, args in [
            ("sts_destroy", [vp]), ("sts_set_field", [vp, c_int, _DP]), ("sts_get_field", [vp, c_int, _DP]),
            ("sts_predict_stats", [vp, ctypes.POINTER(ctypes.c_longlong)]),
            ("sts_set_scalars", [vp, c_double, c_double]), ("sts_save_previous", [vp]), ("sts_assemble", [vp]),
            ("sts_pt_jacobi", [vp, c_int, _DP]), ("sts_correct", [vp, _DP]), ("sts_properties", [vp]),
            ("sts_rho_in_middle_points", [vp]), ("sts_boundary_conditions", [vp, c_int, _DP]),
            ("sts_fluxes_diffusion", [vp]), ("sts_apply_bodies", [vp]), ("sts_jump_overwrite", [vp]),
            ("sts_set_inflow_u", [vp, c_double]), ("sts_continuity", [vp, _DP, _DP, _DP]),
            ("sts_outer_iteration", [vp, c_int, _DP, _DP, _DP]), ("sts_stream_info", [vp, ctypes.POINTER(vp)]),
            ("sts_time_step", [vp, ctypes.POINTER(LoopControl), ctypes.POINTER(c_int), _DP, _DP]),
            ("sts_time_step_noalloc", [vp, ctypes.POINTER(LoopControl), ctypes.POINTER(c_int), _DP, _DP]),
            ("sts_time_step_mixed", [vp, ctypes.POINTER(LoopControl), ctypes.POINTER(MixedControl),
                                     ctypes.POINTER(c_int), ctypes.POINTER(c_int), _DP, _DP]),
            ("sts_time_kernel", [vp, c_int]),
            ("sts_batch_create", [ctypes.POINTER(vp), c_int, ctypes.POINTER(vp)]), ("sts_batch_destroy", [vp]),
            ("sts_batch_time_step", [vp, ctypes.POINTER(LoopControl), ctypes.POINTER(MixedControl), _IP, _IP]),
            ("sts_batch_set_tiles", [vp, c_int, c_int]),
            ("sts_batch_autotune", [vp, c_int, _IP, _IP, _DP]),
            ("sts_unet_create", [c_int, c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(vp)]),
            ("sts_unet_destroy", [vp]), ("sts_unet_set_params", [vp, _FP]), ("sts_unet_get_params", [vp, _FP]),
            ("sts_unet_get_grads", [vp, _FP]), ("sts_unet_forward", [vp, c_int, _FP, _FP]),
            ("sts_unet_infer", [vp, c_int, _FP, _FP]),
            ("sts_unet_backward_mse", [vp, c_int, _FP, _BP, _DP]), ("sts_unet_backward", [vp, c_int, _FP]),
            ("sts_unet_adam", [vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
            ("sts_unet_dataset", [vp, c_int]),
            ("sts_unet_record", [vp, c_int, vp, ctypes.POINTER(UNetNorm)]),
            ("sts_unet_train_step", [vp, _IP, c_int, ctypes.c_float, _DP]),
            ("sts_unet_eval", [vp, _IP, c_int, _DP]),
            ("sts_time_step_ws", [vp, vp, ctypes.POINTER(UNetNorm), ctypes.POINTER(LoopControl),
                                  ctypes.POINTER(MixedControl), ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                                  ctypes.POINTER(c_int), _DP, _DP]),
            ("sts_kernel_time", [vp, c_int, ctypes.POINTER(ctypes.c_longlong), _DP]),
            ("sts_keep_predicted_start", [vp, c_int]),
            ("sts_unet_record_correction", [vp, c_int, vp, ctypes.POINTER(CorrNorm)]),
            ("sts_set_corrector", [vp, vp, ctypes.POINTER(CorrNorm)]),
            ("sts_corrector_stats", [vp, ctypes.POINTER(ctypes.c_longlong)]),
            ("sts_set_corrector_pause", [vp, c_int]),
            ("sts_corrector_skipped", [vp, ctypes.POINTER(ctypes.c_longlong)]),
        ]:
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = c_int
        L.sts_unet_param_count.argtypes = [vp]
        L.sts_unet_param_count.restype = ctypes.c_longlong
        _lib = L
    return _lib


class StsError(RuntimeError):
    pass


def _check(rc, what):
    if rc != STS_OK:
        msg = lib().sts_last_error().decode(errors="replace")
        raise StsError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")


def _dp(a):
    return a.ctypes.data_as(_DP)


class Context:
    """Device state of one case (sts_ctx)."""

    def __init__(self, problem, precision=64):
        self.problem = problem
        self._grid = problem.fill(Grid())
        h = ctypes.c_void_p()
        _check(lib().sts_create(ctypes.byref(self._grid), precision, ctypes.byref(h)), "sts_create")
        self.h = h

    def close(self):
        if self.h:
            lib().sts_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # fields -------------------------------------------------------------
    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        if a.size != self.problem.Na:
            raise ValueError(f"{name}: expected {self.problem.Na} values")
        _check(lib().sts_set_field(self.h, FIELD_ID[name], _dp(a)), f"sts_set_field({name})")

    def get(self, name):
        a = np.empty(self.problem.Na, dtype=np.float64)
        _check(lib().sts_get_field(self.h, FIELD_ID[name], _dp(a)), f"sts_get_field({name})")
        return a

    def set_scalars(self, cT, cPch):
        _check(lib().sts_set_scalars(self.h, cT, cPch), "sts_set_scalars")

    # stages ---------------------------------------------------------------
    def save_previous(self):  # S:3233-3239
        _check(lib().sts_save_previous(self.h), "sts_save_previous")

    def assemble(self):  # S:3269-3756
        _check(lib().sts_assemble(self.h), "sts_assemble")

    def pt_jacobi(self, n):  # S:3763-4131
        r = np.zeros(2 * n)
        _check(lib().sts_pt_jacobi(self.h, n, _dp(r)), "sts_pt_jacobi")
        return r.reshape(n, 2)

    def correct(self):  # S:4170-4330, S:4368
        r = np.zeros(2)
        _check(lib().sts_correct(self.h, _dp(r)), "sts_correct")
        return r

    def properties(self):  # S:4323-4330
        _check(lib().sts_properties(self.h), "sts_properties")

    def rho_in_middle_points(self):  # S:10382
        _check(lib().sts_rho_in_middle_points(self.h), "sts_rho_in_middle_points")

    def boundary_conditions(self, parts=BC_ALL):  # S:4460-4527
        pr = np.zeros(1)
        _check(lib().sts_boundary_conditions(self.h, parts, _dp(pr)), "sts_boundary_conditions")
        return float(pr[0])

    def fluxes_diffusion(self):  # S:4531-4532
        _check(lib().sts_fluxes_diffusion(self.h), "sts_fluxes_diffusion")

    def apply_bodies(self):  # S:6470
        _check(lib().sts_apply_bodies(self.h), "sts_apply_bodies")

    def jump_overwrite(self):  # S:5566-5567
        _check(lib().sts_jump_overwrite(self.h), "sts_jump_overwrite")

    def set_inflow_u(self, u_in):  # S:8102-8144
        _check(lib().sts_set_inflow_u(self.h, u_in), "sts_set_inflow_u")

    def continuity(self):  # S:10621-10646
        n = self.problem.Na
        cs, cf, nr = np.zeros(n), np.zeros(n), np.zeros(3)
        _check(lib().sts_continuity(self.h, _dp(cs), _dp(cf), _dp(nr)), "sts_continuity")
        return cs, cf, nr

    def time_step(self, N_I, N_I_p_c, tol_uv, tol_p, tol_T, batch=0, wait=True):  # S:3230-4608
        """One time step with loop 2 controlled on the GPU; returns (Iter,
        residual log [Iter, 6], p probe).  wait=False: no log and no probe
        (None, None), and the call returns without waiting for the step's
        last launches (calls that read device data wait for them)."""
        lc = LoopControl(N_I, N_I_p_c, tol_uv, tol_p, tol_T, batch)
        it = ctypes.c_int()
        if not wait:
            _check(lib().sts_time_step(self.h, ctypes.byref(lc), ctypes.byref(it), None, None), "sts_time_step")
            return it.value, None, None
        log, pr = np.zeros(6 * N_I), np.zeros(1)
        _check(lib().sts_time_step(self.h, ctypes.byref(lc), ctypes.byref(it), _dp(log), _dp(pr)), "sts_time_step")
        return it.value, log[:6 * it.value].reshape(-1, 6), float(pr[0])

    def time_step_noalloc(self, N_I, N_I_p_c, tol_uv, tol_p, tol_T, batch=0):
        """time_step through sts_time_step_noalloc (no allocation; loop
        buffers in the library's fixed per-GPU array, N_I <= 4096)."""
        lc = LoopControl(N_I, N_I_p_c, tol_uv, tol_p, tol_T, batch)
        it = ctypes.c_int()
        log, pr = np.zeros(6 * N_I), np.zeros(1)
        _check(lib().sts_time_step_noalloc(self.h, ctypes.byref(lc), ctypes.byref(it), _dp(log), _dp(pr)),
               "sts_time_step_noalloc")
        return it.value, log[:6 * it.value].reshape(-1, 6), float(pr[0])

    def time_step_mixed(self, N_I, N_I_p_c, tol_uv, tol_p, tol_T, N_I32, tol32, stall_w=4, stall_ratio=0.7,
                        batch=0, extrapolate=0, fp32_above=0, guard=0, anderson=0, anderson_every=1, anderson32=0,
                        wait=True):
        """Time step through sts_time_step_mixed: an FP32 phase (N_I32 > 0;
        needs precision=32), then the FP64 phase with the reference's test.
        tol32 = (uv, p, T) tolerances of the FP32 phase; extrapolate = degree
        of the predicted start, fp32_above = FP32 gate, guard = restart and
        pause bound of predicted starts, anderson = history depth of the
        Anderson acceleration of the FP64 outer iterations (every
        anderson_every of them), anderson32 = the same for the FP32 phase.  Returns (FP32 iterations, FP64 iterations,
        residual log [n32 + n64, 6], p probe); wait=False as in time_step."""
        lc = LoopControl(N_I, N_I_p_c, tol_uv, tol_p, tol_T, batch)
        mc = MixedControl(N_I32, tol32[0], tol32[1], tol32[2], stall_w, stall_ratio, extrapolate, fp32_above, guard,
                          anderson, anderson_every, anderson32)
        n32, n64 = ctypes.c_int(), ctypes.c_int()
        if not wait:
            _check(lib().sts_time_step_mixed(self.h, ctypes.byref(lc), ctypes.byref(mc), ctypes.byref(n32),
                                             ctypes.byref(n64), None, None), "sts_time_step_mixed")
            return n32.value, n64.value, None, None
        log, pr = np.zeros(6 * (N_I + N_I32)), np.zeros(1)
        _check(lib().sts_time_step_mixed(self.h, ctypes.byref(lc), ctypes.byref(mc), ctypes.byref(n32),
                                         ctypes.byref(n64), _dp(log), _dp(pr)), "sts_time_step_mixed")
        n = n32.value + n64.value
        return n32.value, n64.value, log[:6 * n].reshape(-1, 6), float(pr[0])

    def time_step_ws(self, net, norm, N_I, N_I_p_c, tol_uv, tol_p, tol_T, mixed=None, batch=0):
        """Time step warm-started by the U-Net `net` (UNet) with normalisation
        `norm` (UNetNorm); mixed = None or (N_I32, tol32, stall_w,
        stall_ratio).  Returns (used, n32, n64, residual log, p probe)."""
        lc = LoopControl(N_I, N_I_p_c, tol_uv, tol_p, tol_T, batch)
        mc, n_i32 = None, 0
        if mixed is not None:
            n_i32, tol32, w, r = mixed[:4]
            mc = MixedControl(n_i32, tol32[0], tol32[1], tol32[2], w, r,
                              *(mixed[q] if len(mixed) > q else 0 for q in range(4, 10)))
        used, n32, n64 = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        log, pr = np.zeros(6 * (N_I + n_i32)), np.zeros(1)
        _check(lib().sts_time_step_ws(self.h, net.h, ctypes.byref(norm), ctypes.byref(lc),
                                      ctypes.byref(mc) if mc is not None else None, ctypes.byref(used),
                                      ctypes.byref(n32), ctypes.byref(n64), _dp(log), _dp(pr)), "sts_time_step_ws")
        n = n32.value + n64.value
        return used.value, n32.value, n64.value, log[:6 * n].reshape(-1, 6), float(pr[0])

    def keep_predicted_start(self, on=True):
        _check(lib().sts_keep_predicted_start(self.h, 1 if on else 0), "sts_keep_predicted_start")

    def set_corrector(self, net, norm=None):
        """Attach the corrector network (UNet with cin 9, cout 4; None
        detaches) with its CorrNorm."""
        _check(lib().sts_set_corrector(self.h, net.h if net is not None else None,
                                       ctypes.byref(norm) if norm is not None else None), "sts_set_corrector")

    def corrector_stats(self):
        """(refinements kept, refinements rejected)."""
        out = (ctypes.c_longlong * 2)()
        _check(lib().sts_corrector_stats(self.h, out), "sts_corrector_stats")
        return tuple(out)

    def set_corrector_pause(self, max_pause):
        """After a refinement that does not pay (rejected, or kept and its
        step does not need at least one outer iteration fewer than the
        running mean of predicted steps without one) the corrector sits out 1, 2, 4, ...
        predicted steps, at most max_pause (0: every predicted step)."""
        _check(lib().sts_set_corrector_pause(self.h, int(max_pause)), "sts_set_corrector_pause")

    def corrector_skipped(self):
        """Predicted steps the corrector sat out."""
        out = ctypes.c_longlong()
        _check(lib().sts_corrector_skipped(self.h, ctypes.byref(out)), "sts_corrector_skipped")
        return out.value

    def predict_stats(self):
        """(predicted steps, restarted ones, FP32 and FP64 iterations of the
        abandoned attempts, steps run from x^n by the guard)."""
        out = (ctypes.c_longlong * 5)()
        _check(lib().sts_predict_stats(self.h, out), "sts_predict_stats")
        return tuple(out)

    def outer_iteration(self, n_pc):  # S:3263-4608
        r, r0, pr = np.zeros(4), np.zeros(2), np.zeros(1)
        _check(lib().sts_outer_iteration(self.h, n_pc, _dp(r), _dp(r0), _dp(pr)), "sts_outer_iteration")
        return r, r0, float(pr[0])


class Batch:
    """Several contexts stepped together in shared launches (sts_batch)."""

    NVARIANT = 4

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self.contexts))(*[c.h.value for c in self.contexts])
        h = ctypes.c_void_p()
        _check(lib().sts_batch_create(arr, len(self.contexts), ctypes.byref(h)), "sts_batch_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().sts_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def time_step(self, N_I, N_I_p_c, tol_uv, tol_p, tol_T, mixed=None, batch=0):
        """One time step of every case; mixed = None or (N_I32, tol32,
        stall_w, stall_ratio[, extrapolate[, fp32_above[, guard[, anderson[,
        anderson_every]]]]]).
        Returns (FP32 counts, FP64 counts)."""
        n = len(self.contexts)
        lc = LoopControl(N_I, N_I_p_c, tol_uv, tol_p, tol_T, batch)
        mc = None
        if mixed is not None:
            n_i32, tol32, w, r = mixed[:4]
            mc = MixedControl(n_i32, tol32[0], tol32[1], tol32[2], w, r,
                              *(mixed[q] if len(mixed) > q else 0 for q in range(4, 10)))
        a, b = np.zeros(n, np.int32), np.zeros(n, np.int32)
        _check(lib().sts_batch_time_step(self.h, ctypes.byref(lc), ctypes.byref(mc) if mc is not None else None,
                                         a.ctypes.data_as(_IP), b.ctypes.data_as(_IP)), "sts_batch_time_step")
        return a.tolist(), b.tolist()

    def set_tiles(self, cell_variant, sweep_variant):
        _check(lib().sts_batch_set_tiles(self.h, cell_variant, sweep_variant), "sts_batch_set_tiles")

    def autotune(self, reps=5):
        cv, sv = np.zeros(1, np.int32), np.zeros(1, np.int32)
        ms = np.zeros(2 * self.NVARIANT)
        _check(lib().sts_batch_autotune(self.h, reps, cv.ctypes.data_as(_IP), sv.ctypes.data_as(_IP), _dp(ms)),
               "sts_batch_autotune")
        return int(cv[0]), int(sv[0]), ms


# ---------------------------------------------------------------------------
# Warm-start U-Net (include/sts_unet.h)
# ---------------------------------------------------------------------------
def _fp(a):
    return a.ctypes.data_as(_FP)


class UNet:
    """The BF16 U-Net on the GPU.  Arrays are float32 NHWC numpy arrays."""

    def __init__(self, cin, cout, base, max_n, H, W):
        self.cin, self.cout, self.base, self.max_n, self.H, self.W = cin, cout, base, max_n, H, W
        h = ctypes.c_void_p()
        _check(lib().sts_unet_create(cin, cout, base, max_n, H, W, ctypes.byref(h)), "sts_unet_create")
        self.h = h
        self.n_params = int(lib().sts_unet_param_count(h))

    def close(self):
        if getattr(self, "h", None):
            lib().sts_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        assert flat.size == self.n_params
        _check(lib().sts_unet_set_params(self.h, _fp(flat)), "sts_unet_set_params")

    def params(self):
        out = np.zeros(self.n_params, np.float32)
        _check(lib().sts_unet_get_params(self.h, _fp(out)), "sts_unet_get_params")
        return out

    def grads(self):
        out = np.zeros(self.n_params, np.float32)
        _check(lib().sts_unet_get_grads(self.h, _fp(out)), "sts_unet_get_grads")
        return out

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[0]
        y = np.zeros((n, self.H, self.W, self.cout), np.float32)
        _check(lib().sts_unet_forward(self.h, n, _fp(x), _fp(y)), "sts_unet_forward")
        return y

    def infer(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[0]
        y = np.zeros((n, self.H, self.W, self.cout), np.float32)
        _check(lib().sts_unet_infer(self.h, n, _fp(x), _fp(y)), "sts_unet_infer")
        return y

    def backward_mse(self, t, mask):
        t = np.ascontiguousarray(t, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        loss = np.zeros(1)
        _check(lib().sts_unet_backward_mse(self.h, t.shape[0], _fp(t), m.ctypes.data_as(_BP), _dp(loss)),
               "sts_unet_backward_mse")
        return float(loss[0])

    def backward(self, dy):
        dy = np.ascontiguousarray(dy, dtype=np.float32)
        _check(lib().sts_unet_backward(self.h, dy.shape[0], _fp(dy)), "sts_unet_backward")

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        _check(lib().sts_unet_adam(self.h, lr, b1, b2, eps), "sts_unet_adam")

    def dataset(self, slots):
        _check(lib().sts_unet_dataset(self.h, slots), "sts_unet_dataset")

    def record(self, slot, ctx, norm):
        _check(lib().sts_unet_record(self.h, slot, ctx.h, ctypes.byref(norm)), "sts_unet_record")

    def record_correction(self, slot, ctx, norm):
        _check(lib().sts_unet_record_correction(self.h, slot, ctx.h, ctypes.byref(norm)),
               "sts_unet_record_correction")

    def train_step(self, slots, lr):
        s = np.ascontiguousarray(slots, dtype=np.int32)
        loss = np.zeros(1)
        _check(lib().sts_unet_train_step(self.h, s.ctypes.data_as(_IP), s.size, lr, _dp(loss)),
               "sts_unet_train_step")
        return float(loss[0])

    def eval(self, slots):
        s = np.ascontiguousarray(slots, dtype=np.int32)
        loss = np.zeros(1)
        _check(lib().sts_unet_eval(self.h, s.ctypes.data_as(_IP), s.size, _dp(loss)), "sts_unet_eval")
        return float(loss[0])


# ---------------------------------------------------------------------------
# Host setup (libsimple_ts_host.so, include/simple_ts_host.h)
# ---------------------------------------------------------------------------
_hlib = None


def host_lib():
    global _hlib
    if _hlib is None:
        if not os.path.exists(LIB_HOST):
            raise RuntimeError(f"{LIB_HOST} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_HOST)
        vp = ctypes.c_void_p
        L.sts_host_last_error.restype = ctypes.c_char_p
        L.sts_case_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
        L.sts_case_load.restype = ctypes.c_int
        L.sts_case_free.argtypes = [vp]
        L.sts_case_free.restype = None
        L.sts_case_dims.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.sts_case_grid.argtypes = [vp, ctypes.POINTER(Grid)]
        L.sts_case_mesh.argtypes = [vp] + [_DP] * 6
        L.sts_case_vol.argtypes = [vp, _UP]
        L.sts_case_value.argtypes = [vp, ctypes.c_char_p]
        L.sts_case_value.restype = ctypes.c_double
        L.sts_case_write_config.argtypes = [vp, ctypes.c_char_p]
        _hlib = L
    return _hlib


class Case:
    """A case directory (EnteredData.txt + body files) after the reference's
    setup: mesh, body mask, solve boxes and constants."""

    def __init__(self, directory):
        L = host_lib()
        h = ctypes.c_void_p()
        rc = L.sts_case_load(os.fsencode(directory), ctypes.byref(h))
        if rc != STS_OK:
            raise StsError(f"sts_case_load({directory}): {L.sts_host_last_error().decode()}")
        self.h = h
        nx, ny = ctypes.c_int(), ctypes.c_int()
        L.sts_case_dims(h, ctypes.byref(nx), ctypes.byref(ny))
        self.Nx, self.Ny = nx.value, ny.value
        self.x_f, self.y_f = np.zeros(self.Nx + 1), np.zeros(self.Ny + 1)
        self.hx, self.hy = np.zeros(self.Nx), np.zeros(self.Ny)
        self.x_v, self.y_v = np.zeros(self.Nx), np.zeros(self.Ny)
        L.sts_case_mesh(h, *[_dp(a) for a in (self.x_f, self.y_f, self.hx, self.hy, self.x_v, self.y_v)])
        self.vol = np.zeros(self.Nx * self.Ny, dtype=np.uint32)
        L.sts_case_vol(h, self.vol.ctypes.data_as(_UP))
        g = Grid()
        L.sts_case_grid(h, ctypes.byref(g))
        self.bodies = [dict(u_body=g.bodies[k].u_body, v_body=g.bodies[k].v_body, p_body=g.bodies[k].p_body,
                            T_body=g.bodies[k].T_body, is_slip=g.bodies[k].is_slip, F_slip=g.bodies[k].F_slip,
                            is_jump=g.bodies[k].is_jump, F_jump=g.bodies[k].F_jump,
                            dTdn_on_wall_0=g.bodies[k].dTdn_on_wall_0) for k in range(g.nbodies)]
        self.boxes = {k: tuple(getattr(g, k)) for k in ("Nx_u", "Ny_u", "Nx_v", "Ny_v", "Nx_o", "Ny_o")}
        self.scalars = {k: getattr(g, k) for k in Problem.SCALARS}
        self.flags = dict(given_xb=g.given_xb, given_xe=g.given_xe, kn_local=g.kn_local, periodic_x=g.periodic_x)
        self.bc = {k: getattr(g, k) for k in BC_FLAGS}

    def value(self, name):
        return host_lib().sts_case_value(self.h, name.encode())

    def problem(self):
        return Problem(self.hx, self.hy, self.vol, self.bodies, self.boxes, self.scalars, bc=self.bc, **self.flags)

    def write_config(self, path):
        rc = host_lib().sts_case_write_config(self.h, os.fsencode(path))
        if rc != STS_OK:
            raise StsError(host_lib().sts_host_last_error().decode())

    def close(self):
        if self.h:
            host_lib().sts_case_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
