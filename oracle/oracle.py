"""ctypes binding of the CPU FP64 restatement (libsts_oracle.so).

Test infrastructure: the checker, loaded only by tests/, the smoke check and
the benchmark's CPU-baseline leg.  The product never imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libsts_oracle.so")
REF_BIN = os.path.join(HERE, "_ref", "sts_ref")
REF_DBG = os.path.join(HERE, "_ref", "sts_ref_dbg")

_DP = ctypes.POINTER(ctypes.c_double)

# ora_fields member order (sts_oracle.h), the reference's array names
ORA_FIELDS = ["u", "v", "p", "Temper", "rho", "sqrt_T", "u_pr", "v_pr", "p_pr", "Temper_pr", "rho_pr",
              "rho_u", "rho_v", "Fx", "Fy", "Gamma_xf", "Gamma_yf", "D_ux", "D_uy", "D_vx", "D_vy", "DTx",
              "DTy", "u_pseudo", "v_pseudo", "d_p_c_12", "d_p_c_34", "apx", "apy", "bpx", "bpy", "ScT1",
              "Temper_of_fluid_on_the_wall_u", "Temper_of_fluid_on_the_wall_v", "p_old", "Temper_old"]


def _grid_struct():
    import importlib
    import sys

    root = os.path.dirname(HERE)
    if root not in sys.path:
        sys.path.insert(0, root)
    pkg = importlib.import_module("simple-ts_amd")

    class OraGrid(ctypes.Structure):
        _fields_ = pkg.grid_fields(True)

    return OraGrid


class OraFields(ctypes.Structure):
    _fields_ = [(n, _DP) for n in ORA_FIELDS]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} not built (make -f oracle/Makefile)")
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        for name, extra in [("ora_save_previous", []), ("ora_assemble_u", []), ("ora_assemble_v", []),
                            ("ora_energy_source", []), ("ora_pt_jacobi", [ctypes.c_int, _DP]),
                            ("ora_correct_velocities", [_DP]), ("ora_properties", []),
                            ("ora_rho_in_middle_points", []), ("ora_bc_slip_u", []), ("ora_bc_slip_v", []),
                            ("ora_bc_velocities", []), ("ora_bc_temperature_jump", []),
                            ("ora_bc_temperature", []), ("ora_bc_pressure", []), ("ora_jump_overwrite", []),
                            ("ora_fluxes", []), ("ora_diffusion", []), ("ora_apply_bodies", []),
                            ("ora_outer_iteration", [ctypes.c_int, _DP])]:
            fn = getattr(L, name)
            fn.argtypes = [vp, vp] + extra
            fn.restype = None
        L.ora_continuity.argtypes = [vp, vp, _DP, _DP]
        L.ora_continuity.restype = None
        _lib = L
    return _lib


class Oracle:
    """Host state of one case for the CPU restatement: every field is a numpy
    array in self.f (by the reference's name)."""

    def __init__(self, problem):
        self.problem = problem
        self._grid = problem.fill(_grid_struct()())
        n = problem.Na
        self.f = {name: np.zeros(n) for name in ORA_FIELDS}
        self._fs = OraFields()
        self._sync_ptrs()

    def _sync_ptrs(self):
        for name in ORA_FIELDS:
            setattr(self._fs, name, self.f[name].ctypes.data_as(_DP))

    def _call(self, name, *extra):
        getattr(lib(), name)(ctypes.byref(self._grid), ctypes.byref(self._fs), *extra)

    def set(self, name, arr):
        self.f[name][:] = np.asarray(arr, dtype=np.float64).reshape(-1)

    def get(self, name):
        return self.f[name].copy()

    def set_scalars(self, cT, cPch):
        self._grid.cT, self._grid.cPch = cT, cPch

    def save_previous(self):
        self._call("ora_save_previous")

    def assemble(self):
        self._call("ora_assemble_u")
        self._call("ora_assemble_v")
        self._call("ora_energy_source")

    def pt_jacobi(self, n):
        r = np.zeros(2 * n)
        self._call("ora_pt_jacobi", n, r.ctypes.data_as(_DP))
        return r.reshape(n, 2)

    def correct(self):
        r = np.zeros(2)
        self._call("ora_correct_velocities", r.ctypes.data_as(_DP))
        self._call("ora_properties")
        self._call("ora_rho_in_middle_points")
        return r

    def properties(self):
        self._call("ora_properties")

    def rho_in_middle_points(self):
        self._call("ora_rho_in_middle_points")

    def boundary_conditions(self, parts=63):
        order = [(1, "ora_bc_slip_u"), (2, "ora_bc_slip_v"), (4, "ora_bc_velocities"),
                 (8, "ora_bc_temperature_jump"), (16, "ora_bc_temperature"), (32, "ora_bc_pressure")]
        for bit, name in order:
            if parts & bit:
                self._call(name)

    def fluxes(self):
        self._call("ora_fluxes")

    def fluxes_diffusion(self):
        self._call("ora_fluxes")
        self._call("ora_diffusion")

    def apply_bodies(self):
        self._call("ora_apply_bodies")

    def jump_overwrite(self):
        self._call("ora_jump_overwrite")

    def continuity(self):
        n = self.problem.Na
        cs, cf = np.zeros(n), np.zeros(n)
        lib().ora_continuity(ctypes.byref(self._grid), ctypes.byref(self._fs), cs.ctypes.data_as(_DP),
                             cf.ctypes.data_as(_DP))
        return cs, cf

    def outer_iteration(self, n_pc):
        r = np.zeros(4)
        self._call("ora_outer_iteration", n_pc, r.ctypes.data_as(_DP))
        return r
