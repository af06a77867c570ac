"""CPU oracle of the ATJUP time step — TEST INFRASTRUCTURE ONLY.

Two checkers live here:

* ``Oracle``: ctypes face of ``libatjup_oracle.so``, the from-scratch C++/OpenMP restatement of the
  reference path (cpu_ref.cpp), at any grid size;
* ``Reference``: ctypes face of ``_ref/libatjup_ref_O2.so``, the reference model itself compiled
  from its own sources (Makefile target ``ref``) behind ref_harness.cpp. One knob profile per
  process (the reference caches its knobs in function statics), so callers run it in a
  subprocess (``run_reference``).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import this package; the
product path (atjup_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from atjup_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libatjup_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libatjup_ref_O2.so")


def ref_so(opt: str = "O2") -> str:
    """The compiled reference at optimisation level opt ("O2" or "O0")."""
    return os.path.join(HERE, "_ref", f"libatjup_ref_{opt}.so")

_D = C.POINTER(C.c_double)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_D)


_orc = None


def _orc_lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} is not built (make -C oracle)")
        L = C.CDLL(ORACLE_SO)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(abi.Grid), C.POINTER(abi.Consts), C.POINTER(abi.Knobs)]
        L.orc_destroy.argtypes = [C.c_void_p]
        for n in ("orc_upload_field", "orc_download_field"):
            getattr(L, n).argtypes = [C.c_void_p, C.c_int, _D]
        for n in ("orc_upload_mask", "orc_upload_vel_star", "orc_download_vel_star",
                  "orc_download_buoy_ref_level"):
            getattr(L, n).argtypes = [C.c_void_p, _D]
        for n in ("orc_rk4_step", "orc_buoy_ref_level", "orc_hydrostatic", "orc_bc_after_step",
                  "orc_bc_radius", "orc_bc_theta", "orc_bc_phi", "orc_bc_solid",
                  "orc_mixture_density", "orc_clamp_species", "orc_swap_nlevel",
                  "orc_scalar_sync", "orc_turbulence_run", "orc_turb_wall", "orc_flux_limiter",
                  "orc_check_finite", "orc_radiation", "orc_precipitation", "orc_chemistry",
                  "orc_mca", "orc_nh4sh_eq"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.orc_pressure_solve.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_rb_colour_pass.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
        L.orc_satadj.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_gas.argtypes = [C.c_void_p, C.c_int, C.POINTER(abi.GasParams)]
        L.orc_set_radiation.argtypes = [C.c_void_p, C.POINTER(abi.RadParams)]
        L.orc_set_precipitation.argtypes = [C.c_void_p, C.POINTER(abi.PrecipParams)]
        L.orc_set_chemistry.argtypes = [C.c_void_p, C.POINTER(abi.ChemParams)]
        L.orc_set_mca.argtypes = [C.c_void_p, C.POINTER(abi.MCAParams)]
        L.orc_download_precip_surface.argtypes = [C.c_void_p, _D]
        L.orc_shapiro.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int]
        L.orc_step.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_set_threads.argtypes = [C.c_int]
        _orc = L
    return _orc


class Oracle:
    """The CPU restatement, with the same operators as the HIP library."""

    def __init__(self, grid, consts, knobs, threads: int = 0):
        self.L = _orc_lib()
        if threads:
            self.L.orc_set_threads(threads)
        self.shape = (grid.im, grid.jm, grid.km)
        self.h = self.L.orc_create(C.byref(grid), C.byref(consts), C.byref(knobs))
        if not self.h:
            raise RuntimeError("orc_create failed")
        for gas in abi.GAS_PARAMS:
            gp = abi.gas_params(gas)
            self._chk(self.L.orc_set_gas(self.h, gas, C.byref(gp)), "set_gas")

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    __del__ = close

    def _chk(self, rc, what):
        if rc < 0:
            raise RuntimeError(f"oracle {what} failed: {rc}")
        return rc

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(self.L.orc_upload_field(self.h, abi.FIELD[name], _dp(a)), f"upload {name}")

    def get(self, name):
        a = np.empty(self.shape)
        self._chk(self.L.orc_download_field(self.h, abi.FIELD[name], _dp(a)), f"download {name}")
        return a

    def set_mask(self, m):
        a = np.ascontiguousarray(m, dtype=np.float64)
        self._chk(self.L.orc_upload_mask(self.h, _dp(a)), "mask")

    def set_vel_star(self, vs):
        a = np.ascontiguousarray(vs, dtype=np.float64)
        self._chk(self.L.orc_upload_vel_star(self.h, _dp(a)), "vel_star")

    def buoy_ref_level(self):
        a = np.empty(self.shape[0])
        self._chk(self.L.orc_download_buoy_ref_level(self.h, _dp(a)), "brl")
        return a

    def load_state(self, st):
        self.set_mask(st.mask)
        for n, v in st.fields.items():
            self.set(n, v)

    def set_radiation(self, params):
        self._chk(self.L.orc_set_radiation(self.h, C.byref(params)), "set_radiation")

    def set_chemistry(self, params):
        self._chk(self.L.orc_set_chemistry(self.h, C.byref(params)), "set_chemistry")

    def set_mca(self, params):
        self._chk(self.L.orc_set_mca(self.h, C.byref(params)), "set_mca")

    def set_precipitation(self, params):
        self._chk(self.L.orc_set_precipitation(self.h, C.byref(params)), "set_precipitation")

    def precip_surface(self):
        jm, km = self.shape[1], self.shape[2]
        a = np.empty((5, jm, km))
        self._chk(self.L.orc_download_precip_surface(self.h, _dp(a)), "precip_surface")
        return {n: a[i] for i, n in enumerate(abi.PRECIP_SURFACE)}

    def call(self, op, *args):
        return self._chk(getattr(self.L, "orc_" + op)(self.h, *args), op)


# ---- the reference, compiled from its own sources -----------------------------------------
class Reference:
    """The compiled reference model behind ref_harness.cpp (one profile per process)."""

    def __init__(self, env: dict, quiet: bool = True, opt: str = "O2"):
        """opt: the optimisation level the reference was compiled at ("O2", or "O0", the level of
        the reference's own Makefile); each lives in its own library under _ref/."""
        so = ref_so(opt)
        if not os.path.exists(so):
            raise RuntimeError(f"{so} is not built (make -C oracle ref OPT=-{opt})")
        L = C.CDLL(so)
        L.ref_init.argtypes = [C.c_char_p]
        L.ref_set_field.argtypes = [C.c_char_p, _D]
        L.ref_get_field.argtypes = [C.c_char_p, _D]
        L.ref_set_scalar.argtypes = [C.c_char_p, C.c_double]
        L.ref_get_scalar.argtypes = [C.c_char_p, _D]
        L.ref_set_mask.argtypes = [_D]
        L.ref_call.argtypes = [C.c_char_p]
        L.ref_get_field2d.argtypes = [C.c_char_p, _D]
        L.ref_save_state.argtypes = [C.c_char_p, C.c_int]
        L.ref_load_state.argtypes = [C.c_char_p, C.c_int]
        L.ref_buoy_ref_level.restype = C.c_double
        L.ref_buoy_ref_level.argtypes = [C.c_int]
        L.ref_set_vel_star.argtypes = [_D]
        L.ref_get_vel_star.argtypes = [_D]
        L.ref_quiet.argtypes = [C.c_int]
        L.ref_step.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        self.L = L
        L.ref_quiet(1 if quiet else 0)
        if L.ref_init(abi.env_string(env).encode()) != 0:
            raise RuntimeError("ref_init failed (one profile per process)")
        im, jm, km = C.c_int(), C.c_int(), C.c_int()
        L.ref_dims(C.byref(im), C.byref(jm), C.byref(km))
        self.shape = (im.value, jm.value, km.value)

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if self.L.ref_set_field(name.encode(), _dp(a)) != 0:
            raise KeyError(name)

    def get(self, name):
        a = np.empty(self.shape)
        if self.L.ref_get_field(name.encode(), _dp(a)) != 0:
            raise KeyError(name)
        return a

    def save_state(self, directory, iteration):
        """The reference's save_state: writes directory/jup_restart_<iteration>.bin."""
        self.L.ref_save_state(str(directory).encode(), int(iteration))

    def load_state_file(self, directory, iteration):
        return self.L.ref_load_state(str(directory).encode(), int(iteration)) == 0

    def get2d(self, name):
        a = np.empty(self.shape[1:])
        if self.L.ref_get_field2d(name.encode(), _dp(a)) != 0:
            raise KeyError(name)
        return a

    def set_scalar(self, name, v):
        if self.L.ref_set_scalar(name.encode(), float(v)) != 0:
            raise KeyError(name)

    def get_res(self):
        """The continuity residual's cell (i_res, j_res, k_res) the last steady_query left."""
        a = (C.c_int * 3)()
        self.L.ref_get_res.argtypes = [C.POINTER(C.c_int)]
        self.L.ref_get_res(a)
        return tuple(a)

    def get_scalar(self, name):
        v = C.c_double()
        if self.L.ref_get_scalar(name.encode(), C.byref(v)) != 0:
            raise KeyError(name)
        return v.value

    def set_mask(self, m):
        a = np.ascontiguousarray(m, dtype=np.float64)
        self.L.ref_set_mask(_dp(a))

    def set_vel_star(self, vs):
        a = np.ascontiguousarray(vs, dtype=np.float64)
        self.L.ref_set_vel_star(_dp(a))

    def get_vel_star(self):
        a = np.empty(self.shape[1:])
        self.L.ref_get_vel_star(_dp(a))
        return a

    def buoy_ref_level(self):
        return np.array([self.L.ref_buoy_ref_level(i) for i in range(self.shape[0])])

    def call(self, op):
        if self.L.ref_call(op.encode()) != 0:
            raise KeyError(op)

    def load_state(self, st, consts=None):
        self.set_mask(st.mask)
        for n, v in st.fields.items():
            self.set(n, v)
        if consts is not None:
            for n, _ in abi.Consts._fields_:
                try:
                    self.set_scalar(n, getattr(consts, n))
                except KeyError:
                    pass

    def step(self, iter_n, first, satadj=True, turb=False):
        """The loop body's on-path functions in the reference's order (cJupiterModel.cpp:1010-1555)."""
        if self.L.ref_step(iter_n, 1 if first else 0, 1 if satadj else 0, 1 if turb else 0) != 0:
            raise RuntimeError("ref_step failed")

    def step_physics(self, iter_n, first, satadj=True, turb=False, radiation=False, precip=False,
                     chemistry=False, nh4sh_eq=False):
        """The loop body with the physics block's radiation, precipitation, chemistry sequence and
        NH4SH equilibrium as the reference runs them (cJupiterModel.cpp:1032-1161)."""
        flags = ((1 if satadj else 0) | (2 if turb else 0) | (4 if radiation else 0) | (8 if precip else 0)
                 | (16 if chemistry else 0) | (32 if nh4sh_eq else 0))
        self.L.ref_step_physics.argtypes = [C.c_int, C.c_int, C.c_int]
        if self.L.ref_step_physics(iter_n, 1 if first else 0, flags) != 0:
            raise RuntimeError("ref_step_physics failed")

    def set_threads(self, n: int):
        """OpenMP threads of the reference's loops."""
        self.L.ref_set_threads.argtypes = [C.c_int]
        self.L.ref_set_threads(int(n))

    def max_threads(self) -> int:
        return int(self.L.ref_max_threads())


def reference_available(opt: str = "O2") -> bool:
    return os.path.exists(ref_so(opt))
