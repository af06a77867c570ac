"""ctypes face of libatjup_hip.so, the HIP implementation of the ATJUP time step.

``Context`` owns one device context (atjup_hip_ctx_create) and exposes the operators of
include/atjup_hip.h under the names of the reference entry points they replace. There is no CPU
fallback: if the library is missing or no GPU is present, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
# ATJUP_HIP_LIB: load another build of the same library (a tuning variant) instead of the in-tree one
LIB_PATH = os.environ.get("ATJUP_HIP_LIB", os.path.join(HERE, "libatjup_hip.so"))
HEADER = os.path.join(os.path.dirname(HERE), "include", "atjup_hip.h")

_D = C.POINTER(C.c_double)
_lib = None


class AtjupError(RuntimeError):
    pass


def load():
    """Load the library (build it first with atjup_amd.build.build() or __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AtjupError(f"{LIB_PATH} is not built: run `python -m atjup_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.atjup_hip_ctx_create.restype = C.c_void_p
    L.atjup_hip_ctx_create.argtypes = [C.POINTER(abi.Grid), C.POINTER(abi.Consts),
                                       C.POINTER(abi.Knobs), C.c_int]
    L.atjup_hip_ctx_destroy.argtypes = [C.c_void_p]
    L.atjup_hip_default_knobs.argtypes = [C.POINTER(abi.Knobs)]
    for n in ("atjup_hip_upload_field", "atjup_hip_download_field"):
        getattr(L, n).argtypes = [C.c_void_p, C.c_int, _D]
    for n in ("atjup_hip_upload_field_rows", "atjup_hip_download_field_rows"):
        getattr(L, n).argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    for n in ("atjup_hip_upload_mask", "atjup_hip_upload_vel_star", "atjup_hip_download_vel_star",
              "atjup_hip_download_buoy_ref_level"):
        getattr(L, n).argtypes = [C.c_void_p, _D]
    for n in ("atjup_hip_rk4_step", "atjup_hip_turbulence_run", "atjup_hip_turb_wall",
              "atjup_hip_bc_after_step",
              "atjup_hip_mixture_density", "atjup_hip_flux_limiter", "atjup_hip_clamp_species",
              "atjup_hip_swap_nlevel", "atjup_hip_scalar_sync", "atjup_hip_check_finite",
              "atjup_hip_sync", "atjup_hip_pitch", "atjup_hip_radiation", "atjup_hip_precipitation",
              "atjup_hip_chemistry", "atjup_hip_mca", "atjup_hip_nh4sh_equilibrium"):
        getattr(L, n).argtypes = [C.c_void_p]
    L.atjup_hip_device_count.argtypes = [C.POINTER(C.c_int)]
    for n in ("atjup_hip_upload_tangent", "atjup_hip_adj_upload"):
        getattr(L, n).argtypes = [C.c_void_p, C.c_int, _D]
    for n in ("atjup_hip_download_tangent", "atjup_hip_adj_download"):
        getattr(L, n).argtypes = [C.c_void_p, C.c_int, _D]
    for n in ("atjup_hip_adj_init", "atjup_hip_adj_zero"):
        getattr(L, n).argtypes = [C.c_void_p]
    L.atjup_hip_adj_op.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
    L.atjup_hip_adj_forward_checkpointed.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.atjup_hip_adj_backward.argtypes = [C.c_void_p, _D, _D]
    L.atjup_hip_pressure_solve.argtypes = [C.c_void_p, C.c_int, C.c_double]
    L.atjup_hip_satadj.argtypes = [C.c_void_p, C.c_int, C.POINTER(abi.GasParams)]
    L.atjup_hip_set_gas.argtypes = [C.c_void_p, C.c_int, C.POINTER(abi.GasParams)]
    L.atjup_hip_set_radiation.argtypes = [C.c_void_p, C.POINTER(abi.RadParams)]
    L.atjup_hip_default_rad_params.argtypes = [C.POINTER(abi.RadParams)]
    L.atjup_hip_set_precipitation.argtypes = [C.c_void_p, C.POINTER(abi.PrecipParams)]
    L.atjup_hip_default_precip_params.argtypes = [C.POINTER(abi.PrecipParams)]
    L.atjup_hip_download_precip_surface.argtypes = [C.c_void_p, _D]
    L.atjup_hip_set_chemistry.argtypes = [C.c_void_p, C.POINTER(abi.ChemParams)]
    L.atjup_hip_default_chem_params.argtypes = [C.POINTER(abi.ChemParams)]
    L.atjup_hip_set_mca.argtypes = [C.c_void_p, C.POINTER(abi.MCAParams)]
    for n in ("atjup_hip_search_minmax_3d", "atjup_hip_search_minmax_2d"):
        getattr(L, n).argtypes = [C.c_void_p, C.c_int, _D, C.POINTER(C.c_int)]
    L.atjup_hip_continuity_residual.argtypes = [C.c_void_p, _D, C.POINTER(C.c_int)]
    L.atjup_hip_steady_track.argtypes = [C.c_void_p, C.c_int]
    L.atjup_hip_steady_changes.argtypes = [C.c_void_p, _D, C.POINTER(C.c_int)]
    L.atjup_hip_step_block.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.atjup_hip_step_rest.argtypes = [C.c_void_p, C.c_int]
    L.atjup_hip_search_minmax_3d_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                                     C.c_void_p, _D, C.POINTER(C.c_int)]
    L.atjup_hip_download_slice.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _D]
    L.atjup_hip_default_mca_params.argtypes = [C.POINTER(abi.MCAParams)]
    L.atjup_hip_save_state.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int]
    L.atjup_hip_load_state.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.atjup_hip_shapiro.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int]
    L.atjup_hip_step.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.atjup_hip_set_stream.argtypes = [C.c_void_p, C.c_void_p]
    L.atjup_hip_field_device_ptr.restype = C.c_void_p
    L.atjup_hip_field_device_ptr.argtypes = [C.c_void_p, C.c_int]
    L.atjup_hip_last_error_string.restype = C.c_char_p
    L.atjup_hip_set_timing.argtypes = [C.c_void_p, C.c_int]
    L.atjup_hip_read_timers.argtypes = [C.c_void_p, _D]
    L.atjup_hip_libm_probe.argtypes = [C.c_int, _D, _D, _D, C.c_long]
    L.atjup_hip_set_smooth_obstacle.argtypes = [C.c_void_p, C.POINTER(abi.Obstacle)]
    L.atjup_hip_adj_obstacle_gradient.argtypes = [C.c_void_p, _D]
    for n in ("atjup_hip_satadj_tangent", "atjup_hip_satadj_adjoint"):
        getattr(L, n).argtypes = [C.c_void_p, C.c_int, _D, _D]
    for n in ("atjup_hip_flux_limiter_tangent", "atjup_hip_flux_limiter_adjoint"):
        getattr(L, n).argtypes = [C.c_void_p, _D, _D]
    _lib = L
    return L


def device_count() -> int:
    """HIP devices the library's runtime sees (0 when there is none or the library is missing)."""
    try:
        L = load()
    except (AtjupError, OSError):
        return 0
    n = C.c_int(0)
    return n.value if L.atjup_hip_device_count(C.byref(n)) == 0 else 0


def exported_symbols():
    """Names of every function include/atjup_hip.h declares."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(atjup_hip_[a-z_0-9]+)\s*\(", txt)))


def _dp(a):
    return a.ctypes.data_as(_D)


class Context:
    """One device-resident ATJUP state on the current HIP device."""

    def __init__(self, grid, consts, knobs, precision: int = abi.PRECISION_FP64):
        self.L = load()
        self.grid = grid
        self.shape = (grid.im, grid.jm, grid.km)
        self.precision = precision
        self.h = self.L.atjup_hip_ctx_create(C.byref(grid), C.byref(consts), C.byref(knobs), precision)
        if not self.h:
            raise AtjupError("atjup_hip_ctx_create: " + self.error())
        if not knobs.dry:
            for gas in abi.GAS_PARAMS:
                self.set_gas(gas)

    def error(self):
        return self.L.atjup_hip_last_error_string().decode()

    def close(self):
        if getattr(self, "h", None):
            self.L.atjup_hip_ctx_destroy(self.h)
            self.h = None

    __del__ = close

    def _chk(self, rc, what):
        if rc < 0:
            raise AtjupError(f"{what}: {self.error()} ({rc})")
        return rc

    # ---- transfers ---------------------------------------------------------------------
    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.shape, (a.shape, self.shape)
        self._chk(self.L.atjup_hip_upload_field(self.h, abi.FIELD[name], _dp(a)), f"upload {name}")

    def get(self, name):
        a = np.empty(self.shape)
        self._chk(self.L.atjup_hip_download_field(self.h, abi.FIELD[name], _dp(a)), f"download {name}")
        return a

    def set_mask(self, m):
        a = np.ascontiguousarray(m, dtype=np.float64)
        self._chk(self.L.atjup_hip_upload_mask(self.h, _dp(a)), "upload mask")

    def set_vel_star(self, vs):
        a = np.ascontiguousarray(vs, dtype=np.float64)
        self._chk(self.L.atjup_hip_upload_vel_star(self.h, _dp(a)), "upload vel_star")

    def get_vel_star(self):
        a = np.empty(self.shape[1:])
        self._chk(self.L.atjup_hip_download_vel_star(self.h, _dp(a)), "download vel_star")
        return a

    def buoy_ref_level(self):
        a = np.empty(self.shape[0])
        self._chk(self.L.atjup_hip_download_buoy_ref_level(self.h, _dp(a)), "buoy_ref_level")
        return a

    def load_state(self, st, skip=()):
        self.set_mask(st.mask)
        for n, v in st.fields.items():
            if n in skip:
                continue
            self.set(n, v)

    # ---- operators (names follow the reference entry points) --------------------------
    def rk4_step(self):
        self._chk(self.L.atjup_hip_rk4_step(self.h), "rk4_step")

    def pressure_solve(self, sweeps=0, omega=1.0):
        self._chk(self.L.atjup_hip_pressure_solve(self.h, sweeps, omega), "pressure_solve")

    def set_gas(self, gas, params=None):
        gp = params if params is not None else abi.gas_params(gas)
        self._chk(self.L.atjup_hip_set_gas(self.h, gas, C.byref(gp)), "set_gas")

    def satadj(self, gas, params=None):
        p = C.byref(params) if params is not None else None
        self._chk(self.L.atjup_hip_satadj(self.h, gas, p), "satadj")

    def set_radiation(self, params=None):
        rp = params if params is not None else abi.rad_params()
        self._chk(self.L.atjup_hip_set_radiation(self.h, C.byref(rp)), "set_radiation")

    def radiation(self):
        self._chk(self.L.atjup_hip_radiation(self.h), "radiation")

    def set_precipitation(self, params=None):
        pp = params if params is not None else abi.precip_params()
        self._chk(self.L.atjup_hip_set_precipitation(self.h, C.byref(pp)), "set_precipitation")

    def precipitation(self):
        self._chk(self.L.atjup_hip_precipitation(self.h), "precipitation")

    def precip_surface(self):
        """{name: [jm, km]} of the surface precipitation maps the last call left."""
        jm, km = self.shape[1], self.shape[2]
        a = np.empty((5, jm, km))
        self._chk(self.L.atjup_hip_download_precip_surface(self.h, a.ctypes.data_as(_D)), "precip_surface")
        return {n: a[i] for i, n in enumerate(abi.PRECIP_SURFACE)}

    def set_chemistry(self, params=None):
        cp = params if params is not None else abi.chem_params()
        self._chk(self.L.atjup_hip_set_chemistry(self.h, C.byref(cp)), "set_chemistry")

    def chemistry(self):
        self._chk(self.L.atjup_hip_chemistry(self.h), "chemistry")

    def set_mca(self, params=None):
        mp = params if params is not None else abi.mca_params()
        self._chk(self.L.atjup_hip_set_mca(self.h, C.byref(mp)), "set_mca")

    def mca(self):
        self._chk(self.L.atjup_hip_mca(self.h), "mca")

    def nh4sh_equilibrium(self):
        """ChemistryJup::EquilibriumNH4SH("loop") on every fluid cell."""
        self._chk(self.L.atjup_hip_nh4sh_equilibrium(self.h), "nh4sh_equilibrium")

    def search_minmax_3d(self, name):
        """(max, (i, j, k), min, (i, j, k)) of a field, as Reporting::searchMinMax_3D finds them."""
        out, loc = (C.c_double * 2)(), (C.c_int * 6)()
        self._chk(self.L.atjup_hip_search_minmax_3d(self.h, abi.FIELD[name], out, loc), "search_minmax_3d")
        return out[0], tuple(loc[0:3]), out[1], tuple(loc[3:6])

    def search_minmax_2d(self, map_id):
        """(max, (j, k), min, (j, k)) of surface map 0-4 or vel_star (5), as searchMinMax_2D."""
        out, loc = (C.c_double * 2)(), (C.c_int * 4)()
        self._chk(self.L.atjup_hip_search_minmax_2d(self.h, int(map_id), out, loc), "search_minmax_2d")
        return out[0], tuple(loc[0:2]), out[1], tuple(loc[2:4])

    def continuity_residual(self):
        """(max |residual|, (i, j, k)) of steadyQuery's continuity check."""
        v, loc = C.c_double(), (C.c_int * 3)()
        self._chk(self.L.atjup_hip_continuity_residual(self.h, C.byref(v), loc), "continuity_residual")
        return v.value, tuple(loc)

    STEADY_FIELDS = ("p_dyn", "u", "v", "w", "t", "h2o", "h2o_cloud", "h2o_ice", "nh3", "nh3_cloud", "nh3_ice", "nh4sh")

    def steady_track(self, on=True):
        """Keep (or stop keeping) the n-level copies steadyQuery's per-field changes difference against."""
        self._chk(self.L.atjup_hip_steady_track(self.h, 1 if on else 0), "steady_track")

    def steady_changes(self):
        """steadyQuery's per-field changes: [(name, max |f - f_n|, (i, j, k))] in the report's order."""
        vals, cells = (C.c_double * 12)(), (C.c_int * 36)()
        self._chk(self.L.atjup_hip_steady_changes(self.h, vals, cells), "steady_changes")
        return [(n, vals[e], tuple(cells[3 * e:3 * e + 3])) for e, n in enumerate(self.STEADY_FIELDS)]

    def slice(self, name, axis, index):
        im, jm, km = self.shape
        shape = [(jm, km), (im, km), (im, jm)][axis]
        a = np.empty(shape, dtype=np.float64)
        self._chk(self.L.atjup_hip_download_slice(self.h, abi.FIELD[name], axis, index, a.ctypes.data_as(_D)), "slice")
        return a

    def save_restart(self, path, iteration, carried=True):
        """cJupiterModel::save_state: the reference's JUP2 (or, carried=False, JUP1) file."""
        self._chk(self.L.atjup_hip_save_state(self.h, str(path).encode(), int(iteration), 1 if carried else 0),
                  "save_state")

    def load_restart(self, path):
        """cJupiterModel::load_state: returns (iteration, carried)."""
        it, car = C.c_int(0), C.c_int(0)
        self._chk(self.L.atjup_hip_load_state(self.h, str(path).encode(), C.byref(it), C.byref(car)), "load_state")
        return it.value, bool(car.value)

    def turbulence_run(self):
        self._chk(self.L.atjup_hip_turbulence_run(self.h), "turbulence_run")

    def turb_wall(self):
        self._chk(self.L.atjup_hip_turb_wall(self.h), "turb_wall")

    def bc_after_step(self):
        self._chk(self.L.atjup_hip_bc_after_step(self.h), "bc_after_step")

    def shapiro(self, name, order=2, strength=1.0, passes=1, axes=7):
        self._chk(self.L.atjup_hip_shapiro(self.h, abi.FIELD[name], order, strength, passes, axes), "shapiro")

    def mixture_density(self):
        self._chk(self.L.atjup_hip_mixture_density(self.h), "mixture_density")

    def flux_limiter(self):
        self._chk(self.L.atjup_hip_flux_limiter(self.h), "flux_limiter")

    def clamp_species(self):
        self._chk(self.L.atjup_hip_clamp_species(self.h), "clamp_species")

    def swap_nlevel(self):
        self._chk(self.L.atjup_hip_swap_nlevel(self.h), "swap_nlevel")

    def scalar_sync(self):
        self._chk(self.L.atjup_hip_scalar_sync(self.h), "scalar_sync")

    def step(self, iter_n, first=False):
        self._chk(self.L.atjup_hip_step(self.h, iter_n, 1 if first else 0), "step")

    def step_block(self, iter_n, first=False):
        """The iteration's physics block (step = step_block + step_rest)."""
        self._chk(self.L.atjup_hip_step_block(self.h, iter_n, 1 if first else 0), "step_block")

    def step_rest(self, iter_n):
        self._chk(self.L.atjup_hip_step_rest(self.h, iter_n), "step_rest")

    def check_finite(self):
        return self._chk(self.L.atjup_hip_check_finite(self.h), "check_finite") == 0

    def sync(self):
        self._chk(self.L.atjup_hip_sync(self.h), "sync")

    def set_stream(self, stream_ptr):
        self._chk(self.L.atjup_hip_set_stream(self.h, C.c_void_p(stream_ptr)), "set_stream")

    def device_ptr(self, name):
        return self.L.atjup_hip_field_device_ptr(self.h, abi.FIELD[name])

    # ---- tangent form (PRECISION_TANGENT) ------------------------------------------------
    def set_tangent(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(self.L.atjup_hip_upload_tangent(self.h, abi.FIELD[name], _dp(a)), f"upload tangent {name}")

    def get_tangent(self, name):
        a = np.empty(self.shape)
        self._chk(self.L.atjup_hip_download_tangent(self.h, abi.FIELD[name], _dp(a)), f"download tangent {name}")
        return a

    # ---- adjoint of the FP64 dynamics step ----------------------------------------------
    ADJ_OPS = {"bc": 1, "shapiro_pdyn": 2, "mixture": 3, "rk4": 4, "pressure": 5, "step": 6}

    def adj_init(self):
        self._chk(self.L.atjup_hip_adj_init(self.h), "adj_init")

    def adj_zero(self):
        self._chk(self.L.atjup_hip_adj_zero(self.h), "adj_zero")

    def adj_set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(self.L.atjup_hip_adj_upload(self.h, abi.FIELD[name], _dp(a)), f"adj upload {name}")

    def adj_get(self, name):
        a = np.empty(self.shape)
        self._chk(self.L.atjup_hip_adj_download(self.h, abi.FIELD[name], _dp(a)), f"adj download {name}")
        return a

    def adj_op(self, op, iarg=0, darg=0.0):
        self._chk(self.L.atjup_hip_adj_op(self.h, self.ADJ_OPS[op], int(iarg), float(darg)), f"adj_op {op}")

    def set_smooth_obstacle(self, obst=None):
        """The smooth drag region (abi.Obstacle), or None to remove it."""
        self._chk(self.L.atjup_hip_set_smooth_obstacle(self.h, None if obst is None else C.byref(obst)),
                  "set_smooth_obstacle")

    def adj_obstacle_gradient(self):
        """dL/d(a_phi_deg, b_the_deg, c_r_km) of the last adj_backward."""
        out = np.zeros(3)
        self._chk(self.L.atjup_hip_adj_obstacle_gradient(self.h, _dp(out)), "adj_obstacle_gradient")
        return out

    def _lin(self, fn, what, nout, arrays, *args):
        src = np.ascontiguousarray(np.stack([np.asarray(a, dtype=np.float64) for a in arrays]))
        assert src.shape[1:] == self.shape, (src.shape, self.shape)
        out = np.zeros((nout,) + self.shape)
        self._chk(fn(self.h, *args, _dp(src), _dp(out)), what)
        return list(out)

    def satadj_tangent(self, gas, dx):
        """Frozen-branch tangent of one gas's saturation adjustment at the current state:
        dx = d(vapour, cloud, ice, t, p_stat, rho_mix) -> d(vapour, cloud, ice, t)."""
        return self._lin(self.L.atjup_hip_satadj_tangent, "satadj_tangent", 4, dx, gas)

    def satadj_adjoint(self, gas, ly):
        """Its transpose: the outputs' adjoints (vapour, cloud, ice, t) -> the inputs' six."""
        return self._lin(self.L.atjup_hip_satadj_adjoint, "satadj_adjoint", 6, ly, gas)

    def flux_limiter_tangent(self, dx):
        """Frozen-branch tangent of the NH4SH flux limiter: d(nh4sh, u, v, w) -> d fluxlim_nh4sh."""
        return self._lin(self.L.atjup_hip_flux_limiter_tangent, "flux_limiter_tangent", 1, dx)[0]

    def flux_limiter_adjoint(self, ly):
        """Its transpose: the adjoint of fluxlim_nh4sh -> those of (nh4sh, u, v, w)."""
        return self._lin(self.L.atjup_hip_flux_limiter_adjoint, "flux_limiter_adjoint", 4, [ly])

    def adj_forward(self, nsteps, every, first_iter=1):
        self._chk(self.L.atjup_hip_adj_forward_checkpointed(self.h, nsteps, every, first_iter), "adj_forward")

    def adj_backward(self, lam_final=None, download=True):
        """lam_final: adjoint of the final (t, u, v, w), shape (4, im, jm, km), or None to start from
        the adjoint fields already on the device (adj_set); returns the initial adjoint, or None
        when download is False (it then stays in the device's adjoint fields, adj_get)."""
        out = np.empty((4,) + self.shape) if download else None
        lf = None
        if lam_final is not None:
            lf = np.ascontiguousarray(lam_final, dtype=np.float64)
            assert lf.shape == (4,) + self.shape
        self._chk(self.L.atjup_hip_adj_backward(self.h, None if lf is None else _dp(lf),
                                                None if out is None else _dp(out)), "adj_backward")
        return out

    TIMERS = ("poisson_sweeps", "rk_stages", "pressure", "step", "adj_sweeps", "adj_step", "adj_recompute",
              "radiation", "precipitation", "chemistry", "mca")

    def set_timing(self, on=True):
        self._chk(self.L.atjup_hip_set_timing(self.h, 1 if on else 0), "set_timing")

    def read_timers(self):
        """{phase: (milliseconds, kernel launches)} since the last read (atjup_hip_read_timers)."""
        buf = (C.c_double * (2 * len(self.TIMERS)))()
        self._chk(self.L.atjup_hip_read_timers(self.h, buf), "read_timers")
        return {n: (buf[2 * i], int(buf[2 * i + 1])) for i, n in enumerate(self.TIMERS)}

    @property
    def pitch(self):
        return self.L.atjup_hip_pitch(self.h)
