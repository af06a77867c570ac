# cython: language_level=3
"""pyatjup: the Python binding of the MI355X time step, named after the reference's own Cython
binding (python/pyatjup.pyx.template:32-97, python/jupiter_pxd.pxi).

``Jupiter`` keeps the reference binding's face -- ``Jupiter()``, ``load_config(filename)``, ``run()``,
the configuration's scalar properties by their reference names (nm, L_atm, re, pr, g, omega, u_0,
t_ref, Coriolis, centrifugal, buoyancy, chemical_reaction, turb_model, ...), and the context-manager
protocol that frees the model deterministically -- and adds what the reference binding has no way to
do: field access (``get_field`` / ``set_field`` on the device-resident state), single iterations
(``step``) and the hand-written adjoint of the dynamics step (FP64, or FP32 in the FP32-refined form).
``JetMisfit`` is a
``torch.autograd.Function`` over it: L(w0), the cos(latitude)-weighted squared misfit of the zonal-mean
zonal wind at the top integrated level against a target jet profile after N iterations, whose backward
is the checkpointed reverse sweep of the library.

Every call goes straight to the library's C ABI (include/atjup_hip.h); the knob profiles come from
atjup_amd.abi, whose ctypes structures are the ABI's structures, passed here by address.
"""
import ctypes
import xml.etree.ElementTree as ET

import numpy as np

from atjup_amd import abi, inputs


cdef extern from "atjup_hip.h":
    ctypedef struct atjup_grid:
        int im, jm, km
    ctypedef struct atjup_consts:
        pass
    ctypedef struct atjup_knobs:
        pass
    ctypedef struct atjup_gas_params:
        pass
    ctypedef struct atjup_obstacle:
        pass
    int atjup_hip_set_smooth_obstacle(atjup_ctx* ctx, const atjup_obstacle* params)
    int atjup_hip_adj_obstacle_gradient(atjup_ctx* ctx, double* out)
    ctypedef struct atjup_ctx:
        pass
    atjup_ctx* atjup_hip_ctx_create(const atjup_grid* grid, const atjup_consts* consts,
                                    const atjup_knobs* knobs, int precision)
    void atjup_hip_ctx_destroy(atjup_ctx* ctx)
    int atjup_hip_upload_field(atjup_ctx* ctx, int field, const double* src)
    int atjup_hip_download_field(atjup_ctx* ctx, int field, double* dst)
    int atjup_hip_upload_mask(atjup_ctx* ctx, const double* seamount)
    int atjup_hip_set_gas(atjup_ctx* ctx, int gas, const atjup_gas_params* params)
    int atjup_hip_bc_after_step(atjup_ctx* ctx)
    int atjup_hip_swap_nlevel(atjup_ctx* ctx)
    int atjup_hip_mixture_density(atjup_ctx* ctx)
    int atjup_hip_step(atjup_ctx* ctx, int iter, int first)
    int atjup_hip_check_finite(atjup_ctx* ctx)
    int atjup_hip_sync(atjup_ctx* ctx)
    int atjup_hip_adj_init(atjup_ctx* ctx)
    int atjup_hip_adj_zero(atjup_ctx* ctx)
    int atjup_hip_adj_forward_checkpointed(atjup_ctx* ctx, int nsteps, int every, int first_iter)
    int atjup_hip_adj_backward(atjup_ctx* ctx, const double* lam_final, double* lam_init)
    const char* atjup_hip_last_error_string()


class AtjupError(RuntimeError):
    pass


cdef double* _ptr(arr):
    # a C-contiguous float64 numpy array's data (the caller keeps the array alive)
    cdef size_t p = arr.ctypes.data
    return <double*>p


cdef _err(int rc, str what):
    if rc < 0:
        raise AtjupError(f"{what}: {atjup_hip_last_error_string().decode()} ({rc})")
    return rc


# The reference configuration's parameters this binding reads from a config file or accepts as
# properties: the <jupiter> section's names (cli/config_atjup.xml) and their types. The ones the
# time step uses map onto the library's profile; the others are kept and reported back unchanged.
_PARAMS = {
    "nm": int, "checkpoint": int, "panorama_print": int, "checkpoint_save_iter": int, "restart_from_iter": int,
    "Coriolis": float, "centrifugal": float, "buoyancy": float, "chemical_reaction": float,
    "coord_stretching": bool, "turb_model": str, "L_atm": float, "tropopause_pole": float,
    "tropopause_equator": float, "re": float, "ec": float, "ep": float, "hp": float, "lv": float, "ls": float,
    "cp_l": float, "pr": float, "g": float, "omega": float, "gam": float, "u_0": float, "r_0_water": float,
    "ua": float, "va": float, "wa": float, "pa": float, "ca": float, "ta": float, "t_ref": float,
    "t_equator": float, "t_pole": float, "p_ref": float, "R_ref": float, "t_tropopause": float,
    "t_react_onset": float, "p_tropopause": float, "h2_tropopause": float, "he_tropopause": float,
    "h2o_tropopause": float, "h2s_tropopause": float, "nh3_tropopause": float, "nh4sh_tropopause": float,
    "ch4_tropopause": float, "verbose": bool, "output_path": str, "paraview_panorama_vts_flag": bool,
}
# the time step's constants among them (atjup_consts fields of the same name)
_CONSTS = ("re", "pr", "g", "omega", "u_0", "t_ref", "Coriolis", "centrifugal", "buoyancy", "chemical_reaction")

_PRECISION = {"fp64": abi.PRECISION_FP64, "fp32_refined": abi.PRECISION_FP32_REF}

# the fields the adjoint of the dynamics step carries
_ADJ_STATE = ("t", "u", "v", "w", "p_dyn", "aux_u", "aux_v", "aux_w", "rho_mix")


def _parse(text, typ):
    text = (text or "").strip()
    if typ is bool:
        return text.lower() in ("1", "true", "yes")
    return typ(text)


cdef class Jupiter:
    """The model: the reference binding's Jupiter (python/pyatjup.pyx.template:32-97) over the device
    library. The configuration starts at the reference's defaults for the parameters the time step
    uses (cli/config_atjup.xml) with the closure set to k-omega; ``profile`` picks the physics
    ("full": moist, closure, radiation, precipitation, chemistry, NH4SH equilibrium; "dyn": the dry
    dynamics), ``precision`` the form ("fp64", the drop-in, or "fp32_refined")."""
    cdef atjup_ctx* _ctx
    cdef dict _params
    cdef str _profile, _precision
    cdef object _grid, _G, _Cn, _K
    cdef public object state            # the initial state run() starts from (atjup_amd.inputs.State)
    cdef public int iteration           # iterations run on the device-resident state
    cdef readonly int generation        # changes of the device state; a backward must match its forward's
    cdef object _adj_init               # the adjoint's initial state (adjoint_initial_state)
    cdef object _obst                   # the smooth drag region (abi.Obstacle) or None
    cdef int _fwd_gen                   # generation right after the last adjoint_forward

    def __cinit__(self):
        self._ctx = NULL
        self.generation = 0
        self._fwd_gen = -1

    def __init__(self, profile="full", precision="fp64", grid=None, sweeps=200):
        if profile not in ("full", "dyn"):
            raise ValueError("profile: 'full' or 'dyn'")
        if precision not in _PRECISION:
            raise ValueError("precision: 'fp64' or 'fp32_refined'")
        self._profile, self._precision = profile, precision
        self._grid = grid or inputs.Grid()
        self._params = dict(nm=224, checkpoint=8, panorama_print=32, checkpoint_save_iter=-1, restart_from_iter=-1,
                            Coriolis=1.0, centrifugal=1.0, buoyancy=1.0, chemical_reaction=1.0,
                            coord_stretching=False, turb_model="k_omega", L_atm=140.0, re=1000.0, pr=0.69,
                            g=25.92, omega=0.000176, u_0=100.0, t_ref=165.0, sweeps=int(sweeps))
        self.state = None
        self._adj_init = None
        self._obst = None
        self.iteration = 0

    # ---- the reference binding's calls -------------------------------------------------------
    def load_config(self, str filename):
        """Read the <common> and <jupiter> sections of a reference config file (cli/config_atjup.xml
        layout). Call before run(); parameters the file leaves out keep their values."""
        root = ET.parse(filename).getroot()
        for sec in ("common", "jupiter"):
            node = root.find(sec)
            if node is None:
                continue
            for child in node:
                typ = _PARAMS.get(child.tag)
                if typ is not None:
                    self._params[child.tag] = _parse(child.text, typ)
        return None

    def run(self):
        """nm iterations of the time loop on the device (iteration 1 a run's first), from ``state``
        (the seeded full-physics or dynamics state when none is set); raises if the state is left
        non-finite (the NaN watch)."""
        self._check_alive(create=True)
        nm = int(self._params["nm"])
        for _ in range(nm):
            self.step()
        _err(atjup_hip_sync(self._ctx), "sync")
        if atjup_hip_check_finite(self._ctx) != 0:
            raise AtjupError("the state holds a non-finite value after run()")
        return None

    def __enter__(self):
        return self

    def __exit__(self, exc_tp, exc_val, exc_tb):
        self.close()
        return False

    def __dealloc__(self):
        if self._ctx != NULL:
            atjup_hip_ctx_destroy(self._ctx)
            self._ctx = NULL

    def close(self):
        if self._ctx != NULL:
            atjup_hip_ctx_destroy(self._ctx)
            self._ctx = NULL

    # ---- the configuration's properties ------------------------------------------------------
    def __getattr__(self, name):
        p = self._params
        if name in p:
            return p[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in _PARAMS or name == "sweeps":
            if self._ctx != NULL and name in _CONSTS + ("turb_model", "L_atm", "sweeps"):
                raise AtjupError(f"{name} is fixed once the model is on the device (close() first)")
            typ = _PARAMS.get(name, int)
            self._params[name] = typ(value) if typ is not bool else bool(value)
        elif name == "state":
            self.state = value
            self._adj_init = None
        elif name == "iteration":
            self.iteration = value
        else:
            raise AttributeError(f"Jupiter has no parameter {name!r}")

    @property
    def profile(self):
        return self._profile

    @property
    def precision_name(self):
        return self._precision

    @property
    def dthe(self):
        return self._grid.dthe

    # ---- the device state ----------------------------------------------------------------------
    cdef int _check_alive(self, bint create=False) except -1:
        if self._ctx != NULL:
            return 0
        if not create:
            raise RuntimeError("the model is not on the device (run(), step() or set_field() first)")
        p = self._params
        turb = "k_omega" if self._profile == "dyn" else str(p["turb_model"])
        if turb in ("none", "laminar"):
            raise AtjupError("the full profile carries a closure (k_omega, k_omega_SST or k_epsilon)")
        kw = dict(radiation=True, precip=True, chemistry=True, nh4sh_eq=True) if self._profile == "full" else {}
        G, Cn, K, _ = abi.make_profile(self._profile, sweeps=int(p["sweeps"]), grid=self._grid,
                                       turb=turb, **kw)
        G.L_atm = float(p["L_atm"])
        for n in _CONSTS:
            setattr(Cn, n, float(p[n]))
        Cn.chemical_reaction = 0.0 if self._profile == "dyn" else float(p["chemical_reaction"])
        self._G, self._Cn, self._K = G, Cn, K
        cdef size_t ag = ctypes.addressof(G), ac = ctypes.addressof(Cn), ak = ctypes.addressof(K)
        self._ctx = atjup_hip_ctx_create(<atjup_grid*>ag, <atjup_consts*>ac, <atjup_knobs*>ak,
                                         _PRECISION[self._precision])
        if self._ctx == NULL:
            raise AtjupError("atjup_hip_ctx_create: " + atjup_hip_last_error_string().decode())
        cdef size_t agp
        if not K.dry:
            for gas in abi.GAS_PARAMS:
                gp = abi.gas_params(gas)
                agp = ctypes.addressof(gp)
                _err(atjup_hip_set_gas(self._ctx, gas, <atjup_gas_params*>agp), "set_gas")
        st = self.state
        if st is None:
            st = inputs.make_state(self._profile, seed=11, g=self._grid,
                                   moisture="subsaturated" if self._profile == "full" else "seeded")
            self.state = st
        self.set_mask(st.mask)
        for n, v in st.fields.items():
            if self._profile == "dyn" and n in inputs.SPECIES:
                continue
            self.set_field(n, v)
        _err(atjup_hip_bc_after_step(self._ctx), "bc_after_step")
        _err(atjup_hip_swap_nlevel(self._ctx), "swap_nlevel")
        self.iteration = 0
        return 0

    @property
    def shape(self):
        g = self._grid
        return (g.im, g.jm, g.km)

    def set_mask(self, seamount):
        """The obstacle (SeaMount, 1.0 = solid), [im][jm][km]."""
        self._check_alive(create=True)
        a = np.ascontiguousarray(seamount, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"mask: shape {a.shape}, expected {self.shape}")
        self.generation += 1
        _err(atjup_hip_upload_mask(self._ctx, _ptr(a)), "upload_mask")

    def set_field(self, str name, values):
        """Upload one field (reference name, e.g. "w" the zonal wind, "h2o_cloud"), [im][jm][km]."""
        self._check_alive(create=True)
        arr = np.ascontiguousarray(values, dtype=np.float64)
        if arr.shape != self.shape:
            raise ValueError(f"{name}: shape {arr.shape}, expected {self.shape}")
        self.generation += 1
        _err(atjup_hip_upload_field(self._ctx, abi.FIELD[name], _ptr(arr)), f"upload {name}")

    def get_field(self, str name):
        """Download one field as a new [im][jm][km] float64 array."""
        self._check_alive(create=True)
        out = np.empty(self.shape, dtype=np.float64)
        _err(atjup_hip_download_field(self._ctx, abi.FIELD[name], _ptr(out)), f"download {name}")
        return out

    def step(self):
        """One iteration of the time loop on the device."""
        self._check_alive(create=True)
        self.iteration += 1
        self.generation += 1
        _err(atjup_hip_step(self._ctx, self.iteration, 1 if self.iteration == 1 else 0), "step")

    # ---- the adjoint of the dynamics step (the model's precision) ------------------------------
    def adjoint_initial_state(self):
        """The state the misfit's iterations start from, as the reference's run starts (the
        boundary passes, restoreVar, the mixture density) from ``state``: the fields the adjoint
        carries (t, u, v, w, p_dyn, aux_u, aux_v, aux_w, rho_mix), formed once and kept."""
        if self._adj_init is not None:
            return self._adj_init
        self._check_alive(create=True)
        for n, v in self.state.fields.items():
            if self._profile == "dyn" and n in inputs.SPECIES:
                continue
            self.set_field(n, v)
        _err(atjup_hip_bc_after_step(self._ctx), "bc_after_step")
        _err(atjup_hip_swap_nlevel(self._ctx), "swap_nlevel")
        _err(atjup_hip_mixture_density(self._ctx), "mixture_density")
        self._adj_init = {n: self.get_field(n) for n in _ADJ_STATE}
        self.iteration = 0
        return self._adj_init

    def adjoint_forward(self, int nsteps, int every):
        """nsteps iterations from the current state keeping checkpoints every `every` iterations."""
        self._check_alive(create=True)
        _err(atjup_hip_adj_init(self._ctx), "adj_init")
        self.generation += 1                      # the earlier forward's checkpoints are gone
        _err(atjup_hip_adj_forward_checkpointed(self._ctx, nsteps, every, 1), "adj_forward_checkpointed")
        self._fwd_gen = self.generation

    def adjoint_backward(self, lam_final):
        """The reverse sweep: the adjoint of the final t, u, v, w ([4][im][jm][km]) carried back to the
        initial ones. Needs the checkpoints of an adjoint_forward with no step, upload or other
        reverse sweep since; raises otherwise."""
        self._check_alive()
        if self._fwd_gen != self.generation:
            raise RuntimeError("adjoint_backward: the device state changed after the last adjoint_forward "
                               "(or its reverse sweep already ran); run adjoint_forward again")
        lf = np.ascontiguousarray(lam_final, dtype=np.float64)
        if lf.shape != (4,) + self.shape:
            raise ValueError(f"lam_final: shape {lf.shape}, expected {(4,) + self.shape}")
        out = np.empty((4,) + self.shape, dtype=np.float64)
        self.generation += 1                      # the reverse sweep consumes the checkpoints
        _err(atjup_hip_adj_backward(self._ctx, _ptr(lf), _ptr(out)), "adj_backward")
        return out

    def set_smooth_obstacle(self, a_phi_deg=20.0, b_the_deg=10.0, c_r_km=30.0, width=0.25, tau=0.05,
                            j0=None, k0=None):
        """The smooth drag region (no counterpart in the reference, whose obstacle is the binary
        SeaMount mask): -chi u / tau in the momentum equations with chi = 1 / (1 + exp(-(1 - s) /
        width)), s the ellipsoid's normalised squared distance; for gradients with respect to its
        semi-axes (adjoint_obstacle_gradient). The centre defaults to the benchmark obstacle's."""
        self._check_alive(create=True)
        o = abi.obstacle(self._grid, a_phi_deg, b_the_deg, c_r_km, width=width, tau=tau, j0=j0, k0=k0)
        cdef size_t ao = ctypes.addressof(o)
        self.generation += 1
        _err(atjup_hip_set_smooth_obstacle(self._ctx, <atjup_obstacle*>ao), "set_smooth_obstacle")
        self._obst = o

    def set_obstacle_axes(self, axes):
        """The drag region's semi-axes (a_phi_deg, b_the_deg, c_r_km), its other parameters kept."""
        if self._obst is None:
            raise AtjupError("no smooth drag region is set (set_smooth_obstacle)")
        o = self._obst
        a, b, c = (float(x) for x in axes)
        self.set_smooth_obstacle(a, b, c, o.width, o.tau, o.j0, o.k0)

    def adjoint_obstacle_gradient(self):
        """dL/d(a_phi_deg, b_the_deg, c_r_km) of the loss the last adjoint_backward carried back."""
        self._check_alive()
        out = np.zeros(3)
        _err(atjup_hip_adj_obstacle_gradient(self._ctx, _ptr(out)), "adj_obstacle_gradient")
        return out


def jet_misfit(Jupiter jup, w0, target, int nsteps=20, int every=5, i_top=None, axes=None):
    """L(w0) and its seed for the reverse sweep: sets the zonal wind w to w0 on the model's initial
    state (and, with axes, the drag region's semi-axes), runs the checkpointed forward and forms the
    misfit at level i_top (default im-2)."""
    im, jm, km = jup.shape
    i_top = im - 2 if i_top is None else int(i_top)
    if jup.profile != "dyn":
        raise AtjupError("the adjoint covers the dynamics profile: Jupiter(profile='dyn')")
    if axes is not None:
        jup.set_obstacle_axes(axes)
    for n, v in jup.adjoint_initial_state().items():
        jup.set_field(n, w0 if n == "w" else v)
    jup.adjoint_forward(nsteps, every)
    w = jup.get_field("w")
    # colatitudes as the reference forms them, cumulative sums of dthe from 0 (lib/Array_1D.cpp:49-60)
    the = np.zeros(jm)
    for j in range(1, jm):
        the[j] = the[j - 1] + jup.dthe
    cos_lat = np.abs(np.sin(the))
    r = w[i_top, :, :-1].mean(axis=1) - np.asarray(target, dtype=np.float64)
    lam = np.zeros((4, im, jm, km))
    lam[3, i_top, :, :-1] = (2.0 * cos_lat * r / (km - 1))[:, None]
    return float(np.sum(cos_lat * r * r)), lam


try:
    import torch

    class JetMisfit(torch.autograd.Function):
        """L(w0) through pyatjup with the library's checkpointed adjoint as its backward (FP64, CPU
        tensors; the device work stays in the library's own HIP runtime). One forward at a time per
        model: a backward whose forward was followed by another forward, a step or an upload on the
        same model raises, as does a second backward of one forward."""

        @staticmethod
        def forward(fctx, w0, Jupiter jup, target, nsteps=20, every=5, axes=None, i_top=None):
            ax = None if axes is None else axes.detach().cpu().numpy().astype(np.float64)
            loss, lam = jet_misfit(jup, w0.detach().cpu().numpy().astype(np.float64), target, nsteps, every,
                                   i_top=i_top, axes=ax)
            fctx.jup, fctx.lam = jup, lam
            fctx.gen = jup.generation
            fctx.has_axes = axes is not None
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(fctx, grad_out):
            if fctx.jup.generation != fctx.gen:
                raise RuntimeError("JetMisfit.backward: the model's state changed after this forward (a later "
                                   "forward replaced its checkpoints); run the forward again")
            g = fctx.jup.adjoint_backward(float(grad_out) * fctx.lam)[3]
            ga = torch.from_numpy(fctx.jup.adjoint_obstacle_gradient()) if fctx.has_axes else None
            return torch.from_numpy(g), None, None, None, None, ga, None
except ImportError:  # torch is optional for the rest of the binding
    JetMisfit = None
