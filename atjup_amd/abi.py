"""ctypes mirror of include/atjup_hip.h: field ids, the three context structs and the knob profiles.

A profile is the set of reference environment knobs a configuration runs under, expressed twice:
as the ``atjup_knobs`` struct the library takes and as the environment string the reference would
read (``ATJUP_*`` variables, planet/cJupiterModel.h:725-1216 and the shared headers).
"""
from __future__ import annotations

import ctypes as C
import math

from . import inputs

# ---- field ids (enum atjup_field) ----------------------------------------------------------
PROG = ("t", "u", "v", "w", "h2o", "h2o_cloud", "h2o_ice", "h2s", "nh3", "nh3_cloud", "nh3_ice",
        "ch4", "ch4_cloud", "ch4_ice", "nh4sh", "tke", "dis")
N_PROG = len(PROG)
INPUTS = ("p_dyn", "p_stat", "p_hydro", "rho_mix", "wall_nue", "nue", "prod", "tke_source",
          "dis_source", "aux_u", "aux_v", "aux_w", "Q_rad", "Q_precip",
          "S_precip_h2o", "S_precip_h2o_cloud", "S_precip_h2o_ice",
          "S_precip_nh3", "S_precip_nh3_cloud", "S_precip_nh3_ice",
          "S_precip_ch4", "S_precip_ch4_cloud", "S_precip_ch4_ice",
          "massflux_h2s", "massflux_nh3", "massflux_nh4sh", "fluxlim_nh4sh",
          "epsilon", "radiation",
          "P_rain", "P_snow", "P_graupel", "P_nh3_rain", "P_nh3_snow", "P_nh3_graupel",
          "P_ch4_rain", "P_ch4_snow", "P_ch4_graupel", "P_nh4sh",
          "difflux_h2s", "difflux_nh3", "thermalmassflux", "difflux_nh4sh")
NLEVEL = 64

FIELD = {n: i for i, n in enumerate(PROG)}
FIELD.update({n: 20 + i for i, n in enumerate(INPUTS)})
FIELD.update({n + "n": NLEVEL + i for i, n in enumerate(PROG)})   # tn, un, ..., disn
F_END = 20 + len(INPUTS)

OK, ERR_ARG, ERR_HIP, ERR_PROFILE, ERR_NOMEM, ERR_STATE, NONFINITE = 0, -1, -2, -3, -4, -5, 1
PRECISION_FP64, PRECISION_FP32_REF, PRECISION_TANGENT = 0, 1, 2


class Grid(C.Structure):
    _fields_ = [("im", C.c_int), ("jm", C.c_int), ("km", C.c_int),
                ("r0", C.c_double), ("dr", C.c_double), ("dthe", C.c_double), ("dphi", C.c_double),
                ("L_atm", C.c_double)]


class Consts(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "dt", "re", "pr", "g", "omega", "u_0", "t_ref",
        "Coriolis", "centrifugal", "buoyancy", "chemical_reaction",
        "r_mix", "R_mix", "cp_mix", "mue_mix",
        "sc_h2o", "sc_h2s", "sc_nh3", "sc_nh4sh", "sc_ch4",
        "rg_nh4sh", "r_p_nh4sh", "re_turb")]


class Knobs(C.Structure):
    _fields_ = [("turb_model", C.c_int), ("press_sweeps", C.c_int), ("press_omega", C.c_double),
                ("pdyn_cap", C.c_double), ("damp_pdyn", C.c_int), ("rho_pdyn", C.c_double),
                ("sinthe_min", C.c_double), ("t_diff_scale", C.c_double),
                ("species_diff_scale", C.c_double), ("rad_coupling", C.c_double),
                ("precip_coupling", C.c_double), ("turb_coupling", C.c_double),
                ("tke_max_phys", C.c_double), ("nue_max_phys", C.c_double),
                ("wall_nue_factor", C.c_double), ("wall_nue_layers", C.c_int),
                ("satadj", C.c_int), ("rk_scalar_sync", C.c_int), ("dry", C.c_int),
                ("turb_k_bg_phys", C.c_double), ("radiation", C.c_int), ("precip", C.c_int),
                ("chemistry", C.c_int), ("mca", C.c_int), ("nh4sh_eq", C.c_int)]


CHEM_DOUBLES = ("m_nh3", "m_h2s", "m_nh4sh", "m_h2", "m_he", "m_h2o", "X_h2", "X_he", "X_h2o", "X_h2s",
                "X_nh3", "X_nh4sh", "mue_nh3", "mue_h2s", "mue_nh4sh", "rg_nh3", "rg_h2s", "k_mix",
                "cp_nh3", "cp_h2s", "cp_nh4sh", "k_nh3", "k_h2s", "k_nh4sh", "dh_nh4sh", "t_0_nh4sh",
                "t_00_nh4sh")


class ChemParams(C.Structure):
    """atjup_chem_params: the chemistry's constants (planet/cJupiterModel.h:172-401) and knobs."""
    _fields_ = [(n, C.c_double) for n in CHEM_DOUBLES] + [
        ("gate_zero", C.c_int), ("molar_conc", C.c_int), ("enthalpy", C.c_int)]


# the values the reference's start-up leaves (read from the compiled reference's members)
CHEM_DEFAULTS = dict(zip(CHEM_DOUBLES, (
    17.03052, 34.08088, 51.1114, 2.016, 4.0026, 18.01588, 0.864, 0.136, 0.0017, 7.7e-05, 0.0002, 0.0,
    9.2e-06, 1.3e-05, 0.0, 817.0, 949.0, 0.12478002943205087, 2190.0, 2240.0, 2000.0, 0.02102, 0.013,
    0.0, 1768700.0, 230.0, 200.0)), gate_zero=1, molar_conc=0, enthalpy=0)


def chem_params(**over) -> ChemParams:
    d = dict(CHEM_DEFAULTS)
    d.update(over)
    return ChemParams(**d)


class MCAParams(C.Structure):
    """atjup_mca_params: ATJUP_MCA_RELAX / _LEDOUX / _LOADING / _PASSES, ep_h2s and the NH4SH heat."""
    _fields_ = [("relax", C.c_double), ("ep_h2s", C.c_double), ("l_nh4sh", C.c_double),
                ("ledoux", C.c_int), ("loading", C.c_int), ("max_pass", C.c_int)]


# the knobs' defaults (planet/ConvectiveAdjustment.h:409-413) and ep_h2s = m_h2s / M_mix as the
# reference's start-up leaves it (setCompositionEp, planet/InitValues_Jup.cpp:940-946)
MCA_DEFAULTS = dict(relax=1.0, ep_h2s=14.474287352261996, l_nh4sh=0.0, ledoux=1, loading=1, max_pass=64)
MCA_ENV = {"relax": "ATJUP_MCA_RELAX", "ledoux": "ATJUP_MCA_LEDOUX", "loading": "ATJUP_MCA_LOADING",
           "max_pass": "ATJUP_MCA_PASSES"}


def mca_params(**over) -> MCAParams:
    d = dict(MCA_DEFAULTS)
    d.update(over)
    return MCAParams(**d)


class PrecipParams(C.Structure):
    """atjup_precip_params: ATJUP_PRECIP_SCALE and the NH4SH crystal band [K]."""
    _fields_ = [("rate_scale", C.c_double), ("t_0_nh4sh", C.c_double), ("t_00_nh4sh", C.c_double)]


class Obstacle(C.Structure):
    """atjup_obstacle: the smooth drag region standing in for the obstacle in gradients with
    respect to its semi-axes (no counterpart in the reference)."""
    _fields_ = [("j0", C.c_double), ("k0", C.c_double), ("a_phi_deg", C.c_double), ("b_the_deg", C.c_double),
                ("c_r_km", C.c_double), ("width", C.c_double), ("tau", C.c_double)]


def obstacle(g: inputs.Grid, a_phi_deg=20.0, b_the_deg=10.0, c_r_km=30.0, width=0.25, tau=0.05,
             j0=None, k0=None) -> Obstacle:
    """The drag region of the benchmark's obstacle (22 S, 20 deg x 10 deg x 30 km, centred as
    inputs.ellipsoid_mask centres it): chi = 1 / (1 + exp(-(1 - s) / width)) and a drag time tau
    of ten time steps."""
    j0 = round(112 * (g.jm - 1) / 180) if j0 is None else j0
    k0 = round(180 * (g.km - 1) / 360) if k0 is None else k0
    return Obstacle(j0=float(j0), k0=float(k0), a_phi_deg=a_phi_deg, b_the_deg=b_the_deg, c_r_km=c_r_km,
                    width=width, tau=tau)


def precip_params(**over) -> PrecipParams:
    d = dict(rate_scale=1.0, t_0_nh4sh=230.0, t_00_nh4sh=200.0)
    d.update(over)
    return PrecipParams(**d)


# the surface precipitation maps atjup_hip_download_precip_surface returns, in its order
PRECIP_SURFACE = ("precip_srf_h2o", "precip_srf_nh3", "precip_srf_ch4", "precip_srf_nh4sh", "precip_srf_total")


class RadParams(C.Structure):
    """atjup_rad_params: Jupiter's radiation properties (planet/cJupiterModel.h:1109-1118) and the
    ATJUP_ knobs Radiation<P>::run reads (planet/Radiation.h:270-295)."""
    _fields_ = [(n, C.c_double) for n in (
        "F_int", "S_solar", "albedo_bond", "x_H2", "x_He",
        "cia_strength", "opacity_strength", "solar_strength", "sw_tau_per_bar")] + [
        ("solar", C.c_int), ("bottom", C.c_int)]


def rad_params(**over) -> RadParams:
    """The defaults of a new context (Jupiter, every knob at its default), with overrides."""
    d = dict(F_int=5.4, S_solar=50.5, albedo_bond=0.343, x_H2=0.863, x_He=0.134, cia_strength=3.9,
             opacity_strength=1.0, solar_strength=1.0, sw_tau_per_bar=1.0, solar=1, bottom=0)
    d.update(over)
    return RadParams(**d)


# the reference environment variable of each radiation knob (planet/Radiation.h:270-295)
RAD_ENV = {"cia_strength": "ATJUP_CIA_STRENGTH", "opacity_strength": "ATJUP_OPACITY_STRENGTH",
           "solar": "ATJUP_SOLAR", "solar_strength": "ATJUP_SOLAR_STRENGTH",
           "sw_tau_per_bar": "ATJUP_SW_TAU_PER_BAR", "bottom": "ATJUP_RAD_BOTTOM"}


class GasParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "t_0", "t_00", "ep", "lv", "ls", "C", "L0", "R", "del_alf", "del_bet",
        "C_i", "L0_i", "del_alf_i", "del_bet_i")]


# The per-gas constants the reference's time loop hands the saturation adjustment
# (planet/cJupiterModel.cpp:1060-1080) after its start-up with the default composition and
# vapour-pressure fits (setSvpFit, setCompositionEp: planet/InitValues_Jup.cpp:898-950).
GAS_PARAMS = {
    0: (273.15, 210.15, 7.651416982890988, 2500900.0, 2833900.0, 25.096, 3148.2, 461.52, 0.0, -0.0087,
        -1.962499, 2641.4013, 1.629479, -0.003361398),
    1: (195.5, 140.0, 7.232930612074716, 1372000.0, 1832000.0, 196.077451, 3586.5634, 488.21, -15.978352,
        0.02837146, 16.086062, 1857.9898, 0.098768, -0.001179989),
    2: (90.69, 67.36, 6.813102176498581, 511000.0, 511000.0, 1.627, 553.1, 518.28, 1.002, -0.0041,
        1.627, 553.1, 1.002, -0.0041),
}
GAS_FIELDS = {0: ("h2o", "h2o_cloud", "h2o_ice"), 1: ("nh3", "nh3_cloud", "nh3_ice"),
              2: ("ch4", "ch4_cloud", "ch4_ice")}


def gas_params(gas: int) -> GasParams:
    return GasParams(*GAS_PARAMS[gas])


# Reference environment shared by every profile: the balanced start, vortex, cone projection and
# continuity init are off (they are outside this path), the metric radius is 71492 km, dt 0.005.
# make_profile sets the moist convective adjustment and the NH4SH equilibrium per profile.
BASE_ENV = {
    "ATJUP_GEO_BALANCE": "0", "ATJUP_VORTEX": "0", "ATJUP_MCA": "0",
    "ATJUP_V_MASSBAL_STRIDE": "0", "ATJUP_CONE_PROJ": "0", "ATJUP_CONE_PROJ_STRIDE": "0",
    "ATJUP_U_INIT_CONT": "0", "ATJUP_NH4SH_EQ": "0", "ATJUP_METRIC_RADIUS": "71492",
    "ATJUP_DT": "0.005", "ATJUP_BC_PHI_PERIODIC": "1",
}
# the NH4SH heat of formation the equilibrium adjustment and the moist convective adjustment use:
# the van't Hoff enthalpy of the Lewis K_p, 4705 ln(10) R* / m_nh4sh [J/kg] (ChemistryJup.h:282-283)
NH4SH_FORMATION_HEAT = 4705.0 * math.log(10.0) * 8314.46 / 51.1114

TURB_MODELS = {"off": -1, "k_epsilon": 0, "k_omega": 1, "k_omega_SST": 2}


def make_profile(config: str = "dyn", sweeps: int = 200, omega: float = 1.0,
                 grid: inputs.Grid | None = None, radiation: bool = False, precip: bool = False,
                 precip_coupling: float = 1.0, chemistry: bool = False, mca: int = 0,
                 turb: str = "k_omega", nh4sh_eq: bool = False):
    """(Grid, Consts, Knobs, env) of a configuration.

    "dyn"  : dynamics only, dry, closure/radiation/precipitation off (the 5-step benchmark case).
    "full" : k-omega closure, moist, saturation adjustment on, radiation/precipitation as frozen
             inputs (their modules are outside this path).
    radiation: the grey radiation runs in the physics block (ATJUP_RADIATION=1, the reference's
             default), so Q_rad follows the state instead of entering as a frozen input.
    precip:  the precipitation microphysics runs after the saturation adjustments (ATJUP_PRECIP=1,
             the reference's default); with precip_coupling != 0 it reports S_precip_* rates.
    chemistry: the physics block's DiffMassFluxJup / ChemMassRateJup and their smoothing run in the
             step (the reference always runs them), so massflux_* follow the state.
    mca:     the moist convective adjustment every mca iterations after the step (ATJUP_MCA, the
             reference's default 1); 0 off.
    turb:    the closure of the "full" profile (ATJUP_TURB_MODEL): "k_omega" (the benchmark case),
             "k_omega_SST" (the reference's default) or "k_epsilon".
    nh4sh_eq: the NH4SH equilibrium adjustment after the saturation adjustments (ATJUP_NH4SH_EQ,
             the reference's default 1), which also zeroes the NH4SH rate law of the chemistry.
    """
    if config != "dyn" and turb not in ("k_omega", "k_omega_SST", "k_epsilon"):
        raise ValueError(f"closure {turb!r}")
    g = grid or inputs.Grid()
    G = Grid(im=g.im, jm=g.jm, km=g.km, r0=g.radius_km / g.L_atm, dr=g.dr, dthe=g.dthe,
             dphi=g.dphi, L_atm=g.L_atm)
    Cn = Consts(dt=0.005, re=1000.0, pr=0.69, g=25.92, omega=0.000176, u_0=100.0, t_ref=165.0,
                Coriolis=1.0, centrifugal=1.0, buoyancy=1.0,
                chemical_reaction=0.0 if config == "dyn" else 1.0,
                r_mix=inputs.R_MIX_DENSITY, R_mix=inputs.R_MIX, cp_mix=inputs.CP_MIX,
                mue_mix=inputs.MUE_MIX, sc_h2o=0.61, sc_h2s=0.94, sc_nh3=0.61, sc_nh4sh=0.7,
                sc_ch4=0.99, rg_nh4sh=1170.0, r_p_nh4sh=1.0e-5,
                # Turbulence<P>::init sets the model's re_turb = vel_star_ref z_0 / nue_gas of its
                # default constructor (Turbulence.h:245-251, :286); only the k-epsilon sources read it
                re_turb=0.4 * 0.1 / 1.8e-5)
    turb = "off" if config == "dyn" else turb
    K = Knobs(turb_model=TURB_MODELS[turb], press_sweeps=sweeps, press_omega=omega,
              pdyn_cap=0.2, damp_pdyn=1, rho_pdyn=1.0, sinthe_min=0.55, t_diff_scale=0.01,
              species_diff_scale=0.01, rad_coupling=1.0, precip_coupling=precip_coupling, turb_coupling=1.0,
              tke_max_phys=1000.0, nue_max_phys=1.0e5, wall_nue_factor=4.0, wall_nue_layers=3,
              satadj=1, rk_scalar_sync=2, dry=1 if config == "dyn" else 0,
              turb_k_bg_phys=1.0e-4 if config != "dyn" else 1.0, radiation=1 if radiation else 0,
              precip=1 if precip else 0, chemistry=1 if chemistry else 0, mca=int(mca),
              nh4sh_eq=1 if nh4sh_eq else 0)
    env = dict(BASE_ENV)
    env["ATJUP_PRESS_SWEEPS"] = str(sweeps)
    env["ATJUP_RADIATION"] = "1" if radiation else "0"
    env["ATJUP_PRECIP"] = "1" if precip else "0"
    env["ATJUP_PRECIP_COUPLING"] = repr(float(precip_coupling))
    env["ATJUP_MCA"] = str(int(mca))
    env["ATJUP_NH4SH_EQ"] = "1" if nh4sh_eq else "0"
    if config == "dyn":
        env["ATJUP_TURB"] = "0"
    else:
        env["ATJUP_TURB_MODEL"] = turb
        env["ATJUP_TURB_K_BG"] = "1e-4"
    return G, Cn, K, env


def env_string(env: dict) -> str:
    return ";".join(f"{k}={v}" for k, v in env.items())


# The restart file's arrays in the reference's order (planet/FileIO_Jup.cpp:281-305): the fixed list
# (n-level copies by their reference names) and, in JUP2 files, the carried list, then vel_star.
RESTART_FIXED = ("t", "u", "v", "w", "tn", "un", "vn", "wn",
                 "h2o", "h2o_cloud", "h2o_ice", "h2on", "h2o_cloudn", "h2o_icen", "h2s", "h2sn",
                 "nh3", "nh3_cloud", "nh3_ice", "nh3n", "nh3_cloudn", "nh3_icen",
                 "ch4", "ch4_cloud", "ch4_ice", "ch4n", "ch4_cloudn", "ch4_icen", "nh4sh", "nh4shn",
                 "p_dyn", "p_dynn", "p_stat", "tke", "dis", "tken", "disn", "nue")
RESTART_CARRIED = ("aux_u", "aux_v", "aux_w", "difflux_h2s", "difflux_nh3", "massflux_h2s", "massflux_nh3",
                   "fluxlim_nh4sh", "thermalmassflux", "Q_precip", "Q_rad", "rho_mix",
                   "S_precip_h2o", "S_precip_h2o_cloud", "S_precip_h2o_ice", "S_precip_nh3", "S_precip_nh3_ice")
