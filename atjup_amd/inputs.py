"""Seeded synthetic initial states for the ATJUP time step.

The reference builds its initial state from measured jet profiles and a long initialisation chain
(planet/cJupiterModel.cpp:316-745). The benchmark and parity runs instead start from a seeded,
reproducible state with the same physical shape:

* temperature: the parabolic pole-to-equator surface profile with a constant lapse rate and an
  isothermal clamp, as planet/InitValues_Jup.cpp:186-264 (init_temperature) builds it;
* static pressure: polytropic below the clamp, isothermal above, as
  planet/InitValues_Jup.cpp:633-693 (init_PressureStatic);
* zonal wind (the reference's ``w``, the phi component): a sum of 12 Gaussian jets in latitude
  drawn from ``numpy.random.default_rng(seed)``, scaled so that max|w| = 150 m/s;
* radial ``u`` and meridional ``v``: N(0, 0.01 m/s) noise;
* the obstacle: an ellipsoid centred at j = 112 (22 S), k = 180, with semi-axes of 20 cells in
  phi, 10 cells in theta and 30 km (8.57 layers) in r above i = 0.

Everything is nondimensional in the reference's units (velocities / u_0, temperatures / t_ref).
Coordinates are built as the reference builds them, by cumulative sums (lib/Array_1D.cpp:49-60),
so the metric tables agree bit for bit.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Reference grid (planet/cJupiterModel.h:166) and spacing (planet/cJupiterModel.cpp:160-177).
IM, JM, KM = 41, 181, 361
PI180 = 180.0 / math.pi
DR = 0.025

# Mixture constants the reference computes at start-up with its default composition knobs
# (ChemistryJup::ThermalPropertiesJup after setComposition, planet/ChemistryJup.h:545-578).
R_MIX = 3531.1847352206946
CP_MIX = 11810.757033551794
R_MIX_DENSITY = 1.2843999999999998   # r_mix [kg/m3]
MUE_MIX = 1.0819029080091497e-05


@dataclass
class Grid:
    im: int = IM
    jm: int = JM
    km: int = KM
    L_atm: float = 140.0           # shell depth [km]
    radius_km: float = 71492.0     # metric radius (ATJUP_METRIC_RADIUS)
    the_degree: float = 1.0        # theta spacing [deg]
    phi_degree: float = 1.0        # phi spacing [deg]

    @property
    def dr(self) -> float:
        return 1.0 / (self.im - 1)

    @property
    def dthe(self) -> float:
        return self.the_degree / PI180

    @property
    def dphi(self) -> float:
        return self.phi_degree / PI180

    def coords(self):
        """rad.z, the.z, phi.z as the reference's cumulative sums."""
        def cum(n, z0, dz):
            z = np.empty(n)
            acc = z0
            z[0] = acc
            for l in range(1, n):
                acc = acc + dz
                z[l] = acc
            return z
        return (cum(self.im, self.radius_km / self.L_atm, self.dr),
                cum(self.jm, 0.0, self.dthe),
                cum(self.km, 0.0, self.dphi))

    def layer_heights(self):
        """get_layer_height(i): float32 i * (L_atm * dr), in km (planet/cJupiterModel.h:611-624)."""
        h = np.float32(self.L_atm * self.dr)
        return np.array([np.float32(i) * h for i in range(self.im)], dtype=np.float32)

    @property
    def shape(self):
        return (self.im, self.jm, self.km)


def fine_grid() -> Grid:
    """81 x 361 x 721 shell: half the spacing in every direction, same depth and radius."""
    return Grid(im=81, jm=361, km=721, the_degree=0.5, phi_degree=0.5)


@dataclass
class State:
    grid: Grid
    fields: dict = field(default_factory=dict)   # name -> float64 array (im, jm, km)
    mask: np.ndarray | None = None                # float64 0/1, the reference's SeaMount


SPECIES = ("h2o", "h2o_cloud", "h2o_ice", "h2s", "nh3", "nh3_cloud", "nh3_ice",
           "ch4", "ch4_cloud", "ch4_ice", "nh4sh")


def ellipsoid_mask(g: Grid, j0=None, k0=None, a_phi=None, b_the=None, c_r_km=30.0):
    """Binary ellipsoid obstacle on the grid; semi-axes in cells (phi, theta) and km (r)."""
    if j0 is None:
        j0 = round(112 * (g.jm - 1) / 180)
    if k0 is None:
        k0 = round(180 * (g.km - 1) / 360)
    if a_phi is None:
        a_phi = 20.0 / g.phi_degree
    if b_the is None:
        b_the = 10.0 / g.the_degree
    c_r = c_r_km / (g.L_atm * g.dr)            # in layers
    i = np.arange(g.im)[:, None, None]
    j = np.arange(g.jm)[None, :, None]
    k = np.arange(g.km)[None, None, :]
    q = ((k - k0) / a_phi) ** 2 + ((j - j0) / b_the) ** 2 + (i / c_r) ** 2
    return (q <= 1.0).astype(np.float64)


def temperature_pressure(g: Grid, t_equator=350.0, t_pole=330.0, t_ref=165.0, gam=2.0,
                         t_min=110.0, p_ref=1.0, grav=25.92, R_mix=R_MIX):
    """t (nondimensional) and p_stat [bar], zonally uniform (init_temperature / init_PressureStatic)."""
    im, jm = g.im, g.jm
    j_half = (jm - 1) // 2
    hts = g.layer_heights()
    t = np.empty((im, jm))
    p = np.empty((im, jm))
    t_h2_eff = t_pole - t_equator
    R_kJ = R_mix * 1.0e-3
    exp_p = grav / (gam * R_kJ)
    R_spec = R_kJ * 1.0e3
    for j in range(jm):
        ratio = j / float(j_half)
        t0 = (t_h2_eff * (ratio * ratio - 2 * ratio) + t_pole) / t_ref
        t[0, j] = t0
        isb = im
        for i in range(1, im):
            t_lapse = -gam * float(hts[i]) + t0 * t_ref
            if t_lapse < t_min:
                t[i, j] = t_min / t_ref
                if isb == im:
                    isb = i
            else:
                t[i, j] = t_lapse / t_ref
        for i in range(0, min(isb, im)):
            p[i, j] = p_ref * math.pow(t[i, j] * t_ref / t_ref, exp_p)
        if 0 < isb < im:
            t_iso = t[isb, j] * t_ref
            H = R_spec * t_iso / grav
            p_c = p[isb - 1, j]
            z_c = float(hts[isb - 1]) * 1.0e3
            for i in range(isb, im):
                z = float(hts[i]) * 1.0e3
                p[i, j] = p_c * math.exp(-(z - z_c) / H)
    return t, p


def zonal_jets(g: Grid, rng, n_jets=12, w_max=150.0):
    """A sum of Gaussian jets in latitude [m/s], max |w| = w_max."""
    _, the, _ = g.coords()
    lat = 90.0 - the * PI180
    amp = rng.uniform(-1.0, 1.0, n_jets)
    ctr = rng.uniform(-75.0, 75.0, n_jets)
    wid = rng.uniform(3.0, 12.0, n_jets)
    prof = np.zeros(g.jm)
    for a, c, s in zip(amp, ctr, wid):
        prof += a * np.exp(-0.5 * ((lat - c) / s) ** 2)
    prof *= w_max / np.max(np.abs(prof))
    return prof


def make_state(config: str = "dyn", seed: int = 1234, g: Grid | None = None,
               with_obstacle: bool = True, u_0=100.0, moisture: str = "seeded") -> State:
    """The seeded initial state of one configuration.

    config "dyn": dry dynamics (species zero, k = 0, dis = 1e-10 as resetArrays leaves them).
    config "full": moist with the k-omega closure; moisture "seeded" (``_moisten``: vapour at
    0.6-1.25 x saturation and condensate seeds, so that every branch of the adjustment runs) or
    "subsaturated" (``_moisten_subsaturated``: vapour at 0.5-1.0 x saturation, no condensate, the
    full-physics benchmark case).
    """
    g = g or Grid()
    rng = np.random.default_rng(seed)
    shp = g.shape
    t2, p2 = temperature_pressure(g)
    f = {}
    f["t"] = np.repeat(t2[:, :, None], g.km, axis=2)
    f["p_stat"] = np.repeat(p2[:, :, None], g.km, axis=2)
    prof = zonal_jets(g, rng)
    f["w"] = np.broadcast_to((prof / u_0)[None, :, None], shp).copy()
    f["u"] = rng.normal(0.0, 0.01, shp) / u_0
    f["v"] = rng.normal(0.0, 0.01, shp) / u_0
    f["p_dyn"] = np.zeros(shp)
    for s in SPECIES:
        f[s] = np.zeros(shp)
    f["tke"] = np.zeros(shp)
    f["dis"] = np.full(shp, 1.0e-10)
    mask = ellipsoid_mask(g) if with_obstacle else np.zeros(shp)
    solid = mask == 1.0
    for n in ("u", "v", "w"):
        f[n][solid] = 0.0
    # the seam meridian k = km-1 is the copy of k = 0 everywhere (planet/ATPhys.h:103-139)
    for n in f:
        f[n][:, :, g.km - 1] = f[n][:, :, 0]
    st = State(grid=g, fields=f, mask=mask)
    if config == "full":
        if moisture == "seeded":
            _moisten(st, rng)
        elif moisture == "subsaturated":
            _moisten_subsaturated(st, rng)
        else:
            raise ValueError(f"moisture recipe {moisture!r}")
    return st


def _saturation_mixing(gas, T, P, rho):
    """q_sat over the liquid [kg/m3] of one gas, capped at 0.05, from the Sanchez-Lavega fit."""
    from .abi import GAS_PARAMS
    t_0, t_00, ep, lv, ls, C, L0, R, da, db = GAS_PARAMS[gas][:10]
    e = np.exp(C + (-L0 / T + da * np.log(T) + db * T) / (1e-3 * R))   # [bar] (planet/ATPhys.h:49-52)
    return np.minimum(rho * ep * e / P, 0.05)


def _moisten_subsaturated(st: State, rng):
    """The full-physics benchmark's moist state: H2O, NH3 and CH4 vapour at seeded U(0.5, 1.0) x
    saturation over the liquid, no condensate (clouds form only where the flow or the radiation
    brings a cell to saturation), H2S and NH4SH traces, and the k-omega start values."""
    from .abi import GAS_FIELDS
    g = st.grid
    f = st.fields
    shp = g.shape
    T = f["t"] * 165.0
    P = f["p_stat"]
    rho = P * 1.0e5 / (R_MIX * T)
    for gas, (vap, cld, ice) in GAS_FIELDS.items():
        f[vap] = _saturation_mixing(gas, T, P, rho) * rng.uniform(0.5, 1.0, shp)
        f[cld] = np.zeros(shp)
        f[ice] = np.zeros(shp)
    f["h2s"] = 0.0004 * rng.uniform(0.5, 1.0, shp)
    f["nh4sh"] = 1.0e-5 * rng.uniform(0.0, 1.0, shp)
    f["tke"] = np.full(shp, 1.0e-4 / 100.0 ** 2)
    f["dis"] = np.full(shp, 1.0e-3 * 140.0e3 / 100.0)
    solid = st.mask == 1.0
    for s in SPECIES + ("tke",):
        f[s][solid] = 0.0
    for n in f:
        f[n][:, :, g.km - 1] = f[n][:, :, 0]


def _moisten(st: State, rng):
    """Moist state of the full-physics case, seeded.

    Vapour of H2O, NH3 and CH4 at U(0.6, 1.25) x saturation over the liquid (so that part of the
    shell is supersaturated and the saturation adjustment condenses there), a condensate seed of
    cloud and ice in a random tenth of the cells of each gas (so that the evaporation and
    Bergeron branches run), H2S and NH4SH traces, and the k-omega start values.
    """
    from .abi import GAS_PARAMS, GAS_FIELDS
    g = st.grid
    f = st.fields
    shp = g.shape
    T = f["t"] * 165.0
    P = f["p_stat"]
    rho = P * 1.0e5 / (R_MIX * T)
    for gas, (vap, cld, ice) in GAS_FIELDS.items():
        t_0, t_00, ep, lv, ls, C, L0, R, da, db = GAS_PARAMS[gas][:10]
        # Sanchez-Lavega vapour pressure over the liquid [bar] (planet/ATPhys.h:49-52)
        e = np.exp(C + (-L0 / T + da * np.log(T) + db * T) / (1e-3 * R))
        q_sat = np.minimum(rho * ep * e / P, 0.05)
        f[vap] = q_sat * rng.uniform(0.6, 1.25, shp)
        seed = rng.uniform(size=shp) < 0.1
        f[cld] = np.where(seed, 0.05 * q_sat * rng.uniform(size=shp), 0.0)
        f[ice] = np.where(seed & (T < t_0), 0.05 * q_sat * rng.uniform(size=shp), 0.0)
    f["h2s"] = 0.0004 * rng.uniform(0.5, 1.0, shp)
    f["nh4sh"] = 1.0e-5 * rng.uniform(0.0, 1.0, shp)
    # k = 1e-4 m2/s2 -> 1e-8 in u_0^2; omega = 1e-3 1/s -> 1e-3 * L / u_0 = 1.4
    f["tke"] = np.full(shp, 1.0e-4 / 100.0 ** 2)
    f["dis"] = np.full(shp, 1.0e-3 * 140.0e3 / 100.0)
    solid = st.mask == 1.0
    for s in SPECIES + ("tke",):
        f[s][solid] = 0.0
    for n in f:
        f[n][:, :, g.km - 1] = f[n][:, :, 0]
