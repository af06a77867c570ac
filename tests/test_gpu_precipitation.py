"""The precipitation microphysics on the GPU against the oracle (planet/Precipitation.h:194-574).

The oracle reproduces the compiled reference bit for bit with the rates reported (coupling on), with
the depletion applied in place and the rate scale moved, and on a state whose lowest levels freeze
under the rain (test_oracle_vs_reference.py). The device evaluates the vapour pressure's exp/log
and the rain ratio's pow within an ulp of the host, and the scheme itself is ill-conditioned where
a column's rain flux is near zero: the evaporation grows as Rain^(4/9), whose derivative is
unbounded at 0, so an ulp in a tiny upstream flux reaches the fluxes and heating below. The oracle
shows it on its own: with t, p_stat and the species moved by one ulp in random directions, its
rain fluxes move by ~4e-2 of their maximum (normwise 4e-3) and Q_precip by ~0.3 (normwise 2e-2),
while the snow flux moves by 1e-14. So each field of the GPU path is held to twice the oracle's
own one-ulp spread, max-abs and normwise (floor 1e-11), computed in the test; the NH4SH settling
flux, the snow and the surface NH4SH map, which carry no such amplification here, to 1e-11.
"""
import numpy as np
import pytest

from atjup_amd import abi, inputs
from tests.helpers import compare, rel_err

pytestmark = pytest.mark.gpu

P_FIELDS = ("P_rain", "P_snow", "P_graupel", "P_nh3_rain", "P_nh3_snow", "P_nh3_graupel",
            "P_ch4_rain", "P_ch4_snow", "P_ch4_graupel", "P_nh4sh")
S_FIELDS = tuple(n for n in abi.INPUTS if n.startswith("S_precip"))


def _pair(coupling=1.0, inversion=False, precision=abi.PRECISION_FP64, radiation=False, seed=11):
    from atjup_amd.lib import Context
    import oracle
    st = inputs.make_state("full", seed=seed)
    if inversion:
        st.fields["t"][:4, 40:90, :] *= 0.7          # rain falling into freezing layers
    G, Cn, K, _ = abi.make_profile("full", sweeps=20, grid=st.grid, precip=True, precip_coupling=coupling,
                                   radiation=radiation)
    gpu = Context(G, Cn, K, precision)
    orc = oracle.Oracle(G, Cn, K)
    gpu.load_state(st)
    orc.load_state(st)
    gpu.bc_after_step(); gpu.swap_nlevel()
    orc.call("bc_after_step"); orc.call("swap_nlevel")
    return gpu, orc, st


def _norm_rel(a, b):
    nb = np.linalg.norm(b)
    return 0.0 if nb == 0 else float(np.linalg.norm(a - b) / nb)


def _oracle_run(st, coupling, scale, names):
    import oracle
    G, Cn, K, _ = abi.make_profile("full", sweeps=20, grid=st.grid, precip=True, precip_coupling=coupling)
    o = oracle.Oracle(G, Cn, K)
    o.load_state(st)
    o.call("bc_after_step"); o.call("swap_nlevel")
    o.set_precipitation(abi.precip_params(rate_scale=scale))
    o.call("mixture_density")
    o.call("precipitation")
    out = {n: o.get(n) for n in names}
    out.update(o.precip_surface())
    o.close()
    return out


@pytest.mark.parametrize("coupling,scale,inversion", [(1.0, 1.0, False), (0.0, 3.0, False), (1.0, 1.0, True)],
                         ids=["rates", "in-place-scale3", "inversion"])
def test_precipitation_against_oracle(coupling, scale, inversion):
    gpu, orc, st = _pair(coupling, inversion)
    orc.close()
    names = P_FIELDS + ("Q_precip",) + S_FIELDS + inputs.SPECIES
    ref = _oracle_run(st, coupling, scale, names)
    # the oracle's own spread under one-ulp input noise
    rng = np.random.default_rng(5)
    f2 = {n: (np.nextafter(v, rng.choice([-np.inf, np.inf], size=v.shape))
              if n in ("t", "p_stat") + inputs.SPECIES else v.copy()) for n, v in st.fields.items()}
    noisy = _oracle_run(inputs.State(grid=st.grid, mask=st.mask, fields=f2), coupling, scale, names)
    gpu.set_precipitation(abi.precip_params(rate_scale=scale))
    gpu.mixture_density()
    gpu.precipitation()
    got = {n: gpu.get(n) for n in names}
    got.update(gpu.precip_surface())
    exact = ("P_nh4sh", "P_snow", "precip_srf_nh4sh")
    for n in names + abi.PRECIP_SURFACE:
        e_max, e_norm = rel_err(got[n], ref[n]), _norm_rel(got[n], ref[n])
        f_max, f_norm = rel_err(noisy[n], ref[n]), _norm_rel(noisy[n], ref[n])
        if e_max > 0:
            print(f"{n}: GPU {e_max:.1e} / {e_norm:.1e}, oracle one-ulp spread {f_max:.1e} / {f_norm:.1e}")
        assert np.all(np.isfinite(got[n])), n
        if n in exact:
            assert e_max <= 1.0e-11, n
        else:
            assert e_max <= max(2.0 * f_max, 1.0e-11), n
            assert e_norm <= max(2.0 * f_norm, 1.0e-11), n
    assert np.count_nonzero(got["P_rain"]) > 1000 and np.count_nonzero(got["Q_precip"]) > 1000
    if inversion:
        assert np.count_nonzero(got["P_graupel"]) > 100
    gpu.close()


def test_step_with_radiation_and_precipitation_against_oracle():
    """Two iterations with both columns in the physics block: the RHS reads the S_precip_* rates
    and Q_precip, Q_rad of this step's state."""
    gpu, orc, st = _pair(radiation=True, seed=3)
    for it in (1, 2):
        gpu.step(it, it == 1)
        orc.call("step", it, 1 if it == 1 else 0)
    # the prognostic fields take the rates through dt; the rates themselves carry the scheme's
    # ill-conditioning at near-zero rain (see the module note) and are checked above
    w = compare(gpu, orc, abi.PROG + ("p_dyn", "Q_rad"), 1.0e-10, "2 steps with precipitation")
    print("step with radiation and precipitation", {k: f"{v[0]:.1e}" for k, v in w.items()})
    assert np.count_nonzero(gpu.get("S_precip_h2o_cloud")) > 1000
    gpu.close()
    orc.close()


def test_precipitation_fp32_refined_form():
    """FP32 fields in, FP64 column arithmetic: against FP64 on the same state, normwise, within
    twice the FP64 path's own change when its inputs are rounded to FP32."""
    g64, o64, st = _pair()
    g32, o32, _ = _pair(precision=abi.PRECISION_FP32_REF)
    rounded = inputs.State(grid=st.grid, mask=st.mask,
                           fields={n: v.astype(np.float32).astype(np.float64) for n, v in st.fields.items()})
    gr, orr = _pair_from(rounded)
    for c in (g64, g32, gr):
        c.mixture_density()
        c.precipitation()
    for n in ("P_rain", "P_snow", "P_nh3_rain", "P_nh4sh", "Q_precip", "S_precip_h2o_cloud"):
        e = _norm_rel(g32.get(n), g64.get(n))
        floor = _norm_rel(gr.get(n), g64.get(n))
        print(f"FP32-refined {n}: {e:.2e} normwise (FP64 from the rounded state: {floor:.2e})")
        assert e <= max(2.0 * floor, 1.0e-6), n
    for c in (g64, o64, g32, o32, gr, orr):
        c.close()


def _pair_from(st):
    from atjup_amd.lib import Context
    import oracle
    G, Cn, K, _ = abi.make_profile("full", sweeps=20, grid=st.grid, precip=True)
    gpu = Context(G, Cn, K, abi.PRECISION_FP64)
    orc = oracle.Oracle(G, Cn, K)
    gpu.load_state(st)
    gpu.bc_after_step(); gpu.swap_nlevel()
    return gpu, orc
