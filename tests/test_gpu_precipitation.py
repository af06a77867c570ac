"""The precipitation microphysics on the GPU against the oracle (planet/Precipitation.h:194-574).

The oracle reproduces the compiled reference bit for bit with the rates reported (coupling on), with
the depletion applied in place and the rate scale moved, and on a state whose lowest levels freeze
under the rain (test_oracle_vs_reference.py). The saturation vapour pressure's exp/log and the
rain ratio's pow come from the device library, within an ulp of the host's, so the fluxes, rates and
heating carry ulp-level differences: the bar is 1e-11 of each field's maximum. The step with the
radiation and the precipitation in its physics block is held to the drop-in's 1e-10.
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


@pytest.mark.parametrize("coupling,scale,inversion", [(1.0, 1.0, False), (0.0, 3.0, False), (1.0, 1.0, True)],
                         ids=["rates", "in-place-scale3", "inversion"])
def test_precipitation_against_oracle(coupling, scale, inversion):
    gpu, orc, st = _pair(coupling, inversion)
    pp = abi.precip_params(rate_scale=scale)
    gpu.set_precipitation(pp)
    orc.set_precipitation(pp)
    for c in (gpu,):
        c.mixture_density()
    orc.call("mixture_density")
    gpu.precipitation()
    orc.call("precipitation")
    names = P_FIELDS + ("Q_precip",) + S_FIELDS + inputs.SPECIES
    w = compare(gpu, orc, names, 1.0e-11, f"precipitation coupling={coupling}")
    print("precipitation", coupling, scale, inversion,
          {k: f"{v[0]:.1e}" for k, v in w.items() if v[0] > 0})
    a, b = gpu.precip_surface(), orc.precip_surface()
    for n in abi.PRECIP_SURFACE:
        assert rel_err(a[n], b[n]) <= 1.0e-11, n
    assert np.count_nonzero(gpu.get("P_rain")) > 1000 and np.count_nonzero(gpu.get("Q_precip")) > 1000
    if inversion:
        assert np.count_nonzero(gpu.get("P_graupel")) > 100
    gpu.close()
    orc.close()


def test_step_with_radiation_and_precipitation_against_oracle():
    """Two iterations with both columns in the physics block: the RHS reads the S_precip_* rates
    and Q_precip, Q_rad of this step's state."""
    gpu, orc, st = _pair(radiation=True, seed=3)
    for it in (1, 2):
        gpu.step(it, it == 1)
        orc.call("step", it, 1 if it == 1 else 0)
    w = compare(gpu, orc, abi.PROG + ("p_dyn", "Q_rad", "Q_precip") + S_FIELDS, 1.0e-10, "2 steps with precipitation")
    print("step with radiation and precipitation", {k: f"{v[0]:.1e}" for k, v in w.items()})
    assert np.count_nonzero(gpu.get("S_precip_h2o_cloud")) > 1000
    gpu.close()
    orc.close()


def test_precipitation_fp32_refined_form():
    """FP32 fields in, FP64 column arithmetic: the fluxes and rates against FP64 on the same state."""
    g64, o64, st = _pair()
    g32, o32, _ = _pair(precision=abi.PRECISION_FP32_REF)
    for c in (g64, g32):
        c.mixture_density()
        c.precipitation()
    for n in ("P_rain", "P_snow", "P_nh3_rain", "P_nh4sh", "S_precip_h2o_cloud"):
        e = rel_err(g32.get(n), g64.get(n))
        print(f"FP32-refined {n}: {e:.2e}")
        assert e < 1.0e-4, n
    for c in (g64, o64, g32, o32):
        c.close()
