"""The grey two-stream radiation on the GPU against the oracle (planet/Radiation.h:245-487).

The oracle's radiation reproduces the compiled reference bit for bit at the knob defaults and with
every knob moved (test_oracle_vs_reference.py). The layer emissivity 1 - exp(-tau) and the
shortwave beam exp(-k p) come from the device library's exp, within an ulp of the host's, so the
outputs carry ulp-level differences: the bar is 1e-12 of each field's maximum, and the columns'
untouched cells (none here) and the solid cells' zero heating are exact. The step with the
radiation in its physics block (ATJUP_RADIATION=1, the reference's default) is held to the
drop-in's 1e-10 over two iterations with the physics block.
"""
import numpy as np
import pytest

from atjup_amd import abi, inputs
from tests.helpers import compare, make_pair, rel_err

pytestmark = pytest.mark.gpu

OUT = ("epsilon", "radiation", "Q_rad")


@pytest.fixture(scope="module")
def pair():
    st = inputs.make_state("full", seed=7)
    gpu, orc, st = make_pair("full", sweeps=20, state=st)
    gpu.mixture_density()
    orc.call("mixture_density")
    yield gpu, orc, st
    gpu.close()
    orc.close()


@pytest.mark.parametrize("over", [{}, dict(solar=0, bottom=1, cia_strength=2.5),
                                  dict(opacity_strength=3.0, solar_strength=0.5, sw_tau_per_bar=2.0)],
                         ids=["defaults", "no-sun-bottom1", "opacity-sw"])
def test_radiation_against_oracle(pair, over):
    gpu, orc, st = pair
    p = abi.rad_params(**over)
    gpu.set_radiation(p)
    orc.set_radiation(p)
    gpu.radiation()
    orc.call("radiation")
    w = compare(gpu, orc, OUT, 1.0e-12, f"radiation {over}")
    print("radiation", over, {k: f"{v[0]:.1e}/{v[1]}" for k, v in w.items()})
    q = gpu.get("Q_rad")
    solid = st.mask == 1.0
    below = np.zeros_like(solid)
    # solid cells under each column's first fluid cell carry zero heating
    first = np.argmax(~solid, axis=0)
    i = np.arange(st.grid.im)[:, None, None]
    below = i < first[None, :, :]
    assert below.any() and np.all(q[below] == 0.0)
    assert np.count_nonzero(q) > q.size // 2


def test_radiation_fp32_refined_form():
    """The FP32-refined form's radiation (FP32 fields in, FP64 column arithmetic) against FP64 on
    the same state, next to the FP64 radiation of the state rounded to FP32: Q_rad is a flux
    divergence over thin layers, so the rounding of t and p_stat to FP32 alone moves it by about
    1e-5 of its maximum; the form must stay within twice that floor and within 1e-6 on the
    emissivity and the net flux."""
    st = inputs.make_state("full", seed=7)
    rounded = inputs.State(grid=st.grid, mask=st.mask,
                           fields={n: v.astype(np.float32).astype(np.float64) for n, v in st.fields.items()})
    g64, orc, _ = make_pair("full", sweeps=20, state=st)
    g32, orc2, _ = make_pair("full", sweeps=20, state=st, precision=abi.PRECISION_FP32_REF)
    gr, orc3, _ = make_pair("full", sweeps=20, state=rounded)
    for c in (g64, g32, gr):
        c.mixture_density()
        c.radiation()
    for n in OUT:
        e = rel_err(g32.get(n), g64.get(n))
        floor = rel_err(gr.get(n), g64.get(n))
        print(f"FP32-refined {n}: {e:.2e} (FP64 from the rounded state: {floor:.2e})")
        assert e <= max(2.0 * floor, 1.0e-6), n
    for c in (g64, g32, gr, orc, orc2, orc3):
        c.close()


def test_step_with_radiation_against_oracle():
    """Two iterations of the full-physics step with the radiation in the physics block: Q_rad then
    follows the state, and the RHS's radiative heating (rad_coupling 1) reads it."""
    st = inputs.make_state("full", seed=3)
    from atjup_amd.lib import Context
    import oracle
    G, Cn, K, _ = abi.make_profile("full", sweeps=20, grid=st.grid, radiation=True)
    gpu = Context(G, Cn, K, abi.PRECISION_FP64)
    orc = oracle.Oracle(G, Cn, K)
    gpu.load_state(st)
    orc.load_state(st)
    gpu.bc_after_step(); gpu.swap_nlevel()
    orc.call("bc_after_step"); orc.call("swap_nlevel")
    q0 = gpu.get("Q_rad").copy()
    for it in (1, 2):
        gpu.step(it, it == 1)
        orc.call("step", it, 1 if it == 1 else 0)
    w = compare(gpu, orc, abi.PROG + ("p_dyn",) + OUT, 1.0e-10, "2 steps with radiation")
    print("step with radiation", {k: f"{v[0]:.1e}" for k, v in w.items()})
    assert not np.array_equal(gpu.get("Q_rad"), q0)
    gpu.close()
    orc.close()
