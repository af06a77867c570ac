"""The moist convective adjustment on the GPU against the oracle (MoistConvectiveAdjustment<P>::run,
planet/ConvectiveAdjustment.h:403-849, with the gases and tracers of planet/cJupiterModel.cpp:1301-1322).

The oracle reproduces the compiled reference bit for bit at the knob defaults, with LEDOUX 0, with
LOADING 0 and RELAX 0.5, and with a two-sweep cap, over two calls on a state with a warm lower band
and a cold layer aloft and over two steps with the adjustment in the loop
(test_oracle_vs_reference.py). The same cases run here through the C-ABI: the first call mixes
nearly every column of the seeded state, the second works on the adjusted one. The equilibrium
solve evaluates the vapour pressure through the device library's exp and log (within an ulp of the
host's), and its regula falsi stops on a 1e-10 K bracket, so temperatures agree to that bracket and
the mixing ratios to the same relative order, not bit for bit.
"""
import numpy as np
import pytest

from atjup_amd import abi, inputs
from tests.helpers import make_pair, rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("t", "h2o", "h2o_cloud", "h2o_ice", "nh3", "nh3_cloud", "nh3_ice", "h2s", "nh4sh")
TOL = 1e-9          # max |gpu - oracle| / max |oracle| per field


def _state():
    st = inputs.make_state("full", seed=11)
    t = st.fields["t"].copy()
    t[2:7, 60:100, 100:200] *= 1.25
    t[20:24, 100:140, 200:300] *= 0.8
    st.fields["t"] = t
    return st


def _columns_apart(a, b, tol):
    s = max(np.max(np.abs(b)), 1e-300)
    return int((np.abs(a - b) > tol * s).any(axis=0).sum())


@pytest.mark.parametrize("over", [{}, {"ledoux": 0}, {"loading": 0, "relax": 0.5}, {"max_pass": 2}],
                         ids=["defaults", "no-ledoux", "no-loading-relax", "two-sweeps"])
def test_mca_against_oracle(over):
    gpu, orc, st = make_pair("full", sweeps=20, state=_state())
    gpu.set_mca(abi.mca_params(**over))
    orc.set_mca(abi.mca_params(**over))
    gpu.mixture_density(); orc.call("mixture_density")
    for call in (1, 2):
        before = orc.get("t").copy()
        gpu.mca(); orc.call("mca")
        changed = int((orc.get("t") != before).any(axis=0).sum())
        worst = {}
        for n in FIELDS:
            a, b = gpu.get(n), orc.get(n)
            worst[n] = (rel_err(a, b), _columns_apart(a, b, TOL))
        print(f"mca {over} call {call}: {changed} columns adjusted;",
              {n: f"{e:.1e}/{c}" for n, (e, c) in worst.items()})
        if call == 1:
            assert changed > 1000
        for n, (e, c) in worst.items():
            assert e <= TOL, (call, n, e, c)
    gpu.close(); orc.close()


def test_step_with_mca_against_oracle():
    st = inputs.make_state("full", seed=11)
    G, Cn, K, _ = abi.make_profile("full", sweeps=20, grid=st.grid, mca=1)
    from atjup_amd.lib import Context
    import oracle
    gpu = Context(G, Cn, K, abi.PRECISION_FP64)
    orc = oracle.Oracle(G, Cn, K)
    gpu.load_state(st); orc.load_state(st)
    gpu.bc_after_step(); gpu.swap_nlevel()
    orc.call("bc_after_step"); orc.call("swap_nlevel")
    for it in range(1, 4):
        gpu.step(it, it == 1); orc.call("step", it, 1 if it == 1 else 0)
    worst = {n: rel_err(gpu.get(n), orc.get(n)) for n in list(abi.PROG) + ["p_dyn"]}
    print("step with mca", {n: f"{e:.1e}" for n, e in worst.items()})
    for n, e in worst.items():
        assert e <= 1e-10, (n, e)
    gpu.close(); orc.close()


def test_mca_fp32_refined_form():
    """The FP32-refined form's adjustment (FP32 species and pressure, t as hi + lo, FP64 column
    arithmetic) against the FP64 form on the same state, next to the FP64 form on the state
    rounded to FP32: the parcel test compares virtual temperatures to 1e-7 K, so the rounding of
    the inputs alone moves some decisions and the mixed columns with them; the form must stay
    within twice that floor (and within 1e-6 where the floor is smaller)."""
    st = _state()
    rounded = inputs.State(grid=st.grid, mask=st.mask,
                           fields={n: v.astype(np.float32).astype(np.float64) for n, v in st.fields.items()})
    g64, o1, _ = make_pair("full", sweeps=20, state=st)
    g32, o2, _ = make_pair("full", sweeps=20, state=st, precision=abi.PRECISION_FP32_REF)
    gr, o3, _ = make_pair("full", sweeps=20, state=rounded)
    for c in (g64, g32, gr):
        c.mixture_density()
        c.mca()
    for n in FIELDS:
        e = rel_err(g32.get(n), g64.get(n))
        floor = rel_err(gr.get(n), g64.get(n))
        print(f"mca FP32-refined {n}: {e:.2e} (FP64 from the rounded state: {floor:.2e})")
        assert e <= max(2.0 * floor, 1.0e-6), n
    for c in (g64, g32, gr, o1, o2, o3):
        c.close()


def test_mca_deep_columns_against_oracle():
    """81 levels with a strongly unstable lower part: segments longer than a wavefront's 64 lanes,
    so a lane settles two cells of a segment."""
    from tests.helpers import small_state
    g = inputs.Grid(im=81, jm=33, km=49, the_degree=180.0 / 32, phi_degree=360.0 / 48)
    st = small_state("full", seed=5, g=g)
    t = st.fields["t"].copy()
    t[1:60] *= np.linspace(1.6, 1.0, 59)[:, None, None]
    st.fields["t"] = t
    gpu, orc, _ = make_pair("full", sweeps=20, state=st)
    gpu.mixture_density(); orc.call("mixture_density")
    before = orc.get("t").copy()
    gpu.mca(); orc.call("mca")
    cells = (orc.get("t") != before).sum(axis=0)
    assert (cells > 64).sum() > 100
    worst = {n: rel_err(gpu.get(n), orc.get(n)) for n in FIELDS}
    print("mca deep columns", int(cells.max()), "cells at most;", {n: f"{e:.1e}" for n, e in worst.items()})
    for n, e in worst.items():
        assert e <= TOL, (n, e)
    gpu.close(); orc.close()
