"""The physics block's chemistry on the GPU against the oracle (planet/ChemistryJup.h, the sequence
of planet/cJupiterModel.cpp:1146-1161).

The oracle reproduces the compiled reference bit for bit at the knob defaults and with GATE_ZERO 0
and MOLAR_CONC 1, over two calls with the temperature moved in between
(test_oracle_vs_reference.py). The diffusive fluxes and the multicomponent flux assembly have no
transcendental function and agree bit for bit here; the rate law's pow and exp come from the
device library, within an ulp of the host's, and the law subtracts a back reaction from a forward
one, so massflux_* and thermalmassflux are held to twice the oracle's own spread when t and the
species move by one ulp (1e-12 where that spread is smaller). The full-physics step with the
radiation, the precipitation and the chemistry in its physics block is held to the drop-in's 1e-10
on the prognostic fields.
"""
import numpy as np
import pytest

from atjup_amd import abi, inputs
from tests.helpers import compare, rel_err

pytestmark = pytest.mark.gpu

EXACT = ("difflux_h2s", "difflux_nh3", "difflux_nh4sh")
RATED = ("thermalmassflux", "massflux_h2s", "massflux_nh3", "massflux_nh4sh", "fluxlim_nh4sh")


def _norm_rel(a, b):
    nb = np.linalg.norm(b)
    return 0.0 if nb == 0 else float(np.linalg.norm(a - b) / nb)


def _run(ctx_or_orc, is_gpu, st):
    c = ctx_or_orc
    if is_gpu:
        c.mixture_density(); c.chemistry()
        t = st.fields["t"] * 1.03
        c.set("t", t); c.chemistry()
    else:
        c.call("mixture_density"); c.call("chemistry")
        t = st.fields["t"] * 1.03
        c.set("t", t); c.call("chemistry")
    return {n: c.get(n) for n in EXACT + RATED}


def _make(st, gpu):
    from atjup_amd.lib import Context
    import oracle
    G, Cn, K, _ = abi.make_profile("full", sweeps=20, grid=st.grid)
    c = Context(G, Cn, K, abi.PRECISION_FP64) if gpu else oracle.Oracle(G, Cn, K)
    c.load_state(st)
    if gpu:
        c.bc_after_step(); c.swap_nlevel()
    else:
        c.call("bc_after_step"); c.call("swap_nlevel")
    return c


def test_chemistry_against_oracle():
    st = inputs.make_state("full", seed=11)
    rng = np.random.default_rng(6)
    f2 = {n: (np.nextafter(v, rng.choice([-np.inf, np.inf], size=v.shape))
              if n in ("t", "h2s", "nh3", "nh4sh") else v.copy()) for n, v in st.fields.items()}
    noisy_st = inputs.State(grid=st.grid, mask=st.mask, fields=f2)
    o = _make(st, False); ref = _run(o, False, st); o.close()
    o = _make(noisy_st, False); noisy = _run(o, False, noisy_st); o.close()
    g = _make(st, True); got = _run(g, True, st); g.close()
    for n in EXACT:
        assert np.array_equal(got[n], ref[n]), n
    for n in RATED:
        e_max, e_norm = rel_err(got[n], ref[n]), _norm_rel(got[n], ref[n])
        f_max, f_norm = rel_err(noisy[n], ref[n]), _norm_rel(noisy[n], ref[n])
        print(f"{n}: GPU {e_max:.1e} / {e_norm:.1e}, oracle one-ulp spread {f_max:.1e} / {f_norm:.1e}")
        assert e_max <= max(2.0 * f_max, 1.0e-12), n
        assert e_norm <= max(2.0 * f_norm, 1.0e-12), n
    assert np.count_nonzero(got["massflux_nh4sh"]) > 1000


def test_step_with_all_physics_columns_against_oracle():
    """Two iterations of the full-physics step with the radiation, the precipitation and the
    chemistry in the physics block."""
    from atjup_amd.lib import Context
    import oracle
    st = inputs.make_state("full", seed=3)
    G, Cn, K, _ = abi.make_profile("full", sweeps=20, grid=st.grid, radiation=True, precip=True, chemistry=True)
    gpu = Context(G, Cn, K, abi.PRECISION_FP64)
    orc = oracle.Oracle(G, Cn, K)
    gpu.load_state(st); orc.load_state(st)
    gpu.bc_after_step(); gpu.swap_nlevel()
    orc.call("bc_after_step"); orc.call("swap_nlevel")
    for it in (1, 2):
        gpu.step(it, it == 1)
        orc.call("step", it, 1 if it == 1 else 0)
    w = compare(gpu, orc, abi.PROG + ("p_dyn",), 1.0e-10, "2 steps with every physics column")
    print("step with radiation, precipitation, chemistry", {k: f"{v[0]:.1e}" for k, v in w.items()})
    assert np.count_nonzero(gpu.get("massflux_nh4sh")) > 1000
    gpu.close()
    orc.close()


def test_chemistry_fp32_refined_form():
    """FP32 fields in, FP64 arithmetic: against FP64 on the same state, normwise, within twice the
    FP64 path's own change when its inputs are rounded to FP32."""
    from atjup_amd.lib import Context
    st = inputs.make_state("full", seed=11)
    rounded = inputs.State(grid=st.grid, mask=st.mask,
                           fields={n: v.astype(np.float32).astype(np.float64) for n, v in st.fields.items()})
    G, Cn, K, _ = abi.make_profile("full", sweeps=20, grid=st.grid)
    out = {}
    for tag, s_, prec in (("64", st, abi.PRECISION_FP64), ("32", st, abi.PRECISION_FP32_REF),
                          ("r", rounded, abi.PRECISION_FP64)):
        c = Context(G, Cn, K, prec)
        c.load_state(s_); c.bc_after_step(); c.swap_nlevel()
        c.mixture_density(); c.chemistry()
        out[tag] = {n: c.get(n) for n in EXACT + RATED}
        c.close()
    for n in ("difflux_h2s", "massflux_nh3", "massflux_nh4sh", "thermalmassflux"):
        e = _norm_rel(out["32"][n], out["64"][n])
        floor = _norm_rel(out["r"][n], out["64"][n])
        print(f"FP32-refined {n}: {e:.2e} normwise (FP64 from the rounded state: {floor:.2e})")
        assert e <= max(2.0 * floor, 1.0e-6), n
