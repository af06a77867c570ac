"""The NH4SH equilibrium adjustment on the GPU (ChemistryJup::EquilibriumNH4SH("loop"),
planet/ChemistryJup.h:293-370; the reference's default ATJUP_NH4SH_EQ 1) against the oracle, whose
restatement equals the compiled reference bit for bit
(test_oracle_vs_reference.test_nh4sh_equilibrium_bit_exact_against_compiled_reference).

K_p = 10^(14.82 - 4705 / T) comes from the device's pow, within an ulp of the host's, and the
bisection stops on a bracket of 1e-15 of the largest concentration, so the GPU's extent can differ
from the oracle's by that bracket: the bar is the drop-in's 1e-10 of each field's maximum. Cells
that form and that decompose NH4SH are both present in the state."""
import numpy as np
import pytest

from atjup_amd import abi, inputs
from tests.helpers import compare, make_pair, small_state

pytestmark = pytest.mark.gpu

TOL = 1.0e-10


def _state(seed=17, g=None):
    st = inputs.make_state("full", seed=seed) if g is None else small_state("full", seed=seed)
    st.fields["nh4sh"] *= 30.0
    return st


def test_equilibrium_against_oracle():
    gpu, orc, _ = make_pair("full", sweeps=10, state=_state(), chemistry=True, nh4sh_eq=True)
    try:
        gpu.mixture_density()
        orc.call("mixture_density")
        s0 = orc.get("nh4sh")
        gpu.nh4sh_equilibrium()
        orc.call("nh4sh_eq")
        s1 = orc.get("nh4sh")
        formed, decomposed = int(np.count_nonzero(s1 > s0)), int(np.count_nonzero(s1 < s0))
        print("formed", formed, "decomposed", decomposed, flush=True)
        assert formed > 100 and decomposed > 100
        w = compare(gpu, orc, ("t", "nh3", "h2s", "nh4sh"), TOL, "equilibrium")
        print({k: f"{v[0]:.1e}/{v[1]}" for k, v in w.items()})
    finally:
        gpu.close()
        orc.close()


STEP_FIELDS = abi.PROG + ("p_dyn", "massflux_h2s", "massflux_nh3", "massflux_nh4sh", "fluxlim_nh4sh",
                          "difflux_h2s", "thermalmassflux")


def _oracle_steps(st, n):
    import oracle
    G, Cn, K, _ = abi.make_profile("full", sweeps=20, grid=st.grid, radiation=True, precip=True, chemistry=True,
                                   nh4sh_eq=True)
    o = oracle.Oracle(G, Cn, K)
    o.load_state(st)
    o.call("bc_after_step"); o.call("swap_nlevel")
    for it in range(1, n + 1):
        o.call("step", it, 1 if it == 1 else 0)
    out = {f: o.get(f) for f in STEP_FIELDS}
    o.close()
    return out


def test_full_physics_steps_with_equilibrium_and_chemistry():
    """Three iterations with the whole physics block (radiation, adjustments, equilibrium,
    precipitation, closure, chemistry with its rate law off), on the 41 x 181 x 361 shell.

    The equilibrium's "all decomposed" test f(-s) <= 0 reads K_p through pow, so a cell on its
    border ends at exactly zero NH4SH on one side and at a residue of ~1e-16 of its concentration on
    the other; the reference's Superbee limiter is discontinuous where a gradient is zero (its
    denominator df +- 1e-12 changes sign with df, and 1 - superbee(r) with it), so next to such a
    cell the NH4SH flux correction flips, and the smoothing spreads it. The oracle shows the same
    response to one-ulp noise in its inputs, so every field is held to twice the oracle's own
    one-ulp spread after the same steps, max-abs and normwise, and at least to the drop-in's 1e-10."""
    st = _state()
    n = 3
    ref = _oracle_steps(st, n)
    rng = np.random.default_rng(7)
    noisy_fields = {k: (np.nextafter(v, rng.choice([-np.inf, np.inf], size=v.shape))
                        if k in ("t",) + inputs.SPECIES else v.copy()) for k, v in st.fields.items()}
    spread = _oracle_steps(inputs.State(grid=st.grid, mask=st.mask, fields=noisy_fields), n)
    gpu, orc, _ = make_pair("full", sweeps=20, state=st, radiation=True, precip=True, chemistry=True,
                            nh4sh_eq=True)
    orc.close()
    try:
        for it in range(1, n + 1):
            print("step", it, flush=True)
            gpu.step(it, it == 1)
        assert gpu.check_finite()
        for f in STEP_FIELDS:
            a, b, c = gpu.get(f), ref[f], spread[f]
            mb, nb = float(np.max(np.abs(b))), float(np.linalg.norm(b))
            e_max = float(np.max(np.abs(a - b))) / mb if mb > 0 else 0.0
            e_norm = float(np.linalg.norm(a - b)) / nb if nb > 0 else 0.0
            s_max = float(np.max(np.abs(c - b))) / mb if mb > 0 else 0.0
            s_norm = float(np.linalg.norm(c - b)) / nb if nb > 0 else 0.0
            print(f"{f}: GPU {e_max:.1e} / {e_norm:.1e}, oracle one-ulp spread {s_max:.1e} / {s_norm:.1e}")
            assert e_max <= max(2.0 * s_max, TOL), f
            assert e_norm <= max(2.0 * s_norm, TOL), f
    finally:
        gpu.close()


def test_equilibrium_fp32_refined_form():
    """The FP32-refined form with the equilibrium on the small shell: within 1e-5 of FP64 on the
    fields the equilibrium writes (its arithmetic is FP64 on the FP32 inputs)."""
    st = _state(g="small")
    res = {}
    for prec in (abi.PRECISION_FP64, abi.PRECISION_FP32_REF):
        gpu, orc, _ = make_pair("full", sweeps=10, state=st, precision=prec, chemistry=True, nh4sh_eq=True)
        orc.close()
        gpu.mixture_density()
        gpu.nh4sh_equilibrium()
        res[prec] = {n: gpu.get(n) for n in ("t", "nh3", "h2s", "nh4sh")}
        gpu.close()
    for n, b in res[abi.PRECISION_FP64].items():
        a = res[abi.PRECISION_FP32_REF][n]
        e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        print(n, f"{e:.2e}")
        assert e < 1e-5, (n, e)
