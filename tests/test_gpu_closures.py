"""The k-omega SST (the reference's default closure) and k-epsilon closures on the GPU against the
oracle, whose closures equal the compiled reference bit for bit
(test_oracle_vs_reference.test_other_closures_bit_exact_against_compiled_reference).

Turbulence<P>::run (compute_k_omega_SST / compute_k_epsilon, Turbulence.h:886-1106) and the
right-hand side's transport coefficients and sources of each model (RHS_Jup_Turb.cpp:763-998) on the
41 x 181 x 361 moist state with the obstacle, then three iterations of the full-physics step. The
SST blending (tanh of a power) and the k-epsilon damping (exp) come from the device library, so the
bar is the drop-in's 1e-10 of each field's maximum, as for k-omega."""
import numpy as np
import pytest

from atjup_amd import abi, inputs
from tests.helpers import compare, make_pair, small_state

pytestmark = pytest.mark.gpu

TOL = 1.0e-10
TURB_OUT = ("nue", "prod", "tke_source", "dis_source")


@pytest.mark.parametrize("model", ["k_omega_SST", "k_epsilon"])
def test_closure_and_steps_against_oracle(model):
    st = inputs.make_state("full", seed=13)
    gpu, orc, _ = make_pair("full", sweeps=20, state=st, turb=model)
    try:
        gpu.mixture_density()
        orc.call("mixture_density")
        gpu.turbulence_run()
        orc.call("turbulence_run")
        w = compare(gpu, orc, ("tke", "dis") + TURB_OUT, TOL, model + " closure")
        print(model, "closure", {k: f"{v[0]:.1e}/{v[1]}" for k, v in w.items()}, flush=True)
        assert np.count_nonzero(orc.get("prod")) > 1000
        for it in range(1, 4):
            print(model, "step", it, flush=True)
            gpu.step(it, it == 1)
            orc.call("step", it, 1 if it == 1 else 0)
        w = compare(gpu, orc, abi.PROG + ("p_dyn",) + TURB_OUT, TOL, model + " steps")
        print(model, "3 steps", {k: f"{v[0]:.1e}" for k, v in w.items()}, flush=True)
        assert gpu.check_finite()
    finally:
        gpu.close()
        orc.close()


@pytest.mark.parametrize("model", ["k_omega_SST", "k_epsilon"])
def test_closure_fp32_refined_form(model):
    """The FP32-refined form with the other closures on the small shell: finite, and within the
    FP32 form's usual distance of the FP64 path after three steps on the dynamics fields."""
    st = small_state("full")
    res = {}
    for prec in (abi.PRECISION_FP64, abi.PRECISION_FP32_REF):
        gpu, orc, _ = make_pair("full", sweeps=10, state=st, precision=prec, turb=model)
        orc.close()
        for it in range(1, 4):
            gpu.step(it, it == 1)
        assert gpu.check_finite()
        res[prec] = {n: gpu.get(n) for n in ("t", "u", "v", "w", "tke", "dis")}
        gpu.close()
    for n in ("t", "u", "v", "w"):
        a, b = res[abi.PRECISION_FP32_REF][n], res[abi.PRECISION_FP64][n]
        e = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        print(model, n, f"{e:.2e}")
        assert e < 1e-5, (model, n, e)
