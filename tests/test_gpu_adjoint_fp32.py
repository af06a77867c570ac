"""The adjoint of the FP32-refined dynamics step against the FP64 one.

The FP32-refined form's adjoint keeps its fields in FP32 and transposes what the refined forward
runs: the stencils, boundary passes and RK4 stages, and the pressure solve as its first FP64
sweep followed by rounds of FP32 correction sweeps inside the FP64 refinement (the rounds' sweeps
transposed in FP32, their residuals and corrections in FP64). Linearised about the same state
(FP64 fields rounded to FP32, so that both forms start from equal values), each transposed
operator must give the FP64 adjoint's result to FP32 resolution, and the dot-product identity
must hold against the FP64 tangent to that resolution. The misfit gradient through the
checkpointed reverse sweep must agree with the FP64 adjoint's gradient and with central
differences of the FP64 forward (the FP64 reference gradient).

An FP32 result carries errors relative to the size of the whole vector, not to each of its
projections: a dot product of random vectors can cancel to a small fraction of the norms, and a
random direction can be nearly orthogonal to the gradient. The dot-product identity is therefore
held relative to the norms (|<J dx, ly> - <dx, J^T ly>| against the sum of ||dx_f|| ||J^T ly_f||),
and a directional derivative relative to the root mean square of the directional derivatives
tested (the gradient's scale along a random direction).
"""
import numpy as np
import pytest

from atjup_amd import abi, inputs
from tests.test_gpu_adjoint import CASES, STATE, _ctx, _forward, _load, _loss_and_grad_final, _snapshot, _state

pytestmark = pytest.mark.gpu

# the refined solve with several rounds of FP32 correction sweeps (rounds of 50: 50, 50, 19)
CASES32 = CASES + [("pressure", 120, 1.0, ("p_dyn", "aux_u", "aux_v", "aux_w"), ("p_dyn", "aux_u", "aux_v", "aux_w")),
                   ("pressure", 120, 1.5, ("p_dyn", "aux_u", "aux_v", "aux_w"), ("p_dyn", "aux_u", "aux_v", "aux_w"))]


def _rounded(st):
    out = inputs.State(grid=st.grid, mask=st.mask,
                       fields={n: v.astype(np.float32).astype(np.float64) for n, v in st.fields.items()})
    return out


def _snap_rounded(c):
    return {n: c.get(n).astype(np.float32).astype(np.float64) for n in STATE}


@pytest.mark.parametrize("op,iarg,darg,ins,outs", CASES32, ids=[f"{c[0]}-{c[1]}-{c[2]}" for c in CASES32])
def test_fp32_operator_adjoint_against_fp64(op, iarg, darg, ins, outs):
    st = _rounded(_state(seed=5))
    shp = st.grid.shape
    sweeps = iarg if op == "pressure" else 6
    c32 = _ctx(st, abi.PRECISION_FP32_REF, sweeps=sweeps)
    c64 = _ctx(st, abi.PRECISION_FP64, sweeps=sweeps)
    ctg = _ctx(st, abi.PRECISION_TANGENT, sweeps=sweeps)
    # the FP32 context's state (its set-up ran in FP32), in all three forms
    base = _snap_rounded(c32)
    for c in (c32, c64, ctg):
        _load(c, base)
    rng = np.random.default_rng(7)
    dx = {n: rng.standard_normal(shp) * (1e-3 + np.abs(base[n]).max() * 1e-3) for n in ins}
    for n, v in dx.items():
        ctg.set_tangent(n, v)
    _forward(ctg, op, iarg, darg)
    ly = {n: rng.standard_normal(shp) for n in outs}
    lhs = sum(float(np.sum(ctg.get_tangent(n) * ly[n])) for n in outs)
    res = {}
    for name, c in (("fp32", c32), ("fp64", c64)):
        c.adj_init()
        c.adj_zero()
        for n, v in ly.items():
            c.adj_set(n, v.astype(np.float32).astype(np.float64))
        c.adj_op(op, iarg, darg)
        res[name] = {n: c.adj_get(n) for n in ins}
    rhs32 = sum(float(np.sum(dx[n] * res["fp32"][n])) for n in ins)
    norms = sum(float(np.linalg.norm(dx[n]) * np.linalg.norm(res["fp64"][n])) for n in ins)
    rel_dot = abs(lhs - rhs32) / max(norms, 1e-300)
    worst = 0.0
    for n in ins:
        a, b = res["fp32"][n], res["fp64"][n]
        nb = float(np.linalg.norm(b))
        if nb == 0.0:
            assert float(np.linalg.norm(a)) == 0.0, n
            continue
        worst = max(worst, float(np.linalg.norm(a - b)) / nb)
    print(f"FP32 adjoint {op}({iarg}, {darg}): dot product against the FP64 tangent, relative to the norms, {rel_dot:.2e}, "
          f"worst normwise difference from the FP64 adjoint {worst:.2e}")
    assert rel_dot < 1e-5
    assert worst < 2e-5
    for c in (c32, c64, ctg):
        c.close()


def _misfit_case(seed=9):
    st = _rounded(_state(seed=seed))
    g = st.grid
    _, theta, _ = g.coords()
    cos_lat = np.abs(np.sin(theta))
    rng = np.random.default_rng(11)
    target = st.fields["w"][g.im - 2, :, 0] + 0.01 * rng.standard_normal(g.jm)
    return st, g, cos_lat, target


def test_fp32_misfit_gradient_against_fp64_and_central_differences():
    """6 steps with checkpoints every 2 on the small shell, 120-sweep solves (three refinement
    rounds): the FP32 gradient against the FP64 adjoint's and against central differences of the
    FP64 forward along seeded directions."""
    st, g, cos_lat, target = _misfit_case()
    i_top = g.im - 2
    nsteps, every = 6, 2
    grads = {}
    init = None
    for prec in (abi.PRECISION_FP32_REF, abi.PRECISION_FP64):
        c = _ctx(st, prec, sweeps=120)
        if init is None:
            init = _snap_rounded(c)
        _load(c, init)
        c.adj_init()
        c.adj_forward(nsteps, every, 1)
        loss, lam = _loss_and_grad_final(c, target, cos_lat, i_top)
        grads[prec] = c.adj_backward(lam)[3]
        c.close()
    g32, g64 = grads[abi.PRECISION_FP32_REF], grads[abi.PRECISION_FP64]
    rel = float(np.linalg.norm(g32 - g64)) / float(np.linalg.norm(g64))
    print(f"FP32 gradient against the FP64 gradient: normwise {rel:.2e}")
    assert rel < 1e-5
    c = _ctx(st, abi.PRECISION_FP64, sweeps=120)
    fds, ads = [], []
    for d_seed in range(4):
        d = np.random.default_rng(100 + d_seed).standard_normal(g.shape)
        d[:, :, -1] = d[:, :, 0]
        eps = 1e-7
        vals = []
        for sgn in (+1, -1):
            _load(c, init)
            c.set("w", init["w"] + sgn * eps * d)
            for it in range(1, nsteps + 1):
                c.step(it, it == 1)
            vals.append(_loss_and_grad_final(c, target, cos_lat, i_top)[0])
        fds.append((vals[0] - vals[1]) / (2 * eps))
        ads.append(float(np.sum(g32 * d)))
    scale = float(np.sqrt(np.mean(np.square(fds))))
    for d_seed, (fd, ad) in enumerate(zip(fds, ads)):
        rel = abs(fd - ad) / max(abs(fd), abs(ad), 1e-300)
        print(f"direction {d_seed}: FD (FP64 forward) {fd:.10e}  FP32 adjoint {ad:.10e}  rel {rel:.2e}, "
              f"relative to the rms {abs(fd - ad) / scale:.2e}")
        assert abs(fd - ad) < 1e-4 * max(abs(fd), scale)
    c.close()


def test_fp32_misfit_gradient_on_the_full_shell():
    """The 20-step misfit on the 41 x 181 x 361 shell with checkpoints every 5 and 200-sweep solves,
    through JetMisfit on an FP32-refined context: its gradient against the FP64 adjoint's and
    against central differences of the FP64 forward along a seeded direction."""
    torch = pytest.importorskip("torch")
    from atjup_amd.adjoint import JetMisfit, MisfitRunner
    from atjup_amd.lib import Context
    g = inputs.Grid()
    st = _rounded(inputs.make_state("dyn", seed=11, g=g))
    G, Cn, K, _ = abi.make_profile("dyn", sweeps=200, grid=g)
    target = inputs.zonal_jets(g, np.random.default_rng(12)) / 100.0
    grads, runs, ctxs = {}, {}, {}
    init = None
    for prec in (abi.PRECISION_FP32_REF, abi.PRECISION_FP64):
        c = Context(G, Cn, K, prec)
        c.load_state(st, skip=set(inputs.SPECIES))
        c.bc_after_step()
        c.swap_nlevel()
        c.mixture_density()
        if init is None:
            init = _snap_rounded(c)
        _load(c, init)
        run = MisfitRunner(c, init, target, nsteps=20, every=5)
        w0 = torch.from_numpy(init["w"].copy()).requires_grad_(True)
        loss = JetMisfit.apply(w0, run)
        loss.backward()
        grads[prec] = w0.grad.numpy().copy()
        runs[prec], ctxs[prec] = run, c
        print(f"precision {prec}: loss {float(loss):.10e}")
    g32, g64 = grads[abi.PRECISION_FP32_REF], grads[abi.PRECISION_FP64]
    assert np.isfinite(g32).all() and np.abs(g32).max() > 0
    rel = float(np.linalg.norm(g32 - g64)) / float(np.linalg.norm(g64))
    print(f"full shell: FP32 gradient against the FP64 gradient, normwise {rel:.2e}")
    assert rel < 1e-5
    eps = 1.0e-6
    d = np.random.default_rng(1000).standard_normal(g.shape)
    d[:, :, -1] = d[:, :, 0]
    r64 = runs[abi.PRECISION_FP64]
    fd = (r64.evaluate(init["w"] + eps * d) - r64.evaluate(init["w"] - eps * d)) / (2 * eps)
    ad = float(np.sum(g32 * d))
    rel = abs(fd - ad) / max(abs(fd), abs(ad), 1e-300)
    print(f"full shell: FD (FP64 forward) {fd:.10e} FP32 adjoint {ad:.10e} rel {rel:.2e}")
    assert rel < 1e-4
    for c in ctxs.values():
        c.close()


def test_fp32_adjoint_on_the_fine_shell():
    """The 81 x 361 x 721 shell: the misfit after 50 steps with checkpoints every 10 and 200-sweep
    solves, the FP32-refined forward and its FP32 adjoint against the FP64 adjoint's gradient and
    against a central difference of the FP64 forward along a seeded direction."""
    from atjup_amd.adjoint import MisfitRunner
    from atjup_amd.lib import Context
    g = inputs.fine_grid()
    st = _rounded(inputs.make_state("dyn", seed=11, g=g))
    G, Cn, K, _ = abi.make_profile("dyn", sweeps=200, grid=g)
    target = inputs.zonal_jets(g, np.random.default_rng(12)) / 100.0
    grads, runs, ctxs = {}, {}, {}
    init = None
    for prec in (abi.PRECISION_FP32_REF, abi.PRECISION_FP64):
        c = Context(G, Cn, K, prec)
        c.load_state(st, skip=set(inputs.SPECIES))
        c.bc_after_step()
        c.swap_nlevel()
        c.mixture_density()
        if init is None:
            init = _snap_rounded(c)
        _load(c, init)
        run = MisfitRunner(c, init, target, nsteps=50, every=10)
        loss = run.forward(init["w"])
        grads[prec] = run.backward()
        runs[prec], ctxs[prec] = run, c
        print(f"fine shell, precision {prec}: loss {loss:.10e}", flush=True)
    g32, g64 = grads[abi.PRECISION_FP32_REF], grads[abi.PRECISION_FP64]
    assert np.isfinite(g32).all() and np.abs(g32).max() > 0
    rel = float(np.linalg.norm(g32 - g64)) / float(np.linalg.norm(g64))
    print(f"fine shell: FP32 gradient against the FP64 gradient, normwise {rel:.2e}")
    assert rel < 1e-5
    ctxs[abi.PRECISION_FP32_REF].close()
    eps = 1.0e-6
    d = np.random.default_rng(2000).standard_normal(g.shape)
    d[:, :, -1] = d[:, :, 0]
    r64 = runs[abi.PRECISION_FP64]
    fd = (r64.evaluate(init["w"] + eps * d) - r64.evaluate(init["w"] - eps * d)) / (2 * eps)
    ad = float(np.sum(g32 * d))
    rel = abs(fd - ad) / max(abs(fd), abs(ad), 1e-300)
    print(f"fine shell: FD (FP64 forward) {fd:.10e} FP32 adjoint {ad:.10e} rel {rel:.2e}")
    assert rel < 1e-4
    ctxs[abi.PRECISION_FP64].close()
