"""The hand-written adjoint of the FP64 dynamics step against the tangent-linear step.

The tangent form (PRECISION_TANGENT) runs the same forward kernels on (value, derivative) pairs,
so it is the exact forward-mode derivative of the drop-in; its values must equal the FP64 path's
bit for bit, and its derivatives must agree with central finite differences of the FP64 path.
The adjoint of every operator is then held to the dot-product identity <J dx, ly> = <dx, J^T ly>
against it, operator by operator and for whole loop iterations with and without the physics
block, on a small shell with an obstacle; and the gradient of the zonal-jet misfit through the
checkpointed reverse sweep is held to central differences along seeded directions.
"""
import numpy as np
import pytest

from atjup_amd import abi, inputs
from tests.helpers import small_state

pytestmark = pytest.mark.gpu

STATE = ("t", "u", "v", "w", "p_dyn", "aux_u", "aux_v", "aux_w", "rho_mix")


def _state(seed=3, g=None):
    st = small_state("dyn", seed=seed, g=g)
    rng = np.random.default_rng(seed)
    g = st.grid
    # a live pressure field and aux_* so that every term of the step has a derivative
    st.fields["p_dyn"] = 0.01 * rng.standard_normal(g.shape)
    for n in ("aux_u", "aux_v", "aux_w"):
        st.fields[n] = 0.01 * rng.standard_normal(g.shape)
    for n in ("u", "v"):
        st.fields[n] = st.fields[n] + 0.001 * rng.standard_normal(g.shape)
    for n in st.fields:
        st.fields[n][:, :, -1] = st.fields[n][:, :, 0]
    return st


def _ctx(st, precision, sweeps=6):
    from atjup_amd.lib import Context
    G, Cn, K, _ = abi.make_profile("dyn", sweeps=sweeps, grid=st.grid)
    c = Context(G, Cn, K, precision)
    c.load_state(st, skip=set(inputs.SPECIES))
    c.bc_after_step()
    c.swap_nlevel()
    c.mixture_density()
    return c


def _snapshot(c):
    return {n: c.get(n) for n in STATE}


def _load(c, snap):
    for n, v in snap.items():
        c.set(n, v)


def test_tangent_values_are_the_fp64_step_and_derivatives_match_finite_differences():
    """Two loop iterations (the second one's solve sees the first one's aux_*), every prognostic
    field perturbed along a seeded direction scaled to the field."""
    st = _state()
    c64 = _ctx(st, abi.PRECISION_FP64)
    ctg = _ctx(st, abi.PRECISION_TANGENT)
    base = _snapshot(c64)
    rng = np.random.default_rng(1)
    dirs = {}
    for n in ("t", "u", "v", "w"):
        d = rng.standard_normal(st.grid.shape) * np.abs(base[n]).max()
        d[:, :, -1] = d[:, :, 0]
        dirs[n] = d
        ctg.set_tangent(n, d)
    for it in (1, 2):
        c64.step(it, it == 1)
        ctg.step(it, it == 1)
    outs_n = ("t", "u", "v", "w", "p_dyn")
    for n in outs_n:
        assert np.array_equal(ctg.get(n), c64.get(n)), n
    eps = 1e-6
    outs = {}
    for sgn in (+1, -1):
        _load(c64, base)
        for n, d in dirs.items():
            c64.set(n, base[n] + sgn * eps * d)
        for it in (1, 2):
            c64.step(it, it == 1)
        outs[sgn] = {n: c64.get(n) for n in outs_n}
    for n in outs_n:
        fd = (outs[1][n] - outs[-1][n]) / (2 * eps)
        tg = ctg.get_tangent(n)
        # the rounding of the two forward values bounds what the difference quotient resolves
        noise = 4e-16 * np.max(np.abs(outs[1][n])) / (2 * eps)
        scale = np.max(np.abs(tg))
        err = np.max(np.abs(tg - fd)) / scale
        print(f"tangent vs FD {n}: {err:.2e} (derivative {scale:.1e}, difference-quotient noise {noise:.1e})")
        assert scale > 1e3 * noise, n
        assert err < 1e-5, n
    c64.close()
    ctg.close()


CASES = [
    # op, iarg, darg, inputs perturbed, outputs weighted
    ("bc", 0, 0.0, ("t", "u", "v", "w", "p_dyn"), ("t", "u", "v", "w", "p_dyn")),
    ("shapiro_pdyn", 0, 0.0, ("p_dyn",), ("p_dyn",)),
    ("mixture", 0, 0.0, ("t", "p_dyn"), ("rho_mix",)),
    ("rk4", 0, 0.0, ("t", "u", "v", "w", "p_dyn", "rho_mix", "aux_u", "aux_v", "aux_w"),
     ("t", "u", "v", "w", "aux_u", "aux_v", "aux_w")),
    ("pressure", 6, 1.0, ("p_dyn", "aux_u", "aux_v", "aux_w"), ("p_dyn", "aux_u", "aux_v", "aux_w")),
    ("pressure", 4, 1.5, ("p_dyn", "aux_u", "aux_v", "aux_w"), ("p_dyn", "aux_u", "aux_v", "aux_w")),
    ("step", 2, 0.0, STATE, STATE),
    ("step", 3, 0.0, STATE, STATE),
]


def _forward(c, op, iarg, darg):
    if op == "bc":
        c.bc_after_step()
    elif op == "shapiro_pdyn":
        c.shapiro("p_dyn", 2, 1.0, 1, 7)
    elif op == "mixture":
        c.mixture_density()
    elif op == "rk4":
        c.rk4_step()
    elif op == "pressure":
        c.pressure_solve(iarg, darg)
    elif op == "step":
        c.step(iarg, bool(darg))


# The transposed colour pass covers a row two cells per thread, flattened over (row, pair), on the
# 48-meridian shell above (12 pairs per row); one cell per thread where row-aligned pairs would
# fill the wavefronts (256 meridians: 64 pairs) and on an odd count (50 meridians: 25 cells).
OTHER_RINGS = [(km, op) for km in (51, 257) for op in (
    ("pressure", 6, 1.0, ("p_dyn", "aux_u", "aux_v", "aux_w"), ("p_dyn", "aux_u", "aux_v", "aux_w")),
    ("step", 2, 0.0, STATE, STATE))]


def _ring_grid(km):
    return inputs.Grid(im=13, jm=33, km=km, the_degree=180.0 / 32, phi_degree=360.0 / (km - 1))


@pytest.mark.parametrize("op,iarg,darg,ins,outs", CASES, ids=[f"{c[0]}-{c[1]}-{c[2]}" for c in CASES])
def test_dot_product(op, iarg, darg, ins, outs):
    _dot_product(op, iarg, darg, ins, outs, None)


@pytest.mark.parametrize("km,case", OTHER_RINGS, ids=[f"km{k}-{c[0]}-{c[1]}" for k, c in OTHER_RINGS])
def test_dot_product_other_rings(km, case):
    _dot_product(*case, _ring_grid(km))


def _dot_product(op, iarg, darg, ins, outs, g):
    st = _state(seed=5, g=g)
    rng = np.random.default_rng(7)
    shp = st.grid.shape
    c64 = _ctx(st, abi.PRECISION_FP64)
    ctg = _ctx(st, abi.PRECISION_TANGENT)
    base = _snapshot(c64)
    dx = {n: rng.standard_normal(shp) * (1e-3 + np.abs(base[n]).max() * 1e-3) for n in ins}
    for n, v in dx.items():
        ctg.set_tangent(n, v)
    _forward(ctg, op, iarg, darg)
    ly = {n: rng.standard_normal(shp) for n in outs}
    lhs = sum(float(np.sum(ctg.get_tangent(n) * ly[n])) for n in outs)
    # the adjoint, linearised about the same state
    c64.adj_init()
    c64.adj_zero()
    for n, v in ly.items():
        c64.adj_set(n, v)
    c64.adj_op(op, iarg, darg)
    rhs = sum(float(np.sum(dx[n] * c64.adj_get(n))) for n in ins)
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs), 1e-300)
    print(f"dot product {op}({iarg}, {darg}): <J dx, ly> = {lhs:.15e}, <dx, J^T ly> = {rhs:.15e}, rel {rel:.2e}")
    assert rel < 1e-10
    c64.close()
    ctg.close()


def _loss_and_grad_final(c, target, cos_lat, i_top):
    """Zonal-jet misfit at the top level: sum_j cos(lat_j) (mean_k w[i_top, j, :-1] - target_j)^2."""
    w = c.get("w")
    km = w.shape[2]
    zm = w[i_top, :, :-1].mean(axis=1)
    r = zm - target
    loss = float(np.sum(cos_lat * r * r))
    lam = np.zeros((4,) + w.shape)
    lam[3, i_top, :, :-1] = (2.0 * cos_lat * r / (km - 1))[:, None]
    return loss, lam


def test_gradient_of_the_jet_misfit_against_central_differences():
    st = _state(seed=9)
    g = st.grid
    _, theta, _ = g.coords()
    cos_lat = np.abs(np.sin(theta))                 # theta is the colatitude
    rng = np.random.default_rng(11)
    target = st.fields["w"][g.im - 2, :, 0] + 0.01 * rng.standard_normal(g.jm)
    i_top = g.im - 2
    nsteps, every = 6, 2
    c = _ctx(st, abi.PRECISION_FP64)
    init = _snapshot(c)
    c.adj_init()
    c.adj_forward(nsteps, every, 1)
    loss, lam = _loss_and_grad_final(c, target, cos_lat, i_top)
    grad = c.adj_backward(lam)[3]                   # dL / dw0
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
        fd = (vals[0] - vals[1]) / (2 * eps)
        ad = float(np.sum(grad * d))
        rel = abs(fd - ad) / max(abs(fd), abs(ad), 1e-300)
        print(f"direction {d_seed}: FD {fd:.10e}  adjoint {ad:.10e}  rel {rel:.2e}")
        assert rel < 1e-4
    c.close()


def test_jet_misfit_gradient_on_the_full_shell():
    """The 20-step misfit gradient on the 41 x 181 x 361 shell with checkpoints every 5 steps,
    against central differences with eps = 1e-6 along two seeded directions, through the
    torch.autograd.Function wrapper."""
    torch = pytest.importorskip("torch")
    from atjup_amd.adjoint import JetMisfit, MisfitRunner
    from atjup_amd.lib import Context
    g = inputs.Grid()
    st = inputs.make_state("dyn", seed=11, g=g)
    G, Cn, K, _ = abi.make_profile("dyn", sweeps=200, grid=g)
    c = Context(G, Cn, K, abi.PRECISION_FP64)
    c.load_state(st, skip=set(inputs.SPECIES))
    c.bc_after_step()
    c.swap_nlevel()
    c.mixture_density()
    init = _snapshot(c)
    target = inputs.zonal_jets(g, np.random.default_rng(12)) / 100.0
    run = MisfitRunner(c, init, target, nsteps=20, every=5)
    w0 = torch.from_numpy(init["w"].copy()).requires_grad_(True)
    loss = JetMisfit.apply(w0, run)
    loss.backward()
    grad = w0.grad.numpy()
    assert np.isfinite(grad).all() and np.abs(grad).max() > 0
    eps = 1.0e-6
    for d_seed in range(2):
        d = np.random.default_rng(1000 + d_seed).standard_normal(g.shape)
        d[:, :, -1] = d[:, :, 0]
        fd = (run.evaluate(init["w"] + eps * d) - run.evaluate(init["w"] - eps * d)) / (2 * eps)
        ad = float(np.sum(grad * d))
        rel = abs(fd - ad) / max(abs(fd), abs(ad), 1e-300)
        print(f"full shell, direction {d_seed}: loss {float(loss):.10e} FD {fd:.10e} adjoint {ad:.10e} rel {rel:.2e}")
        assert rel < 1e-4
    c.close()


def test_semi_axis_gradient_of_the_smooth_drag_region():
    """dL/d(a_phi_deg, b_the_deg, c_r_km) of the jet misfit with the smooth drag region in place of
    the binary obstacle (all-fluid mask), through JetMisfit with the axes as a tensor, against
    central differences of the same forward; dL/dw0 of that forward against central differences
    along a seeded direction. The drag region has no counterpart in the reference (its obstacle
    is binary); the forward checked here is the one the library computes with the region set."""
    torch = pytest.importorskip("torch")
    from atjup_amd.adjoint import JetMisfit, MisfitRunner
    st = _state(seed=9)
    g = st.grid
    st.mask = np.zeros(g.shape)
    c = _ctx(st, abi.PRECISION_FP64)
    init = _snapshot(c)
    rng = np.random.default_rng(11)
    target = st.fields["w"][g.im - 2, :, 0] + 0.01 * rng.standard_normal(g.jm)
    axes0 = np.array([30.0, 17.0, 60.0])
    obst = abi.obstacle(g, *axes0, width=0.25, tau=0.05, j0=20, k0=24)
    # the misfit at level 3, inside the drag region's height (c = 60 km is five levels here), where
    # the loss responds to the semi-axes well above the rounding of its difference quotients
    target = st.fields["w"][3, :, 0] + 0.01 * rng.standard_normal(g.jm)
    run = MisfitRunner(c, init, target, nsteps=6, every=2, obstacle=obst, i_top=3)
    w0 = torch.from_numpy(init["w"].copy()).requires_grad_(True)
    axes = torch.from_numpy(axes0.copy()).requires_grad_(True)
    loss = JetMisfit.apply(w0, run, axes)
    loss.backward()
    ga, gw = axes.grad.numpy(), w0.grad.numpy()
    assert np.all(np.isfinite(ga)) and np.abs(ga).max() > 0
    for p in range(3):
        eps = 1.0e-4 * axes0[p]
        vals = []
        for sgn in (+1, -1):
            ax = axes0.copy()
            ax[p] += sgn * eps
            run.set_axes(ax)
            vals.append(run.evaluate(init["w"]))
        fd = (vals[0] - vals[1]) / (2 * eps)
        rel = abs(fd - ga[p]) / max(abs(fd), abs(ga[p]), 1e-300)
        print(f"semi-axis {p}: FD {fd:.10e} adjoint {ga[p]:.10e} rel {rel:.2e}", flush=True)
        assert rel < 1e-4
    run.set_axes(axes0)
    d = np.random.default_rng(100).standard_normal(g.shape)
    d[:, :, -1] = d[:, :, 0]
    eps = 1e-7
    fd = (run.evaluate(init["w"] + eps * d) - run.evaluate(init["w"] - eps * d)) / (2 * eps)
    ad = float(np.sum(gw * d))
    rel = abs(fd - ad) / max(abs(fd), abs(ad), 1e-300)
    print(f"w0 direction: FD {fd:.10e} adjoint {ad:.10e} rel {rel:.2e}")
    assert rel < 1e-4
    c.close()
