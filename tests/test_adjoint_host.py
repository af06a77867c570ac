"""The host side of the misfit gradient (atjup_amd/adjoint.py) on a stand-in context whose "step"
is a known linear map, so that the loss, the seed of the reverse sweep, the gradient through
torch.autograd and the refusal of a stale backward are checked without a GPU."""
import numpy as np
import pytest

from atjup_amd import inputs
from atjup_amd.adjoint import MisfitRunner

torch = pytest.importorskip("torch")


class LinearContext:
    """w_N = a^N w_0 per cell (each step scales w by a); the reverse sweep multiplies by a^N."""

    def __init__(self, g, a=0.9):
        self.grid, self.shape, self.a = g, g.shape, a
        self.f = {}
        self.ckpt = None

    def adj_init(self):
        pass

    def set(self, n, v):
        self.f[n] = np.array(v, dtype=np.float64)

    def get(self, n):
        return self.f[n].copy()

    def step(self, it, first):
        self.f["w"] = self.a * self.f["w"]
        self.ckpt = None

    def adj_forward(self, nsteps, every, first_iter):
        self.ckpt = (nsteps, self.f["w"].copy())
        for _ in range(nsteps):
            self.f["w"] = self.a * self.f["w"]

    def adj_backward(self, lam):
        nsteps, _ = self.ckpt
        out = np.array(lam) * self.a ** nsteps
        self.ckpt = None
        return out


def _runner(g):
    rng = np.random.default_rng(4)
    w = rng.standard_normal(g.shape)
    ctx = LinearContext(g)
    target = rng.standard_normal(g.jm)
    return MisfitRunner(ctx, {"w": w, "t": np.ones(g.shape)}, target, nsteps=3, every=1), w


def test_gradient_through_autograd_equals_the_closed_form():
    from atjup_amd.adjoint import JetMisfit
    g = inputs.Grid(im=5, jm=7, km=9, the_degree=30.0, phi_degree=45.0)
    run, w = _runner(g)
    w0 = torch.from_numpy(w.copy()).requires_grad_(True)
    loss = JetMisfit.apply(w0, run)
    loss.backward()
    a3 = 0.9 ** 3
    the = np.cumsum(np.r_[0.0, np.full(g.jm - 1, g.dthe)])
    c = np.abs(np.sin(the))
    r = a3 * w[g.im - 2, :, :-1].mean(axis=1) - run.target
    assert float(loss.detach()) == pytest.approx(float(np.sum(c * r * r)), rel=1e-14)
    expect = np.zeros(g.shape)
    expect[g.im - 2, :, :-1] = (2.0 * c * r * a3 / (g.km - 1))[:, None]
    assert np.allclose(w0.grad.numpy(), expect, rtol=1e-13, atol=1e-300)


def test_stale_backward_is_refused():
    from atjup_amd.adjoint import JetMisfit
    g = inputs.Grid(im=5, jm=7, km=9, the_degree=30.0, phi_degree=45.0)
    run, w = _runner(g)
    a = JetMisfit.apply(torch.from_numpy(w.copy()).requires_grad_(True), run)
    b = JetMisfit.apply(torch.from_numpy(2 * w).requires_grad_(True), run)
    with pytest.raises(RuntimeError, match="replaced"):
        a.backward()
    b.backward()
    # an evaluate() between a forward and its backward also invalidates the forward
    c = JetMisfit.apply(torch.from_numpy(w.copy()).requires_grad_(True), run)
    run.evaluate(w)
    with pytest.raises(RuntimeError, match="replaced"):
        c.backward()
    # a forward's backward runs once
    d = JetMisfit.apply(torch.from_numpy(w.copy()).requires_grad_(True), run)
    d.backward(retain_graph=True)
    with pytest.raises(RuntimeError):
        d.backward()
