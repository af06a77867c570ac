"""Gradients of the dynamics step through the hand-written adjoint, as a torch.autograd.Function.

The loss of the adjoint benchmark: the cos(latitude)-weighted squared misfit of the zonal-mean
zonal wind at the top integrated level against a target jet profile,

    L(w0) = sum_j cos(lat_j) (mean_k w_N[i_top, j, k] - target_j)^2,   k over the ring 0..km-2,

after N loop iterations from an initial state whose zonal wind w (the reference's phi component)
is w0. ``JetMisfit.apply(w0, runner)`` returns L as a torch scalar; its backward runs the
checkpointed reverse sweep (atjup_hip_adj_backward) and returns dL/dw0.

torch is used on the CPU only: the device work runs in the library's own HIP runtime, and the
tensors cross as host arrays (the torch wheel's bundled HIP runtime is a different one).

The reference's obstacle is a binary mask, whose derivative with respect to its semi-axes is zero
almost everywhere. For gradients with respect to the semi-axes the runner takes a smooth drag region
instead (abi.Obstacle, atjup_hip_set_smooth_obstacle: -chi u / tau in the momentum equations with
chi a logistic function of the ellipsoid's normalised distance), with an all-fluid binary mask; the
forward that difference quotients differentiate is then that model, and ``JetMisfit.apply(w0,
runner, axes)`` returns L with dL/dw0 and dL/d(a_phi_deg, b_the_deg, c_r_km) as its gradients.
"""
from __future__ import annotations

import numpy as np

from . import abi

try:  # torch is optional for the numpy interface
    import torch
except Exception:  # pragma: no cover
    torch = None


class MisfitRunner:
    """One context holding the initial state, and the loss definition. An FP64 context gives the
    FP64 adjoint; an FP32-refined one the adjoint of its refined forward with FP32 adjoint fields."""

    def __init__(self, ctx, init_state: dict, target, nsteps=20, every=5, i_top=None, obstacle=None):
        self.ctx = ctx
        # the smooth drag region (abi.Obstacle) whose semi-axes set_axes() changes, or None
        self.obstacle = obstacle
        if obstacle is not None:
            ctx.set_smooth_obstacle(obstacle)
        self.init = {n: np.array(v, dtype=np.float64) for n, v in init_state.items()}
        im, jm, km = ctx.shape
        self.target = np.asarray(target, dtype=np.float64)
        self.i_top = im - 2 if i_top is None else i_top
        self.nsteps, self.every = nsteps, every
        _, theta, _ = _coords(ctx)
        self.cos_lat = np.abs(np.sin(theta))          # theta is the colatitude
        # The checkpoints and the adjoint seed of a forward live in the context until the next
        # forward overwrites them; each forward gets a new generation number so that a backward
        # belonging to an earlier forward is refused instead of returning another one's gradient.
        self.generation = 0
        self._lam = None
        self.axes_grad = None
        ctx.adj_init()

    def loss_final(self):
        w = self.ctx.get("w")
        km = w.shape[2]
        r = w[self.i_top, :, :-1].mean(axis=1) - self.target
        lam = np.zeros((4,) + w.shape)
        lam[3, self.i_top, :, :-1] = (2.0 * self.cos_lat * r / (km - 1))[:, None]
        return float(np.sum(self.cos_lat * r * r)), lam

    def set_axes(self, axes):
        """The drag region's semi-axes (a_phi_deg, b_the_deg, c_r_km) for the next forward."""
        if self.obstacle is None:
            raise RuntimeError("MisfitRunner: no smooth drag region (pass obstacle=abi.obstacle(...))")
        o = self.obstacle
        o.a_phi_deg, o.b_the_deg, o.c_r_km = (float(x) for x in axes)
        self.ctx.set_smooth_obstacle(o)

    def load(self, w0):
        for n, v in self.init.items():
            self.ctx.set(n, w0 if n == "w" else v)

    def forward(self, w0):
        """L(w0) through the checkpointed forward sweep, ready for backward()."""
        self.generation += 1
        self._lam = None
        self.load(w0)
        self.ctx.adj_forward(self.nsteps, self.every, 1)
        loss, self._lam = self.loss_final()
        return loss

    def evaluate(self, w0):
        """L(w0) by plain forward steps (no checkpoints), for difference quotients. The context's
        state then no longer belongs to the last forward(): its backward is refused."""
        self.generation += 1
        self._lam = None
        self.load(w0)
        for n in range(self.nsteps):
            self.ctx.step(1 + n, n == 0)
        return self.loss_final()[0]

    def backward(self, scale=1.0, generation=None):
        """dL/dw0 of the last forward(). generation: the number forward() had when the caller
        ran it; a backward for any other forward raises (the checkpoints are gone)."""
        if self._lam is None or (generation is not None and generation != self.generation):
            raise RuntimeError("MisfitRunner.backward: the checkpoints of that forward were replaced by a later "
                               "forward or evaluate on the same runner; run the forward again (one runner "
                               "holds one forward at a time)")
        g = self.ctx.adj_backward(scale * self._lam)[3]
        # dL/d(a_phi_deg, b_the_deg, c_r_km) of the drag region, from the same reverse sweep
        self.axes_grad = self.ctx.adj_obstacle_gradient() if self.obstacle is not None else None
        self._lam = None          # the reverse sweep consumed the checkpoints' state
        return g


def _coords(ctx):
    """The colatitudes, cumulative sums of dthe from 0 as the reference forms them."""
    g = ctx.grid
    the = np.zeros(g.jm)
    for j in range(1, g.jm):
        the[j] = the[j - 1] + g.dthe
    return None, the, None


if torch is not None:
    class JetMisfit(torch.autograd.Function):
        """L(w0[, axes]) with the adjoint as its backward (FP64 CPU tensors). axes: the drag
        region's (a_phi_deg, b_the_deg, c_r_km) as a tensor of 3, for a runner built with an
        obstacle; its gradient comes from the same reverse sweep."""

        @staticmethod
        def forward(fctx, w0, runner: MisfitRunner, axes=None):
            if axes is not None:
                runner.set_axes(axes.detach().cpu().numpy().astype(np.float64))
            loss = runner.forward(w0.detach().cpu().numpy().astype(np.float64))
            fctx.runner = runner
            fctx.generation = runner.generation
            fctx.has_axes = axes is not None
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(fctx, grad_out):
            g = fctx.runner.backward(float(grad_out), generation=fctx.generation)
            ga = torch.from_numpy(fctx.runner.axes_grad.copy()) if fctx.has_axes else None
            return torch.from_numpy(g), None, ga
