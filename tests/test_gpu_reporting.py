"""The reporting reductions on the GPU against the reporting oracle (planet/Reporting.h:261-398):
searchMinMax_3D / _2D and steadyQuery's continuity residual, and the writers' slices. The oracle
follows the reference's loops literally and the residual's cell equals the compiled reference's
(test_oracle_reporting.py). Values and cells agree exactly: the reductions compare and select, and
the residual is the same expression on the same tables."""
import numpy as np
import pytest

from atjup_amd import abi, inputs
from oracle import reporting
from tests.helpers import make_pair

pytestmark = pytest.mark.gpu


def _same(a, b):
    for x, y in zip(a, b):
        if isinstance(x, tuple):
            assert tuple(x) == tuple(y), (a, b)
        else:
            assert x == y or (np.isnan(x) and np.isnan(y)), (a, b)


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64, abi.PRECISION_FP32_REF], ids=["fp64", "fp32-refined"])
def test_reporting_against_oracle(precision):
    gpu, orc, st = make_pair("full", sweeps=20, precision=precision)
    for it in range(1, 3):
        gpu.step(it, it == 1)
    gpu.mixture_density()
    gpu.precipitation()
    for n in ("t", "u", "v", "w", "p_dyn", "h2o", "nh3_cloud", "nh4sh", "tke", "Q_rad"):
        _same(gpu.search_minmax_3d(n), reporting.search_minmax_3d(gpu.get(n)))
    maps = gpu.precip_surface()
    for m, name in enumerate(abi.PRECIP_SURFACE):
        _same(gpu.search_minmax_2d(m), reporting.search_minmax_2d(maps[name]))
    G = gpu.grid
    rad, the = reporting.grid_axes(G)
    res = gpu.continuity_residual()
    ref = reporting.continuity_residual(gpu.get("u"), gpu.get("v"), gpu.get("w"), rad, the, G.dr, G.dthe, G.dphi)
    print("continuity residual", res, ref)
    assert res == ref
    t = gpu.get("t")
    for axis, index in ((0, 17), (1, 90), (2, 0), (2, G.km - 1)):
        sl = gpu.slice("t", axis, index)
        want = t[index] if axis == 0 else (t[:, index] if axis == 1 else t[:, :, index])
        assert np.array_equal(sl, want), (axis, index)
    gpu.close(); orc.close()


def test_reporting_ties_and_nans():
    """A field of few distinct values with NaNs, and one whose starting cell is NaN."""
    gpu, orc, st = make_pair("full", sweeps=20)
    rng = np.random.default_rng(3)
    f = rng.integers(-2, 3, size=gpu.shape).astype(np.float64)
    f[rng.random(f.shape) < 0.05] = np.nan
    gpu.set("h2s", f)
    _same(gpu.search_minmax_3d("h2s"), reporting.search_minmax_3d(f))
    f[0, 0, 0] = np.nan
    gpu.set("h2s", f)
    _same(gpu.search_minmax_3d("h2s"), reporting.search_minmax_3d(f))
    vs = rng.integers(0, 3, size=(gpu.shape[1], gpu.shape[2])).astype(np.float64)
    gpu.set_vel_star(vs)
    _same(gpu.search_minmax_2d(5), reporting.search_minmax_2d(vs))
    gpu.close(); orc.close()


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64, abi.PRECISION_FP32_REF], ids=["fp64", "fp32-refined"])
def test_steady_changes_against_oracle(precision):
    """steadyQuery's per-field changes after a physics iteration and after a plain one: the
    library's tracked n level (u, v, w, p_dyn from the iteration's start; t and the condensables
    from after the physics block's adjustments, as RK_SCALAR_SYNC leaves them) against the
    restatement on the fields downloaded at those points; values and cells exact."""
    gpu, orc, st = make_pair("full", sweeps=20, precision=precision)
    gpu.steady_track(True)
    names = reporting.STEADY_FIELDS
    synced = ("t", "h2o", "h2o_cloud", "h2o_ice", "nh3", "nh3_cloud", "nh3_ice")
    for it, first in ((1, True), (2, False), (3, False)):
        start = {n: gpu.get(n) for n in names}
        gpu.step_block(it, first)
        physics = it % 2 == 0 or first
        nlevel = dict(start)
        if physics:
            nlevel.update({n: gpu.get(n) for n in synced})
        gpu.step_rest(it)
        now = {n: gpu.get(n) for n in names}
        got = gpu.steady_changes()
        want = reporting.steady_changes(now, nlevel)
        for (n, v, c), (n2, v2, c2) in zip(got, want):
            print(f"iteration {it} {n}: {v:.6e} {c} (restatement {v2:.6e} {c2})")
            assert n == n2 and v == v2 and tuple(c) == tuple(c2), (it, n)
        assert any(v > 0.0 for _, v, _ in got)
    gpu.steady_track(False)
    gpu.close()
    orc.close()
