"""Closed-form and invariant checks of the oracle, independent of the reference run.

Each check is one the mathematics fixes, chosen so that a plausible slip in the oracle (a wrong
coefficient, a swapped neighbour, a missing seam copy) fails it:
  * the Shapiro filter's Fourier response on the periodic ring, 1 - s sin^2(m pi / N) for order 2
    and 1 - s sin^4(m pi / N) for order 4, with the 2-grid mode removed at s = 1 and constants kept;
  * the top sponge keeps the zonal mean of v and w on the layers it damps;
  * boundary identities: the seam copy, the pole rows, the rigid radial walls;
  * the hydrostatic column integral of v^2 / r is v^2 ln(r / r_foot) to second order in dr;
  * a constant pressure with a zero source is a fixed point of the red-black sweep.
"""
import math

import numpy as np
import pytest

import oracle
from atjup_amd import abi, inputs
from tests.helpers import small_grid, small_state


def _orc(st, sweeps=5, omega=1.0):
    G, Cn, K, _ = abi.make_profile("dyn", sweeps=sweeps, omega=omega, grid=st.grid)
    o = oracle.Oracle(G, Cn, K)
    o.load_state(st)
    return o


@pytest.mark.parametrize("order,power", [(2, 2), (4, 4)])
@pytest.mark.parametrize("m", [1, 3, 12, 24])
def test_shapiro_fourier_response_on_the_ring(order, power, m):
    g = small_grid()
    st = inputs.make_state("dyn", g=g, with_obstacle=False)
    o = _orc(st)
    N = g.km - 1
    k = np.arange(g.km)
    f = np.broadcast_to(np.cos(2 * np.pi * m * (k % N) / N), g.shape).copy()
    o.set("p_dyn", f)
    s = 0.7
    o.call("shapiro", abi.FIELD["p_dyn"], order, s, 1, 4)        # the phi pass alone
    out = o.get("p_dyn")
    gain = 1.0 - s * math.sin(math.pi * m / N) ** power
    np.testing.assert_allclose(out, gain * f, atol=1e-13)
    assert np.array_equal(out[:, :, -1], out[:, :, 0])


def test_shapiro_removes_the_two_grid_mode_and_keeps_constants():
    g = small_grid()
    st = inputs.make_state("dyn", g=g, with_obstacle=False)
    o = _orc(st)
    k = np.arange(g.km) % (g.km - 1)
    f = np.broadcast_to((-1.0) ** k, g.shape).copy()
    o.set("p_dyn", f)
    o.call("shapiro", abi.FIELD["p_dyn"], 2, 1.0, 1, 4)
    np.testing.assert_allclose(o.get("p_dyn"), 0.0, atol=1e-15)
    o.set("p_dyn", np.full(g.shape, 3.25))
    o.call("shapiro", abi.FIELD["p_dyn"], 4, 1.0, 2, 7)
    np.testing.assert_allclose(o.get("p_dyn"), 3.25, rtol=0, atol=1e-14)


def test_top_sponge_keeps_the_zonal_mean():
    st = small_state()
    rng = np.random.default_rng(0)
    for n in ("v", "w"):
        st.fields[n] = rng.normal(size=st.grid.shape) * 0.1
    o = _orc(st)
    im = st.grid.im
    before = {n: o.get(n)[im - 3:im - 1, :, :-1].mean(axis=2) for n in ("v", "w")}
    o.call("bc_radius")
    for n in ("v", "w"):
        a = o.get(n)
        np.testing.assert_allclose(a[im - 3:im - 1, :, :-1].mean(axis=2), before[n], atol=1e-15)
        np.testing.assert_allclose(a[im - 1, :, 0], before[n][1], atol=1e-15)   # the lid takes the im-2 mean
        assert np.all(a[im - 1] == a[im - 1, :, :1])


def test_boundary_identities():
    st = small_state()
    rng = np.random.default_rng(1)
    for n in ("t", "u", "v", "w"):
        st.fields[n] = st.fields[n] + rng.normal(size=st.grid.shape) * 1e-3
    o = _orc(st)
    o.call("bc_after_step")
    t, u, v = o.get("t"), o.get("u"), o.get("v")
    fluid = st.mask != 1.0
    assert np.array_equal(t[:, :, -1], t[:, :, 0])
    assert np.array_equal(t[:, 0, :], t[:, 3, :]) and np.array_equal(t[:, 1, :], t[:, 3, :])
    assert np.all(v[:, 0, :] == 0) and np.all(v[:, -1, :] == 0)
    np.testing.assert_allclose(v[:, 1, :], v[:, 3, :] / 3.0, rtol=1e-15)
    assert np.all(u[0] == 0) and np.all(u[-1] == 0)
    assert np.all(u[~fluid] == 0)
    # the 2-point Neumann form on the bottom plane: f0 = (4/3) f1 - (1/3) f2 (before the pole rows)
    # (solid cells at i = 0 are rewritten afterwards by the solid-ground pass)
    j = slice(3, -3)
    col = fluid[0, j, :]
    np.testing.assert_allclose(t[0, j, :][col], (4.0 / 3.0 * t[1, j, :] - 1.0 / 3.0 * t[2, j, :])[col], rtol=1e-14)


def test_hydrostatic_integral_is_a_logarithm():
    g = small_grid()
    st = inputs.make_state("dyn", g=g, with_obstacle=False)
    vv = 0.3
    st.fields["v"] = np.full(g.shape, vv)
    st.fields["w"] = np.zeros(g.shape)
    o = _orc(st)
    o.call("hydrostatic")
    ph = o.get("p_hydro")[:, 10, 5]
    r, _, _ = g.coords()
    exact = vv * vv * np.log(r / r[0])
    # trapezoid error of 1/r: dr^2 / 12 * int r^-3 ~ 1e-10 relative here
    np.testing.assert_allclose(ph, exact, rtol=1e-8, atol=1e-16)


def test_constant_pressure_is_a_fixed_point():
    st = small_state()
    o = _orc(st, sweeps=7)
    o.set("p_dyn", np.full(st.grid.shape, -0.125))
    for n in ("aux_u", "aux_v", "aux_w"):
        o.set(n, np.zeros(st.grid.shape))
    o.call("pressure_solve", 7, 1.0)
    np.testing.assert_allclose(o.get("p_dyn"), -0.125, rtol=1e-14)
    o.call("pressure_solve", 7, 1.6)
    np.testing.assert_allclose(o.get("p_dyn"), -0.125, rtol=1e-13)


def test_sor_relaxes_the_vertical_mode_faster_than_gauss_seidel():
    """With a zero source the relaxation drives a smooth vertical mode cos(pi i / (im-1)) towards a
    constant (the null space of the all-Neumann operator; the strongly coupled radial direction
    makes this a 1-D relaxation). Plain Gauss-Seidel must damp it by an order of magnitude in 60
    sweeps and over-relaxation at omega = 1.5 by a further order: a reversed blend
    (1.5 p - 0.5 p_GS) diverges and a dropped or mis-weighted blend does not accelerate."""
    g = small_grid()
    st = inputs.make_state("dyn", g=g, with_obstacle=False)
    i = np.arange(g.im)
    p0 = np.broadcast_to(np.cos(np.pi * i / (g.im - 1))[:, None, None], g.shape).copy()
    spread = {}
    for omega in (1.0, 1.5):
        o = _orc(st, sweeps=1)
        o.set("p_dyn", p0)
        for n in ("aux_u", "aux_v", "aux_w"):
            o.set(n, np.zeros(g.shape))
        o.call("pressure_solve", 60, omega)
        p = o.get("p_dyn")[1:-1, 1:-1, :-1]
        spread[omega] = float(p.max() - p.min())
        o.close()
    assert spread[1.0] < 0.1 * 2.0
    assert spread[1.5] < 0.1 * spread[1.0]


def _rb_state(seed=5):
    """The small shell with its obstacle, random p_dyn and aux_* (the sources are O(1), so the clamp
    at +-0.2 den cuts some of them)."""
    st = small_state()
    g = st.grid
    rng = np.random.default_rng(seed)
    o = _orc(st, sweeps=1)
    for n in ("p_dyn", "aux_u", "aux_v", "aux_w"):
        a = rng.standard_normal(g.shape) * (1.0 if n == "p_dyn" else 0.02)
        a[:, :, -1] = a[:, :, 0]
        o.set(n, a)
    return o, g


def _colour_mask(g, colour):
    i, j, k = np.meshgrid(np.arange(g.im), np.arange(g.jm), np.arange(g.km), indexing="ij")
    return ((i + j + k) % 2 == colour) & (i >= 1) & (i <= g.im - 2) & (j >= 1) & (j <= g.jm - 2) & (k <= g.km - 2)


def test_colour_passes_compose_to_the_pinned_sweep():
    """A Gauss-Seidel sweep is colour pass 0, colour pass 1, then the outer planes: the two passes
    alone reproduce pressure_solve(1, 1.0), which equals the compiled reference bit for bit
    (test_oracle_vs_reference), and at omega = 1.5 they reproduce pressure_solve(1, 1.5)."""
    for omega in (1.0, 1.5):
        o, g = _rb_state()
        p0 = o.get("p_dyn")
        o.call("pressure_solve", 1, omega)
        sweep = o.get("p_dyn")
        o.set("p_dyn", p0)
        o.call("rb_colour_pass", 0, omega, 0)
        o.call("rb_colour_pass", 1, omega, 1)
        assert np.array_equal(o.get("p_dyn"), sweep), omega
        o.close()


@pytest.mark.parametrize("omega", [0.6, 1.3, 1.5, 1.9])
@pytest.mark.parametrize("colour", [0, 1])
def test_sor_blend_is_the_gauss_seidel_value_relaxed(colour, omega):
    """Over-relaxation as a closed form of the pinned Gauss-Seidel pass: on the same input, a colour
    pass at omega leaves every cell of the other colour and every boundary cell as it was and sets
    each cell of its colour to (1 - omega) p_old + omega p_GS, where p_GS is that cell after the
    omega = 1 pass. Colour 1 is checked on the state after a colour-0 pass, as a sweep reaches it.
    A blend with another weight (an effective omega of 1.3 for 1.5), applied to a neighbour's or an
    updated old value, or to the other colour fails this."""
    o, g = _rb_state(seed=11 + colour)
    if colour == 1:
        o.call("rb_colour_pass", 0, omega, 0)
    p_old = o.get("p_dyn")
    o.call("rb_colour_pass", colour, 1.0, 0)
    p_gs = o.get("p_dyn")
    o.set("p_dyn", p_old)
    o.call("rb_colour_pass", colour, omega, 0)
    p_sor = o.get("p_dyn")
    o.close()
    mine = _colour_mask(g, colour)
    assert np.array_equal(p_sor[~mine], p_old[~mine])
    assert np.count_nonzero(p_gs[mine] != p_old[mine]) > 0.9 * np.count_nonzero(mine)
    expect = (1.0 - omega) * p_old[mine] + omega * p_gs[mine]
    scale = np.maximum(np.abs(p_old[mine]), np.abs(p_gs[mine]))
    assert np.all(np.abs(p_sor[mine] - expect) <= 4.0 * np.finfo(float).eps * scale)
    # and not the blend with the weight of a neighbouring omega
    wrong = (1.0 - (omega - 0.2)) * p_old[mine] + (omega - 0.2) * p_gs[mine]
    assert np.max(np.abs(p_sor[mine] - wrong)) > 1e3 * np.finfo(float).eps * np.max(scale)
