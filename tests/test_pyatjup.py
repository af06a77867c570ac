"""The Cython binding pyatjup (atjup_amd/pyatjup.pyx), the reference binding's face
(python/pyatjup.pyx.template:32-97) over the device library.

CPU tests: the configuration file in the reference's layout (cli/config_atjup.xml) sets the
properties by their reference names, unknown tags are ignored, and a model that cannot reach a GPU
refuses with the library's own message. GPU tests: run() equals the same iterations through the
ctypes face bit for bit, set_field / get_field move values unchanged, and JetMisfit's gradient
through the binding agrees with central differences; a backward whose forward was replaced raises.
"""
import os

import numpy as np
import pytest

from atjup_amd import abi, inputs
from tests.helpers import small_grid, small_state

HERE = os.path.dirname(os.path.abspath(__file__))
CONFIG = os.path.join(HERE, "golden", "config_atjup_small.xml")

pyatjup = pytest.importorskip("atjup_amd.pyatjup", reason="the Cython binding is not built")


def test_load_config_sets_the_reference_properties():
    jup = pyatjup.Jupiter()
    assert jup.nm == 224 and jup.turb_model == "k_omega"
    jup.load_config(CONFIG)
    assert jup.nm == 3 and jup.turb_model == "k_omega_SST" and jup.coord_stretching is False
    assert jup.g == 25.92 and jup.re == 1000.0 and jup.output_path == "output/"
    jup.u_0 = 90.0
    assert jup.u_0 == 90.0
    with pytest.raises(AttributeError):
        jup.not_a_parameter
    with pytest.raises(AttributeError):
        jup.not_a_parameter = 1


def test_context_manager_and_device_refusal():
    from atjup_amd import lib
    with pyatjup.Jupiter(profile="dyn", grid=small_grid()) as jup:
        jup.state = small_state("dyn")
        if lib.device_count() == 0:
            with pytest.raises(pyatjup.AtjupError, match="no HIP device"):
                jup.get_field("t")


@pytest.mark.gpu
def test_run_equals_the_ctypes_path():
    from atjup_amd.lib import Context
    st = small_state("full")
    with pyatjup.Jupiter(profile="full", grid=st.grid, sweeps=10) as jup:
        jup.state = st
        jup.load_config(CONFIG)                   # nm 3, SST closure
        jup.run()
        G, Cn, K, _ = abi.make_profile("full", sweeps=10, grid=st.grid, turb="k_omega_SST", radiation=True,
                                       precip=True, chemistry=True, nh4sh_eq=True)
        ctx = Context(G, Cn, K)
        ctx.load_state(st)
        ctx.bc_after_step()
        ctx.swap_nlevel()
        for it in range(1, 4):
            ctx.step(it, it == 1)
        for n in abi.PROG + ("p_dyn", "nue"):
            assert np.array_equal(jup.get_field(n), ctx.get(n)), n
        ctx.close()
        a = np.random.default_rng(1).standard_normal(jup.shape)
        jup.set_field("p_dyn", a)
        assert np.array_equal(jup.get_field("p_dyn"), a)
        with pytest.raises(ValueError):
            jup.set_field("t", a[:, :, :-1])


@pytest.mark.gpu
def test_jet_misfit_gradient_through_the_binding():
    torch = pytest.importorskip("torch")
    st = small_state("dyn")
    jup = pyatjup.Jupiter(profile="dyn", grid=st.grid, sweeps=20)
    jup.state = st
    g = st.grid
    target = inputs.zonal_jets(g, np.random.default_rng(12)) / 100.0
    w0 = torch.from_numpy(st.fields["w"].copy()).requires_grad_(True)
    loss = pyatjup.JetMisfit.apply(w0, jup, target, 8, 4)
    loss.backward()
    grad = w0.grad.numpy()
    eps = 1.0e-6
    for seed in range(3):
        d = np.random.default_rng(100 + seed).standard_normal(g.shape)
        d[:, :, -1] = d[:, :, 0]
        lp, _ = pyatjup.jet_misfit(jup, st.fields["w"] + eps * d, target, 8, 4)
        lm, _ = pyatjup.jet_misfit(jup, st.fields["w"] - eps * d, target, 8, 4)
        fd = (lp - lm) / (2 * eps)
        ad = float(np.sum(grad * d))
        print("FD", fd, "adjoint", ad)
        assert abs(fd - ad) <= 1e-4 * max(abs(fd), abs(ad))
    # two outstanding forwards: the earlier one's backward is refused
    a = pyatjup.JetMisfit.apply(torch.from_numpy(st.fields["w"].copy()).requires_grad_(True), jup, target, 8, 4)
    b = pyatjup.JetMisfit.apply(torch.from_numpy(st.fields["w"].copy()).requires_grad_(True), jup, target, 8, 4)
    with pytest.raises(RuntimeError, match="replaced"):
        a.backward()
    b.backward()
    jup.close()


@pytest.mark.gpu
def test_semi_axis_gradient_through_the_binding():
    """JetMisfit with the drag region's semi-axes as a tensor input (all-fluid binary mask): the
    axes' gradient against central differences of the same forward."""
    torch = pytest.importorskip("torch")
    st = small_state("dyn")
    st.mask = np.zeros(st.grid.shape)
    jup = pyatjup.Jupiter(profile="dyn", grid=st.grid, sweeps=20)
    jup.state = st
    jup.set_smooth_obstacle(30.0, 17.0, 60.0, width=0.25, tau=0.05, j0=20, k0=24)
    target = st.fields["w"][3, :, :-1].mean(axis=1) * 0.9
    w0 = torch.from_numpy(st.fields["w"].copy()).requires_grad_(True)
    axes0 = np.array([30.0, 17.0, 60.0])
    axes = torch.from_numpy(axes0.copy()).requires_grad_(True)
    loss = pyatjup.JetMisfit.apply(w0, jup, target, 6, 2, axes, 3)
    loss.backward()
    ga = axes.grad.numpy()
    for p in range(3):
        eps = 1.0e-4 * axes0[p]
        lp, _ = pyatjup.jet_misfit(jup, st.fields["w"], target, 6, 2, i_top=3,
                                   axes=axes0 + eps * np.eye(3)[p])
        lm, _ = pyatjup.jet_misfit(jup, st.fields["w"], target, 6, 2, i_top=3,
                                   axes=axes0 - eps * np.eye(3)[p])
        fd = (lp - lm) / (2 * eps)
        print("semi-axis", p, "FD", fd, "adjoint", ga[p])
        assert abs(fd - ga[p]) <= 1e-4 * max(abs(fd), abs(ga[p]))
    jup.close()
