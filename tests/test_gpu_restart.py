"""The restart file through the library (cJupiterModel::save_state / load_state,
planet/FileIO_Jup.cpp:584-684).

oracle/restart.py restates the format and is pinned to the compiled reference byte for byte, both
ways, for JUP2 and JUP1 (test_oracle_vs_reference.py). Here the library's file of a device-resident
state is held to the restatement's bytes for the same values; a fresh context loads it back bit for
bit; a run resumed from a file written mid-run reproduces the straight run bit for bit (the
reference's own reproducibility check for JUP2); and a wrong grid, a short file and n-level copies
that differ from their fields are refused.
"""
import os

import numpy as np
import pytest

from atjup_amd import abi, inputs
from oracle import restart
from tests.helpers import small_state

pytestmark = pytest.mark.gpu


def _ctx(st, precision=abi.PRECISION_FP64, mask=True):
    from atjup_amd.lib import Context
    G, Cn, K, _ = abi.make_profile("full", sweeps=20, grid=st.grid)
    c = Context(G, Cn, K, precision)
    if mask:
        c.load_state(st)
        c.bc_after_step()
        c.swap_nlevel()
    return c


def _seed_carried(c, st, seed=4):
    rng = np.random.default_rng(seed)
    for n in abi.RESTART_CARRIED:
        c.set(n, rng.standard_normal(st.grid.shape))
    c.set_vel_star(rng.standard_normal(st.grid.shape[1:]))


def _values(c):
    base = {"p_dynn": "p_dyn"}
    out = {}
    for n in abi.RESTART_FIXED + abi.RESTART_CARRIED:
        b = base.get(n, n[:-1] if n.endswith("n") and n[:-1] in abi.PROG else n)
        out[n] = c.get(b)
    return out


@pytest.mark.parametrize("carried", [True, False], ids=["JUP2", "JUP1"])
def test_file_bytes_and_reload(tmp_path, carried):
    st = small_state("full", seed=5)
    c = _ctx(st)
    _seed_carried(c, st)
    path = tmp_path / "jup_restart_17.bin"
    c.save_restart(path, 17, carried=carried)
    vals = _values(c)
    ref = tmp_path / "restated.bin"
    restart.write(ref, vals, c.get_vel_star(), 17, carried=carried)
    assert path.read_bytes() == ref.read_bytes()
    d = _ctx(st, mask=False)
    d.set_mask(st.mask)
    it, car = d.load_restart(path)
    assert it == 17 and car == carried
    names = abi.RESTART_FIXED + (abi.RESTART_CARRIED if carried else ())
    back = _values(d)
    for n in names:
        assert np.array_equal(back[n], vals[n]), n
    if carried:
        assert np.array_equal(d.get_vel_star(), c.get_vel_star())
    c.close()
    d.close()


def test_resumed_run_reproduces_the_straight_run(tmp_path):
    st = small_state("full", seed=6)
    a = _ctx(st)
    for it in range(1, 5):
        a.step(it, it == 1)
    path = tmp_path / "jup_restart_4.bin"
    a.save_restart(path, 4)
    b = _ctx(st, mask=False)
    b.set_mask(st.mask)
    it0, _ = b.load_restart(path)
    for it in range(it0 + 1, it0 + 5):
        a.step(it, False)
        b.step(it, False)
    for n in abi.PROG + ("p_dyn", "nue", "rho_mix"):
        assert np.array_equal(a.get(n), b.get(n)), n
    a.close()
    b.close()


def test_fp32_refined_form_writes_its_values_widened(tmp_path):
    st = small_state("full", seed=7)
    c32 = _ctx(st, abi.PRECISION_FP32_REF)
    for it in range(1, 3):
        c32.step(it, it == 1)
    path = tmp_path / "jup_restart_2.bin"
    c32.save_restart(path, 2)
    _, _, got, _ = restart.read(path, st.grid.shape)
    for n in ("t", "u", "h2o", "p_dyn", "rho_mix"):
        assert np.array_equal(got[n], c32.get(n)), n
    c32.close()


def test_refusals(tmp_path):
    st = small_state("full", seed=8)
    c = _ctx(st)
    path = tmp_path / "jup_restart_3.bin"
    c.save_restart(path, 3)
    from atjup_amd.lib import AtjupError
    # a short file
    short = tmp_path / "short.bin"
    short.write_bytes(path.read_bytes()[:-8])
    with pytest.raises(AtjupError):
        c.load_restart(short)
    # another grid
    g2 = inputs.Grid(im=13, jm=33, km=51, the_degree=180.0 / 32, phi_degree=360.0 / 50)
    st2 = inputs.make_state("full", seed=8, g=g2)
    c2 = _ctx(st2)
    with pytest.raises(AtjupError):
        c2.load_restart(path)
    # n-level copies that differ from their fields
    _, car, vals, vs = restart.read(path, st.grid.shape)
    vals = dict(vals)
    vals["tn"] = vals["tn"] + 1.0
    bad = tmp_path / "bad.bin"
    restart.write(bad, vals, vs, 3, carried=car)
    with pytest.raises(AtjupError):
        c.load_restart(bad)
    c.close()
    c2.close()
