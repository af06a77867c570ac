"""Shared set-up of the parity tests: one seeded state loaded into the HIP context and the oracle."""
from __future__ import annotations

import numpy as np

from atjup_amd import abi, inputs

DYN_FIELDS = ("t", "u", "v", "w", "p_dyn", "p_hydro", "rho_mix", "aux_u", "aux_v", "aux_w")


def small_grid():
    """A 13 x 33 x 49 shell (ring of 48 meridians), small enough for pure-CPU runs in a second."""
    return inputs.Grid(im=13, jm=33, km=49, the_degree=180.0 / 32, phi_degree=360.0 / 48)


def small_state(config="dyn", seed=3, g=None):
    g = g if g is not None else small_grid()
    st = inputs.make_state(config, seed=seed, g=g, with_obstacle=False)
    # a small ellipsoid obstacle that still has flanks in all three directions
    st.mask = inputs.ellipsoid_mask(g, j0=20, k0=24, a_phi=4.0, b_the=3.0, c_r_km=40.0)
    solid = st.mask == 1.0
    for n in ("u", "v", "w"):
        st.fields[n][solid] = 0.0
    return st


def make_pair(config="dyn", sweeps=20, omega=1.0, state=None, precision=abi.PRECISION_FP64, **profile):
    """The HIP context and the oracle on the same state; `profile` goes to abi.make_profile (the
    closure, the physics-block components)."""
    from atjup_amd.lib import Context
    import oracle
    st = state if state is not None else inputs.make_state(config, seed=7)
    G, Cn, K, env = abi.make_profile(config, sweeps=sweeps, omega=omega, grid=st.grid, **profile)
    gpu = Context(G, Cn, K, precision)
    orc = oracle.Oracle(G, Cn, K)
    skip = set(inputs.SPECIES) if config == "dyn" else set()
    gpu.load_state(st, skip=skip)
    orc.load_state(st)
    # the reference's start: boundary passes, then restoreVar (cJupiterModel.cpp:702-739)
    gpu.bc_after_step()
    gpu.swap_nlevel()
    orc.call("bc_after_step")
    orc.call("swap_nlevel")
    return gpu, orc, st


def rel_err(a, b):
    """max |a - b| / max |b| (0 when both are zero)."""
    s = np.max(np.abs(b))
    d = np.max(np.abs(a - b))
    return 0.0 if d == 0 else d / (s if s > 0 else 1.0)


def compare(gpu, orc, names, tol, tag=""):
    worst = {}
    for n in names:
        a, b = gpu.get(n), orc.get(n)
        e = rel_err(a, b)
        worst[n] = (e, int(np.count_nonzero(a != b)))
        assert np.all(np.isfinite(a)), f"{tag} {n}: non-finite values"
        assert e <= tol, f"{tag} {n}: relative error {e:.3e} > {tol:.1e} ({worst[n][1]} cells differ)"
    return worst
