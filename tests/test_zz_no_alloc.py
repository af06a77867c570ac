"""The allocation-free searchMinMax_3D (atjup_hip_search_minmax_3d_noalloc) against the context's own
search (atjup_hip_search_minmax_3d), both through the C ABI: the same values and cells on two calls
in a row, the second on a smaller field, then on a field full of ties and a NaN; the partial results
live in the library's fixed scratch array, so the second call must not see the first one's."""
import ctypes as C

import numpy as np
import pytest

from atjup_amd import abi, inputs, lib
from atjup_amd.lib import Context
from oracle import reporting
from tests.helpers import small_grid

pytestmark = pytest.mark.gpu


def _noalloc(ctx, name, **over):
    L = lib.load()
    g = ctx.grid
    dims = dict(im=g.im, jm=g.jm, km=g.km, pitch=L.atjup_hip_pitch(ctx.h))
    dims.update(over)
    out, loc = (C.c_double * 2)(), (C.c_int * 6)()
    ptr = L.atjup_hip_field_device_ptr(ctx.h, abi.FIELD[name])
    rc = L.atjup_hip_search_minmax_3d_noalloc(ptr, dims["im"], dims["jm"], dims["km"], dims["pitch"], None, out, loc)
    if rc != 0:
        return rc, L.atjup_hip_last_error_string().decode()
    return out[0], tuple(loc[0:3]), out[1], tuple(loc[3:6])


def _ctx(g, seed):
    G, Cn, K, _ = abi.make_profile("dyn", sweeps=5, grid=g)
    c = Context(G, Cn, K, abi.PRECISION_FP64)
    st = inputs.make_state("dyn", seed=seed, g=g, with_obstacle=False)
    c.load_state(st, skip=set(inputs.SPECIES))
    return c


def test_noalloc_search_matches_the_context_search():
    big = _ctx(inputs.Grid(), 21)
    small = _ctx(small_grid(), 22)
    try:
        # two calls in a row, the second with a smaller input
        for ctx, name in ((big, "t"), (small, "t"), (big, "u"), (small, "w")):
            a = _noalloc(ctx, name)
            b = ctx.search_minmax_3d(name)
            print(name, ctx.grid.im, a, flush=True)
            assert a == b, (name, a, b)
            assert a == reporting.search_minmax_3d(ctx.get(name))
        # ties everywhere, a NaN away from the starting cell: the reference's first cell wins
        g = small.grid
        f = np.round(np.random.default_rng(5).standard_normal((g.im, g.jm, g.km)), 1)
        f[3, 4, 5] = np.nan
        small.set("p_dyn", f)
        a, b = _noalloc(small, "p_dyn"), small.search_minmax_3d("p_dyn")
        assert a == b == reporting.search_minmax_3d(f), (a, b)
    finally:
        big.close()
        small.close()


def test_noalloc_refuses_fields_above_its_maximum():
    ctx = _ctx(small_grid(), 23)
    try:
        rc, msg = _noalloc(ctx, "t", im=5000)
        assert rc == abi.ERR_ARG and "maximum" in msg
        rc, msg = _noalloc(ctx, "t", pitch=3)
        assert rc == abi.ERR_ARG
        # still usable afterwards
        assert _noalloc(ctx, "t") == ctx.search_minmax_3d("t")
    finally:
        ctx.close()
