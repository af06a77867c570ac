"""The drop-in's own transfer and status entry points on the GPU.

* atjup_hip_upload_field_rows / download_field_rows take the reference's Array storage unchanged:
  `double*** x`, one separately allocated row of km doubles per (i, j) (lib/Array.cpp:43-73). Built
  here from ctypes pointer tables over separate numpy rows, they must move exactly the values the
  contiguous calls move, in both precision forms, and a state loaded row by row must step exactly as
  one loaded contiguously.
* atjup_hip_check_finite is the NaN watch (ATJUP_NANCHECK, planet/FileIO_Jup.cpp:326-373): a NaN or
  an infinity uploaded into a prognostic field is reported as ATJUP_NONFINITE, a clean state as 0.
"""
import ctypes as C

import numpy as np
import pytest

from atjup_amd import abi, inputs, lib
from atjup_amd.lib import Context
from tests.helpers import small_state

pytestmark = pytest.mark.gpu


class RowArray:
    """A field in the reference's layout: rows[i][j] is its own allocation of km doubles, reached
    through a table of im pointers to tables of jm row pointers (a double***)."""

    def __init__(self, arr):
        im, jm, km = arr.shape
        self.rows = [[np.array(arr[i, j], dtype=np.float64, copy=True) for j in range(jm)] for i in range(im)]
        self.tabs = [(C.c_void_p * jm)(*[r.ctypes.data for r in self.rows[i]]) for i in range(im)]
        self.top = (C.c_void_p * im)(*[C.addressof(t) for t in self.tabs])

    @property
    def ptr(self):
        return C.cast(self.top, C.c_void_p)

    def array(self):
        return np.array([[r for r in row] for row in self.rows])


def _ctx(st, precision=abi.PRECISION_FP64, config="full"):
    G, Cn, K, _ = abi.make_profile(config, sweeps=10, grid=st.grid)
    return Context(G, Cn, K, precision)


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64, abi.PRECISION_FP32_REF], ids=["fp64", "fp32-refined"])
def test_field_rows_move_what_the_contiguous_calls_move(precision):
    st = small_state("full")
    ctx = _ctx(st, precision)
    L = lib.load()
    try:
        rng = np.random.default_rng(3)
        shape = (st.grid.im, st.grid.jm, st.grid.km)
        for name in ("t", "h2o", "p_dyn", "tn"):
            a = rng.standard_normal(shape) + 2.0
            ra = RowArray(a)
            assert L.atjup_hip_upload_field_rows(ctx.h, abi.FIELD[name], ra.ptr) == 0, ctx.error()
            got = ctx.get(name)
            if precision == abi.PRECISION_FP64:
                assert np.array_equal(got, a), name
            else:
                assert np.array_equal(got, a.astype(np.float32).astype(np.float64)), name
            b = rng.standard_normal(shape)
            ctx.set(name, b)
            out = RowArray(np.zeros(shape))
            assert L.atjup_hip_download_field_rows(ctx.h, abi.FIELD[name], out.ptr) == 0, ctx.error()
            assert np.array_equal(out.array(), ctx.get(name)), name
    finally:
        ctx.close()


def test_state_loaded_row_by_row_steps_like_the_contiguous_one():
    st = small_state("full")
    L = lib.load()
    a, b = _ctx(st), _ctx(st)
    try:
        a.load_state(st)
        b.set_mask(st.mask)
        keep = []
        for name, arr in st.fields.items():
            ra = RowArray(arr)
            keep.append(ra)
            assert L.atjup_hip_upload_field_rows(b.h, abi.FIELD[name], ra.ptr) == 0, b.error()
        for c in (a, b):
            c.bc_after_step()
            c.swap_nlevel()
            for it in range(1, 4):
                c.step(it, it == 1)
        for name in abi.PROG + ("p_dyn",):
            out = RowArray(np.zeros((st.grid.im, st.grid.jm, st.grid.km)))
            assert L.atjup_hip_download_field_rows(b.h, abi.FIELD[name], out.ptr) == 0
            assert np.array_equal(out.array(), a.get(name)), name
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_check_finite_reports_an_uploaded_nonfinite_value(bad):
    st = small_state("full")
    ctx = _ctx(st)
    L = lib.load()
    try:
        ctx.load_state(st)
        assert L.atjup_hip_check_finite(ctx.h) == 0
        for name in ("t", "nh3_cloud", "dis"):
            arr = st.fields[name].copy()
            arr[4, 7, 11] = bad
            ctx.set(name, arr)
            assert L.atjup_hip_check_finite(ctx.h) == abi.NONFINITE, name
            ctx.set(name, st.fields[name])
            assert L.atjup_hip_check_finite(ctx.h) == 0, name
    finally:
        ctx.close()
