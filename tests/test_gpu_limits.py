"""Sizes beyond the launch arithmetic's limits are refused before any kernel
reads the operands (the arrays passed here are far smaller than the sizes
claimed, so a launch would read out of bounds)."""
from __future__ import annotations

import numpy as np
import pytest

from smaxkernels_amd import capi, matgen

pytestmark = pytest.mark.gpu

BIG = 1 << 40


def test_products_refuse_sizes_beyond_the_launch_limits(ctx):
    A = matgen.lomax_band(2000, half=200)
    rp, col, v = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val)
    X = ctx.zeros(64, np.float64)
    Y = ctx.upload(np.full(64, 5.0))
    lib = ctx.lib
    # CRS SpMM, with and without the no-allocation entry point
    for fn in (lib.smax_hip_crs_spmm, lib.smax_hip_crs_spmm_noalloc):
        for row_major in (1, 0):
            st = fn(ctx.handle, capi.F64, A.n_rows, A.n_cols, A.nnz, rp.ptr, col.ptr, v.ptr,
                    BIG, row_major, X.ptr, BIG, Y.ptr, BIG, None)
            assert st == capi.ERROR_OVERFLOW, (fn, row_major)
    # SELL: a slot count whose product overflows, and k through the handle
    s = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, 8, 64, 64)
    i = s.info
    st = lib.smax_hip_scs_spmv(ctx.handle, capi.F64, 1 << 32, 1 << 32, i.chunk_ptr,
                               i.chunk_lengths, i.col, i.val, X.ptr, Y.ptr)
    assert st == capi.ERROR_INVALID_VALUE
    st = lib.smax_hip_scs_apply_spmm(ctx.handle, capi.F64, s.handle, BIG, X.ptr, Y.ptr)
    assert st == capi.ERROR_OVERFLOW
    st = lib.smax_hip_scs_sddmm(ctx.handle, capi.F64, s.handle, BIG, X.ptr, X.ptr, Y.ptr)
    assert st == capi.ERROR_OVERFLOW
    st = lib.smax_hip_scs_spmm(ctx.handle, capi.F64, i.C, i.n_chunks, i.n_rows, i.chunk_ptr,
                               i.chunk_lengths, i.col, i.val, i.slot_row, BIG, X.ptr, Y.ptr)
    assert st == capi.ERROR_OVERFLOW
    s.destroy()
    # BCRS: block rows beyond int32, and k through a converted matrix
    B = matgen.block7(4, 4)
    brp, bcol, bv = ctx.upload(B.row_ptr), ctx.upload(B.col), ctx.upload(B.val)
    b = ctx.bcrs_from_crs(B.n_rows, B.n_cols, B.nnz, brp, bcol, bv, 8, 4, 4, 4, 4, False)
    bi = b.info
    st = lib.smax_hip_bcrs_spmm(ctx.handle, capi.F64, (1 << 31) + 1, 4, 4, 4, 4, bi.col,
                                bi.row_ptr, bi.val, 0, 1, X.ptr, Y.ptr)
    assert st == capi.ERROR_OVERFLOW
    st = lib.smax_hip_bcrs_spmm(ctx.handle, capi.F64, bi.n_block_rows, 4, 4, 4, 4, bi.col,
                                bi.row_ptr, bi.val, 0, BIG, X.ptr, Y.ptr)
    assert st == capi.ERROR_OVERFLOW
    st = lib.smax_hip_bcrs_sddmm(ctx.handle, capi.F64, b.handle, BIG, X.ptr, X.ptr, Y.ptr)
    assert st == capi.ERROR_OVERFLOW
    b.destroy()
    # row and column counts beyond int32: SpMV and the three conversions
    huge = (1 << 31) + 4
    st = lib.smax_hip_crs_spmv(ctx.handle, capi.F64, huge, A.n_cols, A.nnz, rp.ptr, col.ptr,
                               v.ptr, X.ptr, Y.ptr, None)
    assert st == capi.ERROR_OVERFLOW
    h = capi.vp()
    st = lib.smax_hip_scs_from_crs(ctx.handle, A.n_rows, huge, A.nnz, rp.ptr, col.ptr, v.ptr, 8,
                                   64, 64, capi.C.byref(h))
    assert st == capi.ERROR_OVERFLOW and h.value is None
    st = lib.smax_hip_scs_transpose_from_crs(ctx.handle, A.n_rows, huge, A.nnz, rp.ptr,
                                             col.ptr, v.ptr, 8, 64, 64, capi.C.byref(h))
    assert st == capi.ERROR_OVERFLOW and h.value is None
    st = lib.smax_hip_bcrs_from_crs(ctx.handle, huge, huge, B.nnz, brp.ptr, bcol.ptr, bv.ptr, 8,
                                    4, 4, 4, 4, 0, capi.C.byref(h))
    assert st == capi.ERROR_OVERFLOW and h.value is None
    ctx.sync()
    assert (ctx.download(Y) == 5.0).all()
    for d in (rp, col, v, X, Y, brp, bcol, bv):
        d.free()
