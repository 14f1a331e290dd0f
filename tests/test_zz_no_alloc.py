"""smax_hip_crs_spmm_noalloc: the allocation-free column-major CRS SpMM
(scratch in a fixed device array), called through the C ABI and compared bit
for bit with smax_hip_crs_spmm (and the oracle) on two calls in a row, the
second with a smaller input, plus its size limit and the row-major path."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from smaxkernels_amd import capi, matgen
from tests.helpers import assert_bitwise

pytestmark = pytest.mark.gpu

SCRATCH = 32 << 20


def operands(vt, A, k, seed):
    if vt == capi.F64:
        return A.val, matgen.uniform_f64(seed, A.n_cols * k)
    if vt == capi.F32:
        return A.val.astype(np.float32), matgen.uniform_f32(seed, A.n_cols * k)
    return matgen.to_bf16(A.val), matgen.uniform_f32(seed, A.n_cols * k)


def run(ctx, entry, vt, A, dA, k, row_major, dX, ldx, ldy, dtype, plan=None):
    rp, col, v = dA
    n = A.n_rows
    dY = ctx.upload(np.full((n if row_major else ldy) * (ldy if row_major else k), 3.0, dtype))
    st = getattr(ctx.lib, entry)(ctx.handle, vt, n, A.n_cols, A.nnz, rp.ptr, col.ptr, v.ptr, k,
                                 int(row_major), dX.ptr, ldx, dY.ptr, ldy, plan)
    assert st == capi.SUCCESS, ctx.lib.smax_hip_last_error(ctx.handle)
    out = ctx.download(dY)
    dY.free()
    return out


@pytest.mark.parametrize("vt", [capi.F32, capi.F64, capi.BF16])
def test_noalloc_matches_workspace_entry_on_two_calls(ctx, vt):
    """Call 1: 200k rows, k = 32, padded leading dimensions: F32/BF16 run two
    passes of 16 columns, F64 four passes (10, 10, 10, 2).  Call 2: a smaller
    product (30k rows, k = 7, one pass), which must not see call 1's scratch
    contents.  Both equal smax_hip_crs_spmm and the oracle bit for bit."""
    for n, k, seed in ((200_000, 32, 5), (30_000, 7, 6)):
        A = matgen.lomax_band(n)
        val, X = operands(vt, A, k, seed)
        dtype = X.dtype
        ldx, ldy = A.n_cols + 3, A.n_rows + 5
        Xp = np.zeros((k, ldx), dtype)
        Xp[:, :A.n_cols] = X.reshape(k, A.n_cols)
        dA = (ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(val))
        dX = ctx.upload(Xp.ravel())
        plan = ctx.crs_plan(A.n_rows, A.nnz, dA[0])
        got = run(ctx, "smax_hip_crs_spmm_noalloc", vt, A, dA, k, False, dX, ldx, ldy, dtype,
                  plan)
        want = run(ctx, "smax_hip_crs_spmm", vt, A, dA, k, False, dX, ldx, ldy, dtype, plan)
        assert_bitwise(got, want, f"n={n} k={k}")
        got = got.reshape(k, ldy)
        assert (got[:, A.n_rows:] == 3.0).all(), "padding between columns written"
        for c in (0, k // 2, k - 1):
            assert_bitwise(got[c, :A.n_rows],
                           oracle.crs_spmv(A.row_ptr, A.col, val, X.reshape(k, -1)[c]),
                           f"column {c} vs oracle")
        ctx.crs_plan_destroy(plan)
        for d in (*dA, dX):
            d.free()


def test_noalloc_row_major_and_single_column(ctx):
    """Row-major products and k = 1 need no scratch and equal the other entry."""
    A = matgen.lomax_band(20_000, half=2500)
    dA = (ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val))
    for k, row_major in ((32, True), (5, True), (1, False)):
        X = matgen.uniform_f64(9, A.n_cols * k)
        dX = ctx.upload(X)
        ld = k if row_major else A.n_rows
        got = run(ctx, "smax_hip_crs_spmm_noalloc", capi.F64, A, dA, k, row_major, dX, ld, ld,
                  np.float64)
        want = run(ctx, "smax_hip_crs_spmm", capi.F64, A, dA, k, row_major, dX, ld, ld,
                   np.float64)
        assert_bitwise(got, want, f"k={k} row_major={row_major}")
        dX.free()
    for d in dA:
        d.free()


def test_noalloc_refuses_products_beyond_its_scratch(ctx):
    """One FP64 column of X and Y over n_rows + n_cols = 4,400,000 entries is
    35.2 MB, more than the 32 MiB scratch: refused before any launch, with
    an error naming the limit; Y is left untouched."""
    n = 2_200_000
    rp = np.arange(n + 1, dtype=np.int32)
    col = np.arange(n, dtype=np.int32)
    A = matgen.CRS(n, n, rp, col, np.ones(n))
    dA = (ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val))
    k = 2
    dX = ctx.upload(np.ones(n * k))
    dY = ctx.upload(np.full(n * k, 4.0))
    st = ctx.lib.smax_hip_crs_spmm_noalloc(ctx.handle, capi.F64, n, n, n, dA[0].ptr, dA[1].ptr,
                                           dA[2].ptr, k, 0, dX.ptr, n, dY.ptr, n, None)
    assert st == capi.ERROR_INVALID_VALUE
    assert b"scratch" in ctx.lib.smax_hip_last_error(ctx.handle)
    ctx.sync()
    assert (ctx.download(dY) == 4.0).all()
    # the workspace entry point takes the same product
    st = ctx.lib.smax_hip_crs_spmm(ctx.handle, capi.F64, n, n, n, dA[0].ptr, dA[1].ptr,
                                   dA[2].ptr, k, 0, dX.ptr, n, dY.ptr, n, None)
    assert st == capi.SUCCESS
    assert (ctx.download(dY) == 1.0).all()
    for d in (*dA, dX, dY):
        d.free()
