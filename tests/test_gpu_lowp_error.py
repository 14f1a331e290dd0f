"""Error of every FP32 / BF16-value result against the FP64 oracle on the
FP64 matrix (values stored in FP64, operands widened exactly), per output:

    |y_lo - y_64| <= (u_v + gamma_n (1 + u_v)) (|A| |x|)

u_v is the relative rounding of a stored value (FP32: 2^-24; BF16, stored
as bfloat16(float(a)), 8 significand bits: 2^-8 + 2^-24 + 2^-32), gamma_n = n u / (1 - n u)
with u = 2^-24 is the bound of the n fma roundings of the FP32 chain, n the
number of terms of that output (row length, or 4 per block of a BCRS block
row).  The SDDMM rounds no stored value: its bound is its reduction tree's,
(q + log2 g) u (sum |g x|).  Each test also reports the largest observed
ratio |y_lo - y_64| / (|A| |x|), the figure the benchmark states per
result."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle
from smaxkernels_amd import capi, matgen
from tests.helpers import gamma, sddmm_tree

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
U_VALUE = {capi.F32: U32, capi.BF16: 2.0 ** -8 + 2.0 ** -24 + 2.0 ** -32}


def stored(vt, val64):
    return val64.astype(np.float32) if vt == capi.F32 else matgen.to_bf16(val64)


def check(y_lo, y64, scale, n_terms, u_v, what):
    bound = (u_v + gamma(n_terms, U32) * (1 + u_v)) * scale
    err = np.abs(np.asarray(y_lo, np.float64) - y64)
    assert (err <= bound).all(), what
    ratio = float(np.max(err / np.where(scale > 0, scale, 1)))
    assert ratio <= u_v + float(gamma(np.max(n_terms), U32)) * (1 + u_v), what
    return ratio


def abs_csr(A):
    return sp.csr_matrix((np.abs(A.val), A.col, A.row_ptr), shape=(A.n_rows, A.n_cols))


@pytest.mark.parametrize("vt", [capi.F32, capi.BF16])
def test_sell_spmv_error_vs_fp64(ctx, vt):
    A = matgen.lomax_band(200_000)
    x = matgen.uniform_f32(matgen.SEED_X, A.n_cols)
    y64 = oracle.crs_spmv(A.row_ptr, A.col, A.val, x.astype(np.float64))
    scale = abs_csr(A) @ np.abs(x.astype(np.float64))
    val = stored(vt, A.val)
    rp, col, v, dx = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(val), ctx.upload(x)
    s = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, val.itemsize, 64, 4096)
    dy = ctx.zeros(A.n_rows, np.float32)
    s.apply(vt, dx, dy, unpermuted=True)
    check(ctx.download(dy), y64, scale, np.diff(A.row_ptr), U_VALUE[vt], f"SELL vt={vt}")
    s.destroy()
    for d in (rp, col, v, dx, dy):
        d.free()


def test_spmm_adjoint_sddmm_error_vs_fp64(ctx):
    """FP32 values, k = 32: Y = A X (CRS SpMM), A^T W (SELL layout of A^T)
    and the SDDMM, each against FP64."""
    k = 32
    A = matgen.lomax_band(100_000)
    v32 = A.val.astype(np.float32)
    X = matgen.uniform_f32(matgen.SEED_X, A.n_cols * k)
    W = matgen.uniform_f32(matgen.SEED_W, A.n_rows * k)
    X64, W64 = X.astype(np.float64).reshape(-1, k), W.astype(np.float64).reshape(-1, k)
    rp, col, v = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(v32)
    dX, dW = ctx.upload(X), ctx.upload(W)
    dY = ctx.zeros(A.n_rows * k, np.float32)
    ctx.crs_spmm(capi.F32, A.n_rows, A.n_cols, A.nnz, rp, col, v, k, True, dX, k, dY, k)
    Y64 = oracle.crs_spmm(A.row_ptr, A.col, A.val, k, True, X64.ravel()).reshape(-1, k)
    scale = abs_csr(A) @ np.abs(X64)
    check(ctx.download(dY).reshape(-1, k), Y64, scale, np.diff(A.row_ptr)[:, None], U32,
          "SpMM")
    T = ctx.scs_transpose_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, 4, 64, 4096)
    dG = ctx.zeros(A.n_cols * k, np.float32)
    T.apply_spmm(capi.F32, k, dW, dG)
    trp, tcol, tval = oracle.crs_transpose(A.n_rows, A.n_cols, A.row_ptr, A.col, A.val)
    At = matgen.CRS(A.n_cols, A.n_rows, trp, tcol, tval)
    G64 = oracle.crs_spmm(trp, tcol, tval, k, True, W64.ravel()).reshape(-1, k)
    check(ctx.download(dG).reshape(-1, k), G64, abs_csr(At) @ np.abs(W64),
          np.diff(trp)[:, None], U32, "A^T W")
    S = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, 4, 64, 4096)
    ddv = ctx.zeros(A.nnz, np.float32)
    S.sddmm(capi.F32, k, dW, dX, ddv)
    got = ctx.download(ddv).astype(np.float64)
    want = oracle.sddmm_f64(A.row_ptr, A.col, k, W64.ravel(), X64.ravel())
    rows = np.repeat(np.arange(A.n_rows), np.diff(A.row_ptr))
    sc = np.einsum("ij,ij->i", np.abs(W64[rows]), np.abs(X64[A.col]))
    g, q = sddmm_tree(k)
    # the FP64 sum's own rounding (k terms) is added to the FP32 tree's bound
    assert (np.abs(got - want) <= ((q + np.log2(g)) * U32 * (1 + 4 * U32)
                                   + float(gamma(k, 2.0 ** -53))) * sc).all()
    for d in (rp, col, v, dX, dW, dY, dG, ddv):
        d.free()
    S.destroy()
    T.destroy()


def test_bcrs_bf16_spmm_error_vs_fp64(ctx):
    """4x4 BF16 blocks of a 7-point block adjacency (16^3 blocks), k = 8."""
    k = 8
    A = matgen.block7(16, 4)
    v16 = matgen.to_bf16(A.val)
    X = matgen.uniform_f32(matgen.SEED_X, A.n_rows * k)
    rp, col, v = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(v16)
    B = ctx.bcrs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, 2, 4, 4, 4, 4, False)
    dX, dY = ctx.upload(X), ctx.zeros(A.n_rows * k, np.float32)
    B.spmm(capi.BF16, k, dX, dY)
    X64 = X.astype(np.float64)
    Y64 = oracle.crs_spmm(A.row_ptr, A.col, A.val, k, True, X64).reshape(-1, k)
    check(ctx.download(dY).reshape(-1, k), Y64, abs_csr(A) @ np.abs(X64.reshape(-1, k)),
          np.diff(A.row_ptr)[:, None], U_VALUE[capi.BF16], "BCRS BF16 SpMM")
    for d in (rp, col, v, dX, dY):
        d.free()
    B.destroy()
