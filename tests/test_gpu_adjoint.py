"""Differentiation: SELL layout of A^T, adjoint SpMV/SpMM, SDDMM value
gradient, and a thin torch.autograd.Function over the C ABI checked against
FP64 central finite differences."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from smaxkernels_amd import capi, matgen
from tests.helpers import assert_bitwise, fd_gradient_check, sddmm_tree

pytestmark = pytest.mark.gpu


def upload_crs(ctx, A, val=None):
    val = A.val if val is None else val
    return ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(val)


@pytest.mark.parametrize("C_,sigma", [(64, 1), (64, 4096), (32, 1 << 30), (1, 1)])
def test_transpose_layout_matches_explicit_transpose(ctx, C_, sigma):
    A = matgen.lomax_band(20000, half=2500)
    rp, col, v = upload_crs(ctx, A)
    T = ctx.scs_transpose_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, 8, C_, sigma)
    trp, tcol, tval = oracle.crs_transpose(A.n_rows, A.n_cols, A.row_ptr, A.col, A.val)
    want = oracle.crs_to_scs(trp, tcol, tval, C_, sigma)
    got = T.arrays(np.float64)
    for f in ("chunk_ptr", "chunk_lengths", "col", "perm"):
        assert_bitwise(got[f], want[f], f)
    assert_bitwise(got["val"], want["val"], "val")
    real = got["src_index"] >= 0
    assert_bitwise(got["val"][real], A.val[got["src_index"][real]], "src_index")
    # adjoint SpMV through the transpose layout == reference CRS SpMV of A^T
    g = matgen.uniform_f64(5, A.n_rows)
    dg, dz = ctx.upload(g), ctx.zeros(A.n_cols, np.float64)
    T.spmv(capi.F64, dg, dz, unpermuted=True)
    assert_bitwise(ctx.download(dz), oracle.crs_spmv(trp, tcol, tval, g), "A^T g")
    T.destroy()


def test_transpose_refresh_follows_new_values(ctx):
    A = matgen.lomax_band(5000, half=600)
    rp, col, v = upload_crs(ctx, A)
    T = ctx.scs_transpose_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, 8, 64, 256)
    new = matgen.uniform_f64(99, A.nnz)
    T.refresh_values(ctx.upload(new))
    trp, tcol, tval = oracle.crs_transpose(A.n_rows, A.n_cols, A.row_ptr, A.col, new)
    assert_bitwise(T.arrays(np.float64)["val"],
                   oracle.crs_to_scs(trp, tcol, tval, 64, 256)["val"])
    T.destroy()


@pytest.mark.parametrize("vt,k", [(capi.F32, 32), (capi.F64, 32), (capi.F32, 8),
                                  (capi.BF16, 64), (capi.F64, 2), (capi.F64, 16),
                                  (capi.F32, 4), (capi.F64, 64), (capi.F32, 1)])
def test_sddmm_row_and_element_kernels_agree(ctx, vt, k):
    """With X 16-byte aligned a power-of-two k takes the row kernel (16-byte
    loads); shifted by one entry it takes the per-element kernel.  Both
    evaluate the same tree, so dvals agree bit for bit."""
    A = matgen.lomax_band(6000, half=700)
    val = {capi.F64: A.val, capi.F32: A.val.astype(np.float32),
           capi.BF16: matgen.to_bf16(A.val)}[vt]
    rp, col, v = upload_crs(ctx, A, val)
    S = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, val.dtype.itemsize, 64, 256)
    vec = np.float64 if vt == capi.F64 else np.float32
    G = matgen.uniform_f64(16, A.n_rows * k).astype(vec)
    X = matgen.uniform_f64(17, A.n_cols * k).astype(vec)
    dG = ctx.upload(G)
    out = []
    for shift in (0, 1):
        # shift 1: X starts one entry (4 or 8 bytes) past a 16-byte boundary
        dX = ctx.upload(np.concatenate([X[:1], X]) if shift else X)
        assert dX.ptr % 16 == 0
        dv = ctx.zeros(A.nnz, vec)
        S.sddmm(vt, k, dG, dX.ptr + shift * X.itemsize, dv)
        out.append(ctx.download(dv))
        dv.free()
        dX.free()
    assert_bitwise(out[0], out[1], "row kernel vs element kernel")
    for d in (dG, rp, col, v):
        d.free()
    S.destroy()


@pytest.mark.parametrize("vt,k,C_", [(capi.F64, 1, 64), (capi.F64, 5, 64), (capi.F64, 32, 64),
                                     (capi.F32, 1, 64), (capi.F32, 32, 64), (capi.BF16, 32, 64),
                                     (capi.F32, 8, 7), (capi.F32, 64, 64), (capi.F64, 3, 32),
                                     (capi.F32, 16, 3)])
def test_sddmm(ctx, vt, k, C_):
    """k a power of two takes the row-ordered kernel, other k the per-element
    one; C = 3 and 7 leave workgroups partly unused."""
    A = matgen.lomax_band(20000, half=2500)
    val = {capi.F64: A.val, capi.F32: A.val.astype(np.float32),
           capi.BF16: matgen.to_bf16(A.val)}[vt]
    rp, col, v = upload_crs(ctx, A, val)
    S = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, val.dtype.itemsize,
                         C_, 512)
    vec = np.float64 if vt == capi.F64 else np.float32
    G = matgen.uniform_f64(6, A.n_rows * k).astype(vec)
    X = matgen.uniform_f64(7, A.n_cols * k).astype(vec)
    dv = ctx.zeros(A.nnz, vec)
    S.sddmm(vt, k, ctx.upload(G), ctx.upload(X), dv)
    got = ctx.download(dv)
    want = oracle.sddmm_f64(A.row_ptr, A.col, k, G.astype(np.float64),
                            X.astype(np.float64))
    rows = np.repeat(np.arange(A.n_rows), np.diff(A.row_ptr))
    # exact enough for an FP64 kernel too: x87 extended precision (64-bit
    # significand) products and sums
    Gl = G.astype(np.longdouble).reshape(-1, k)[rows]
    Xl = X.astype(np.longdouble).reshape(-1, k)[A.col]
    exact = (Gl * Xl).sum(axis=1)
    scale = np.abs(Gl * Xl).sum(axis=1)
    if k == 1:
        # one product, rounded once
        assert_bitwise(got, G[rows] * X[A.col], "k=1")
    u = np.finfo(vec).eps / 2
    g, q = sddmm_tree(k)
    # q - 1 chained fmas per lane, log2(g) butterfly additions, one product rounding
    bound = (q + np.log2(g)) * u * scale * (1 + 4 * u) + 1e-300
    assert (np.abs(got.astype(np.longdouble) - exact) <= bound).all()
    # the oracle's FP64 sums agree within their own rounding and ours
    assert (np.abs(got.astype(np.float64) - want) <=
            (k + np.ceil(np.log2(k)) + 2) * np.finfo(np.float64).eps * scale.astype(np.float64)
            + (u if vec == np.float32 else 0) * 8 * scale.astype(np.float64) + 1e-300).all()
    S.destroy()


# ------------------------------------------------------- autograd wrapper
torch = pytest.importorskip("torch")


class SpMMFunction(torch.autograd.Function):
    """Y = A X for a fixed CRS pattern, differentiable in A's values and X.

    forward: CRS SpMM (row-major X).  backward: dX = A^T dY through the SELL
    layout of A^T (values refreshed from the current A), dvals = SDDMM on A's
    SELL layout.  All three run in libsmax_hip.so through the C ABI."""

    @staticmethod
    def forward(fctx, vals, X, op):
        k = X.shape[1]
        Y = torch.empty(op.n_rows, k, dtype=X.dtype, device=X.device)
        op.ctx.crs_spmm(op.vt, op.n_rows, op.n_cols, op.nnz, op.rp, op.col, vals,
                        k, True, X, k, Y, k)
        fctx.save_for_backward(vals, X)
        fctx.op = op
        return Y

    @staticmethod
    def backward(fctx, dY):
        vals, X = fctx.saved_tensors
        op = fctx.op
        dY = dY.contiguous()
        k = X.shape[1]
        op.T.refresh_values(vals)
        dX = torch.empty_like(X)
        op.T.apply_spmm(op.vt, k, dY, dX)
        dvals = torch.empty_like(vals)
        op.S.sddmm(op.vt, k, dY, X, dvals)
        return dvals, dX, None


class SpMMOp:
    def __init__(self, ctx, A, vt):
        self.ctx, self.vt = ctx, vt
        self.n_rows, self.n_cols, self.nnz = A.n_rows, A.n_cols, A.nnz
        self.rp, self.col = ctx.upload(A.row_ptr), ctx.upload(A.col)
        v = ctx.upload(A.val.astype(np.float32))
        self.S = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, self.rp, self.col, v, 4,
                                  64, 4096)
        self.T = ctx.scs_transpose_from_crs(A.n_rows, A.n_cols, A.nnz, self.rp,
                                            self.col, v, 4, 64, 4096)


def test_autograd_matches_central_differences(ctx):
    """C4 at reduced N: FP32 values, k = 32, loss = sum(Y o W).  64 value and
    64 X entries; metric |g - g_FD| / max(|g_FD|, sum_v |W_iv X_jv|) <= 1e-6.
    L is linear in each entry, so the FP64 central difference (h = 2^-10
    max(1,|v|)) is exact up to rounding; it is evaluated on the rows the
    entry touches (every other term of L cancels identically)."""
    N, k = 100_000, 32
    A = matgen.lomax_band(N)
    dev = torch.device("cuda", 0)
    ctx.sync()
    op = SpMMOp(ctx, A, capi.F32)
    stream = torch.cuda.current_stream(dev)
    ctx.lib.smax_hip_set_stream(ctx.handle, stream.cuda_stream)
    try:
        vals = torch.tensor(A.val.astype(np.float32), device=dev, requires_grad=True)
        Xh = matgen.uniform_f32(matgen.SEED_X, N * k).reshape(N, k)
        Wh = matgen.uniform_f32(matgen.SEED_W, N * k).reshape(N, k)
        X = torch.tensor(Xh, device=dev, requires_grad=True)
        W = torch.tensor(Wh, device=dev)
        Y = SpMMFunction.apply(vals, X, op)
        (Y * W).sum().backward()
        torch.cuda.synchronize()
        gv = vals.grad.cpu().numpy().astype(np.float64)
        gx = X.grad.cpu().numpy().astype(np.float64)
        Yh = Y.detach().cpu().numpy()
    finally:
        ctx.lib.smax_hip_set_stream(ctx.handle, None)
    np.testing.assert_allclose(
        Yh, oracle.crs_spmm(A.row_ptr, A.col, A.val.astype(np.float32), k, True,
                            Xh.ravel()).reshape(N, k), rtol=0, atol=0)
    fd = fd_gradient_check(A, A.val.astype(np.float32), Xh, Wh, k, gv, gx)
    assert fd["max_rel_err"] <= 1e-6, fd


@pytest.mark.parametrize("vt,k", [(capi.F32, 32), (capi.F64, 16), (capi.F32, 5)])
def test_sddmm_on_transpose_layout_equals_sddmm_on_a(ctx, vt, k):
    """The SDDMM on A^T's SELL layout (its src_index points back at A's CRS
    positions, not consecutively) with G and X exchanged forms, for every
    stored element, the same products in the same reduction tree as the SDDMM
    on A's own layout: dvals agree bit for bit (the row kernel without and
    with consecutive destinations; k = 5 takes the per-element kernel)."""
    A = matgen.lomax_band(30000, half=4000)
    val = A.val if vt == capi.F64 else A.val.astype(np.float32)
    rp, col, v = upload_crs(ctx, A, val)
    S = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, val.dtype.itemsize, 64, 4096)
    T = ctx.scs_transpose_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, val.dtype.itemsize,
                                   64, 4096)
    vec = np.float64 if vt == capi.F64 else np.float32
    W = matgen.uniform_f64(31, A.n_rows * k).astype(vec)
    X = matgen.uniform_f64(32, A.n_cols * k).astype(vec)
    dW, dX = ctx.upload(W), ctx.upload(X)
    d1, d2 = ctx.zeros(A.nnz, vec), ctx.zeros(A.nnz, vec)
    S.sddmm(vt, k, dW, dX, d1)
    T.sddmm(vt, k, dX, dW, d2)
    assert_bitwise(ctx.download(d2), ctx.download(d1), "A^T layout vs A layout")
    for d in (rp, col, v, dW, dX, d1, d2):
        d.free()
    S.destroy()
    T.destroy()
