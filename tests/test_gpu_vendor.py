"""The rocSPARSE comparison (smaxkernels_amd/vendor/rocsparse_cmp.cpp, the
counterpart of the reference's cuSPARSE benches,
examples/benchmarks/cusparse_benchmarks/cusparse_crs_spmv_bench.cpp:92 and
the BSR/SpMM benches) run on the same device arrays as the HIP kernels.
rocSPARSE sums in its own order, so it agrees with the reference's fma
chains only to within the rounding of two summation orders: per output,
|y_lib - y_ours| <= 2 gamma_n (|A| |x|), gamma_n = n u / (1 - n u), n the
number of terms.  FP64 CSR SpMV must also agree to 1e-13 relative to the
largest |y|; FP32 BSR SpMM of 4x4 blocks agrees exactly (the same
sequential order within a block row)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from smaxkernels_amd import capi, matgen, vendor
from tests.helpers import assert_bitwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    return vendor.load()  # built by build(); a missing library fails the test


def gamma(n, u):
    n = np.asarray(n, np.float64)
    return n * u / (1 - n * u)


def abs_product(A, X, k):
    """|A| |X| in FP64 (row-major X with k columns); its own rounding is far
    below the bounds it scales."""
    import scipy.sparse as sp
    M = sp.csr_matrix((np.abs(A.val), A.col, A.row_ptr), shape=(A.n_rows, A.n_cols))
    return M @ np.abs(X.astype(np.float64)).reshape(-1, k)


def test_csr_spmv_f64_agrees_with_reference_chain(ctx, lib):
    A = matgen.lomax_band(200_000)
    x = matgen.uniform_f64(matgen.SEED_X, A.n_cols)
    rp, col, val, dx = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val), ctx.upload(x)
    dy = ctx.zeros(A.n_rows, np.float64)
    ctx.crs_spmv(capi.F64, A.n_rows, A.n_cols, A.nnz, rp, col, val, dx, dy, None)
    ours = ctx.download(dy)
    assert_bitwise(ours, oracle.crs_spmv(A.row_ptr, A.col, A.val, x), "ours vs oracle")
    bound = 2 * gamma(np.diff(A.row_ptr), 2.0 ** -53) * abs_product(A, x, 1)[:, 0]
    stream = ctx.lib.smax_hip_get_stream(ctx.handle)
    ran = 0
    for alg, name in vendor.SPMV_ALGS:
        h = lib.rocsparse_cmp_create(alg, A.n_rows, A.n_cols, A.nnz, rp.ptr, col.ptr, val.ptr,
                                     dx.ptr, dy.ptr, stream)
        if not h:
            continue  # an algorithm rocSPARSE does not offer for this input
        ctx._check(ctx.lib.smax_hip_zero(ctx.handle, dy.ptr, dy.nbytes))
        assert lib.rocsparse_cmp_spmv(h) == 0, name
        got = ctx.download(dy)
        lib.rocsparse_cmp_destroy(h)
        d = np.abs(got - ours)
        assert (d <= bound).all(), name
        assert d.max() / np.abs(ours).max() <= 1e-13, name
        ran += 1
    assert ran >= 2
    for a in (rp, col, val, dx, dy):
        a.free()


def test_csr_spmm_f32_agrees_with_reference_chain(ctx, lib):
    k = 32
    A = matgen.lomax_band(100_000)
    v32 = A.val.astype(np.float32)
    X = matgen.uniform_f32(matgen.SEED_X, A.n_cols * k)
    rp, col, val, dX = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(v32), ctx.upload(X)
    dY = ctx.zeros(A.n_rows * k, np.float32)
    ctx.crs_spmm(capi.F32, A.n_rows, A.n_cols, A.nnz, rp, col, val, k, True, dX, k, dY, k)
    ours = ctx.download(dY)
    assert_bitwise(ours, oracle.crs_spmm(A.row_ptr, A.col, v32, k, True, X), "ours vs oracle")
    A32 = matgen.CRS(A.n_rows, A.n_cols, A.row_ptr, A.col, v32.astype(np.float64))
    bound = (2 * gamma(np.diff(A.row_ptr), 2.0 ** -24)[:, None] * abs_product(A32, X, k)).ravel()
    stream = ctx.lib.smax_hip_get_stream(ctx.handle)
    ran = 0
    for alg, name in vendor.CSR_SPMM_ALGS:
        h = lib.rocsparse_cmp_spmm_create(alg, 0, A.n_rows, A.n_cols, A.nnz, rp.ptr, col.ptr,
                                          val.ptr, k, dX.ptr, dY.ptr, stream)
        if not h:
            continue
        ctx._check(ctx.lib.smax_hip_zero(ctx.handle, dY.ptr, dY.nbytes))
        assert lib.rocsparse_cmp_spmm(h) == 0, name
        got = ctx.download(dY)
        lib.rocsparse_cmp_spmm_destroy(h)
        assert (np.abs(got.astype(np.float64) - ours) <= bound).all(), name
        ran += 1
    assert ran >= 2
    for a in (rp, col, val, dX, dY):
        a.free()


@pytest.mark.parametrize("k", [8, 64])
def test_bsr_spmm_f32_equals_bcrs_spmm(ctx, lib, k):
    """4x4 row-major blocks of a 7-point block adjacency (24^3 blocks): our
    BCRS SpMM and rocSPARSE's BSR SpMM give the same Y, bit for bit."""
    A = matgen.block7(24, 4)
    v32 = A.val.astype(np.float32)
    rp, col, v = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(v32)
    B = ctx.bcrs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, 4, 4, 4, 4, 4, False)
    for a in (rp, col, v):
        a.free()
    i = B.info
    X = matgen.uniform_f32(matgen.SEED_X, A.n_rows * k)
    dX, dY = ctx.upload(X), ctx.zeros(A.n_rows * k, np.float32)
    B.spmm(capi.F32, k, dX, dY)
    ours = ctx.download(dY)
    assert_bitwise(ours, oracle.crs_spmm(A.row_ptr, A.col, v32, k, True, X), "ours vs oracle")
    h = lib.rocsparse_cmp_spmm_create(vendor.BSR_SPMM, 4, i.n_block_rows, i.n_block_cols,
                                      i.n_blocks, i.row_ptr, i.col, i.val, k, dX.ptr, dY.ptr,
                                      ctx.lib.smax_hip_get_stream(ctx.handle))
    assert h, "rocSPARSE BSR SpMM not available"
    ctx._check(ctx.lib.smax_hip_zero(ctx.handle, dY.ptr, dY.nbytes))
    assert lib.rocsparse_cmp_spmm(h) == 0
    got = ctx.download(dY)
    lib.rocsparse_cmp_spmm_destroy(h)
    assert_bitwise(got, ours, f"rocSPARSE BSR k={k}")
    for a in (dX, dY):
        a.free()
    B.destroy()
