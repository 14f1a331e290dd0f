"""GPU CRS SpMV (lane-per-row and LDS row-block schedules) and CRS SpMM
against the oracle: every output bit-identical to the reference's fma chain."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from smaxkernels_amd import capi, matgen
from tests.helpers import almost_equal, arr, assert_bitwise, reference_kats

pytestmark = pytest.mark.gpu
VTS = [capi.F64, capi.F32, capi.BF16]


def operands(vt, A, k=1, seed=21):
    if vt == capi.F64:
        return A.val, matgen.uniform_f64(seed, A.n_cols * k)
    if vt == capi.F32:
        return A.val.astype(np.float32), matgen.uniform_f32(seed, A.n_cols * k)
    return matgen.to_bf16(A.val), matgen.uniform_f32(seed, A.n_cols * k)


def run_spmv(ctx, vt, A, val, x, blocked):
    rp, col, v, dx = (ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(val),
                      ctx.upload(x))
    dy = ctx.zeros(A.n_rows, x.dtype)
    plan = ctx.crs_plan(A.n_rows, A.nnz, rp) if blocked else None
    ctx.crs_spmv(vt, A.n_rows, A.n_cols, A.nnz, rp, col, v, dx, dy, plan)
    y = ctx.download(dy)
    if plan:
        ctx.crs_plan_destroy(plan)
    for d in (rp, col, v, dx, dy):
        d.free()
    return y


class Csr:
    def __init__(self, n_rows, n_cols, rp, col, val):
        self.n_rows, self.n_cols = n_rows, n_cols
        self.row_ptr, self.col, self.val = rp, col, val

    @property
    def nnz(self):
        return int(self.row_ptr[-1])


def test_kat_spmv_on_gpu(ctx):
    k = reference_kats()["spmv_test.hpp"]
    A = Csr(3, 3, arr(k, "A->row_ptr", np.int32), arr(k, "A->col", np.int32),
            arr(k, "A->val", np.float64))
    for blocked in (False, True):
        y = run_spmv(ctx, capi.F64, A, A.val, np.ones(3), blocked)
        assert almost_equal(y, arr(k, "y_expected", np.float64)).all()


@pytest.mark.parametrize("vt", VTS)
@pytest.mark.parametrize("blocked", [False, True])
def test_crs_spmv_lomax_bit_exact(ctx, vt, blocked):
    A = matgen.lomax_band(60000, half=7500)
    val, x = operands(vt, A)
    assert_bitwise(run_spmv(ctx, vt, A, val, x, blocked),
                   oracle.crs_spmv(A.row_ptr, A.col, val, x))


@pytest.mark.parametrize("blocked", [False, True])
def test_crs_spmv_config1_poisson(ctx, blocked):
    A = matgen.poisson2d(512)
    x = matgen.uniform_f64(matgen.SEED_X, A.n_cols)
    assert_bitwise(run_spmv(ctx, capi.F64, A, A.val, x, blocked),
                   oracle.crs_spmv(A.row_ptr, A.col, A.val, x))


@pytest.mark.parametrize("vt", VTS)
def test_crs_spmv_long_and_empty_rows(ctx, vt):
    """Rows across the 512-long-row and 1024-tile boundaries, empty rows,
    a dense row, a single-row matrix and an all-empty matrix."""
    rng = np.random.default_rng(9)
    n = 4000
    L = rng.integers(0, 20, n)
    L[rng.random(n) < 0.2] = 0
    L[[5, 6, 7]] = [511, 512, 513]
    L[[100, 1000]] = [1023, 1025]
    L[2000] = n
    L[3000:3300] = 0
    rp = np.concatenate([[0], np.cumsum(L)]).astype(np.int32)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in L]
                         ).astype(np.int32)
    A = Csr(n, n, rp, col, rng.uniform(-1, 1, rp[-1]))
    val, x = operands(vt, A)
    for blocked in (False, True):
        assert_bitwise(run_spmv(ctx, vt, A, val, x, blocked),
                       oracle.crs_spmv(A.row_ptr, A.col, val, x), f"blocked={blocked}")
    one = Csr(1, n, np.array([0, 3000], np.int32),
              np.arange(0, 3000, dtype=np.int32), rng.uniform(-1, 1, 3000))
    val1, x1 = operands(vt, one)
    for blocked in (False, True):
        assert_bitwise(run_spmv(ctx, vt, one, val1, x1, blocked),
                       oracle.crs_spmv(one.row_ptr, one.col, val1, x1))
    Z = Csr(300, 300, np.zeros(301, np.int32), np.zeros(0, np.int32), np.zeros(0))
    valz, xz = operands(vt, Z)
    for blocked in (False, True):
        assert not run_spmv(ctx, vt, Z, valz, xz, blocked).any()


@pytest.mark.parametrize("vt", VTS)
@pytest.mark.parametrize("k", [1, 4, 8, 32, 64])
@pytest.mark.parametrize("row_major", [True, False])
@pytest.mark.parametrize("planned", [False, True])
def test_crs_spmm_bit_exact(ctx, vt, k, row_major, planned):
    """With a plan, row-major products run in the plan's row order (long rows
    first, then by length per window): the same chains, the same Y."""
    A = matgen.lomax_band(20000, half=2500)
    val, X = operands(vt, A, k)
    rp, col, v, dX = (ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(val),
                      ctx.upload(X))
    dY = ctx.zeros(A.n_rows * k, X.dtype)
    ld = k if row_major else A.n_rows
    plan = ctx.crs_plan(A.n_rows, A.nnz, rp) if planned else None
    ctx.crs_spmm(vt, A.n_rows, A.n_cols, A.nnz, rp, col, v, k, row_major, dX, ld,
                 dY, ld, plan)
    assert_bitwise(ctx.download(dY),
                   oracle.crs_spmm(A.row_ptr, A.col, val, k, row_major, X))
    if plan:
        ctx.crs_plan_destroy(plan)
    for d in (rp, col, v, dX, dY):
        d.free()


@pytest.mark.parametrize("vt", VTS)
@pytest.mark.parametrize("pad_x,pad_y", [(5, 3), (4, 8)])
def test_crs_spmm_column_major_padded_and_growing(ctx, vt, pad_x, pad_y):
    """Column-major operands with leading dimensions past n_rows, widths off
    the 32-wide staging tile, and a workspace that grows (k = 64 after 37)
    and is then reused (k = 3): Y equals the reference column by column, and
    the padding between columns is left untouched.  Odd pads stage through
    the element-wise relayout; pads of 4 and 8 (with k = 64) through the
    16-byte one, whose last row tile (9000 = 140 * 64 + 40) is partial."""
    A = matgen.lomax_band(9000, half=1000)
    n = A.n_rows
    rp, col = ctx.upload(A.row_ptr), ctx.upload(A.col)
    for k in (37, 64, 3):
        val, X = operands(vt, A, k)  # X: n x k column-major, ld n
        v = ctx.upload(val)
        ldx, ldy = n + pad_x, n + pad_y
        Xp = np.zeros(ldx * k, X.dtype)
        for c in range(k):
            Xp[c * ldx:c * ldx + n] = X[c * n:(c + 1) * n]
        dX = ctx.upload(Xp)
        sentinel = np.full(ldy * k, 7.0, X.dtype)
        dY = ctx.upload(sentinel)
        ctx.crs_spmm(vt, n, n, A.nnz, rp, col, v, k, False, dX, ldx, dY, ldy)
        got = ctx.download(dY).reshape(k, ldy)
        want = oracle.crs_spmm(A.row_ptr, A.col, val, k, False, X).reshape(k, n)
        assert_bitwise(got[:, :n], want, f"k={k}")
        assert (got[:, n:] == 7.0).all()
        for d in (v, dX, dY):
            d.free()
    for d in (rp, col):
        d.free()


@pytest.mark.parametrize("vt", VTS)
def test_crs_spmm_row_major_unaligned_operands(ctx, vt):
    """k = 32 row-major with X and Y one entry past a 16-byte boundary takes
    the one-column-per-lane kernel instead of the 16-byte one; both run the
    same chains, so both equal the reference."""
    k = 32
    A = matgen.lomax_band(8000, half=900)
    val, X = operands(vt, A, k)
    rp, col, v = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(val)
    dX = ctx.upload(np.concatenate([X[:1], X]))
    dY = ctx.zeros(A.n_rows * k + 1, X.dtype)
    it = X.itemsize
    ctx.crs_spmm(vt, A.n_rows, A.n_cols, A.nnz, rp, col, v, k, True, dX.ptr + it, k,
                 dY.ptr + it, k)
    assert_bitwise(ctx.download(dY)[1:],
                   oracle.crs_spmm(A.row_ptr, A.col, val, k, True, X))
    for d in (rp, col, v, dX, dY):
        d.free()


def test_kat_spmm_on_gpu(ctx):
    kats = reference_kats()
    for name, row_major in (("spmm_test.hpp", False), ("spmm_rowmaj_test.hpp", True)):
        k = kats[name]
        nv = k["consts"]["n_vectors"]["value"]
        rp, col, v = (ctx.upload(arr(k, "A->row_ptr", np.int32)),
                      ctx.upload(arr(k, "A->col", np.int32)),
                      ctx.upload(arr(k, "A->val", np.float64)))
        dX, dY = ctx.upload(np.ones(3 * nv)), ctx.zeros(3 * nv, np.float64)
        ld = nv if row_major else 3
        ctx.crs_spmm(capi.F64, 3, 3, 5, rp, col, v, nv, row_major, dX, ld, dY, ld)
        assert almost_equal(ctx.download(dY), arr(k, "Y_expected", np.float64)).all()


def rectangular(A, n_rows=None, n_cols=None):
    """The first n_rows rows of A with the columns >= n_cols dropped."""
    n_rows = A.n_rows if n_rows is None else n_rows
    n_cols = A.n_cols if n_cols is None else n_cols
    keep = A.col < n_cols
    rows = np.repeat(np.arange(A.n_rows), np.diff(A.row_ptr))
    keep &= rows < n_rows
    lens = np.bincount(rows[keep], minlength=A.n_rows)[:n_rows]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    return Csr(n_rows, n_cols, rp, A.col[keep].copy(), A.val[keep].copy())


@pytest.mark.parametrize("vt", VTS)
@pytest.mark.parametrize("shape", ["wide", "tall"])
def test_crs_spmm_column_major_rectangular(ctx, vt, shape):
    """Column-major products with n_cols != n_rows: X is staged over n_cols
    rows and Y over n_rows, so a swap of the two extents would show here.
    Padded leading dimensions (ldx >= n_cols, ldy > n_rows) leave the padding
    untouched; each column of Y equals the oracle's SpMV of that column of X.
    The reference's own leading dimension (ldx = ldy = n_rows, columns of X
    overlapping when n_cols > n_rows) matches the oracle's column-major SpMM."""
    B = matgen.lomax_band(9000, half=1000)
    A = rectangular(B, n_rows=6000) if shape == "wide" else rectangular(B, n_cols=6000)
    n, m = A.n_rows, A.n_cols
    rp, col = ctx.upload(A.row_ptr), ctx.upload(A.col)
    for k in (5, 32):
        val, _ = operands(vt, A)
        v = ctx.upload(val)
        X = operands(vt, A, k, seed=40 + k)[1].reshape(k, m)
        ldx, ldy = m + 7, n + 3
        Xp = np.zeros((k, ldx), X.dtype)
        Xp[:, :m] = X
        dX = ctx.upload(Xp.ravel())
        dY = ctx.upload(np.full(ldy * k, 7.0, X.dtype))
        ctx.crs_spmm(vt, n, m, A.nnz, rp, col, v, k, False, dX, ldx, dY, ldy)
        got = ctx.download(dY).reshape(k, ldy)
        for c in range(k):
            assert_bitwise(got[c, :n], oracle.crs_spmv(A.row_ptr, A.col, val, X[c]),
                           f"{shape} k={k} column {c}")
        assert (got[:, n:] == 7.0).all()
        # the reference's layout: leading dimension n_rows for X and Y
        need = (k - 1) * n + m
        Xr = (matgen.uniform_f64(90 + k, need) if vt == capi.F64
              else matgen.uniform_f32(90 + k, need))
        dXr = ctx.upload(Xr)
        dYr = ctx.zeros(n * k, X.dtype)
        ctx.crs_spmm(vt, n, m, A.nnz, rp, col, v, k, False, dXr, n, dYr, n)
        assert_bitwise(ctx.download(dYr), oracle.crs_spmm(A.row_ptr, A.col, val, k, False, Xr),
                       f"{shape} k={k} ld = n_rows")
        for d in (v, dX, dY, dXr, dYr):
            d.free()
    for d in (rp, col):
        d.free()


def test_crs_spmm_column_major_refuses_overlapping_y(ctx):
    """A column-major Y whose columns would overlap (ldy < n_rows) is refused
    before anything is launched; with k = 1 the leading dimension is unused."""
    A = matgen.lomax_band(3000, half=300)
    rp, col, v = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val)
    k = 4
    dX = ctx.upload(matgen.uniform_f64(3, A.n_cols * k))
    dY = ctx.upload(np.full(A.n_rows * k, 5.0))
    lib = ctx.lib
    for ldy in (A.n_rows - 1, 0, -A.n_rows):
        st = lib.smax_hip_crs_spmm(ctx.handle, capi.F64, A.n_rows, A.n_cols, A.nnz, rp.ptr,
                                   col.ptr, v.ptr, k, 0, dX.ptr, A.n_rows, dY.ptr, ldy, None)
        assert st == capi.ERROR_INVALID_VALUE, ldy
        assert b"ldy" in lib.smax_hip_last_error(ctx.handle)
    ctx.sync()
    assert (ctx.download(dY) == 5.0).all()
    st = lib.smax_hip_crs_spmm(ctx.handle, capi.F64, A.n_rows, A.n_cols, A.nnz, rp.ptr,
                               col.ptr, v.ptr, 1, 0, dX.ptr, A.n_rows, dY.ptr, 0, None)
    assert st == capi.SUCCESS
    assert_bitwise(ctx.download(dY)[:A.n_rows],
                   oracle.crs_spmv(A.row_ptr, A.col, A.val, ctx.download(dX)[:A.n_cols]))
    for d in (rp, col, v, dX, dY):
        d.free()


def test_crs_spmm_column_major_graph_survives_workspace_growth(ctx):
    """A HIP graph captured with a column-major SpMM keeps its staging
    workspace: a later, larger product grows the context's workspace, and
    every replay of the old graph (eager work and synchronisation between
    replays) still equals the eager result.  A context of its own, so the
    workspace starts empty whatever ran before."""
    torch = pytest.importorskip("torch")
    with capi.Context(0) as own:
        _graph_survives_growth(own, torch)


def _graph_survives_growth(ctx, torch):
    A = matgen.lomax_band(5000, half=600)
    k = 8
    rp, col, v = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val)
    X = matgen.uniform_f64(11, A.n_cols * k)
    dX, dY = ctx.upload(X), ctx.zeros(A.n_rows * k, np.float64)
    want = oracle.crs_spmm(A.row_ptr, A.col, A.val, k, False, X)
    call = lambda: ctx.crs_spmm(capi.F64, A.n_rows, A.n_cols, A.nnz, rp, col, v, k, False,
                                dX, A.n_rows, dY, A.n_rows)
    st = torch.cuda.Stream()
    ctx.set_stream(st.cuda_stream)
    try:
        call()  # sizes the workspace before capture
        ctx.sync()
        assert_bitwise(ctx.download(dY), want, "eager")
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=st):
            call()
        # a larger product (k = 64 on a bigger matrix) grows the workspace
        B = matgen.lomax_band(40000, half=5000)
        kb = 64
        brp, bcol, bv = ctx.upload(B.row_ptr), ctx.upload(B.col), ctx.upload(B.val)
        bX = matgen.uniform_f64(12, B.n_cols * kb)
        dbX, dbY = ctx.upload(bX), ctx.zeros(B.n_rows * kb, np.float64)
        ctx.crs_spmm(capi.F64, B.n_rows, B.n_cols, B.nnz, brp, bcol, bv, kb, False, dbX,
                     B.n_rows, dbY, B.n_rows)
        ctx.sync()
        assert_bitwise(ctx.download(dbY), oracle.crs_spmm(B.row_ptr, B.col, B.val, kb, False, bX),
                       "larger product")
        for rep in range(3):
            ctx._check(ctx.lib.smax_hip_zero(ctx.handle, dY.ptr, dY.nbytes))
            other = torch.ones(1 << 20, device="cuda") * (rep + 1)  # eager work between replays
            with torch.cuda.stream(st):
                g.replay()
            torch.cuda.synchronize()
            assert float(other.sum()) == float((rep + 1) << 20)
            assert_bitwise(ctx.download(dY), want, f"replay {rep}")
        for d in (brp, bcol, bv, dbX, dbY):
            d.free()
    finally:
        ctx.set_stream(None)
    for d in (rp, col, v, dX, dY):
        d.free()
