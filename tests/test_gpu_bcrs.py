"""GPU CRS -> BCRS conversion, BCRS SpMV and BCRS SpMM against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from smaxkernels_amd import capi, matgen
from tests.helpers import almost_equal, arr, assert_bitwise, reference_kats

pytestmark = pytest.mark.gpu
KATS = reference_kats()


def gpu_bcrs(ctx, n_rows, n_cols, rp, col, val, bh, bw, hp, wp, cm):
    d = [ctx.upload(a) for a in (rp, col, val)]
    b = ctx.bcrs_from_crs(n_rows, n_cols, int(rp[-1]), d[0], d[1], d[2],
                          val.dtype.itemsize, bh, bw, hp, wp, cm)
    for a in d:
        a.free()
    return b


@pytest.mark.parametrize("name", ["spmv_bcrs_1_1_test_1.hpp",
                                  "spmv_bcrs_1_1_test_2.hpp"])
def test_kat_bcrs_conversion_on_gpu(ctx, name):
    k = KATS[name]
    c = {n: v["value"] for n, v in k["consts"].items()}
    b = gpu_bcrs(ctx, 4, 4, arr(k, "A_crs->row_ptr", np.int32),
                 arr(k, "A_crs->col", np.int32), arr(k, "A_crs->val", np.float64),
                 c["target_b_height"], c["target_b_width"], c["target_height_pad"],
                 c["target_width_pad"], c["use_blocked_column_major"])
    got = b.arrays(np.float64)
    assert b.info.n_blocks == k["ctors"]["A_bcrs_exp"]["args"][2]
    assert_bitwise(got["col"], arr(k, "A_bcrs_exp->col", np.int32))
    assert_bitwise(got["row_ptr"], arr(k, "A_bcrs_exp->row_ptr", np.int32))
    assert almost_equal(got["val"], arr(k, "A_bcrs_exp->val", np.float64)).all()
    b.destroy()


def test_kat_bcrs_spmv_on_gpu(ctx):
    k = KATS["spmv_bcrs_1_1_test_3.hpp"]
    nbr, nbc, nb, bh, bw, hp, wp = k["ctors"]["A_bcrs"]["args"]
    d = [ctx.upload(arr(k, n, t)) for n, t in (("A_bcrs->col", np.int32),
                                               ("A_bcrs->row_ptr", np.int32),
                                               ("A_bcrs->val", np.float64))]
    dx, dy = ctx.upload(np.ones(nbc * wp)), ctx.zeros(nbr * hp, np.float64)
    ctx.bcrs_spmv(capi.F64, nbr, bh, bw, hp, wp, d[0], d[1], d[2], True, dx, dy)
    assert almost_equal(ctx.download(dy), arr(k, "y_expected", np.float64)).all()


@pytest.mark.parametrize("cm", [True, False])
def test_laplace_blocked_equals_crs_on_gpu(ctx, cm):
    """spmv_bcrs_laplace_test.hpp: 16x16 1-D Laplacian, 4x8 blocks."""
    from tests.test_oracle import laplace1d
    rp, col, val = laplace1d(16)
    b = gpu_bcrs(ctx, 16, 16, rp, col, val, 4, 8, 4, 8, cm)
    x = 1.0 / (((np.arange(16) + 3) % 23) + 1.0)
    dx, dy = ctx.upload(x), ctx.zeros(16, np.float64)
    i = b.info
    ctx.bcrs_spmv(capi.F64, i.n_block_rows, 4, 8, 4, 8, i.col, i.row_ptr, i.val,
                  cm, dx, dy)
    assert almost_equal(ctx.download(dy), oracle.crs_spmv(rp, col, val, x)).all()
    b.destroy()


def vals_for(vt, v64):
    if vt == capi.F64:
        return v64
    if vt == capi.F32:
        return v64.astype(np.float32)
    return matgen.to_bf16(v64)


@pytest.mark.parametrize("vt", [capi.F64, capi.F32, capi.BF16])
@pytest.mark.parametrize("bh,bw,hp,wp,cm", [(4, 4, 4, 4, True), (4, 4, 4, 4, False),
                                            (4, 4, 5, 6, True), (2, 4, 3, 4, False),
                                            (2, 4, 2, 4, True), (2, 2, 2, 2, False)])
def test_block7_conversion_spmv_spmm(ctx, vt, bh, bw, hp, wp, cm):
    A = matgen.block7(10, 4)  # 4000 rows, 4x4 dense blocks
    val = vals_for(vt, A.val)
    b = gpu_bcrs(ctx, A.n_rows, A.n_cols, A.row_ptr, A.col, val, bh, bw, hp, wp, cm)
    want = oracle.crs_to_bcrs(A.n_rows, A.n_cols, A.row_ptr, A.col, val, bh, bw, hp,
                              wp, cm)
    got = b.arrays(val.dtype)
    assert b.info.n_blocks == want["n_blocks"]
    for f in ("col", "row_ptr"):
        assert_bitwise(got[f], want[f], f)
    assert_bitwise(got["val"].view(np.uint8), want["val"].view(np.uint8), "val")
    i = b.info
    xt = np.float64 if vt == capi.F64 else np.float32
    x = matgen.uniform_f64(3, i.n_block_cols * wp).astype(xt)
    dx, dy = ctx.upload(x), ctx.zeros(i.n_block_rows * hp, xt)
    ctx.bcrs_spmv(vt, i.n_block_rows, bh, bw, hp, wp, i.col, i.row_ptr, i.val, cm,
                  dx, dy)
    yw = oracle.bcrs_spmv(i.n_block_rows, bh, bw, hp, wp, want["col"],
                          want["row_ptr"], want["val"], cm, x)
    rows = (np.arange(i.n_block_rows * hp) % hp) < bh
    assert_bitwise(ctx.download(dy)[rows], yw[rows], "bcrs spmv")
    # shift 1: X and Y one entry past a 16-byte boundary (no 16-byte runs; the
    # same chains).  k = 8, 10 and misaligned 16 with 4-byte X entries and
    # dense blocks take column pairs per lane
    for k, shift in ((1, 0), (6, 0), (8, 0), (8, 1), (10, 0), (16, 0), (16, 1), (32, 0),
                     (64, 0)):
        X = matgen.uniform_f64(4, i.n_block_cols * wp * k).astype(xt)
        n_y = i.n_block_rows * hp * k
        dX = ctx.upload(np.concatenate([X[:shift], X]))
        dY = ctx.zeros(n_y + shift, xt)
        ctx.bcrs_spmm(vt, i.n_block_rows, bh, bw, hp, wp, i.col, i.row_ptr, i.val,
                      cm, k, dX.ptr + shift * X.itemsize, dY.ptr + shift * X.itemsize)
        Yw = oracle.bcrs_spmm(i.n_block_rows, bh, bw, hp, wp, want["col"],
                              want["row_ptr"], want["val"], cm, k, X)
        rk = np.repeat(rows, k)
        assert_bitwise(ctx.download(dY)[shift:][rk], Yw[rk], f"bcrs spmm k={k} shift={shift}")
        dX.free()
        dY.free()
    b.destroy()


@pytest.mark.parametrize("vt", [capi.F32, capi.BF16])
@pytest.mark.parametrize("cm", [True, False])
@pytest.mark.parametrize("k", [8, 16, 32, 64])
def test_bcrs_spmm_ragged_block_rows(ctx, vt, cm, k):
    """Dense 4x4 blocks at k = 64 (one value per lane, broadcast across the
    block row's 16 lanes) and k = 8 and 16 (a quarter block per lane,
    broadcast across the row's 4 lanes): empty block rows, rows of 1-9 blocks (every
    remainder of the 4-block groups) and rows of 40 blocks, with a block row
    count that leaves the last workgroup partly empty; bit for bit against
    the oracle."""
    rng = np.random.default_rng(64)
    nbr, nbc = 203, 300
    lens = rng.integers(1, 10, nbr)
    lens[::5] = 0
    lens[3::17] = 40
    row_ptr = np.zeros(nbr + 1, np.int32)
    row_ptr[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(nbc, n, replace=False)) for n in lens]
                         ).astype(np.int32)
    val = vals_for(vt, rng.uniform(-1.0, 1.0, int(row_ptr[-1]) * 16))
    X = rng.uniform(-1.0, 1.0, nbc * 4 * k).astype(np.float32)
    d = [ctx.upload(a) for a in (col, row_ptr, val, X)]
    dY = ctx.upload(np.full(nbr * 4 * k, 7.0, np.float32))
    ctx.bcrs_spmm(vt, nbr, 4, 4, 4, 4, d[0], d[1], d[2], cm, k, d[3], dY)
    want = oracle.bcrs_spmm(nbr, 4, 4, 4, 4, col, row_ptr, val, cm, k, X)
    got = ctx.download(dY)
    assert_bitwise(got, want, f"bcrs spmm k={k}")
    assert (got.reshape(nbr, -1)[lens == 0] == 0).all()
    for a in (*d, dY):
        a.free()


def test_bcrs_scalar_equivalence(ctx):
    """BCRS SpMV of the 7-point block matrix equals the CRS SpMV up to
    summation order (rows of the block matrix padded by zeros)."""
    A = matgen.block7(12, 4)
    b = gpu_bcrs(ctx, A.n_rows, A.n_cols, A.row_ptr, A.col, A.val, 4, 4, 4, 4, True)
    x = matgen.uniform_f64(8, A.n_cols)
    dx, dy = ctx.upload(x), ctx.zeros(A.n_rows, np.float64)
    i = b.info
    ctx.bcrs_spmv(capi.F64, i.n_block_rows, 4, 4, 4, 4, i.col, i.row_ptr, i.val,
                  True, dx, dy)
    want = oracle.crs_spmv(A.row_ptr, A.col, A.val, x)
    np.testing.assert_allclose(ctx.download(dy), want, rtol=1e-13, atol=1e-13)
    b.destroy()


@pytest.mark.parametrize("bh,bw", [(3, 3), (8, 2), (16, 4), (6, 1)])
def test_bcrs_conversion_other_block_heights(ctx, bh, bw):
    """Block heights outside the register-specialised {1, 2, 4} (8 included)
    take the generic block-row merge: arrays equal the oracle's conversion
    and the SpMV equals the oracle's BCRS SpMV."""
    A = matgen.lomax_band(4800, half=300)  # ascending columns; 4800 = 3*16*100
    b = gpu_bcrs(ctx, A.n_rows, A.n_cols, A.row_ptr, A.col, A.val, bh, bw, bh, bw, False)
    want = oracle.crs_to_bcrs(A.n_rows, A.n_cols, A.row_ptr, A.col, A.val, bh, bw, bh, bw,
                              False)
    got = b.arrays(np.float64)
    assert b.info.n_blocks == want["n_blocks"]
    for f in ("col", "row_ptr"):
        assert_bitwise(got[f], want[f], f)
    assert_bitwise(got["val"], want["val"], "val")
    i = b.info
    x = matgen.uniform_f64(9, A.n_cols)
    dx, dy = ctx.upload(x), ctx.zeros(A.n_rows, np.float64)
    ctx.bcrs_spmv(capi.F64, i.n_block_rows, bh, bw, bh, bw, i.col, i.row_ptr, i.val, False,
                  dx, dy)
    assert_bitwise(ctx.download(dy), oracle.bcrs_spmv(i.n_block_rows, bh, bw, bh, bw,
                                                      want["col"], want["row_ptr"],
                                                      want["val"], False, x))
    for d in (dx, dy):
        d.free()
    b.destroy()


def test_bcrs_conversion_duplicate_entries_last_wins(ctx):
    """A CRS row may list a column twice (Matrix Market files can): the
    reference places the values in order, so the later one stays; the GPU
    conversion places each row's values in order too."""
    rp = np.array([0, 3, 4, 6, 7], np.int32)
    col = np.array([0, 1, 1, 2, 2, 2, 3], np.int32)
    val = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0])
    b = gpu_bcrs(ctx, 4, 4, rp, col, val, 2, 2, 2, 2, False)
    want = oracle.crs_to_bcrs(4, 4, rp, col, val, 2, 2, 2, 2, False)
    got = b.arrays(np.float64)
    for f in ("col", "row_ptr", "val"):
        assert_bitwise(got[f], want[f], f)
    b.destroy()


def mixed_row_lengths(n, short, long_from, long_len, seed):
    """n x n CRS: rows below long_from hold `short` ascending columns near the
    diagonal, the others long_len distinct ascending random columns."""
    rng = np.random.default_rng(seed)
    cols = []
    for r in range(n):
        if r < long_from:
            c = np.clip(np.arange(r - short // 2, r - short // 2 + short), 0, n - 1)
            cols.append(np.unique(c))
        else:
            cols.append(np.sort(rng.choice(n, long_len, replace=False)))
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in cols])
    col = np.concatenate(cols).astype(np.int32)
    return rp, col, rng.uniform(-1.0, 1.0, int(rp[-1]))


@pytest.mark.parametrize("vt", [capi.F64, capi.F32, capi.BF16])
@pytest.mark.parametrize("cm", [True, False])
def test_bcrs_conversion_long_rows_read_from_memory(ctx, vt, cm):
    """Workgroups whose CRS range exceeds the LDS tile merge straight from
    memory: rows 0..511 are short (staged), rows 512..1023 hold 600 entries
    each (256 such rows per workgroup, far past the tile)."""
    rp, col, v64 = mixed_row_lengths(1024, 3, 512, 600, 11)
    val = vals_for(vt, v64)
    b = gpu_bcrs(ctx, 1024, 1024, rp, col, val, 4, 4, 4, 4, cm)
    want = oracle.crs_to_bcrs(1024, 1024, rp, col, val, 4, 4, 4, 4, cm)
    got = b.arrays(val.dtype)
    assert b.info.n_blocks == want["n_blocks"]
    for f in ("col", "row_ptr"):
        assert_bitwise(got[f], want[f], f)
    assert_bitwise(got["val"].view(np.uint8), want["val"].view(np.uint8), "val")
    b.destroy()


def test_bcrs_conversion_of_a_matrix_without_entries(ctx):
    """nnz = 0: no blocks, an all-zero block row pointer, and an SpMV that
    writes zeros."""
    rp = np.zeros(9, np.int32)
    b = gpu_bcrs(ctx, 8, 8, rp, np.zeros(0, np.int32), np.zeros(0), 2, 2, 2, 2, False)
    i = b.info
    assert i.n_blocks == 0 and i.n_block_rows == 4
    assert (b.arrays(np.float64)["row_ptr"] == 0).all()
    dx, dy = ctx.upload(np.ones(8)), ctx.upload(np.full(8, 5.0))
    ctx.bcrs_spmv(capi.F64, 4, 2, 2, 2, 2, i.col, i.row_ptr, i.val, False, dx, dy)
    assert (ctx.download(dy) == 0).all()
    for d in (dx, dy):
        d.free()
    b.destroy()


def test_bcrs_conversion_rejects_columns_out_of_range(ctx):
    """A column index outside [0, n_cols) would place a value outside its
    block: the conversion refuses the matrix."""
    rp = np.array([0, 2, 3, 4, 4], np.int32)
    for bad in (4, -1, 1 << 30):
        col = np.array([0, 1, bad, 2], np.int32)
        with pytest.raises(capi.SmaxHipError) as e:
            gpu_bcrs(ctx, 4, 4, rp, col, np.ones(4), 2, 2, 2, 2, False)
        assert e.value.status == capi.ERROR_INVALID_VALUE
        assert "outside" in str(e.value)


def test_bcrs_conversion_rejects_what_the_reference_rejects(ctx):
    A = matgen.poisson2d(6)  # 36 x 36
    for bh, bw, hp, wp in ((5, 4, 5, 4), (4, 5, 4, 5), (4, 4, 3, 4), (4, 4, 4, 3)):
        with pytest.raises(capi.SmaxHipError) as e:
            gpu_bcrs(ctx, 36, 36, A.row_ptr, A.col, A.val, bh, bw, hp, wp, True)
        assert e.value.status == capi.ERROR_INVALID_VALUE
    # a row whose block columns descend (convert_formats.hpp:230-234 throws)
    rp = np.array([0, 2, 2, 2, 2], np.int32)
    col = np.array([3, 0], np.int32)
    with pytest.raises(capi.SmaxHipError) as e:
        gpu_bcrs(ctx, 4, 4, rp, col, np.ones(2), 2, 2, 2, 2, True)
    assert "Blocked Column" in str(e.value)


@pytest.mark.parametrize("vt", [capi.F64, capi.F32, capi.BF16])
@pytest.mark.parametrize("cm", [True, False])
def test_bcrs_adjoint_transpose_spmm_sddmm(ctx, vt, cm):
    """C5 at reduced size: A^T through the transposed block layout equals the
    reference BCRS conversion of the explicit A^T (bytes), its SpMM equals the
    reference CRS SpMM of A^T bit for bit, and the block SDDMM matches FP64."""
    A = matgen.block7(8, 4)  # 2048 rows, dense 4x4 blocks
    val = vals_for(vt, A.val)
    b = gpu_bcrs(ctx, A.n_rows, A.n_cols, A.row_ptr, A.col, val, 4, 4, 4, 4, cm)
    T = b.transpose()
    trp, tcol, tval = oracle.crs_transpose(A.n_rows, A.n_cols, A.row_ptr, A.col, val)
    want = oracle.crs_to_bcrs(A.n_cols, A.n_rows, trp, tcol, tval, 4, 4, 4, 4, not cm)
    got = T.arrays(val.dtype)
    assert T.info.block_col_major == (0 if cm else 1)
    for f in ("col", "row_ptr"):
        assert_bitwise(got[f], want[f], f)
    assert_bitwise(got["val"].view(np.uint8), want["val"].view(np.uint8), "val")
    xt = np.float64 if vt == capi.F64 else np.float32
    for k in (1, 8, 64):
        G = matgen.uniform_f64(11, A.n_rows * k).astype(xt)
        dG, dZ = ctx.upload(G), ctx.zeros(A.n_cols * k, xt)
        T.spmm(vt, k, dG, dZ)
        assert_bitwise(ctx.download(dZ), oracle.crs_spmm(trp, tcol, tval, k, True, G),
                       f"A^T G k={k}")
    # refresh after a value change keeps A^T current
    new = vals_for(vt, matgen.uniform_f64(123, A.nnz))
    b2 = gpu_bcrs(ctx, A.n_rows, A.n_cols, A.row_ptr, A.col, new, 4, 4, 4, 4, cm)
    T.refresh_values(b2.info.val)
    t2 = oracle.crs_transpose(A.n_rows, A.n_cols, A.row_ptr, A.col, new)
    w2 = oracle.crs_to_bcrs(A.n_cols, A.n_rows, *t2, 4, 4, 4, 4, not cm)
    assert_bitwise(T.arrays(val.dtype)["val"].view(np.uint8), w2["val"].view(np.uint8))
    # block SDDMM vs FP64 at every stored scalar position
    k = 8
    G = matgen.uniform_f64(12, A.n_rows * k).astype(xt)
    X = matgen.uniform_f64(13, A.n_cols * k).astype(xt)
    dv = ctx.zeros(b.info.n_blocks * 16, xt)
    b.sddmm(vt, k, ctx.upload(G), ctx.upload(X), dv)
    got = ctx.download(dv).astype(np.float64)
    ref = oracle.sddmm_f64(A.row_ptr, A.col, k, G.astype(np.float64),
                           X.astype(np.float64))
    rows = np.repeat(np.arange(A.n_rows), np.diff(A.row_ptr))
    bi = b.arrays(val.dtype)
    br, bc = rows // 4, A.col // 4
    blk = np.array([bi["row_ptr"][r] + np.searchsorted(
        bi["col"][bi["row_ptr"][r]:bi["row_ptr"][r + 1]], c) for r, c in zip(br, bc)])
    off = (A.col % 4) * 4 + rows % 4 if cm else (rows % 4) * 4 + A.col % 4
    mine = got[blk * 16 + off]
    scale = np.abs(G.astype(np.float64).reshape(-1, k)[rows] *
                   X.astype(np.float64).reshape(-1, k)[A.col]).sum(axis=1)
    u = np.finfo(xt).eps / 2
    assert (np.abs(mine - ref) <= (np.log2(k) + 1) * u * scale + 1e-300).all()
    for h in (T, b, b2):
        h.destroy()
