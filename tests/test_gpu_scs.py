"""GPU CRS -> SELL-C-sigma conversion and SELL SpMV/SpMM against the oracle.

Integer/index outputs must be bit-identical; every SpMV output too (each
slot is the reference's fma chain in the same order)."""
from __future__ import annotations

import hashlib
import json
import subprocess

import numpy as np
import pytest

from oracle import oracle
from smaxkernels_amd import capi, matgen
from tests.helpers import (GOLDEN, almost_equal, arr, assert_bitwise,
                           assert_same_values, reference_kats, std_sort_windows)

pytestmark = pytest.mark.gpu

KATS = reference_kats()
SCS_TESTS = [k for k in KATS if k.startswith("spmv_scs_")]


def gpu_scs(ctx, A, C_, sigma, val=None):
    val = A.val if val is None else val
    rp, col, v = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(val)
    s = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v,
                         val.dtype.itemsize, C_, sigma)
    for d in (rp, col, v):
        d.free()
    return s


def check_against_oracle(s, A, C_, sigma, val=None):
    val = A.val if val is None else val
    want = oracle.crs_to_scs(A.row_ptr, A.col, val, C_, sigma)
    got = s.arrays(val.dtype)
    for f in ("n_rows_padded", "n_chunks", "n_elements", "nnz"):
        assert getattr(s.info, f) == want[f], f
    for f in ("chunk_ptr", "chunk_lengths", "col", "perm"):
        assert_bitwise(got[f], want[f], f"{f} C={C_} sigma={sigma}")
    assert_bitwise(got["val"].view(np.uint8), want["val"].view(np.uint8), "val bytes")
    # slot_row is the inverse of perm; src_index names each element's CRS slot
    assert_bitwise(got["slot_row"][got["perm"]], np.arange(A.n_rows, dtype=np.int32))
    assert_bitwise(np.sort(got["slot_row"]),
                   np.arange(s.info.n_rows_padded, dtype=np.int32), "slot_row")
    want_len = np.zeros(s.info.n_rows_padded, np.int32)
    want_len[got["perm"]] = np.diff(A.row_ptr)
    assert_bitwise(got["slot_len"], want_len, "slot_len")
    src = got["src_index"]
    real = src >= 0
    assert real.sum() == A.nnz
    assert_bitwise(got["col"][real], A.col[src[real]], "src_index -> col")
    return got


class Csr:
    def __init__(self, n_rows, n_cols, rp, col, val):
        self.n_rows, self.n_cols = n_rows, n_cols
        self.row_ptr, self.col, self.val = rp, col, val

    @property
    def nnz(self):
        return int(self.row_ptr[-1])


@pytest.mark.parametrize("name", SCS_TESTS)
def test_kat_conversion_on_gpu(ctx, name):
    k = KATS[name]
    C_, sigma = k["ctors"]["A_scs"]["args"]
    A = Csr(*k["ctors"]["A_crs"]["args"][:2], arr(k, "A_crs->row_ptr", np.int32),
            arr(k, "A_crs->col", np.int32), arr(k, "A_crs->val", np.float64))
    s = gpu_scs(ctx, A, C_, sigma)
    got = s.arrays(np.float64)
    sc = {n.split("->")[1]: v["value"] for n, v in k["scalars"].items()}
    for f in ("n_rows", "n_rows_padded", "n_chunks", "n_elements", "nnz"):
        assert getattr(s.info, f) == sc[f], f
    for f in ("chunk_ptr", "chunk_lengths", "col", "perm"):
        assert_bitwise(got[f], arr(k, f"A_scs_exp->{f}", np.int32), f)
    assert almost_equal(got["val"], arr(k, "A_scs_exp->val", np.float64)).all()
    s.destroy()


@pytest.mark.parametrize("C_", [1, 4, 32, 64])
def test_conversion_matches_oracle_all_sigmas(ctx, C_):
    A = matgen.lomax_band(30000, half=30000 // 8)
    npad = -(-A.n_rows // C_) * C_
    for sigma in (1, 16, 17, 64, 100, 4096, npad, 1 << 40):
        s = gpu_scs(ctx, A, C_, sigma)
        check_against_oracle(s, A, C_, sigma)
        s.destroy()


def test_conversion_ragged_and_empty_rows(ctx):
    n = 5003
    A = ragged(n, 3)
    for C_, sigma in ((64, 4096), (7, 50), (32, n), (128, 1)):
        s = gpu_scs(ctx, A, C_, sigma)
        check_against_oracle(s, A, C_, sigma)
        s.destroy()
    Z = Csr(100, 100, np.zeros(101, np.int32), np.zeros(0, np.int32), np.zeros(0))
    s = gpu_scs(ctx, Z, 64, 64)
    assert s.info.n_elements == 0
    assert_bitwise(s.arrays(np.float64)["chunk_ptr"], np.zeros(3, np.int32))
    s.destroy()


def test_conversion_heap_fallback_matches_libstdcxx(ctx, sort_pin, tmp_path):
    """Row lengths from McIlroy's adversary push the window sort into the
    heap fallback; the GPU must reproduce libstdc++'s order."""
    n = 2000
    out = tmp_path / "k.bin"
    subprocess.run([str(sort_pin), "killer", str(n), str(out)], check=True)
    L = np.fromfile(out, np.int32) % 97
    rp = np.concatenate([[0], np.cumsum(L)]).astype(np.int32)
    col = np.concatenate([np.arange(l) for l in L]).astype(np.int32)
    A = Csr(n, n, rp, col, np.ones(rp[-1]))
    # the adversary is defined on the raw counts; check it still bites here
    cnt = np.fromfile(out, np.int32)
    rp2 = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    col2 = np.concatenate([np.arange(l) % n for l in cnt]).astype(np.int32)
    B = Csr(n, n, rp2, col2, np.ones(rp2[-1]))
    for M in (A, B):
        s = gpu_scs(ctx, M, 1, n)
        got = check_against_oracle(s, M, 1, n)
        order = std_sort_windows(sort_pin, np.diff(M.row_ptr), n, tmp_path)
        assert_bitwise(got["slot_row"], order, "std::sort order")
        s.destroy()


def test_full_size_tie_order_digests(ctx):
    """N = 4M Lomax workload, C in {32, 64} and every sigma of the committed
    digests (1, 17, 64, 4096, N): perm and chunk_lengths digests written from
    libstdc++ std::sort, and n_elements."""
    fx = json.loads((GOLDEN / "sort_digests.json").read_text())
    A = matgen.lomax_band(4_000_000)
    rp, col, v = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val)
    done = 0
    for case in fx["cases"]:
        if case["N"] != A.n_rows:
            continue
        s = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, 8,
                             case["C"], case["sigma"])
        perm = ctx.download_ptr(s.info.perm, A.n_rows, np.int32)
        assert hashlib.sha256(perm.tobytes()).hexdigest() == case["perm_sha256"], case
        assert s.info.n_elements == case["n_elements"]
        cl = ctx.download_ptr(s.info.chunk_lengths, s.info.n_chunks, np.int32)
        assert hashlib.sha256(cl.tobytes()).hexdigest() == case["chunk_lengths_sha256"]
        s.destroy()
        done += 1
    assert done == 10
    for d in (rp, col, v):
        d.free()


# ------------------------------------------------------------------ SpMV
def values_for(vt, A):
    if vt == capi.F64:
        return A.val, matgen.uniform_f64(21, A.n_cols)
    if vt == capi.F32:
        return A.val.astype(np.float32), matgen.uniform_f32(21, A.n_cols)
    return matgen.to_bf16(A.val), matgen.uniform_f32(21, A.n_cols)


@pytest.mark.parametrize("vt", [capi.F64, capi.F32, capi.BF16])
@pytest.mark.parametrize("C_,sigma", [(64, 1), (64, 4096), (32, 64), (4, 16),
                                      (1, 1), (128, 512)])
def test_scs_spmv_bit_exact(ctx, vt, C_, sigma):
    A = matgen.lomax_band(40000, half=5000)
    val, x = values_for(vt, A)
    s = gpu_scs(ctx, A, C_, sigma, val)
    want_s = oracle.crs_to_scs(A.row_ptr, A.col, val, C_, sigma)
    want = oracle.scs_spmv(C_, want_s["n_chunks"], want_s["chunk_ptr"],
                           want_s["chunk_lengths"], want_s["col"], want_s["val"], x)
    dx = ctx.upload(x)
    dy = ctx.zeros(s.info.n_rows_padded, x.dtype)
    s.spmv(vt, dx, dy)
    assert_bitwise(ctx.download(dy), want, "permuted y")
    # original order output == reference CRS SpMV, bit for bit
    dy2 = ctx.zeros(A.n_rows, x.dtype)
    s.spmv(vt, dx, dy2, unpermuted=True)
    assert_bitwise(ctx.download(dy2), oracle.crs_spmv(A.row_ptr, A.col, val, x),
                   "unpermuted y vs CRS")
    for d in (dx, dy, dy2):
        d.free()
    s.destroy()


def ragged(n, seed, max_len=40, empty=0.3):
    rng = np.random.default_rng(seed)
    L = rng.integers(0, max_len, n)
    L[rng.random(n) < empty] = 0
    L[::997] = 300
    rp = np.concatenate([[0], np.cumsum(L)]).astype(np.int32)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in L]
                         ).astype(np.int32)
    return Csr(n, n, rp, col, rng.uniform(-1, 1, rp[-1]))


@pytest.mark.parametrize("vt", [capi.F64, capi.F32, capi.BF16])
@pytest.mark.parametrize("C_,sigma", [(64, 1), (64, 64), (64, 1 << 20), (32, 256),
                                      (32, 1 << 20)])
@pytest.mark.parametrize("x0", [0.5, -0.0, np.inf, np.nan])
def test_scs_apply_matches_full_chunk_chains(ctx, vt, C_, sigma, x0):
    """The handle-level SpMV stops each lane at its row's length; it must
    still equal the full padded chain bit for bit, also when x[0] is -0, inf
    or NaN (the padding terms 0*x[0] then decide the result)."""
    A = ragged(9001, 11)
    val, x = values_for(vt, A)
    x = x.copy()
    x[0] = x0
    s = gpu_scs(ctx, A, C_, sigma, val)
    want_s = oracle.crs_to_scs(A.row_ptr, A.col, val, C_, sigma)
    want = oracle.scs_spmv(C_, want_s["n_chunks"], want_s["chunk_ptr"],
                           want_s["chunk_lengths"], want_s["col"], want_s["val"], x)
    dx = ctx.upload(x)
    dy = ctx.zeros(s.info.n_rows_padded, x.dtype)
    s.apply(vt, dx, dy)
    got = ctx.download(dy)
    assert_same_values(got, want, "apply, permuted")
    # padding terms reach some slots: a non-finite x[0] must show there
    assert np.isfinite(got).all() == bool(np.isfinite(x0))
    dy2 = ctx.zeros(A.n_rows, x.dtype)
    s.apply(vt, dx, dy2, unpermuted=True)
    assert_same_values(ctx.download(dy2), want[want_s["perm"]], "apply, original order")
    for d in (dx, dy, dy2):
        d.free()
    s.destroy()


def short_rows(n, seed, max_len=8, empty=0.2):
    """Rows of 0..max_len entries: SELL-64 chunks of at most max_len."""
    rng = np.random.default_rng(seed)
    L = rng.integers(0, max_len + 1, n)
    L[rng.random(n) < empty] = 0
    rp = np.concatenate([[0], np.cumsum(L)]).astype(np.int32)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in L]
                         ).astype(np.int32)
    return Csr(n, n, rp, col, rng.uniform(-1, 1, rp[-1]))


@pytest.mark.parametrize("vt", [capi.F64, capi.F32, capi.BF16])
@pytest.mark.parametrize("sigma", [1, 64, 1 << 20])
@pytest.mark.parametrize("x0", [0.5, -0.0, np.inf, np.nan])
def test_scs_apply_short_chunks(ctx, vt, sigma, x0):
    """Layouts whose chunks hold at most 8 entries per slot run single-group
    chains; the handle, a handle over the same SELL arrays and the plan path
    must all equal the full padded chains bit for bit (x[0] decides the
    padding terms)."""
    A = short_rows(9001, 23)
    val, x = values_for(vt, A)
    x = x.copy()
    x[0] = x0
    s = gpu_scs(ctx, A, 64, sigma, val)
    want_s = oracle.crs_to_scs(A.row_ptr, A.col, val, 64, sigma)
    assert want_s["chunk_lengths"].max() <= 8
    want = oracle.scs_spmv(64, want_s["n_chunks"], want_s["chunk_ptr"],
                           want_s["chunk_lengths"], want_s["col"], want_s["val"], x)
    dx = ctx.upload(x)
    dy = ctx.zeros(s.info.n_rows_padded, x.dtype)
    s.apply(vt, dx, dy)
    assert_same_values(ctx.download(dy), want, "apply, permuted")
    dy2 = ctx.zeros(A.n_rows, x.dtype)
    s.apply(vt, dx, dy2, unpermuted=True)
    assert_same_values(ctx.download(dy2), want[want_s["perm"]], "apply, original order")
    h = sell_handle(ctx, s, val.dtype.itemsize)
    h.apply(vt, dx, dy)
    assert_same_values(ctx.download(dy), want, "handle over SELL arrays")
    h.destroy()
    i = s.info
    plan = ctx.scs_plan(64, sigma, i.n_rows_padded, i.n_chunks, i.n_elements, A.nnz,
                        i.chunk_lengths)
    ctx.scs_spmv_plan(vt, plan, i.chunk_ptr, i.chunk_lengths, i.col, i.val, dx, dy)
    assert_same_values(ctx.download(dy), want, "plan")
    ctx.scs_plan_destroy(plan)
    for d in (dx, dy, dy2):
        d.free()
    s.destroy()


@pytest.mark.parametrize("C_", [64, 32])
def test_scs_apply_matrix_without_entries(ctx, C_):
    """A matrix with no stored entries (every chunk of length 0) gives y = +0
    through the handle, in both orders, whatever x holds."""
    Z = Csr(100, 100, np.zeros(101, np.int32), np.zeros(0, np.int32), np.zeros(0))
    s = gpu_scs(ctx, Z, C_, 64)
    x = np.full(100, np.nan)
    dx = ctx.upload(x)
    dy = ctx.upload(np.full(s.info.n_rows_padded, 7.0))
    s.apply(capi.F64, dx, dy)
    assert_bitwise(ctx.download(dy), np.zeros(s.info.n_rows_padded))
    dy2 = ctx.upload(np.full(100, 7.0))
    s.apply(capi.F64, dx, dy2, unpermuted=True)
    assert_bitwise(ctx.download(dy2), np.zeros(100))
    for d in (dx, dy, dy2):
        d.free()
    s.destroy()


@pytest.mark.parametrize("C_,sigma", [(64, 1), (64, 256), (64, 1 << 20), (32, 512), (8, 1)])
def test_scs_spmv_plan_on_caller_arrays(ctx, C_, sigma):
    """The plan path (schedule only, caller-owned arrays) equals the raw SELL
    SpMV bit for bit, with long chunks present (rows of 300)."""
    A = ragged(9001, 17)
    val, x = values_for(capi.F64, A)
    s = gpu_scs(ctx, A, C_, sigma, val)
    i = s.info
    plan = ctx.scs_plan(C_, sigma, i.n_rows_padded, i.n_chunks, i.n_elements, A.nnz,
                        i.chunk_lengths)
    dx = ctx.upload(x)
    y1, y2 = (ctx.zeros(i.n_rows_padded, np.float64) for _ in range(2))
    s.spmv(capi.F64, dx, y1)
    ctx.scs_spmv_plan(capi.F64, plan, i.chunk_ptr, i.chunk_lengths, i.col, i.val, dx, y2)
    assert_same_values(ctx.download(y2), ctx.download(y1), "plan vs raw")
    ctx.scs_plan_destroy(plan)
    for d in (dx, y1, y2):
        d.free()
    s.destroy()


@pytest.mark.parametrize("vt", [capi.F64, capi.F32, capi.BF16])
@pytest.mark.parametrize("C_,sigma", [(64, 1), (64, 64), (32, 64), (128, 1)])
@pytest.mark.parametrize("x0", [0.5, np.nan])
def test_scs_apply_tail_waves(ctx, vt, C_, sigma, x0):
    """Heavy-tailed rows in unsorted chunks pad the layout 9-17x: the long
    slots of skewed chunks then run in tail waves of their own (a SELL-64
    whatever C is).  The result is still the full padded chain of every slot,
    bit for bit, in both output orders (a NaN x[0] reaches every padded slot
    through its padding term, tail slots included), and again after a value
    refresh, which must reach the tail waves' copies.  The fully sorted
    layout (sigma = N) of the same matrix gets no tail waves."""
    A = matgen.lomax_band(60000, half=7500)
    val, x = values_for(vt, A)
    x = x.copy()
    x[0] = x0
    s = gpu_scs(ctx, A, C_, sigma, val)
    slots, waves, entries = s.tails()
    assert slots > 0 and 0 < waves <= slots and entries >= 64 * waves
    dx = ctx.upload(x)
    dy = ctx.zeros(s.info.n_rows_padded, x.dtype)
    dy2 = ctx.zeros(A.n_rows, x.dtype)
    for rnd in range(2):
        if rnd:
            new = matgen.uniform_f64(78, A.nnz)
            val = {capi.F64: new, capi.F32: new.astype(np.float32),
                   capi.BF16: matgen.to_bf16(new)}[vt]
            dv = ctx.upload(val)
            s.refresh_values(dv)
            dv.free()
        want_s = oracle.crs_to_scs(A.row_ptr, A.col, val, C_, sigma)
        want = oracle.scs_spmv(C_, want_s["n_chunks"], want_s["chunk_ptr"],
                               want_s["chunk_lengths"], want_s["col"], want_s["val"], x)
        s.apply(vt, dx, dy)
        assert_same_values(ctx.download(dy), want, f"apply with tail waves, round {rnd}")
        s.apply(vt, dx, dy2, unpermuted=True)
        assert_same_values(ctx.download(dy2), want[want_s["perm"]],
                           f"apply with tail waves, original order, round {rnd}")
    for d in (dx, dy, dy2):
        d.free()
    s.destroy()
    s2 = gpu_scs(ctx, A, C_, A.n_rows, val)
    assert s2.tails() == (0, 0, 0)
    s2.destroy()


def test_scs_apply_rejects_mismatched_value_type(ctx):
    A = matgen.poisson2d(16)
    s = gpu_scs(ctx, A, 64, 1)
    dx, dy = ctx.zeros(A.n_cols, np.float32), ctx.zeros(s.info.n_rows_padded, np.float32)
    with pytest.raises(capi.SmaxHipError) as e:
        s.apply(capi.F32, dx, dy)
    assert e.value.status == capi.ERROR_INVALID_VALUE
    for d in (dx, dy):
        d.free()
    s.destroy()


def test_scs_spmv_config1_poisson(ctx):
    A = matgen.poisson2d(512)
    x = matgen.uniform_f64(matgen.SEED_X, A.n_cols)
    s = gpu_scs(ctx, A, 64, 1)
    assert s.info.n_elements == 1309696  # reference conversion, C=64 sigma=1
    dx, dy = ctx.upload(x), ctx.zeros(A.n_rows, np.float64)
    s.spmv(capi.F64, dx, dy, unpermuted=True)
    want = oracle.crs_spmv(A.row_ptr, A.col, A.val, x)
    assert_bitwise(ctx.download(dy), want)
    # the handle runs this layout (chunks of at most 5) with single-group chains
    dy2 = ctx.zeros(A.n_rows, np.float64)
    s.apply(capi.F64, dx, dy2, unpermuted=True)
    assert_bitwise(ctx.download(dy2), want, "apply")
    for d in (dx, dy, dy2):
        d.free()
    s.destroy()


@pytest.mark.parametrize("vt", [capi.F64, capi.F32, capi.BF16])
@pytest.mark.parametrize("k", [1, 2, 3, 8, 32, 64])
def test_scs_spmm_bit_exact(ctx, vt, k):
    A = matgen.lomax_band(20000, half=2500)
    val, _ = values_for(vt, A)
    X = (matgen.uniform_f64(5, A.n_cols * k) if vt == capi.F64
         else matgen.uniform_f32(5, A.n_cols * k))
    s = gpu_scs(ctx, A, 64, 256, val)
    dX, dY = ctx.upload(X), ctx.zeros(A.n_rows * k, X.dtype)
    s.spmm(vt, k, dX, dY)
    want = oracle.crs_spmm(A.row_ptr, A.col, val, k, True, X)
    assert_bitwise(ctx.download(dY), want, "SELL SpMM vs CRS SpMM")
    s.destroy()


@pytest.mark.parametrize("vt", [capi.F64, capi.F32, capi.BF16])
@pytest.mark.parametrize("C_,sigma", [(64, 1), (64, 4096), (32, 1 << 20), (4, 16)])
@pytest.mark.parametrize("k", [1, 3, 32, 64, 70])
def test_scs_apply_spmm_equals_crs_spmm(ctx, vt, C_, sigma, k):
    """Handle-level SpMM runs each output over the row's stored entries only:
    it equals the reference CRS SpMM bit for bit even with inf/NaN in X's
    row 0, which the padding terms of the padded chain would spread."""
    A = ragged(6007, 5)
    val, _ = values_for(vt, A)
    X = (matgen.uniform_f64(5, A.n_cols * k) if vt == capi.F64
         else matgen.uniform_f32(5, A.n_cols * k))
    X[:k] = np.inf
    X[1] = np.nan
    s = gpu_scs(ctx, A, C_, sigma, val)
    dX, dY = ctx.upload(X), ctx.zeros(A.n_rows * k, X.dtype)
    s.apply_spmm(vt, k, dX, dY)
    want = oracle.crs_spmm(A.row_ptr, A.col, val, k, True, X)
    assert_same_values(ctx.download(dY), want, "SELL apply_spmm vs CRS SpMM")
    for d in (dX, dY):
        d.free()
    s.destroy()


def test_refresh_values(ctx):
    A = matgen.lomax_band(10000, half=1250)
    s = gpu_scs(ctx, A, 64, 128)
    new = matgen.uniform_f64(77, A.nnz)
    d = ctx.upload(new)
    s.refresh_values(d)
    want = oracle.crs_to_scs(A.row_ptr, A.col, new, 64, 128)
    assert_bitwise(s.arrays(np.float64)["val"], want["val"], "refreshed")
    s.destroy()


def test_invalid_arguments_are_rejected(ctx):
    A = matgen.poisson2d(8)
    rp, col, v = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val)
    for C_, sigma, vb in ((0, 1, 8), (4, 0, 8), (4, 1, 3)):
        with pytest.raises(capi.SmaxHipError) as e:
            ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, vb, C_, sigma)
        assert e.value.status == capi.ERROR_INVALID_VALUE
    with pytest.raises(capi.SmaxHipError) as e:
        ctx.scs_spmv(9, 64, 1, rp, rp, col, v, v, v)
    assert e.value.status == capi.ERROR_UNSUPPORTED


def few_long_rows(n, sigma, n_long, long_len, short_len, seed):
    """CRS whose sigma-windows hold n_long[w % len(n_long)] rows of long_len
    entries (at random places) among rows of short_len (columns sorted,
    distinct, within a +-n/8 band): after the window sort each window's first
    chunk has that many long lanes over short ones, the chunks the execution
    order starts first."""
    rng = np.random.default_rng(seed)
    L = np.full(n, short_len, np.int64)
    for k, w in enumerate(range(0, n, sigma)):
        m = min(sigma, n - w)
        L[w + rng.choice(m, min(n_long[k % len(n_long)], m), replace=False)] = long_len
    rp = np.concatenate([[0], np.cumsum(L)]).astype(np.int32)
    half = n // 8
    cols = []
    for i in range(n):
        lo, hi = max(0, i - half), min(n, i + half + 1)
        cols.append(lo + np.sort(rng.choice(hi - lo, L[i], replace=False)))
    col = np.concatenate(cols).astype(np.int32)
    val = rng.uniform(-1, 1, rp[-1])
    return Csr(n, n, rp, col, val)


@pytest.mark.parametrize("vt", [capi.F64, capi.F32, capi.BF16])
@pytest.mark.parametrize("n,sigma,n_long", [(70000, 4096, (5,)), (40000, 2048, (3, 17, 16)),
                                            (9000, 9000, (16,)), (30000, 30000, (1,))])
def test_scs_apply_few_long_rows(ctx, vt, n, sigma, n_long):
    """Layouts whose long chunks hold a few long lanes over short ones: the
    long-first execution order (XCD-remapped for sigma <= n/16, global for
    sigma = n) with slot lengths.  Both y orders equal the oracle bit for
    bit."""
    A = few_long_rows(n, sigma, n_long, 700, 6, seed=n + sigma)
    val, x = values_for(vt, A)
    s = gpu_scs(ctx, A, 64, sigma, val)
    want = oracle.crs_spmv(A.row_ptr, A.col, val, x)
    dx = ctx.upload(x)
    dy = ctx.zeros(A.n_rows, x.dtype)
    s.apply(vt, dx, dy, unpermuted=True)
    assert_bitwise(ctx.download(dy), want, "unpermuted")
    yp = ctx.zeros(s.info.n_rows_padded, x.dtype)
    s.apply(vt, dx, yp)
    perm = s.arrays(val.dtype)["perm"]
    assert_bitwise(ctx.download(yp)[perm], want, "permuted")
    for d in (dx, dy, yp):
        d.free()
    s.destroy()


# ------------------------------------------------- handles over SELL arrays
def sell_handle(ctx, S, vb):
    """A SpMV handle over the device arrays of handle S (as a caller that
    registered SELL arrays would hold them)."""
    i = S.info
    return ctx.scs_from_sell(i.C, i.sigma, i.n_rows, i.n_rows_padded, i.n_cols, i.n_chunks,
                             i.n_elements, i.nnz, i.chunk_ptr, i.chunk_lengths, i.col, i.val,
                             vb, i.perm)


@pytest.mark.parametrize("vt", [capi.F64, capi.F32, capi.BF16])
@pytest.mark.parametrize("C_,sigma", [(64, 1), (64, 4096), (32, 64), (32, 1 << 20), (8, 16)])
def test_scs_from_sell_equals_converted_handle(ctx, vt, C_, sigma):
    """A handle built from SELL arrays derives the converted handle's slot
    lengths, rows and tail waves from the arrays alone (no stored entry of
    this matrix is +0.0), and its SpMV equals it bit for bit in both output
    orders."""
    A = matgen.lomax_band(30000, half=3750)
    val, x = values_for(vt, A)
    S = gpu_scs(ctx, A, C_, sigma, val)
    T = sell_handle(ctx, S, val.dtype.itemsize)
    a, b = S.arrays(val.dtype), T.arrays(val.dtype)
    assert b["src_index"] is None
    for f in ("chunk_ptr", "chunk_lengths", "col", "perm", "slot_len"):
        assert np.array_equal(a[f], b[f]), f
    assert np.array_equal(a["val"].view(np.uint8), b["val"].view(np.uint8))
    real = a["slot_row"] < A.n_rows
    assert np.array_equal(a["slot_row"][real], b["slot_row"][real])
    assert (b["slot_row"][~real] >= A.n_rows).all()
    assert S.tails() == T.tails()
    dx = ctx.upload(x)
    for unperm, n in ((False, S.info.n_rows_padded), (True, A.n_rows)):
        y1, y2 = ctx.zeros(n, x.dtype), ctx.zeros(n, x.dtype)
        S.apply(vt, dx, y1, unpermuted=unperm)
        T.apply(vt, dx, y2, unpermuted=unperm)
        assert_same_values(ctx.download(y2), ctx.download(y1), f"unpermuted={unperm}")
        y1.free()
        y2.free()
    dx.free()
    T.destroy()
    S.destroy()


@pytest.mark.parametrize("x0", [1.0, -1.0, -0.0, np.inf, np.nan])
def test_scs_from_sell_trailing_zero_entries(ctx, x0):
    """Rows whose last stored entries sit at column 0 (CRS order need not be
    sorted): a trailing +0.0 there is the padding term exactly and may be cut
    from the slot's length; a trailing -0.0 may not (fma(-0, x, -0) = -0 where
    the padding term gives +0).  The SpMV still equals the reference's full
    padded chain bit for bit for every x[0], also for rows of -0.0 only."""
    rows = [[(5, 1.5), (0, 0.0)], [(0, -0.0)], [(0, 0.0)], [(3, -2.0), (0, -0.0)],
            [], [(7, 0.25), (0, 0.0), (0, 0.0)], [(2, -0.0), (0, 0.0)],
            [(1, -1.0)], [(0, -0.0), (0, 0.0)]] * 23
    n = len(rows)
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.array([c for r in rows for c, _ in r], np.int32)
    val = np.array([v for r in rows for _, v in r], np.float64)
    x = matgen.uniform_f64(5, n)
    x[0] = x0
    for C_, sigma in ((8, 1), (4, 8), (32, 1)):
        want_s = oracle.crs_to_scs(rp, col, val, C_, sigma)
        want = oracle.scs_spmv(C_, want_s["n_chunks"], want_s["chunk_ptr"],
                               want_s["chunk_lengths"], want_s["col"], want_s["val"], x)
        d = {k: ctx.upload(want_s[k]) for k in ("chunk_ptr", "chunk_lengths", "col", "val",
                                               "perm")}
        T = ctx.scs_from_sell(C_, sigma, n, want_s["n_rows_padded"], n, want_s["n_chunks"],
                              want_s["n_elements"], len(col), d["chunk_ptr"],
                              d["chunk_lengths"], d["col"], d["val"], 8, d["perm"])
        dx = ctx.upload(x)
        dy = ctx.zeros(want_s["n_rows_padded"], np.float64)
        T.apply(capi.F64, dx, dy)
        assert_same_values(ctx.download(dy), want, f"C={C_} sigma={sigma}")
        T.destroy()
        for a in (*d.values(), dx, dy):
            a.free()


def test_scs_from_sell_refusals(ctx):
    """Arrays that do not form a SELL layout are refused before any kernel
    reads through them, and a handle over SELL arrays refuses what needs CRS
    positions."""
    A = matgen.lomax_band(5000, half=600)
    S = gpu_scs(ctx, A, 32, 64)
    a = S.arrays(np.float64)
    i = S.info
    lib = ctx.lib

    def build(**over):
        arrs = {k: a[k].copy() for k in ("chunk_ptr", "chunk_lengths", "col", "val", "perm")}
        arrs.update(over)
        d = {k: ctx.upload(v) for k, v in arrs.items()}
        h = capi.vp()
        st = lib.smax_hip_scs_from_sell(ctx.handle, i.C, i.sigma, i.n_rows, i.n_rows_padded,
                                        i.n_cols, i.n_chunks, i.n_elements, i.nnz,
                                        d["chunk_ptr"].ptr, d["chunk_lengths"].ptr,
                                        d["col"].ptr, d["val"].ptr, 8, d["perm"].ptr,
                                        capi.C.byref(h))
        for v in d.values():
            v.free()
        return st, h.value

    bad_col = a["col"].copy()
    bad_col[len(bad_col) // 2] = A.n_cols
    bad_ptr = a["chunk_ptr"].copy()
    bad_ptr[3] += 1
    dup = a["perm"].copy()
    dup[1] = dup[0]
    oob = a["perm"].copy()
    oob[7] = i.n_rows_padded
    for name, over in (("column", {"col": bad_col}), ("chunk_ptr", {"chunk_ptr": bad_ptr}),
                       ("perm duplicate", {"perm": dup}), ("perm range", {"perm": oob})):
        st, h = build(**over)
        assert st == capi.ERROR_INVALID_VALUE and h is None, name
    # sizes that do not fit together, including a chunk count whose product
    # with C would overflow 64 bits
    d = {k: ctx.upload(a[k]) for k in ("chunk_ptr", "chunk_lengths", "col", "val", "perm")}
    for C_, npad, nch in ((i.C, i.n_rows_padded + 1, i.n_chunks), (1 << 40, i.n_rows_padded,
                                                                   1 << 40)):
        h = capi.vp()
        st = lib.smax_hip_scs_from_sell(ctx.handle, C_, i.sigma, i.n_rows, npad, i.n_cols, nch,
                                        i.n_elements, i.nnz, d["chunk_ptr"].ptr,
                                        d["chunk_lengths"].ptr, d["col"].ptr, d["val"].ptr, 8,
                                        d["perm"].ptr, capi.C.byref(h))
        assert st == capi.ERROR_INVALID_VALUE and h.value is None, (C_, npad, nch)
    for v in d.values():
        v.free()
    st, h = build()
    assert st == capi.SUCCESS
    T = capi.Scs(ctx, h)
    dv = ctx.zeros(A.nnz, np.float64)
    G = ctx.zeros(A.n_rows * 4, np.float64)
    with pytest.raises(capi.SmaxHipError):
        T.refresh_values(dv)
    with pytest.raises(capi.SmaxHipError):
        T.sddmm(capi.F64, 4, G, G, dv)
    with pytest.raises(capi.SmaxHipError):
        T.apply_spmm(capi.F64, 4, G, G)
    for d in (dv, G):
        d.free()
    T.destroy()
    S.destroy()
