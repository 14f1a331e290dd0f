"""Pins the oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is independent of the oracle's own code: the reference's known-
answer tests (examples/tests, extracted to golden/reference_kats.json), the
system libstdc++'s real std::sort, exact rational arithmetic, closed forms and
dense brute force.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle
from smaxkernels_amd import matgen
from tests.helpers import (GOLDEN, almost_equal, arr, assert_bitwise,
                           padded_counts, reference_kats, std_sort_windows)

KATS = reference_kats()
SCS_TESTS = [k for k in KATS if k.startswith("spmv_scs_")]


# ----------------------------------------------------------- reference KATs
def test_kat_spmv():
    """examples/tests/spmv_test.hpp: 3x3 CRS, x = 1 -> y = {2.3, 2.2, 6.4}."""
    k = KATS["spmv_test.hpp"]
    rp, col = arr(k, "A->row_ptr", np.int32), arr(k, "A->col", np.int32)
    val = arr(k, "A->val", np.float64)
    y = oracle.crs_spmv(rp, col, val, np.ones(3))
    assert almost_equal(y, arr(k, "y_expected", np.float64)).all()


@pytest.mark.parametrize("name", SCS_TESTS)
def test_kat_scs_conversion(name):
    """The nine spmv_scs_*_test_* conversions (C in {1,2,4}, sigma in {1,4})."""
    k = KATS[name]
    C_, sigma = k["ctors"]["A_scs"]["args"]
    rp, col = arr(k, "A_crs->row_ptr", np.int32), arr(k, "A_crs->col", np.int32)
    val = arr(k, "A_crs->val", np.float64)
    s = oracle.crs_to_scs(rp, col, val, C_, sigma)
    sc = {n.split("->")[1]: v["value"] for n, v in k["scalars"].items()}
    for f in ("n_rows", "n_rows_padded", "n_chunks", "n_elements", "nnz"):
        assert s[f] == sc[f], f
    for f, dt in (("chunk_ptr", np.int32), ("chunk_lengths", np.int32),
                  ("col", np.int32), ("perm", np.int32)):
        assert_bitwise(s[f], arr(k, f"A_scs_exp->{f}", dt), f)
    assert almost_equal(s["val"], arr(k, "A_scs_exp->val", np.float64)).all()


@pytest.mark.parametrize("name", ["spmv_bcrs_1_1_test_1.hpp",
                                  "spmv_bcrs_1_1_test_2.hpp"])
def test_kat_bcrs_conversion(name):
    k = KATS[name]
    c = {n: v["value"] for n, v in k["consts"].items()}
    rp, col = arr(k, "A_crs->row_ptr", np.int32), arr(k, "A_crs->col", np.int32)
    val = arr(k, "A_crs->val", np.float64)
    b = oracle.crs_to_bcrs(4, 4, rp, col, val, c["target_b_height"],
                           c["target_b_width"], c["target_height_pad"],
                           c["target_width_pad"], c["use_blocked_column_major"])
    assert b["n_blocks"] == k["ctors"]["A_bcrs_exp"]["args"][2]
    assert_bitwise(b["col"], arr(k, "A_bcrs_exp->col", np.int32), "col")
    assert_bitwise(b["row_ptr"], arr(k, "A_bcrs_exp->row_ptr", np.int32), "row_ptr")
    assert almost_equal(b["val"], arr(k, "A_bcrs_exp->val", np.float64)).all()


def test_kat_bcrs_spmv():
    """spmv_bcrs_1_1_test_3.hpp: 2x2 blocks, column-major, x = 1."""
    k = KATS["spmv_bcrs_1_1_test_3.hpp"]
    nbr, nbc, nb, bh, bw, hp, wp = k["ctors"]["A_bcrs"]["args"]
    y = oracle.bcrs_spmv(nbr, bh, bw, hp, wp, arr(k, "A_bcrs->col", np.int32),
                         arr(k, "A_bcrs->row_ptr", np.int32),
                         arr(k, "A_bcrs->val", np.float64), True,
                         np.ones(nbc * wp))
    assert almost_equal(y, arr(k, "y_expected", np.float64)).all()


def laplace1d(n: int):
    """The 1-D Laplacian of spmv_bcrs_laplace_test.hpp:139-178."""
    rows, cols, vals = [], [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= j < n:
                rows.append(i); cols.append(j); vals.append(v)
    rp = np.searchsorted(np.array(rows), np.arange(n + 1)).astype(np.int32)
    return rp, np.array(cols, np.int32), np.array(vals)


@pytest.mark.parametrize("col_major", [True, False])
def test_kat_bcrs_laplace(col_major):
    """spmv_bcrs_laplace_test.hpp: 16x16 1-D Laplacian as 4x8 blocks equals
    the CRS product for x_i = 1/((i+3)%23+1)."""
    rp, col, val = laplace1d(16)
    b = oracle.crs_to_bcrs(16, 16, rp, col, val, 4, 8, 4, 8, col_major)
    x = 1.0 / (((np.arange(16) + 3) % 23) + 1.0)
    y = oracle.bcrs_spmv(b["n_block_rows"], 4, 8, 4, 8, b["col"], b["row_ptr"],
                         b["val"], col_major, x)
    assert almost_equal(y, oracle.crs_spmv(rp, col, val, x)).all()


@pytest.mark.parametrize("name,row_major", [("spmm_test.hpp", False),
                                            ("spmm_rowmaj_test.hpp", True)])
def test_kat_spmm(name, row_major):
    k = KATS[name]
    nv = k["consts"]["n_vectors"]["value"]
    rp, col = arr(k, "A->row_ptr", np.int32), arr(k, "A->col", np.int32)
    val = arr(k, "A->val", np.float64)
    Y = oracle.crs_spmm(rp, col, val, nv, row_major, np.ones(3 * nv))
    assert almost_equal(Y, arr(k, "Y_expected", np.float64)).all()


# ------------------------------------------------------ fma-chain semantics
def _fma_exact(a: float, b: float, c: float) -> float:
    # float(Fraction) rounds to nearest-even exactly: a true fused multiply-add.
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def test_crs_spmv_is_sequential_fma_chain():
    """Each row is fma(val_j, x_col, acc) in CRS order: emulated exactly with
    rational arithmetic and compared bit for bit (mul+add or a reordered sum
    differs on these inputs)."""
    rng = np.random.default_rng(7)
    n, L = 64, 24
    rp = np.arange(0, n * L + 1, L, dtype=np.int32)
    col = rng.integers(0, n, n * L).astype(np.int32)
    val = rng.uniform(-1, 1, n * L) * 10.0 ** rng.integers(-8, 8, n * L)
    x = rng.uniform(-1, 1, n)
    y = oracle.crs_spmv(rp, col, val, x)
    want = np.empty(n)
    naive = np.empty(n)
    for r in range(n):
        acc = 0.0
        acc2 = 0.0
        for j in range(rp[r], rp[r + 1]):
            acc = _fma_exact(val[j], x[col[j]], acc)
            acc2 = acc2 + val[j] * x[col[j]]
        want[r] = acc
        naive[r] = acc2
    assert_bitwise(y, want, "fma chain")
    assert (naive != want).any()  # the inputs do discriminate


def test_laplacian_row_sums_closed_form():
    """5-point Laplacian times the ones vector: 0 inside, 1 on edges, 2 in
    corners (exact in any order)."""
    A = matgen.poisson2d(64)
    y = oracle.crs_spmv(A.row_ptr, A.col, A.val, np.ones(A.n_cols))
    g = np.zeros((64, 64))
    g[0, :] += 1; g[-1, :] += 1; g[:, 0] += 1; g[:, -1] += 1
    assert_bitwise(y, g.ravel(), "row sums")


def test_generators_match_reference_shapes():
    """Closed-form sizes of the workloads (5N-4n, (3n-2)^3, 7n^3-6n^2 blocks)."""
    assert matgen.poisson2d(512).nnz == 1308672
    assert matgen._load().smax_gen_stencil27_nnz(160) == 109215352
    assert matgen._load().smax_gen_block7_nnz(96, 4) == 6137856 * 16
    A = matgen.stencil27(6)
    d = A.val[A.col == np.repeat(np.arange(A.n_rows), np.diff(A.row_ptr))]
    assert (d == 26.01).all() and A.nnz == 16 ** 3


# ------------------------------------------------------------ std::sort pin
@pytest.mark.parametrize("seed", range(6))
def test_sell_sort_matches_libstdcxx(sort_pin, tmp_path, seed):
    """Random windows with heavy ties, every size class of the introsort."""
    rng = np.random.default_rng(seed)
    for n, hi in ((1, 3), (16, 4), (17, 3), (40, 2), (257, 5), (5000, 30),
                  (20000, 2000)):
        cnt = rng.integers(0, hi, n).astype(np.int32)
        want = std_sort_windows(sort_pin, cnt, n, tmp_path)
        got, got_cnt = oracle.sell_sort(np.arange(n, dtype=np.int32), cnt)
        assert_bitwise(got, want, f"n={n}")
        assert_bitwise(got_cnt, cnt[want], "counts")


def test_sell_sort_heap_fallback_matches_libstdcxx(sort_pin, tmp_path):
    """McIlroy's adversary forces introsort's heap-sort fallback."""
    import subprocess
    lib = oracle._load()
    heap = ctypes.c_longlong.in_dll(lib, "oracle_heap_segments")
    for n in (100, 1000, 4097):
        out = tmp_path / f"k{n}.bin"
        subprocess.run([str(sort_pin), "killer", str(n), str(out)], check=True)
        cnt = np.fromfile(out, np.int32)
        before = heap.value
        got, _ = oracle.sell_sort(np.arange(n, dtype=np.int32), cnt)
        assert heap.value > before, "adversary did not reach the heap path"
        assert_bitwise(got, std_sort_windows(sort_pin, cnt, n, tmp_path), "heap")


def test_scs_perm_matches_libstdcxx_on_lomax(sort_pin, tmp_path):
    """Conversion perm == inverse of std::sort's window order, sigma 17..N."""
    A = matgen.lomax_band(30000, half=30000 // 8)
    for C_ in (32, 64):
        cnt = padded_counts(A.row_ptr, C_)
        for sigma in (1, 16, 17, 64, 4096, len(cnt)):
            order = std_sort_windows(sort_pin, cnt, sigma, tmp_path)
            s = oracle.crs_to_scs(A.row_ptr, A.col, A.val, C_, sigma)
            slot = np.empty(len(order), np.int64)
            slot[order] = np.arange(len(order))
            assert_bitwise(s["perm"], slot[:A.n_rows].astype(np.int32),
                           f"C={C_} sigma={sigma}")


def test_scs_digests_match_committed_fixture():
    """Tie order at full config size: perm digests written from libstdc++'s
    std::sort (golden/make_sort_digests.py)."""
    fx = json.loads((GOLDEN / "sort_digests.json").read_text())
    for case in fx["cases"]:
        if case["N"] > 400000:
            continue  # full-size cases run in the GPU suite
        A = matgen.lomax_band(case["N"], half=case["N"] // 8)
        s = oracle.crs_to_scs(A.row_ptr, A.col, A.val, case["C"], case["sigma"])
        assert hashlib.sha256(s["perm"].tobytes()).hexdigest() == case["perm_sha256"]
        assert s["n_elements"] == case["n_elements"]


# ----------------------------------------------------- closed-form SELL
@pytest.mark.parametrize("C_,sigma", [(1, 1), (4, 4), (32, 1), (64, 64),
                                      (64, 100), (64, 1 << 30)])
def test_scs_metadata_is_tie_independent(C_, sigma):
    """chunk_lengths = max count per chunk of the sorted counts (a stable
    numpy sort gives the same sorted count sequence)."""
    A = matgen.lomax_band(5000, half=600)
    s = oracle.crs_to_scs(A.row_ptr, A.col, A.val, C_, sigma)
    cnt = padded_counts(A.row_ptr, C_)
    for w in range(0, len(cnt), min(sigma, len(cnt))):
        seg = cnt[w:w + sigma]
        cnt[w:w + sigma] = -np.sort(-seg, kind="stable")
    want = cnt.reshape(-1, C_).max(axis=1)
    assert_bitwise(s["chunk_lengths"], want.astype(np.int32), "chunk_lengths")
    assert_bitwise(s["chunk_ptr"], np.concatenate(
        [[0], np.cumsum(want * C_)]).astype(np.int32), "chunk_ptr")


@pytest.mark.parametrize("C_,sigma", [(1, 1), (8, 32), (64, 4096)])
def test_scs_spmv_equals_crs_spmv(C_, sigma):
    """SELL keeps each row's CRS order, padding adds exact zeros: y_scs[perm]
    is bit-identical to y_crs (mkl_scs_spmv_validate.cpp:87-89 compares the
    same way)."""
    A = matgen.lomax_band(20000, half=2500)
    x = matgen.uniform_f64(3, A.n_cols)
    s = oracle.crs_to_scs(A.row_ptr, A.col, A.val, C_, sigma)
    ys = oracle.scs_spmv(C_, s["n_chunks"], s["chunk_ptr"], s["chunk_lengths"],
                         s["col"], s["val"], x)
    assert_bitwise(ys[s["perm"]], oracle.crs_spmv(A.row_ptr, A.col, A.val, x))


# ------------------------------------------------------- dense brute force
def _dense(n_rows, n_cols, rp, col, val):
    D = np.zeros((n_rows, n_cols))
    for r in range(n_rows):
        for j in range(rp[r], rp[r + 1]):
            D[r, col[j]] += val[j]
    return D


@pytest.mark.parametrize("bh,bw,hp,wp,cm", [(4, 4, 4, 4, True), (4, 4, 4, 4, False),
                                            (2, 3, 3, 4, True), (3, 2, 4, 2, False)])
def test_bcrs_conversion_dense(bh, bw, hp, wp, cm):
    rng = np.random.default_rng(11)
    n_rows, n_cols = 12 * bh, 10 * bw
    D = (rng.random((n_rows, n_cols)) < 0.1) * rng.uniform(-1, 1, (n_rows, n_cols))
    rows, cols = np.nonzero(D)
    rp = np.searchsorted(rows, np.arange(n_rows + 1)).astype(np.int32)
    col, val = cols.astype(np.int32), D[rows, cols]
    b = oracle.crs_to_bcrs(n_rows, n_cols, rp, col, val, bh, bw, hp, wp, cm)
    R = np.zeros_like(D)
    for br in range(n_rows // bh):
        for j in range(b["row_ptr"][br], b["row_ptr"][br + 1]):
            blk = b["val"][j * hp * wp:(j + 1) * hp * wp]
            blk = blk.reshape(wp, hp).T if cm else blk.reshape(hp, wp)
            R[br * bh:(br + 1) * bh, b["col"][j] * bw:(b["col"][j] + 1) * bw] = blk[:bh, :bw]
    assert_bitwise(R, D, "dense")
    x = rng.uniform(-1, 1, (n_cols // bw) * wp)
    xs = x.reshape(-1, wp)[:, :bw].ravel()
    y = oracle.bcrs_spmv(n_rows // bh, bh, bw, hp, wp, b["col"], b["row_ptr"],
                         b["val"], cm, x).reshape(-1, hp)[:, :bh].ravel()
    np.testing.assert_allclose(y, D @ xs, rtol=1e-12, atol=1e-12)


def test_bcrs_rejects_bad_blocking():
    A = matgen.poisson2d(6)
    for bh, bw, hp, wp in ((5, 4, 5, 4), (4, 5, 4, 5), (4, 4, 3, 4), (4, 4, 4, 3)):
        with pytest.raises(ValueError):
            oracle.crs_to_bcrs(36, 36, A.row_ptr, A.col, A.val, bh, bw, hp, wp, True)


def test_transpose_and_sddmm_dense():
    rng = np.random.default_rng(5)
    A = matgen.lomax_band(300, half=40)
    D = _dense(A.n_rows, A.n_cols, A.row_ptr, A.col, A.val)
    trp, tcol, tval = oracle.crs_transpose(A.n_rows, A.n_cols, A.row_ptr, A.col, A.val)
    assert_bitwise(_dense(A.n_cols, A.n_rows, trp, tcol, tval), D.T, "A^T")
    for c in range(A.n_cols):  # rows of A^T keep ascending original rows
        assert (np.diff(tcol[trp[c]:trp[c + 1]]) > 0).all()
    k = 5
    G = rng.uniform(-1, 1, A.n_rows * k)
    X = rng.uniform(-1, 1, A.n_cols * k)
    dv = oracle.sddmm_f64(A.row_ptr, A.col, k, G, X)
    rows = np.repeat(np.arange(A.n_rows), np.diff(A.row_ptr))
    want = (G.reshape(-1, k)[rows] * X.reshape(-1, k)[A.col]).sum(axis=1)
    np.testing.assert_allclose(dv, want, rtol=1e-13, atol=1e-15)


def test_spmm_equals_columnwise_spmv():
    A = matgen.lomax_band(2000, half=250)
    k = 7
    X = matgen.uniform_f64(9, A.n_cols * k)
    Yr = oracle.crs_spmm(A.row_ptr, A.col, A.val, k, True, X).reshape(-1, k)
    Xc = np.ascontiguousarray(X.reshape(-1, k).T).ravel()
    Yc = oracle.crs_spmm(A.row_ptr, A.col, A.val, k, False, Xc).reshape(k, -1)
    for v in range(k):
        yv = oracle.crs_spmv(A.row_ptr, A.col, A.val,
                             np.ascontiguousarray(X.reshape(-1, k)[:, v]))
        assert_bitwise(Yr[:, v], yv, "row-major")
        assert_bitwise(Yc[v], yv, "col-major")


def test_low_precision_paths_widen_exactly():
    """FP32 and BF16 oracles: values widened exactly, fmaf chain; BF16 with
    values that are exact in bfloat16 equals the FP32 result bit for bit."""
    A = matgen.lomax_band(3000, half=400)
    v32 = matgen.bf16_to_f32(matgen.to_bf16(A.val))
    x = matgen.uniform_f32(4, A.n_cols)
    y32 = oracle.crs_spmv(A.row_ptr, A.col, v32, x)
    ybf = oracle.crs_spmv(A.row_ptr, A.col, matgen.to_bf16(A.val), x)
    assert_bitwise(ybf, y32, "bf16 vs f32")
    ref = _dense(A.n_rows, A.n_cols, A.row_ptr, A.col, v32.astype(np.float64)) @ x
    assert np.max(np.abs(y32 - ref)) < 1e-4
