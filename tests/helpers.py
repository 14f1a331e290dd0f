"""Shared comparison helpers for the parity tests."""
from __future__ import annotations

import json
import subprocess
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def reference_kats() -> dict:
    return json.loads((GOLDEN / "reference_kats.json").read_text())


def arr(kat: dict, name: str, dtype) -> np.ndarray:
    return np.asarray(kat["arrays"][name]["values"], dtype=dtype)


def almost_equal(a, b, dtype=np.float64) -> np.ndarray:
    """The reference's test predicate (examples/tests/tests_common.hpp:9-35):
    |a-b| <= 1e-14 (1e-7 for float), else |a-b| <= 4 eps max(|a|,|b|)."""
    a = np.asarray(a, dtype)
    b = np.asarray(b, dtype)
    abs_eps = 1e-14 if dtype == np.float64 else 1e-7
    eps = np.finfo(dtype).eps
    d = np.abs(a - b)
    return (d <= abs_eps) | (d <= 4 * eps * np.maximum(np.abs(a), np.abs(b)))


def assert_bitwise(a: np.ndarray, b: np.ndarray, what: str = "") -> None:
    """Bit-identical (treating +0 == -0, NaN never expected)."""
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.flatnonzero(a != b)
    assert bad.size == 0, (f"{what}: {bad.size} of {a.size} differ; first at "
                           f"{bad[0]}: {a[bad[0]]!r} vs {b[bad[0]]!r}")


def assert_same_values(a: np.ndarray, b: np.ndarray, what: str = "") -> None:
    """Equal values with equal zero signs, NaN where the other has NaN (NaN
    payloads and signs differ between the host and the GPU's default NaN)."""
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    na, nb = np.isnan(a), np.isnan(b)
    bad = np.flatnonzero((na != nb) | (~na & ((a != b) | (np.signbit(a) != np.signbit(b)))))
    assert bad.size == 0, (f"{what}: {bad.size} of {a.size} differ; first at "
                           f"{bad[0]}: {a[bad[0]]!r} vs {b[bad[0]]!r}")


def std_sort_windows(exe, counts: np.ndarray, sigma: int, tmpdir) -> np.ndarray:
    """Index order after libstdc++ std::sort of each sigma window."""
    tmpdir = Path(tmpdir)
    inp, out = tmpdir / "cnt.bin", tmpdir / "idx.bin"
    np.ascontiguousarray(counts, np.int32).tofile(inp)
    subprocess.run([str(exe), "sort", str(sigma), str(inp), str(out)], check=True)
    return np.fromfile(out, np.int32)


def padded_counts(row_ptr: np.ndarray, C: int) -> np.ndarray:
    n = len(row_ptr) - 1
    npad = -(-n // C) * C
    c = np.zeros(npad, np.int32)
    c[:n] = np.diff(row_ptr)
    return c


def fd_gradient_check(A, v32, Xf, Wf, k, gv, gx, n_samples=64, seed=2024) -> dict:
    """FP64 central differences of L = sum(Y o W), Y = A X (row-major X, W
    with k columns), for n_samples value entries and n_samples X entries,
    against the value gradient gv (CRS order, from the SDDMM) and the X
    gradient gx (n_cols x k, from A^T W).  L is linear in each entry, so the
    central difference (h = 2^-10 max(1, |entry|)) is exact up to rounding;
    it is evaluated on the rows the entry touches (every other term of L
    cancels identically).  Metric per entry: |g - g_FD| / max(|g_FD|,
    sum of |terms|)."""
    rng = np.random.default_rng(seed)
    X = np.asarray(Xf, np.float64).reshape(A.n_cols, k)
    W = np.asarray(Wf, np.float64).reshape(A.n_rows, k)
    v64 = np.asarray(v32, np.float64)
    rows_of = np.repeat(np.arange(A.n_rows), np.diff(A.row_ptr))

    def loss(vals, Xm, rows):
        tot = 0.0
        for i in rows:
            s, e = A.row_ptr[i], A.row_ptr[i + 1]
            tot += float((vals[s:e] @ Xm[A.col[s:e]]) @ W[i])
        return tot

    def central(arr, at, rows):
        # perturb arr[at] by +-h in place (arr is v64 or X), then restore
        keep = arr[at]
        h = 2.0 ** -10 * max(1.0, abs(keep))
        arr[at] = keep + h
        lp = loss(v64, X, rows)
        arr[at] = keep - h
        lm = loss(v64, X, rows)
        arr[at] = keep
        return (lp - lm) / (2 * h)

    worst_v = worst_x = 0.0
    for j in rng.choice(A.nnz, n_samples, replace=False):
        i = rows_of[j]
        fd = central(v64, j, [i])
        scale = max(abs(fd), float(np.abs(W[i] * X[A.col[j]]).sum()))
        worst_v = max(worst_v, abs(float(gv[j]) - fd) / scale)
    for idx in rng.choice(A.n_cols * k, n_samples, replace=False):
        jr, vv = divmod(int(idx), k)
        mask = A.col == jr
        rows = np.unique(rows_of[mask])
        fd = central(X, (jr, vv), rows)
        scale = max(abs(fd), float(np.abs(v64[mask] * W[rows_of[mask], vv]).sum()))
        worst_x = max(worst_x, abs(float(gx[jr, vv]) - fd) / scale)
    return {"entries": 2 * n_samples, "max_rel_err": max(worst_v, worst_x),
            "max_rel_err_values": worst_v, "max_rel_err_X": worst_x, "bound": 1e-6,
            "metric": "|g - g_FD| / max(|g_FD|, sum |terms|), FP64 central differences"}


def sddmm_tree(k: int) -> tuple[int, int]:
    """The SDDMM reduction tree for k: g lanes per element, each summing a run
    of q consecutive v, then a butterfly over the g lanes."""
    p = 1
    while p < k:
        p *= 2
    g = min(64, max(1, p // 4))
    return g, -(-k // g)


def gamma(n, u):
    """gamma_n = n u / (1 - n u): the relative error bound of n roundings."""
    n = np.asarray(n, np.float64)
    return n * u / (1 - n * u)
