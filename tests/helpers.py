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
