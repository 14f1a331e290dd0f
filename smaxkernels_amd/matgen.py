"""Seeded synthetic workloads (wrapper of gen/matgen.c).

Every generator is a pure function of its sizes and seed (integer hashing and
correctly rounded IEEE arithmetic only), so tests, the smoke check and the
benchmark rebuild identical matrices on any host.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_build" / "libsmax_matgen.so"
_lib = None

# Lomax scale giving a mean row length of 16.03 (N = 4M) at shape 1.5, cap 2000.
LOMAX_LAMBDA = 8.3
LOMAX_CAP = 2000

SEED_MATRIX = 0x5EED0001
SEED_VALUES = 0x5EED0002
SEED_X = 0x5EED0003
SEED_B = 0x5EED0004
SEED_W = 0x5EED0005


def _load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} not built (python -m smaxkernels_amd.build)")
        lib = C.CDLL(str(LIB_PATH))
        i64, u64, vp, d = C.c_int64, C.c_uint64, C.c_void_p, C.c_double
        sig = {
            "smax_gen_uniform_f64": ([u64, i64, vp], None),
            "smax_gen_uniform_f32": ([u64, i64, vp], None),
            "smax_gen_poisson2d_nnz": ([i64], i64),
            "smax_gen_poisson2d": ([i64, vp, vp, vp], None),
            "smax_gen_lomax_rowptr": ([i64, u64, d, C.c_int32, i64, vp], i64),
            "smax_gen_lomax_cols": ([i64, u64, i64, vp, vp], None),
            "smax_gen_stencil27_nnz": ([i64], i64),
            "smax_gen_stencil27": ([i64, d, vp, vp, vp], None),
            "smax_gen_block7_nnz": ([i64, i64], i64),
            "smax_gen_block7": ([i64, i64, u64, d, vp, vp, vp], None),
            "smax_gen_f64_to_bf16": ([i64, vp, vp], None),
            "smax_gen_f64_to_f32": ([i64, vp, vp], None),
            "smax_mtx_header": ([C.c_char_p, vp, vp, vp, vp], C.c_int),
            "smax_mtx_read": ([C.c_char_p, i64, vp, vp, vp], i64),
        }
        for n, (a, r) in sig.items():
            f = getattr(lib, n)
            f.argtypes, f.restype = a, r
        _lib = lib
    return _lib


@dataclass
class CRS:
    n_rows: int
    n_cols: int
    row_ptr: np.ndarray  # int32[n_rows+1]
    col: np.ndarray      # int32[nnz]
    val: np.ndarray      # float64[nnz]

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])


def uniform_f64(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    _load().smax_gen_uniform_f64(seed, n, out.ctypes.data)
    return out


def uniform_f32(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    _load().smax_gen_uniform_f32(seed, n, out.ctypes.data)
    return out


def to_bf16(a: np.ndarray) -> np.ndarray:
    """float64 -> bfloat16 bit patterns (via float32, round to nearest even)."""
    a = np.ascontiguousarray(a, np.float64)
    out = np.empty(a.shape, np.uint16)
    _load().smax_gen_f64_to_bf16(a.size, a.ctypes.data, out.ctypes.data)
    return out


def bf16_to_f32(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.uint32) << 16).view(np.float32)


def poisson2d(n: int) -> CRS:
    """2-D 5-point Laplacian on an n x n grid (reference row order)."""
    N = n * n
    nnz = _load().smax_gen_poisson2d_nnz(n)
    rp = np.empty(N + 1, np.int32)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    _load().smax_gen_poisson2d(n, rp.ctypes.data, col.ctypes.data, val.ctypes.data)
    return CRS(N, N, rp, col, val)


def lomax_band(N: int, seed: int = SEED_MATRIX, half: int | None = None,
               lam: float = LOMAX_LAMBDA, cap: int = LOMAX_CAP,
               val_seed: int = SEED_VALUES) -> CRS:
    """Lomax(1.5) row lengths, distinct uniform columns within +-half (N/8)."""
    half = N // 8 if half is None else half
    rp = np.empty(N + 1, np.int32)
    nnz = _load().smax_gen_lomax_rowptr(N, seed, lam, cap, half, rp.ctypes.data)
    if nnz < 0:
        raise OverflowError("nnz exceeds int32")
    col = np.empty(nnz, np.int32)
    _load().smax_gen_lomax_cols(N, seed, half, rp.ctypes.data, col.ctypes.data)
    val = uniform_f64(val_seed, nnz)
    return CRS(N, N, rp, col, val)


def stencil27(n: int, diag: float = 26.01) -> CRS:
    N = n ** 3
    nnz = _load().smax_gen_stencil27_nnz(n)
    rp = np.empty(N + 1, np.int32)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    _load().smax_gen_stencil27(n, diag, rp.ctypes.data, col.ctypes.data,
                               val.ctypes.data)
    return CRS(N, N, rp, col, val)


def block7(n: int, b: int = 4, seed: int = SEED_VALUES, shift: float = 8.0) -> CRS:
    """Scalar CRS of a 7-point block adjacency on n^3 blocks of size b x b."""
    N = b * n ** 3
    nnz = _load().smax_gen_block7_nnz(n, b)
    rp = np.empty(N + 1, np.int32)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    _load().smax_gen_block7(n, b, seed, shift, rp.ctypes.data, col.ctypes.data,
                            val.ctypes.data)
    return CRS(N, N, rp, col, val)


MTX_ERRORS = {-1: "cannot open file", -2: "missing or malformed Matrix Market banner",
              -3: "unsupported Matrix Market type (coordinate real/integer/pattern, "
                  "general/symmetric only)",
              -4: "malformed size line", -5: "malformed or out-of-range entry",
              -6: "capacity", -7: "matrix does not fit int32 CRS", -8: "out of memory"}


def read_mtx(path) -> CRS:
    """CRS of a Matrix Market coordinate file, built as the reference's
    benchmarks build it (COOMatrix::read_from_mtx + convert_coo_to_crs,
    examples/matrix_structs.hpp:52-180, 338-388): 0-based indices, pattern
    values 0.01, symmetric entries mirrored, rows stable-sorted (file order
    within a row)."""
    b = str(path).encode()
    n_rows, n_cols, n_ent = C.c_int64(), C.c_int64(), C.c_int64()
    flags = C.c_int()
    rc = _load().smax_mtx_header(b, C.byref(n_rows), C.byref(n_cols), C.byref(n_ent),
                                 C.byref(flags))
    if rc:
        raise ValueError(f"{path}: {MTX_ERRORS.get(rc, rc)}")
    cap = n_ent.value * (2 if flags.value & 1 else 1)
    rp = np.empty(n_rows.value + 1, np.int32)
    col = np.empty(max(cap, 1), np.int32)
    val = np.empty(max(cap, 1), np.float64)
    nnz = _load().smax_mtx_read(b, cap, rp.ctypes.data, col.ctypes.data, val.ctypes.data)
    if nnz < 0:
        raise ValueError(f"{path}: {MTX_ERRORS.get(nnz, nnz)}")
    return CRS(n_rows.value, n_cols.value, rp, col[:nnz].copy(), val[:nnz].copy())
