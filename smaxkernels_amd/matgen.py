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
