"""ORACLE — test infrastructure only (ctypes loader of oracle/_build/liboracle.so).

The C restatement of the reference's CPU algorithms (oracle.c).  Imported by
tests/, by __graft_entry__.smoke() as the checker and by bench.py's CPU
baseline leg; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

i64, vp = C.c_int64, C.c_void_p
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def _load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
        lib = C.CDLL(str(LIB_PATH))
        sig = {
            "oracle_crs_spmv_f64": [i64, vp, vp, vp, vp, vp],
            "oracle_crs_spmv_f32": [i64, vp, vp, vp, vp, vp],
            "oracle_crs_spmv_bf16": [i64, vp, vp, vp, vp, vp],
            "oracle_scs_spmv_f64": [i64, i64, vp, vp, vp, vp, vp, vp],
            "oracle_scs_spmv_f32": [i64, i64, vp, vp, vp, vp, vp, vp],
            "oracle_scs_spmv_bf16": [i64, i64, vp, vp, vp, vp, vp, vp],
            "oracle_bcrs_spmv_f64": [i64, i64, i64, i64, i64, vp, vp, vp, C.c_int, vp, vp],
            "oracle_bcrs_spmv_f32": [i64, i64, i64, i64, i64, vp, vp, vp, C.c_int, vp, vp],
            "oracle_bcrs_spmv_bf16": [i64, i64, i64, i64, i64, vp, vp, vp, C.c_int, vp, vp],
            "oracle_bcrs_spmm_f64": [i64, i64, i64, i64, i64, vp, vp, vp, C.c_int, i64, vp, vp],
            "oracle_bcrs_spmm_f32": [i64, i64, i64, i64, i64, vp, vp, vp, C.c_int, i64, vp, vp],
            "oracle_bcrs_spmm_bf16": [i64, i64, i64, i64, i64, vp, vp, vp, C.c_int, i64, vp, vp],
            "oracle_crs_spmm_f64": [i64, vp, vp, vp, i64, C.c_int, vp, vp],
            "oracle_crs_spmm_f32": [i64, vp, vp, vp, i64, C.c_int, vp, vp],
            "oracle_crs_spmm_bf16": [i64, vp, vp, vp, i64, C.c_int, vp, vp],
            "oracle_sell_sort": [i64, vp, vp],
            "oracle_crs_to_scs_meta": [i64, vp, i64, i64, vp, vp, vp],
            "oracle_crs_to_scs": [i64, vp, vp, vp, i64, i64, i64, vp, vp, vp, vp, vp],
            "oracle_crs_to_bcrs_nblocks": [i64, i64, vp, vp, i64, i64, i64, i64],
            "oracle_crs_to_bcrs": [i64, i64, vp, vp, vp, i64, i64, i64, i64, i64,
                                   C.c_int, vp, vp, vp],
            "oracle_crs_transpose": [i64, i64, vp, vp, vp, i64, vp, vp, vp],
            "oracle_sddmm_f64": [i64, vp, vp, i64, vp, vp, vp],
        }
        for name, args in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = C.c_int if ("meta" in name or name == "oracle_crs_to_scs"
                                    or name == "oracle_crs_to_bcrs") else (
                i64 if name.endswith("nblocks") else None)
        _lib = lib
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


_SUFFIX = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32",
           np.dtype(np.uint16): "bf16"}
_VEC = {"f64": np.float64, "f32": np.float32, "bf16": np.float32}


def crs_spmv(row_ptr, col, val, x, out=None) -> np.ndarray:
    """y = A x; `out` (n_rows entries of the vector type) is written if given."""
    s = _SUFFIX[val.dtype]
    n = len(row_ptr) - 1
    y = np.empty(n, _VEC[s]) if out is None else out
    assert y.dtype == _VEC[s] and y.size >= n
    getattr(_load(), f"oracle_crs_spmv_{s}")(n, _p(row_ptr), _p(col), _p(val),
                                             _p(x), _p(y))
    return y


def scs_spmv(C_, n_chunks, chunk_ptr, chunk_len, col, val, x, out=None) -> np.ndarray:
    s = _SUFFIX[val.dtype]
    y = np.empty(n_chunks * C_, _VEC[s]) if out is None else out
    assert y.dtype == _VEC[s] and y.size >= n_chunks * C_
    getattr(_load(), f"oracle_scs_spmv_{s}")(C_, n_chunks, _p(chunk_ptr),
                                             _p(chunk_len), _p(col), _p(val),
                                             _p(x), _p(y))
    return y


def bcrs_spmv(nbr, b_h, b_w, h_pad, w_pad, col, row_ptr, val, col_major, x,
              y=None) -> np.ndarray:
    s = _SUFFIX[val.dtype]
    if y is None:
        y = np.zeros(nbr * h_pad, _VEC[s])
    getattr(_load(), f"oracle_bcrs_spmv_{s}")(nbr, b_h, b_w, h_pad, w_pad,
                                              _p(col), _p(row_ptr), _p(val),
                                              int(col_major), _p(x), _p(y))
    return y


def bcrs_spmm(nbr, b_h, b_w, h_pad, w_pad, col, row_ptr, val, col_major, k,
              X, out=None) -> np.ndarray:
    s = _SUFFIX[val.dtype]
    Y = np.zeros(nbr * h_pad * k, _VEC[s]) if out is None else out
    assert Y.dtype == _VEC[s] and Y.size >= nbr * h_pad * k
    getattr(_load(), f"oracle_bcrs_spmm_{s}")(nbr, b_h, b_w, h_pad, w_pad,
                                              _p(col), _p(row_ptr), _p(val),
                                              int(col_major), k, _p(X), _p(Y))
    return Y


def crs_spmm(row_ptr, col, val, k, row_major, X, out=None) -> np.ndarray:
    s = _SUFFIX[val.dtype]
    n = len(row_ptr) - 1
    Y = np.empty(n * k, _VEC[s]) if out is None else out
    assert Y.dtype == _VEC[s] and Y.size >= n * k
    getattr(_load(), f"oracle_crs_spmm_{s}")(n, _p(row_ptr), _p(col), _p(val),
                                             k, int(row_major), _p(X), _p(Y))
    return Y


def sell_sort(index: np.ndarray, count: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    idx = np.ascontiguousarray(index, np.int32).copy()
    cnt = np.ascontiguousarray(count, np.int32).copy()
    _load().oracle_sell_sort(len(idx), _p(idx), _p(cnt))
    return idx, cnt


def crs_to_scs(row_ptr, col, val, C_, sigma) -> dict:
    n_rows = len(row_ptr) - 1
    nc, npad, ne = i64(), i64(), i64()
    rc = _load().oracle_crs_to_scs_meta(n_rows, _p(row_ptr), C_, sigma,
                                        C.byref(nc), C.byref(npad), C.byref(ne))
    if rc:
        raise ValueError("invalid SELL arguments")
    out = {
        "C": C_, "sigma": sigma, "n_rows": n_rows, "n_rows_padded": npad.value,
        "n_chunks": nc.value, "n_elements": ne.value, "nnz": int(row_ptr[-1]),
        "chunk_ptr": np.empty(nc.value + 1, np.int32),
        "chunk_lengths": np.empty(nc.value, np.int32),
        "col": np.empty(ne.value, np.int32),
        "val": np.empty(ne.value, val.dtype),
        "perm": np.empty(n_rows, np.int32),
    }
    rc = _load().oracle_crs_to_scs(n_rows, _p(row_ptr), _p(col), _p(val),
                                   val.dtype.itemsize, C_, sigma,
                                   _p(out["chunk_ptr"]), _p(out["chunk_lengths"]),
                                   _p(out["col"]), _p(out["val"]), _p(out["perm"]))
    if rc:
        raise ValueError("conversion failed")
    return out


def crs_to_bcrs(n_rows, n_cols, row_ptr, col, val, b_h, b_w, h_pad, w_pad,
                col_major) -> dict:
    nb = _load().oracle_crs_to_bcrs_nblocks(n_rows, n_cols, _p(row_ptr), _p(col),
                                            b_h, b_w, h_pad, w_pad)
    if nb < 0:
        raise ValueError("block size does not divide the matrix or bad padding")
    nbr = n_rows // b_h
    out = {"n_block_rows": nbr, "n_block_cols": n_cols // b_w, "n_blocks": nb,
           "col": np.empty(nb, np.int32), "row_ptr": np.empty(nbr + 1, np.int32),
           "val": np.empty(nb * h_pad * w_pad, val.dtype)}
    rc = _load().oracle_crs_to_bcrs(n_rows, n_cols, _p(row_ptr), _p(col), _p(val),
                                    val.dtype.itemsize, b_h, b_w, h_pad, w_pad,
                                    int(col_major), _p(out["col"]),
                                    _p(out["row_ptr"]), _p(out["val"]))
    if rc:
        raise ValueError(f"conversion failed ({rc})")
    return out


def crs_transpose(n_rows, n_cols, row_ptr, col, val):
    nnz = int(row_ptr[-1])
    trp = np.empty(n_cols + 1, np.int32)
    tcol = np.empty(nnz, np.int32)
    tval = np.empty(nnz, val.dtype)
    _load().oracle_crs_transpose(n_rows, n_cols, _p(row_ptr), _p(col), _p(val),
                                 val.dtype.itemsize, _p(trp), _p(tcol), _p(tval))
    return trp, tcol, tval


def sddmm_f64(row_ptr, col, k, G, X) -> np.ndarray:
    n = len(row_ptr) - 1
    out = np.empty(int(row_ptr[-1]), np.float64)
    _load().oracle_sddmm_f64(n, _p(row_ptr), _p(col), k, _p(G), _p(X), _p(out))
    return out
