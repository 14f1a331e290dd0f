"""ctypes binding of libsmax_hip.so (declarations in include/smax_hip.h).

The library is the product; this module only marshals arguments.  Loading
fails loudly when the library has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_build" / "libsmax_hip.so"

SUCCESS = 0
ERROR_INVALID_VALUE = 1
ERROR_UNSUPPORTED = 2
ERROR_OUT_OF_MEMORY = 3
ERROR_LAUNCH = 4
ERROR_RUNTIME = 5
ERROR_NO_DEVICE = 6
ERROR_OVERFLOW = 7
ERROR_NOT_CONVERGED = 8

F64, F32, BF16 = 0, 1, 2
INT16, INT32, INT64, UINT16, UINT32, UINT64 = range(6)
COLUMNS_AUTO, COLUMNS_32, COLUMNS_16, COLUMNS_8 = range(4)

VALUE_NP = {F64: np.float64, F32: np.float32, BF16: np.uint16}
VECTOR_NP = {F64: np.float64, F32: np.float32, BF16: np.float32}

i64 = C.c_int64
i32p = C.POINTER(C.c_int32)
vp = C.c_void_p


class SmaxHipError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"smax_hip status {status}: {message}")
        self.status = status


class ScsInfo(C.Structure):
    _fields_ = [(n, i64) for n in ("C", "sigma", "n_rows", "n_rows_padded",
                                   "n_cols", "n_chunks", "n_elements", "nnz",
                                   "val_bytes")] + [
        ("chunk_ptr", vp), ("chunk_lengths", vp), ("col", vp), ("val", vp),
        ("perm", vp), ("slot_row", vp), ("src_index", vp), ("slot_len", vp)]


class BcrsInfo(C.Structure):
    _fields_ = [(n, i64) for n in ("n_block_rows", "n_block_cols", "n_blocks",
                                   "b_h", "b_w", "h_pad", "w_pad", "val_bytes",
                                   "block_col_major")] + [
        ("col", vp), ("row_ptr", vp), ("val", vp), ("src_block", vp)]


class CgParams(C.Structure):
    _fields_ = [("rel_tol", C.c_double), ("max_iters", i64),
                ("check_every", i64), ("inner_rel_tol", C.c_double),
                ("max_inner", i64), ("inner_value_type", C.c_int32),
                ("keep_diag_f32", C.c_int32), ("restart_inner", C.c_int32),
                ("column_coding", C.c_int32)]


class CgReport(C.Structure):
    _fields_ = [("iterations", i64), ("outer_iterations", i64),
                ("rel_residual", C.c_double), ("seconds", C.c_double),
                ("setup_seconds", C.c_double), ("column_coding", C.c_int32)]


_SIGS = {
    "smax_hip_abi_version": ([], C.c_int),
    "smax_hip_check_types": ([C.c_int, C.c_int], C.c_int),
    "smax_hip_ctx_create": ([C.c_int, C.POINTER(vp)], C.c_int),
    "smax_hip_ctx_destroy": ([vp], C.c_int),
    "smax_hip_last_error": ([vp], C.c_char_p),
    "smax_hip_set_stream": ([vp, vp], C.c_int),
    "smax_hip_get_stream": ([vp], vp),
    "smax_hip_sync": ([vp], C.c_int),
    "smax_hip_malloc": ([vp, C.POINTER(vp), C.c_size_t], C.c_int),
    "smax_hip_free": ([vp, vp], C.c_int),
    "smax_hip_memcpy_h2d": ([vp, vp, vp, C.c_size_t], C.c_int),
    "smax_hip_memcpy_d2h": ([vp, vp, vp, C.c_size_t], C.c_int),
    "smax_hip_memcpy_d2d": ([vp, vp, vp, C.c_size_t], C.c_int),
    "smax_hip_zero": ([vp, vp, C.c_size_t], C.c_int),
    "smax_hip_timer_start": ([vp], C.c_int),
    "smax_hip_timer_stop": ([vp, C.POINTER(C.c_float)], C.c_int),
    "smax_hip_crs_spmv": ([vp, C.c_int, i64, i64, i64, vp, vp, vp, vp, vp, vp],
                          C.c_int),
    "smax_hip_crs_plan_create": ([vp, i64, i64, vp, C.POINTER(vp)], C.c_int),
    "smax_hip_crs_plan_destroy": ([vp, vp], C.c_int),
    "smax_hip_crs_spmm": ([vp, C.c_int, i64, i64, i64, vp, vp, vp, i64, C.c_int,
                           vp, i64, vp, i64, vp], C.c_int),
    "smax_hip_crs_spmm_noalloc": ([vp, C.c_int, i64, i64, i64, vp, vp, vp, i64, C.c_int,
                                   vp, i64, vp, i64, vp], C.c_int),
    "smax_hip_scs_spmv": ([vp, C.c_int, i64, i64, vp, vp, vp, vp, vp, vp],
                          C.c_int),
    "smax_hip_scs_spmv_unpermuted": ([vp, C.c_int, i64, i64, i64, vp, vp, vp,
                                      vp, vp, vp, vp], C.c_int),
    "smax_hip_scs_plan_create": ([vp, i64, i64, i64, i64, i64, i64, vp, vp], C.c_int),
    "smax_hip_scs_plan_destroy": ([vp, vp], C.c_int),
    "smax_hip_scs_spmv_plan": ([vp, C.c_int, vp, vp, vp, vp, vp, vp, vp], C.c_int),
    "smax_hip_scs_apply_spmm": ([vp, C.c_int, vp, i64, vp, vp], C.c_int),
    "smax_hip_scs_spmm": ([vp, C.c_int, i64, i64, i64, vp, vp, vp, vp, vp, i64,
                           vp, vp], C.c_int),
    "smax_hip_scs_from_crs": ([vp, i64, i64, i64, vp, vp, vp, i64, i64, i64,
                               C.POINTER(vp)], C.c_int),
    "smax_hip_scs_get_info": ([vp, C.POINTER(ScsInfo)], C.c_int),
    "smax_hip_scs_destroy": ([vp, vp], C.c_int),
    "smax_hip_scs_refresh_values": ([vp, vp, vp], C.c_int),
    "smax_hip_scs_get_tails": ([vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)], C.c_int),
    "smax_hip_scs_from_sell": ([vp, i64, i64, i64, i64, i64, i64, i64, i64, vp, vp, vp, vp, i64,
                                vp, C.POINTER(vp)], C.c_int),
    "smax_hip_scs_apply": ([vp, C.c_int, vp, vp, vp, C.c_int], C.c_int),
    "smax_hip_bcrs_spmv": ([vp, C.c_int, i64, i64, i64, i64, i64, vp, vp, vp,
                            C.c_int, vp, vp], C.c_int),
    "smax_hip_bcrs_spmm": ([vp, C.c_int, i64, i64, i64, i64, i64, vp, vp, vp,
                            C.c_int, i64, vp, vp], C.c_int),
    "smax_hip_bcrs_from_crs": ([vp, i64, i64, i64, vp, vp, vp, i64, i64, i64,
                                i64, i64, C.c_int, C.POINTER(vp)], C.c_int),
    "smax_hip_bcrs_get_info": ([vp, C.POINTER(BcrsInfo)], C.c_int),
    "smax_hip_bcrs_destroy": ([vp, vp], C.c_int),
    "smax_hip_scs_transpose_from_crs": ([vp, i64, i64, i64, vp, vp, vp, i64,
                                         i64, i64, C.POINTER(vp)], C.c_int),
    "smax_hip_scs_sddmm": ([vp, C.c_int, vp, i64, vp, vp, vp], C.c_int),
    "smax_hip_bcrs_transpose": ([vp, vp, C.POINTER(vp)], C.c_int),
    "smax_hip_bcrs_refresh_values": ([vp, vp, vp], C.c_int),
    "smax_hip_bcrs_sddmm": ([vp, C.c_int, vp, i64, vp, vp, vp], C.c_int),
    "smax_hip_cg_f64": ([vp, vp, vp, vp, C.POINTER(CgParams),
                         C.POINTER(CgReport)], C.c_int),
    "smax_hip_ircg": ([vp, vp, vp, vp, C.POINTER(CgParams),
                       C.POINTER(CgReport)], C.c_int),
    "smax_hip_stream_copy": ([vp, vp, vp, i64], C.c_int),
    "smax_hip_stream_triad": ([vp, vp, vp, vp, C.c_double, i64], C.c_int),
    "smax_hip_stream_copy_wide": ([vp, vp, vp, i64, C.c_int], C.c_int),
    "smax_hip_gather_rate": ([vp, vp, vp, i64, i64, C.c_int], C.c_int),
}

_lib = None


def load(path: Path | str | None = None) -> C.CDLL:
    """Loads libsmax_hip.so; raises if it is missing (build it first)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"{p} not found: build the HIP library with "
            "`python -m smaxkernels_amd.build` (there is no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def _ptr(x) -> int | None:
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return x.ptr
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):  # torch tensor
        return x.data_ptr()
    raise TypeError(f"not a device pointer: {type(x)}")


class DeviceArray:
    """A device allocation owned by a Context (freed with it or by free())."""

    def __init__(self, ctx: "Context", ptr: int, shape, dtype, owned=True):
        self.ctx, self.ptr, self.dtype = ctx, ptr, np.dtype(dtype)
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.owned = owned

    @property
    def size(self) -> int:
        return int(np.prod(self.shape)) if self.shape else 1

    @property
    def nbytes(self) -> int:
        return self.size * self.dtype.itemsize

    def numpy(self) -> np.ndarray:
        return self.ctx.download(self)

    def free(self) -> None:
        if self.owned and self.ptr:
            self.ctx._check(self.ctx.lib.smax_hip_free(self.ctx.handle, self.ptr))
            self.ctx._owned.discard(self.ptr)
        self.ptr = 0


class Context:
    def __init__(self, device: int = 0, lib: C.CDLL | None = None):
        self.lib = lib or load()
        h = vp()
        st = self.lib.smax_hip_ctx_create(device, C.byref(h))
        if st:
            raise SmaxHipError(st, self.lib.smax_hip_last_error(None).decode())
        self.handle = h.value
        self._owned: set[int] = set()

    # ------------------------------------------------------------- basics
    def _check(self, st: int) -> None:
        if st:
            raise SmaxHipError(st, self.lib.smax_hip_last_error(self.handle).decode())

    def close(self) -> None:
        if self.handle:
            self.lib.smax_hip_sync(self.handle)
            for p in list(self._owned):
                self.lib.smax_hip_free(self.handle, p)
            self._owned.clear()
            self.lib.smax_hip_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_stream(self, stream) -> None:
        """Run subsequent work on `stream` (a hipStream_t as int); None restores
        the context's own stream."""
        self._check(self.lib.smax_hip_set_stream(self.handle, stream))

    def sync(self) -> None:
        self._check(self.lib.smax_hip_sync(self.handle))

    def empty(self, shape, dtype) -> DeviceArray:
        dtype = np.dtype(dtype)
        n = int(np.prod(shape))
        p = vp()
        self._check(self.lib.smax_hip_malloc(self.handle, C.byref(p),
                                             max(1, n * dtype.itemsize)))
        self._owned.add(p.value)
        return DeviceArray(self, p.value, shape, dtype)

    def zeros(self, shape, dtype) -> DeviceArray:
        a = self.empty(shape, dtype)
        self._check(self.lib.smax_hip_zero(self.handle, a.ptr, a.nbytes))
        return a

    def upload(self, arr: np.ndarray) -> DeviceArray:
        arr = np.ascontiguousarray(arr)
        d = self.empty(arr.shape, arr.dtype)
        if arr.nbytes:
            self._check(self.lib.smax_hip_memcpy_h2d(
                self.handle, d.ptr, arr.ctypes.data, arr.nbytes))
        return d

    def download(self, d: DeviceArray, count: int | None = None) -> np.ndarray:
        n = d.size if count is None else count
        out = np.empty(n, dtype=d.dtype)
        if n:
            self._check(self.lib.smax_hip_memcpy_d2h(
                self.handle, out.ctypes.data, d.ptr, n * d.dtype.itemsize))
        return out.reshape(d.shape) if count is None else out

    def download_ptr(self, ptr: int, count: int, dtype) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if count:
            self._check(self.lib.smax_hip_memcpy_d2h(
                self.handle, out.ctypes.data, ptr, out.nbytes))
        return out

    def timer_start(self) -> None:
        self._check(self.lib.smax_hip_timer_start(self.handle))

    def timer_stop(self) -> float:
        ms = C.c_float()
        self._check(self.lib.smax_hip_timer_stop(self.handle, C.byref(ms)))
        return ms.value

    # ------------------------------------------------------------- kernels
    def crs_spmv(self, vtype, n_rows, n_cols, nnz, row_ptr, col, val, x, y,
                 plan=None) -> None:
        self._check(self.lib.smax_hip_crs_spmv(
            self.handle, vtype, n_rows, n_cols, nnz, _ptr(row_ptr), _ptr(col),
            _ptr(val), _ptr(x), _ptr(y), plan))

    def crs_plan(self, n_rows, nnz, row_ptr) -> int:
        p = vp()
        self._check(self.lib.smax_hip_crs_plan_create(
            self.handle, n_rows, nnz, _ptr(row_ptr), C.byref(p)))
        return p.value

    def crs_plan_destroy(self, plan) -> None:
        self._check(self.lib.smax_hip_crs_plan_destroy(self.handle, plan))

    def crs_spmm(self, vtype, n_rows, n_cols, nnz, row_ptr, col, val, k,
                 row_major, X, ldx, Y, ldy, plan=None) -> None:
        self._check(self.lib.smax_hip_crs_spmm(
            self.handle, vtype, n_rows, n_cols, nnz, _ptr(row_ptr), _ptr(col),
            _ptr(val), k, int(row_major), _ptr(X), ldx, _ptr(Y), ldy, plan))

    def scs_spmv(self, vtype, C_, n_chunks, chunk_ptr, chunk_len, col, val, x,
                 y) -> None:
        self._check(self.lib.smax_hip_scs_spmv(
            self.handle, vtype, C_, n_chunks, _ptr(chunk_ptr), _ptr(chunk_len),
            _ptr(col), _ptr(val), _ptr(x), _ptr(y)))

    def scs_plan(self, C_, sigma, n_rows_padded, n_chunks, n_elements, nnz,
                 chunk_lengths) -> int:
        p = vp()
        self._check(self.lib.smax_hip_scs_plan_create(
            self.handle, C_, sigma, n_rows_padded, n_chunks, n_elements, nnz,
            _ptr(chunk_lengths), C.byref(p)))
        return p.value

    def scs_plan_destroy(self, plan) -> None:
        self._check(self.lib.smax_hip_scs_plan_destroy(self.handle, plan))

    def scs_spmv_plan(self, vtype, plan, chunk_ptr, chunk_len, col, val, x, y) -> None:
        self._check(self.lib.smax_hip_scs_spmv_plan(
            self.handle, vtype, plan, _ptr(chunk_ptr), _ptr(chunk_len), _ptr(col),
            _ptr(val), _ptr(x), _ptr(y)))

    def scs_from_sell(self, C_, sigma, n_rows, n_rows_padded, n_cols, n_chunks, n_elements,
                      nnz, chunk_ptr, chunk_lengths, col, val, val_bytes, perm) -> "Scs":
        """A SpMV-only handle over device SELL arrays (copied)."""
        h = vp()
        self._check(self.lib.smax_hip_scs_from_sell(
            self.handle, C_, sigma, n_rows, n_rows_padded, n_cols, n_chunks, n_elements, nnz,
            _ptr(chunk_ptr), _ptr(chunk_lengths), _ptr(col), _ptr(val), val_bytes, _ptr(perm),
            C.byref(h)))
        return Scs(self, h.value)

    def scs_from_crs(self, n_rows, n_cols, nnz, row_ptr, col, val, val_bytes,
                     C_, sigma) -> "Scs":
        h = vp()
        self._check(self.lib.smax_hip_scs_from_crs(
            self.handle, n_rows, n_cols, nnz, _ptr(row_ptr), _ptr(col),
            _ptr(val), val_bytes, C_, sigma, C.byref(h)))
        return Scs(self, h.value)

    def scs_transpose_from_crs(self, n_rows, n_cols, nnz, row_ptr, col, val,
                               val_bytes, C_, sigma) -> "Scs":
        h = vp()
        self._check(self.lib.smax_hip_scs_transpose_from_crs(
            self.handle, n_rows, n_cols, nnz, _ptr(row_ptr), _ptr(col),
            _ptr(val), val_bytes, C_, sigma, C.byref(h)))
        return Scs(self, h.value)

    def bcrs_spmv(self, vtype, nbr, b_h, b_w, h_pad, w_pad, col, row_ptr, val,
                  col_major, x, y) -> None:
        self._check(self.lib.smax_hip_bcrs_spmv(
            self.handle, vtype, nbr, b_h, b_w, h_pad, w_pad, _ptr(col),
            _ptr(row_ptr), _ptr(val), int(col_major), _ptr(x), _ptr(y)))

    def bcrs_spmm(self, vtype, nbr, b_h, b_w, h_pad, w_pad, col, row_ptr, val,
                  col_major, k, X, Y) -> None:
        self._check(self.lib.smax_hip_bcrs_spmm(
            self.handle, vtype, nbr, b_h, b_w, h_pad, w_pad, _ptr(col),
            _ptr(row_ptr), _ptr(val), int(col_major), k, _ptr(X), _ptr(Y)))

    def bcrs_from_crs(self, n_rows, n_cols, nnz, row_ptr, col, val, val_bytes,
                      b_h, b_w, h_pad, w_pad, col_major) -> "Bcrs":
        h = vp()
        self._check(self.lib.smax_hip_bcrs_from_crs(
            self.handle, n_rows, n_cols, nnz, _ptr(row_ptr), _ptr(col),
            _ptr(val), val_bytes, b_h, b_w, h_pad, w_pad, int(col_major),
            C.byref(h)))
        return Bcrs(self, h.value)

    def stream_copy(self, src, dst, n) -> None:
        self._check(self.lib.smax_hip_stream_copy(self.handle, _ptr(src),
                                                  _ptr(dst), n))

    def stream_copy_wide(self, src, dst, n, loads_per_lane) -> None:
        self._check(self.lib.smax_hip_stream_copy_wide(self.handle, _ptr(src), _ptr(dst),
                                                       n, loads_per_lane))

    def stream_triad(self, a, b, c, s, n) -> None:
        self._check(self.lib.smax_hip_stream_triad(self.handle, _ptr(a),
                                                   _ptr(b), _ptr(c), s, n))

    def gather_rate(self, x, y, n, window, per_lane) -> None:
        self._check(self.lib.smax_hip_gather_rate(self.handle, _ptr(x), _ptr(y), n,
                                                  window, per_lane))

    def cg(self, A: "Scs", b, x, params: CgParams, mixed: bool) -> CgReport:
        rep = CgReport()
        fn = self.lib.smax_hip_ircg if mixed else self.lib.smax_hip_cg_f64
        st = fn(self.handle, A.handle, _ptr(b), _ptr(x), C.byref(params),
                C.byref(rep))
        if st and st != ERROR_NOT_CONVERGED:
            self._check(st)
        rep.status = st
        return rep


class Scs:
    """SELL-C-sigma matrix on the device, built by the GPU conversion."""

    def __init__(self, ctx: Context, handle: int):
        self.ctx, self.handle = ctx, handle
        self.info = ScsInfo()
        ctx._check(ctx.lib.smax_hip_scs_get_info(handle, C.byref(self.info)))

    def __getattr__(self, name):
        if name in ("info", "ctx", "handle"):
            raise AttributeError(name)
        return getattr(self.info, name)

    def arrays(self, dtype) -> dict:
        """Downloads every array of the matrix (for checks)."""
        i, d = self.info, self.ctx.download_ptr
        return {
            "chunk_ptr": d(i.chunk_ptr, i.n_chunks + 1, np.int32),
            "chunk_lengths": d(i.chunk_lengths, i.n_chunks, np.int32),
            "col": d(i.col, i.n_elements, np.int32),
            "val": d(i.val, i.n_elements, dtype),
            "perm": d(i.perm, i.n_rows, np.int32),
            "slot_row": d(i.slot_row, i.n_rows_padded, np.int32),
            "src_index": d(i.src_index, i.n_elements, np.int32) if i.src_index else None,
            "slot_len": d(i.slot_len, i.n_rows_padded, np.int32),
        }

    def spmv(self, vtype, x, y, unpermuted=False, val=None) -> None:
        i, c = self.info, self.ctx
        v = _ptr(val) if val is not None else i.val
        if unpermuted:
            c._check(c.lib.smax_hip_scs_spmv_unpermuted(
                c.handle, vtype, i.C, i.n_chunks, i.n_rows, i.chunk_ptr,
                i.chunk_lengths, i.col, v, i.slot_row, _ptr(x), _ptr(y)))
        else:
            c._check(c.lib.smax_hip_scs_spmv(
                c.handle, vtype, i.C, i.n_chunks, i.chunk_ptr, i.chunk_lengths,
                i.col, v, _ptr(x), _ptr(y)))

    def apply(self, vtype, x, y, unpermuted=False) -> None:
        c = self.ctx
        c._check(c.lib.smax_hip_scs_apply(c.handle, vtype, self.handle, _ptr(x),
                                          _ptr(y), int(unpermuted)))

    def spmm(self, vtype, k, X, Y, val=None) -> None:
        i, c = self.info, self.ctx
        v = _ptr(val) if val is not None else i.val
        c._check(c.lib.smax_hip_scs_spmm(
            c.handle, vtype, i.C, i.n_chunks, i.n_rows, i.chunk_ptr,
            i.chunk_lengths, i.col, v, i.slot_row, k, _ptr(X), _ptr(Y)))

    def apply_spmm(self, vtype, k, X, Y) -> None:
        """Y = A X over the stored entries only (equals the CRS SpMM)."""
        c = self.ctx
        c._check(c.lib.smax_hip_scs_apply_spmm(c.handle, vtype, self.handle, k,
                                               _ptr(X), _ptr(Y)))

    def sddmm(self, vtype, k, G, X, dvals) -> None:
        c = self.ctx
        c._check(c.lib.smax_hip_scs_sddmm(c.handle, vtype, self.handle, k,
                                          _ptr(G), _ptr(X), _ptr(dvals)))

    def tails(self) -> tuple:
        """(slots, waves, stored entries) of the handle's tail waves."""
        out = [i64(), i64(), i64()]
        self.ctx._check(self.ctx.lib.smax_hip_scs_get_tails(self.handle, *map(C.byref, out)))
        return tuple(v.value for v in out)

    def refresh_values(self, crs_val) -> None:
        c = self.ctx
        c._check(c.lib.smax_hip_scs_refresh_values(c.handle, self.handle,
                                                   _ptr(crs_val)))

    def destroy(self) -> None:
        if self.handle:
            self.ctx.lib.smax_hip_scs_destroy(self.ctx.handle, self.handle)
            self.handle = None


class Bcrs:
    """BCRS matrix on the device, built by the GPU conversion."""

    def __init__(self, ctx: Context, handle: int):
        self.ctx, self.handle = ctx, handle
        self.info = BcrsInfo()
        ctx._check(ctx.lib.smax_hip_bcrs_get_info(handle, C.byref(self.info)))

    def arrays(self, dtype) -> dict:
        i, d = self.info, self.ctx.download_ptr
        return {
            "col": d(i.col, i.n_blocks, np.int32),
            "row_ptr": d(i.row_ptr, i.n_block_rows + 1, np.int32),
            "val": d(i.val, i.n_blocks * i.h_pad * i.w_pad, dtype),
        }

    def spmm(self, vtype, k, X, Y) -> None:
        i, c = self.info, self.ctx
        c._check(c.lib.smax_hip_bcrs_spmm(
            c.handle, vtype, i.n_block_rows, i.b_h, i.b_w, i.h_pad, i.w_pad, i.col,
            i.row_ptr, i.val, int(i.block_col_major), k, _ptr(X), _ptr(Y)))

    def spmv(self, vtype, x, y) -> None:
        i, c = self.info, self.ctx
        c._check(c.lib.smax_hip_bcrs_spmv(
            c.handle, vtype, i.n_block_rows, i.b_h, i.b_w, i.h_pad, i.w_pad, i.col,
            i.row_ptr, i.val, int(i.block_col_major), _ptr(x), _ptr(y)))

    def transpose(self) -> "Bcrs":
        h = vp()
        self.ctx._check(self.ctx.lib.smax_hip_bcrs_transpose(
            self.ctx.handle, self.handle, C.byref(h)))
        return Bcrs(self.ctx, h.value)

    def refresh_values(self, src_val) -> None:
        c = self.ctx
        c._check(c.lib.smax_hip_bcrs_refresh_values(c.handle, self.handle,
                                                    _ptr(src_val)))

    def sddmm(self, vtype, k, G, X, dvals) -> None:
        c = self.ctx
        c._check(c.lib.smax_hip_bcrs_sddmm(c.handle, vtype, self.handle, k,
                                           _ptr(G), _ptr(X), _ptr(dvals)))

    def destroy(self) -> None:
        if self.handle:
            self.ctx.lib.smax_hip_bcrs_destroy(self.ctx.handle, self.handle)
            self.handle = None
