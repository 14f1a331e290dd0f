"""Benchmark of the MI355X SpMV path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]

A step is one SpMV over the whole matrix with operands resident in HBM.
Headline workload: FP64 SELL-C-sigma SpMV on the 4M-row Lomax band matrix
(N = 4,000,000, row lengths Lomax(1.5) with mean 16 and cap 2000, columns
uniform within +-N/8, values and x uniform [-1, 1)).  The value is algorithmic
GB/s: B_alg = nnz*(8+4) + (N+1)*4 + N*8 (x) + N*8 (y) per SpMV, divided by
the average kernel time measured with HIP events on the library's stream.

Also reported: every (C, sigma) variant, CRS (lane-per-row and LDS row-block
schedules), rocSPARSE's CSR SpMV algorithms on the same arrays (library
baseline, comparison only), FP32/BF16-value SELL, the 512^2 5-point
Laplacian (warm and cold L2/MALL, and the kernel-only cold duration from a
committed rocprofv3 trace), FP64 CG vs BF16/FP32-value refined CG on the
160^3 27-point stencil (with first-solve setup), FP32 SpMM k=32 with its
adjoint and SDDMM (1M-row Lomax band), BCRS 4x4 BF16 SpMM forward and
adjoint (96^3 7-point block adjacency, k = 8 and 64), the device copy
bandwidth by several methods (the best is the measured stream peak), and the
oracle's OpenMP kernels on the host cores: the CRS SpMV of the headline
matrix (bounded sample) as the CPU baseline, and one kernel per config.
Every FP32/BF16 result carries its error against the FP64 product and the
bound that error must stay under.

    python bench.py --summarize-trace DIR

prints per-kernel durations from rocprofv3 --kernel-trace CSV output instead
(no GPU work); profiles/c1_poisson512_cold_warm_kernels.json is made this way.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

# the CPUs this process may use, read before any OpenMP runtime loads (torch's
# pins the main thread to one core when OMP_PROC_BIND is set)
CPUS_AT_START = len(os.sched_getaffinity(0))

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# the CPU baseline's OpenMP threads: bound close, one per core (read when the
# oracle's OpenMP runtime starts, which is after this point)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

# torch first: it brings its own HIP runtime, and libsmax_hip.so then binds to
# the same one (the C1 warm figure replays a torch-captured HIP graph).
import torch  # noqa: E402,F401

from smaxkernels_amd import capi, matgen  # noqa: E402

METRIC = ("SpMV GB/s & % HBM peak; BF16-refined vs FP64 CG speedup at 1e-10 "
          "residual")
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def b_alg(n_rows, n_cols, nnz, vb=8, ib=4, xb=8, yb=8) -> int:
    return nnz * (vb + ib) + (n_rows + 1) * ib + n_cols * xb + n_rows * yb


U32 = 2.0 ** -24
U_VALUE = {"f32": U32, "bf16": 2.0 ** -8 + 2.0 ** -24 + 2.0 ** -32}


def err_vs_fp64(y_lo, y64, scale, n_max, u_v) -> dict:
    """Largest |y_lo - y_64| / (|A| |x|) over the outputs, next to its bound
    u_v + gamma_n (1 + u_v) (u_v: relative rounding of a stored value;
    gamma_n = n u / (1 - n u), u = 2^-24, n the longest chain).  y_64 is
    the FP64 product of the FP64 matrix with the same (exactly widened)
    operands, computed by the FP64 kernels, which the tests show
    bit-identical to the oracle."""
    y_lo = np.asarray(y_lo, np.float64).ravel()
    y64, scale = np.asarray(y64).ravel(), np.asarray(scale).ravel()
    nz = scale > 0
    ratio = float(np.max(np.abs(y_lo[nz] - y64[nz]) / scale[nz])) if nz.any() else 0.0
    g = n_max * U32 / (1 - n_max * U32)
    return {"max_rel_err_vs_fp64": ratio, "bound": u_v + g * (1 + u_v),
            "metric": "max |y_lo - y_64| / (|A||x|) per output"}


def time_launches(ctx, fn, steps, warmup) -> float:
    """Average ms per launch over `steps` back-to-back launches."""
    for _ in range(warmup):
        fn()
    ctx.sync()
    ctx.timer_start()
    for _ in range(steps):
        fn()
    return ctx.timer_stop() / steps


def graph_ms(ctx, fn, per_graph=200, replays=10, check=None) -> float:
    """Average ms per call of `fn` replayed from a HIP graph that holds
    `per_graph` calls (no host launch cost between kernels).  `check`, if
    given, is called after the eager run and after the replays and must
    return equal results."""
    import torch
    s = torch.cuda.Stream()
    ctx.set_stream(s.cuda_stream)
    try:
        fn()
        torch.cuda.synchronize()
        want = check() if check else None
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            for _ in range(per_graph):
                fn()
        with torch.cuda.stream(s):
            g.replay()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(s):
            e0.record()
            for _ in range(replays):
                g.replay()
            e1.record()
        torch.cuda.synchronize()
        if check is not None and not np.array_equal(check(), want):
            raise RuntimeError("graph replay result differs from the eager run")
        return e0.elapsed_time(e1) / (replays * per_graph)
    finally:
        ctx.set_stream(None)


def stream_peak(ctx, steps=10) -> dict:
    """Device copy bandwidth by several methods over 4 GiB arrays: the copy
    kernel with 1, 2, 4 and 8 16-byte loads in flight per lane, and
    hipMemcpy device-to-device; each the best of `steps` event-timed runs
    (read + write bytes).  The best of them is the measured stream peak that
    every frac_of_measured_stream divides by.  The triad is reported too."""
    n = 1 << 29  # 4 GiB per array
    a = ctx.zeros(n, np.float64)
    b = ctx.zeros(n, np.float64)
    gbs = lambda ms, arrays=2: round(arrays * n * 8 / ms / 1e6, 1)
    methods = {}
    ms = min(time_launches(ctx, lambda: ctx.stream_copy(a, b, n), 1, 2) for _ in range(steps))
    methods["copy_kernel"] = gbs(ms)
    for L in (2, 4, 8):
        ms = min(time_launches(ctx, lambda: ctx.stream_copy_wide(a, b, n, L), 1, 2)
                 for _ in range(steps))
        methods[f"copy_wide_kernel_{L}_loads"] = gbs(ms)
    d2d = lambda: ctx._check(ctx.lib.smax_hip_memcpy_d2d(ctx.handle, b.ptr, a.ptr, n * 8))
    ms = min(time_launches(ctx, d2d, 1, 2) for _ in range(steps))
    methods["hipMemcpy_d2d"] = gbs(ms)
    c = ctx.zeros(n, np.float64)
    ms3 = min(time_launches(ctx, lambda: ctx.stream_triad(a, b, c, 0.5, n), 1, 2)
              for _ in range(steps))
    for d in (a, b, c):
        d.free()
    best = max(methods, key=methods.get)
    return {"copy_GBs": methods[best], "copy_method": best, "copy_methods_GBs": methods,
            "triad_GBs": gbs(ms3, 3), "array_GiB": 4}


def gather_ceiling(ctx, A, steps, warmup) -> dict:
    """Random-gather ceiling for A's x accesses: n_rows lanes each read 16
    uniformly random entries of x from a window as wide as A's band
    (max - min of col - row), with no index stream.  An SpMV over A makes
    the same number of gathers over the same window and cannot beat this
    rate; its own gathers per second over this ceiling is the fraction of
    the gather roofline it reaches."""
    rows = np.repeat(np.arange(A.n_rows, dtype=np.int64), np.diff(A.row_ptr))
    rel = A.col.astype(np.int64) - rows
    n = min(A.n_rows, A.n_cols)
    window = int(min(rel.max() - rel.min() + 1, n))
    del rows, rel
    per_lane = 16
    x = ctx.upload(np.ones(n, np.float64))
    y = ctx.zeros(n, np.float64)
    ms = time_launches(ctx, lambda: ctx.gather_rate(x, y, n, window, per_lane), steps, warmup)
    ok = bool(np.all(ctx.download(y) == per_lane))
    x.free()
    y.free()
    if not ok:
        raise RuntimeError("gather_rate kernel returned wrong sums")
    return {"window_entries": window, "window_MB": round(window * 8 / 1e6, 2),
            "lanes": n, "per_lane": per_lane, "ms": round(ms, 4),
            "G_gathers_per_s": round(n * per_lane / ms / 1e6, 1)}


def hbm_traffic(kernel: str, n_rows: int, fmt: str):
    """Bytes per launch of `kernel` between L2 and the fabric (HBM + MALL) on
    the N-row Lomax workload in format `fmt`, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (calibrated against the copy kernel), or
    None when that kernel, size and format were not profiled."""
    f = ROOT / "profiles" / "hbm_traffic.json"
    if not f.exists():
        return None
    rec = json.loads(f.read_text()).get(f"{kernel}@lomax_band_N{n_rows}@{fmt}")
    return None if rec is None else rec["bytes_per_launch"]


def doubling_time(fn, min_seconds) -> tuple[float, int, float]:
    """Seconds per call of fn: one warm-up call, then a doubling iteration
    count until one timed batch lasts >= min_seconds (the method of the
    reference's benchmarks, examples/benchmarks/benchmarks_common.hpp:60-95).
    Returns (seconds per call, calls in the last batch, its duration)."""
    fn()
    n = 1
    while True:
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        dt = time.perf_counter() - t0
        if dt >= min_seconds or n >= 1 << 20:
            return dt / n, n, dt
        n *= 2


def host_description() -> dict:
    """The host the CPU baseline ran on: lscpu's model, sockets, cores and
    NUMA nodes, the CPUs this process may use, and the OpenMP settings."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            if key.strip() in ("Model name", "Socket(s)", "Core(s) per socket",
                               "Thread(s) per core", "NUMA node(s)", "CPU(s)"):
                info[key.strip()] = val.strip()
    except (OSError, subprocess.SubprocessError) as e:
        info["lscpu_error"] = str(e)
    info["nproc"] = os.cpu_count()
    info["affinity_cpus"] = CPUS_AT_START
    for var in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES"):
        info[var] = os.environ.get(var)
    return info


def cpu_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", CPUS_AT_START))


def cpu_baseline(A, x, seconds) -> dict:
    """The oracle's OpenMP CRS SpMV (the reference's loop order and fma
    chain, spmv_cpu_crs_impl.hpp:9-38) on the host cores, on the headline
    matrix: the bounded sample the headline's cpu_baseline reports."""
    from oracle import oracle
    y = np.empty(A.n_rows)
    t, n_iter, runtime = doubling_time(
        lambda: oracle.crs_spmv(A.row_ptr, A.col, A.val, x, out=y), seconds)
    gbs = b_alg(A.n_rows, A.n_cols, A.nnz) / t / 1e9
    return {"value": round(gbs, 2), "unit": "GB/s", "cores": cpu_threads(),
            "kind": "port",
            "sample": f"{n_iter} CRS SpMVs of the N=4M Lomax matrix "
                      f"({runtime:.1f} s, {t * 1e3:.2f} ms each, "
                      f"{2 * A.nnz / t / 1e9:.2f} GFLOP/s), OpenMP static rows, "
                      f"doubling loop after a warm-up call"}


def cpu_configs(ctx, A, x, seconds, host) -> dict:
    """The oracle's OpenMP kernels (reference loop orders) on every config,
    each timed by the doubling loop for >= `seconds`; GFLOP/s and GB/s on
    the same algorithmic bytes as the GPU lines.
      C1  CRS SpMV, 512^2 5-point Laplacian, FP64
      C2  SELL-64-4096 SpMV of the 4M-row Lomax matrix, FP64 (arrays from
          the GPU conversion, bit-identical to the reference's)
      C3  CRS SpMV of the 160^3 27-point stencil, FP64: the product of one
          CG iteration
      C4  CRS SpMM, FP32, k = 32 row-major, 1M-row Lomax matrix
      C5  BCRS 4x4 SpMM, BF16 values widened to FP32, k = 8, 96^3 block
          adjacency (the reference CPU path runs k BCRS SpMVs; the oracle's
          k-column loop is that, with the columns innermost)"""
    from oracle import oracle
    out = {"host": host, "threads": cpu_threads(), "kind": "port",
           "method": "warm-up call, then doubling iteration counts until one batch "
                     f"lasts >= {seconds} s (benchmarks_common.hpp:60-95)"}

    def record(name, fn, balg, flops, what):
        t, n, dt = doubling_time(fn, seconds)
        out[name] = {"ms": round(t * 1e3, 4), "GBs_alg": round(balg / t / 1e9, 2),
                     "GFLOPs": round(flops / t / 1e9, 2), "calls": n,
                     "batch_s": round(dt, 2), "what": what}
        log(f"CPU {name}: {out[name]}")

    P = matgen.poisson2d(512)
    px = matgen.uniform_f64(matgen.SEED_X, P.n_cols)
    py = np.empty(P.n_rows)
    record("c1_poisson512_crs_f64", lambda: oracle.crs_spmv(P.row_ptr, P.col, P.val, px, out=py),
           b_alg(P.n_rows, P.n_cols, P.nnz), 2 * P.nnz, "CRS SpMV")
    del P, px, py

    rp, col, val = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val)
    s = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, val, 8, 64, 4096)
    arr = s.arrays(np.float64)
    nch = s.info.n_chunks
    s.destroy()
    for d in (rp, col, val):
        d.free()
    sy = np.empty(nch * 64)
    record("c2_lomax4M_sell64_4096_f64",
           lambda: oracle.scs_spmv(64, nch, arr["chunk_ptr"], arr["chunk_lengths"],
                                   arr["col"], arr["val"], x, out=sy),
           b_alg(A.n_rows, A.n_cols, A.nnz), 2 * A.nnz, "SELL-C-sigma SpMV")
    del arr, sy

    S = matgen.stencil27(160)
    sx = matgen.uniform_f64(matgen.SEED_B, S.n_cols)
    sy = np.empty(S.n_rows)
    record("c3_stencil27_160_crs_f64", lambda: oracle.crs_spmv(S.row_ptr, S.col, S.val, sx, out=sy),
           b_alg(S.n_rows, S.n_cols, S.nnz), 2 * S.nnz, "CRS SpMV (one CG iteration's product)")
    del S, sx, sy

    k = 32
    M = matgen.lomax_band(1_000_000)
    v32 = M.val.astype(np.float32)
    X = matgen.uniform_f32(matgen.SEED_X, M.n_cols * k)
    Y = np.empty(M.n_rows * k, np.float32)
    record("c4_lomax1M_crs_spmm_f32_k32",
           lambda: oracle.crs_spmm(M.row_ptr, M.col, v32, k, True, X, out=Y),
           M.nnz * 8 + (M.n_rows + 1) * 4 + (M.n_rows + M.n_cols) * k * 4, 2 * M.nnz * k,
           "CRS SpMM, row-major X/Y")
    del M, v32, X, Y

    k = 8
    B = matgen.block7(96, 4)
    brp, bcol, bval = ctx.upload(B.row_ptr), ctx.upload(B.col), ctx.upload(matgen.to_bf16(B.val))
    H = ctx.bcrs_from_crs(B.n_rows, B.n_cols, B.nnz, brp, bcol, bval, 2, 4, 4, 4, 4, False)
    for d in (brp, bcol, bval):
        d.free()
    h = H.arrays(np.uint16)
    nbr, nb = H.info.n_block_rows, H.info.n_blocks
    H.destroy()
    X = matgen.uniform_f32(matgen.SEED_X, B.n_rows * k)
    Y = np.empty(B.n_rows * k, np.float32)
    record("c5_block7_96_bcrs_spmm_bf16_k8",
           lambda: oracle.bcrs_spmm(nbr, 4, 4, 4, 4, h["col"], h["row_ptr"], h["val"], False, k,
                                    X, out=Y),
           nb * (16 * 2 + 4) + (nbr + 1) * 4 + 2 * B.n_rows * k * 4, 2 * nb * 16 * k,
           "BCRS 4x4 SpMM, BF16 values widened to FP32")
    return out


def cg_section(ctx) -> dict:
    """27-point stencil on 160^3 (N = 4,096,000, nnz = 109,215,352, diagonal
    26.01): FP64 CG vs iterative refinement with BF16 and FP32 values (FP32
    diagonal side array) to ||b - A x|| / ||b|| <= 1e-10, b uniform [-1, 1).
    Each solver runs on a SELL handle of its own, so its first solve pays the
    one-time setup (solver copies of the values, column tables): reported as
    setup_seconds of the first solve, next to the best of three solves.
    Every solution's residual is recomputed on the host in FP64 (scipy's CSR
    product, independent of the GPU kernels), and the refined solutions'
    distance from the FP64 CG solution is stated."""
    import scipy.sparse as sp
    log("C3: 27-point stencil 160^3")
    A = matgen.stencil27(160)
    b = matgen.uniform_f64(matgen.SEED_B, A.n_rows)
    rp, col, val = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val)
    db = ctx.upload(b)
    out = {"n_rows": A.n_rows, "nnz": A.nnz, "format": "SELL-64-1"}
    Ah = sp.csr_matrix((A.val, A.col, A.row_ptr), shape=(A.n_rows, A.n_cols))

    def solve(mixed, lo=capi.BF16, inner=0.1, reps=3):
        S = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, val, 8, 64, 1)
        out["fill"] = round(S.info.n_elements / A.nnz, 4)
        best = first = None
        for _ in range(reps):
            dx = ctx.zeros(A.n_rows, np.float64)
            prm = capi.CgParams(rel_tol=1e-10, max_iters=20000, check_every=25,
                                inner_rel_tol=inner, max_inner=5000,
                                inner_value_type=lo, keep_diag_f32=1)
            rep = ctx.cg(S, db, dx, prm, mixed)
            first = first or (rep.seconds, rep.setup_seconds)
            if best is None or rep.seconds < best[0].seconds:
                best = (rep, ctx.download(dx))
            dx.free()
        S.destroy()
        rep, x = best
        r = b - Ah @ x
        return {"status": rep.status, "seconds": round(rep.seconds, 5),
                "iterations": rep.iterations, "outer": rep.outer_iterations,
                "rel_residual_gpu": rep.rel_residual,
                "rel_residual_host_fp64": float(np.linalg.norm(r) / np.linalg.norm(b)),
                "first_solve_seconds": round(first[0], 5),
                "first_solve_setup_seconds": round(first[1], 5),
                "setup_seconds_cached": round(rep.setup_seconds, 5),
                "column_coding": {capi.COLUMNS_8: "8-bit table indices",
                                  capi.COLUMNS_16: "16-bit offsets",
                                  capi.COLUMNS_32: "32-bit columns"}.get(rep.column_coding)}, x

    fp64, x64 = solve(False)
    bf16, xbf = solve(True, capi.BF16)
    f32, x32 = solve(True, capi.F32)
    n64 = np.linalg.norm(x64)
    for rec, x in ((bf16, xbf), (f32, x32)):
        rec["max_abs_diff_vs_fp64_cg"] = float(np.max(np.abs(x - x64)))
        rec["rel_l2_diff_vs_fp64_cg"] = float(np.linalg.norm(x - x64) / n64)
    with_setup = lambda r: r["first_solve_seconds"] + r["first_solve_setup_seconds"]
    out.update({"fp64_cg": fp64, "ir_bf16": bf16, "ir_f32": f32,
                "speedup_bf16_vs_fp64": round(fp64["seconds"] / bf16["seconds"], 3),
                "speedup_f32_vs_fp64": round(fp64["seconds"] / f32["seconds"], 3),
                "speedup_bf16_vs_fp64_first_solve_incl_setup":
                    round(with_setup(fp64) / with_setup(bf16), 3),
                "speedup_f32_vs_fp64_first_solve_incl_setup":
                    round(with_setup(fp64) / with_setup(f32), 3)})
    log(f"C3 {out}")
    for d in (rp, col, val, db):
        d.free()
    return out


def mtx_section(ctx, path, steps, warmup) -> dict:
    """FP64 SpMV on a Matrix Market file, read as the reference's benchmarks
    read it (matgen.read_mtx); x uniform [-1, 1) from the benchmark seed."""
    M = matgen.read_mtx(path)
    log(f"{path}: {M.n_rows} x {M.n_cols}, nnz {M.nnz}")
    x = matgen.uniform_f64(matgen.SEED_X, M.n_cols)
    ba = b_alg(M.n_rows, M.n_cols, M.nnz)
    rp, col, val, dx = ctx.upload(M.row_ptr), ctx.upload(M.col), ctx.upload(M.val), ctx.upload(x)
    y = ctx.zeros(M.n_rows, np.float64)
    out = {"file": str(path), "n_rows": M.n_rows, "nnz": M.nnz, "B_alg": ba}
    plan = ctx.crs_plan(M.n_rows, M.nnz, rp)
    ms = time_launches(ctx, lambda: ctx.crs_spmv(capi.F64, M.n_rows, M.n_cols, M.nnz,
                                                 rp, col, val, dx, y, plan), steps, warmup)
    out["crs_row_blocks"] = {"ms": round(ms, 4), "GBs_alg": round(ba / ms / 1e6, 1)}
    ctx.crs_plan_destroy(plan)
    for sigma in (1, 4096):
        s = ctx.scs_from_crs(M.n_rows, M.n_cols, M.nnz, rp, col, val, 8, 64, sigma)
        yy = ctx.zeros(s.info.n_rows_padded, np.float64)
        ms = time_launches(ctx, lambda: s.apply(capi.F64, dx, yy), steps, warmup)
        out[f"sell64_s{sigma}"] = {"ms": round(ms, 4), "GBs_alg": round(ba / ms / 1e6, 1),
                                   "fill": round(s.info.n_elements / max(M.nnz, 1), 4)}
        yy.free()
        s.destroy()
    for d in (rp, col, val, dx, y):
        d.free()
    log(f"mtx {out}")
    return out


def vendor_lib():
    """The rocSPARSE comparison library (smaxkernels_amd/vendor/
    rocsparse_cmp.cpp, benchmark only), or an error string."""
    import ctypes as C
    path = ROOT / "smaxkernels_amd" / "_build" / "librocsparse_cmp.so"
    try:
        lib = C.CDLL(str(path))
    except OSError as e:
        return f"rocSPARSE comparison not loaded: {e}"
    lib.rocsparse_cmp_create.restype = C.c_void_p
    lib.rocsparse_cmp_create.argtypes = [C.c_int] + [C.c_int64] * 3 + [C.c_void_p] * 6
    lib.rocsparse_cmp_spmv.argtypes = [C.c_void_p]
    lib.rocsparse_cmp_destroy.argtypes = [C.c_void_p]
    lib.rocsparse_cmp_spmm_create.restype = C.c_void_p
    lib.rocsparse_cmp_spmm_create.argtypes = ([C.c_int] + [C.c_int64] * 4 + [C.c_void_p] * 3
                                              + [C.c_int64] + [C.c_void_p] * 3)
    lib.rocsparse_cmp_spmm.argtypes = [C.c_void_p]
    lib.rocsparse_cmp_spmm_destroy.argtypes = [C.c_void_p]
    return lib


def vendor_spmm(ctx, lib, algs, block_dim, n_rows, n_cols, nnz, rp, col, val, k, X, Y,
                ours, steps, warmup, balg, flops):
    """rocSPARSE FP32 SpMM (CSR or BSR) on the same arrays; per algorithm the
    time and the largest difference from our result `ours`."""
    stream = ctx.lib.smax_hip_get_stream(ctx.handle)
    out = {}
    for alg, name in algs:
        h = lib.rocsparse_cmp_spmm_create(alg, block_dim, n_rows, n_cols, nnz, capi._ptr(rp),
                                          capi._ptr(col), capi._ptr(val), k, capi._ptr(X),
                                          capi._ptr(Y), stream)
        if not h:
            out[name] = None
            continue
        ms = time_launches(ctx, lambda: lib.rocsparse_cmp_spmm(h), steps, warmup)
        got = ctx.download(Y)
        out[name] = {"ms": round(ms, 4), "GBs_alg": round(balg / ms / 1e6, 1),
                     "GFLOPs": round(flops / ms / 1e6, 1),
                     "max_rel_diff_vs_ours":
                         float(np.max(np.abs(got - ours)) / np.max(np.abs(ours)))}
        lib.rocsparse_cmp_spmm_destroy(h)
    return out


def vendor_section(ctx, A, rp, col, val, dx, steps, warmup) -> dict:
    """rocSPARSE CSR SpMV (FP64) on the same device arrays: a library
    baseline for comparison only (smaxkernels_amd/vendor/rocsparse_cmp.cpp)."""
    lib = vendor_lib()
    if isinstance(lib, str):
        return {"error": lib}
    stream = ctx.lib.smax_hip_get_stream(ctx.handle)
    y = ctx.zeros(A.n_rows, np.float64)
    ctx.crs_spmv(capi.F64, A.n_rows, A.n_cols, A.nnz, rp, col, val, dx, y, None)
    ours = ctx.download(y)
    balg = b_alg(A.n_rows, A.n_cols, A.nnz)
    out = {}
    for alg, name in ((1, "csr_adaptive"), (2, "csr_rowsplit"), (3, "csr_lrb"),
                      (4, "csr_nnzsplit")):
        h = lib.rocsparse_cmp_create(alg, A.n_rows, A.n_cols, A.nnz, rp.ptr, col.ptr,
                                     val.ptr, dx.ptr, y.ptr, stream)
        if not h:
            out[name] = None
            continue
        ms = time_launches(ctx, lambda: lib.rocsparse_cmp_spmv(h), steps, warmup)
        got = ctx.download(y)
        out[name] = {"ms": round(ms, 4), "GBs_alg": round(balg / ms / 1e6, 1),
                     "max_rel_diff_vs_reference_chain":
                         float(np.max(np.abs(got - ours)) / np.max(np.abs(ours)))}
        lib.rocsparse_cmp_destroy(h)
    y.free()
    log(f"rocSPARSE {out}")
    return out


def lowp_section(ctx, A, rp, col, val, steps, warmup) -> dict:
    """The 4M-row Lomax matrix with FP32 and BF16 values (FP32 x, y and
    accumulation), SELL-64 at sigma = 4096 and 65536 through the handle path,
    each with its error against the FP64 product."""
    out = {}
    lens = np.diff(A.row_ptr)
    dabs = ctx.upload(np.abs(A.val))
    for name, vt, vals, vb in (("f32", capi.F32, A.val.astype(np.float32), 4),
                               ("bf16", capi.BF16, matgen.to_bf16(A.val), 2)):
        ba = A.nnz * (vb + 4) + (A.n_rows + 1) * 4 + A.n_cols * 4 + A.n_rows * 4
        dv = ctx.upload(vals)
        x32 = matgen.uniform_f32(matgen.SEED_X, A.n_cols)
        dx = ctx.upload(x32)
        # FP64 reference: the FP64 matrix times x widened exactly, and |A||x|
        x64, xa = ctx.upload(x32.astype(np.float64)), ctx.upload(np.abs(x32.astype(np.float64)))
        y64, ya = ctx.zeros(A.n_rows, np.float64), ctx.zeros(A.n_rows, np.float64)
        ctx.crs_spmv(capi.F64, A.n_rows, A.n_cols, A.nnz, rp, col, val, x64, y64, None)
        ctx.crs_spmv(capi.F64, A.n_rows, A.n_cols, A.nnz, rp, col, dabs, xa, ya, None)
        ref, scale = ctx.download(y64), ctx.download(ya)
        for d in (x64, xa, y64, ya):
            d.free()
        for sigma in (4096, 65536):
            s = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, dv, vb, 64, sigma)
            yy = ctx.zeros(s.info.n_rows_padded, np.float32)
            ms = time_launches(ctx, lambda: s.apply(vt, dx, yy), steps, warmup)
            yo = ctx.zeros(A.n_rows, np.float32)
            s.apply(vt, dx, yo, unpermuted=True)
            out[f"{name}_C64_s{sigma}"] = {
                "ms": round(ms, 4), "B_alg": ba, "GBs_alg": round(ba / ms / 1e6, 1),
                **err_vs_fp64(ctx.download(yo), ref, scale, int(lens.max()), U_VALUE[name])}
            yo.free()
            yy.free()
            s.destroy()
        dv.free()
        dx.free()
    dabs.free()
    log(f"SELL low precision {out}")
    return out


def c4_section(ctx, steps, warmup) -> dict:
    """FP32 SpMM with k = 32 row-major columns on the 1M-row Lomax band matrix
    (nnz ~ 16M): forward Y = A X (CRS and SELL-64-N), adjoint A^T W through
    the SELL layout of A^T, and the SDDMM value gradient."""
    k = 32
    A = matgen.lomax_band(1_000_000)
    log(f"C4: lomax_band N={A.n_rows} nnz={A.nnz} FP32 SpMM k={k}")
    v32 = A.val.astype(np.float32)
    X = matgen.uniform_f32(matgen.SEED_X, A.n_cols * k)
    W = matgen.uniform_f32(matgen.SEED_W, A.n_rows * k)
    rp, col, val = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(v32)
    dX, dW = ctx.upload(X), ctx.upload(W)
    dY, dG = ctx.zeros(A.n_rows * k, np.float32), ctx.zeros(A.n_cols * k, np.float32)
    ddv = ctx.zeros(A.nnz, np.float32)
    # sigma = 65536: chunks carry little padding (fill ~1.07) while the rows
    # of a chunk still come from one 65536-row window, whose X band the
    # XCD-local order keeps together (0.40 ms vs 0.44 ms at sigma = N)
    S = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, val, 4, 64, 65536)
    T = ctx.scs_transpose_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, val, 4, 64,
                                   A.n_cols)
    ba = A.nnz * 8 + (A.n_rows + 1) * 4 + (A.n_rows + A.n_cols) * k * 4
    plan = ctx.crs_plan(A.n_rows, A.nnz, rp)
    runs = {
        "forward_crs": lambda: ctx.crs_spmm(capi.F32, A.n_rows, A.n_cols, A.nnz, rp, col,
                                            val, k, True, dX, k, dY, k),
        "forward_crs_plan": lambda: ctx.crs_spmm(capi.F32, A.n_rows, A.n_cols, A.nnz, rp,
                                                 col, val, k, True, dX, k, dY, k, plan),
        # the reference's default layout: column-major X and Y (ld = n_rows)
        "forward_crs_colmajor_plan": lambda: ctx.crs_spmm(
            capi.F32, A.n_rows, A.n_cols, A.nnz, rp, col, val, k, False, dX, A.n_rows, dY,
            A.n_rows, plan),
        "forward_sell": lambda: S.apply_spmm(capi.F32, k, dX, dY),
        "adjoint_sell_AT": lambda: T.apply_spmm(capi.F32, k, dW, dG),
        "sddmm_sell": lambda: S.sddmm(capi.F32, k, dW, dX, ddv),
    }
    out = {"n_rows": A.n_rows, "nnz": A.nnz, "k": k, "B_alg": ba}
    for name, fn in runs.items():
        ms = time_launches(ctx, fn, steps, warmup)
        out[name] = {"ms": round(ms, 4), "GBs_alg": round(ba / ms / 1e6, 1),
                     "GFLOPs": round(2 * A.nnz * k / ms / 1e6, 1)}
    from tests.helpers import fd_gradient_check
    out["fd_check"] = fd_gradient_check(A, v32, X, W, k, ctx.download(ddv).astype(np.float64),
                                        ctx.download(dG).astype(np.float64).reshape(A.n_cols, k))
    out["error_vs_fp64"] = c4_errors(ctx, A, rp, col, X, W, k, runs, dY, dG, ddv)
    lib = vendor_lib()
    if not isinstance(lib, str):
        runs["forward_crs"]()
        ours = ctx.download(dY)
        out["rocsparse_csr_spmm"] = vendor_spmm(
            ctx, lib, ((1, "csr"), (2, "csr_row_split"), (3, "csr_nnz_split"),
                       (4, "csr_merge_path")),
            0, A.n_rows, A.n_cols, A.nnz, rp, col, val, k, dX, dY, ours, steps, warmup, ba,
            2 * A.nnz * k)
    log(f"C4 {out}")
    ctx.crs_plan_destroy(plan)
    for d in (rp, col, val, dX, dW, dY, dG, ddv):
        d.free()
    S.destroy()
    T.destroy()
    return out


def c4_errors(ctx, A, rp, col, X, W, k, runs, dY, dG, ddv) -> dict:
    """FP32 results of C4 against FP64: Y = A X and A^T W against the FP64
    matrix's products (FP64 kernels, operands widened exactly), the SDDMM
    against FP64 sums on 1M sampled entries (the SDDMM rounds no stored
    value; its bound is its reduction tree's, 7u for k = 32)."""
    out = {}
    lens = np.diff(A.row_ptr)
    X64, W64 = X.astype(np.float64), W.astype(np.float64)
    v64, va = ctx.upload(A.val), ctx.upload(np.abs(A.val))
    dX64, dXa = ctx.upload(X64), ctx.upload(np.abs(X64))
    Y64, Ya = ctx.zeros(A.n_rows * k, np.float64), ctx.zeros(A.n_rows * k, np.float64)
    # the planned and column-major CRS products are bit-identical to these
    # two (tests/test_gpu_fullsize.py)
    for name in ("forward_crs", "forward_sell"):
        runs[name]()
        y_lo = ctx.download(dY)
        if name == "forward_crs":
            ctx.crs_spmm(capi.F64, A.n_rows, A.n_cols, A.nnz, rp, col, v64, k, True, dX64, k,
                         Y64, k)
            ctx.crs_spmm(capi.F64, A.n_rows, A.n_cols, A.nnz, rp, col, va, k, True, dXa, k,
                         Ya, k)
            ref, scale = ctx.download(Y64), ctx.download(Ya)
        out[name] = err_vs_fp64(y_lo, ref, scale, int(lens.max()), U32)
    for d in (dX64, dXa, Y64, Ya):
        d.free()
    T64 = ctx.scs_transpose_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v64, 8, 64, A.n_cols)
    dW64, dWa = ctx.upload(W64), ctx.upload(np.abs(W64))
    G64 = ctx.zeros(A.n_cols * k, np.float64)
    runs["adjoint_sell_AT"]()
    g_lo = ctx.download(dG)
    T64.apply_spmm(capi.F64, k, dW64, G64)
    ref = ctx.download(G64)
    T64.refresh_values(va)
    T64.apply_spmm(capi.F64, k, dWa, G64)
    scale = ctx.download(G64)
    col_len = np.bincount(A.col, minlength=A.n_cols)
    out["adjoint_sell_AT"] = err_vs_fp64(g_lo, ref, scale, int(col_len.max()), U32)
    for d in (v64, va, dW64, dWa, G64):
        d.free()
    T64.destroy()
    runs["sddmm_sell"]()
    dv = ctx.download(ddv).astype(np.float64)
    ent = np.random.default_rng(7).choice(A.nnz, 1_000_000, replace=False)
    rows = np.repeat(np.arange(A.n_rows), lens)[ent]
    Wr, Xr = W64.reshape(-1, k)[rows], X64.reshape(-1, k)[A.col[ent]]
    exact = np.einsum("ij,ij->i", Wr, Xr)
    scale = np.einsum("ij,ij->i", np.abs(Wr), np.abs(Xr))
    e = err_vs_fp64(dv[ent], exact, scale, 0, 0.0)
    e["bound"] = 7 * U32 * (1 + 4 * U32) + k * 2.0 ** -53
    e["sample"] = "1M random stored entries"
    out["sddmm_sell"] = e
    return out


def c5_section(ctx, steps, warmup) -> dict:
    """BCRS 4x4 SpMM with BF16 values and FP32 X/Y (row-major), 7-point block
    adjacency on 96^3 blocks: forward and adjoint (A^T through the
    byte-preserving BCRS transpose) for k in {8, 64}."""
    A = matgen.block7(96, 4)
    log(f"C5: block7 96^3 4x4 N={A.n_rows} nnz={A.nnz} BF16 BCRS SpMM")
    rp, col, val = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(matgen.to_bf16(A.val))
    B = ctx.bcrs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, val, 2, 4, 4, 4, 4, False)
    for d in (rp, col, val):
        d.free()
    T = B.transpose()
    nb = B.info.n_blocks
    out = {"n_rows": A.n_rows, "n_blocks": nb, "values": nb * 16}
    # the BCRS SpMV (one right-hand side) on the same matrix
    ba1 = nb * (16 * 2 + 4) + (B.info.n_block_rows + 1) * 4 + 2 * A.n_rows * 4
    x1 = ctx.upload(matgen.uniform_f32(matgen.SEED_X, A.n_rows))
    y1 = ctx.zeros(A.n_rows, np.float32)
    ms = time_launches(ctx, lambda: B.spmv(capi.BF16, x1, y1), steps, warmup)
    out["spmv"] = {"ms": round(ms, 4), "GBs_alg": round(ba1 / ms / 1e6, 1),
                   "GFLOPs": round(2 * nb * 16 / ms / 1e6, 1), "B_alg": ba1}
    x1.free()
    y1.free()
    # the FP64 matrix (and |A|) in the same BCRS layout, for the errors vs FP64
    rp, col = ctx.upload(A.row_ptr), ctx.upload(A.col)
    v64, va = ctx.upload(A.val), ctx.upload(np.abs(A.val))
    B64 = ctx.bcrs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v64, 8, 4, 4, 4, 4, False)
    Ba = ctx.bcrs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, va, 8, 4, 4, 4, 4, False)
    for d in (rp, col, v64, va):
        d.free()
    T64, Ta = B64.transpose(), Ba.transpose()
    n_chain = 4 * int(np.diff(B.arrays(np.uint16)["row_ptr"]).max())
    for k in (8, 64):
        ba = nb * (16 * 2 + 4) + (B.info.n_block_rows + 1) * 4 + 2 * A.n_rows * k * 4
        Xh = matgen.uniform_f32(matgen.SEED_X, A.n_rows * k)
        X = ctx.upload(Xh)
        Y = ctx.zeros(A.n_rows * k, np.float32)
        X64, Xa = ctx.upload(Xh.astype(np.float64)), ctx.upload(np.abs(Xh.astype(np.float64)))
        Y64 = ctx.zeros(A.n_rows * k, np.float64)
        for name, M, M64, Mabs in (("forward", B, B64, Ba), ("adjoint_AT", T, T64, Ta)):
            ms = time_launches(ctx, lambda: M.spmm(capi.BF16, k, X, Y), steps, warmup)
            M64.spmm(capi.F64, k, X64, Y64)
            ref = ctx.download(Y64)
            Mabs.spmm(capi.F64, k, Xa, Y64)
            out[f"{name}_k{k}"] = {"ms": round(ms, 4), "GBs_alg": round(ba / ms / 1e6, 1),
                                   "GFLOPs": round(2 * nb * 16 * k / ms / 1e6, 1),
                                   "B_alg": ba,
                                   **err_vs_fp64(ctx.download(Y), ref, ctx.download(Y64),
                                                 n_chain, U_VALUE["bf16"])}
        for d in (X, Y, X64, Xa, Y64):
            d.free()
    for M in (B, T, B64, Ba, T64, Ta):
        M.destroy()
    # rocSPARSE has no BF16-value x FP32-vector SpMM: compare both in FP32
    lib = vendor_lib()
    if not isinstance(lib, str):
        rp, col = ctx.upload(A.row_ptr), ctx.upload(A.col)
        v32 = ctx.upload(A.val.astype(np.float32))
        B = ctx.bcrs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v32, 4, 4, 4, 4, 4, False)
        for d in (rp, col, v32):
            d.free()
        i = B.info
        for k in (8, 64):
            ba = nb * (16 * 4 + 4) + (i.n_block_rows + 1) * 4 + 2 * A.n_rows * k * 4
            X = ctx.upload(matgen.uniform_f32(matgen.SEED_X, A.n_rows * k))
            Y = ctx.zeros(A.n_rows * k, np.float32)
            ms = time_launches(ctx, lambda: B.spmm(capi.F32, k, X, Y), steps, warmup)
            ours = ctx.download(Y)
            out[f"f32_forward_k{k}"] = {"ms": round(ms, 4), "GBs_alg": round(ba / ms / 1e6, 1),
                                        "GFLOPs": round(2 * nb * 16 * k / ms / 1e6, 1)}
            out[f"rocsparse_bsr_f32_k{k}"] = vendor_spmm(
                ctx, lib, ((5, "bsr"),), 4, i.n_block_rows, i.n_block_cols, i.n_blocks,
                i.row_ptr, i.col, i.val, k, X, Y, ours, steps, warmup, ba, 2 * nb * 16 * k)
            X.free()
            Y.free()
        B.destroy()
    log(f"C5 {out}")
    return out


def summarize_trace(directory) -> dict:
    """Per-kernel durations (us) from rocprofv3 --kernel-trace CSV files
    under `directory`.  A launch that directly follows a copy_kernel on the
    GPU timeline is 'cold' (the benchmark's C1 cold launches each follow a
    1 GiB copy that evicts L2 and MALL); the others are 'warm'."""
    import csv
    rows = []
    for f in sorted(Path(directory).rglob("*kernel_trace.csv")):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                             r["Kernel_Name"]))
    rows.sort()
    stats: dict = {}
    prev = ""
    for t0, t1, name in rows:
        short = name.split("(")[0].replace("void ", "").strip()
        kind = "cold" if prev.startswith("smaxhip::copy_kernel") else "warm"
        stats.setdefault(short, {"cold": [], "warm": []})[kind].append((t1 - t0) / 1e3)
        prev = short
    out = {}
    for name, d in stats.items():
        out[name] = {k: {"count": len(v), "median_us": round(float(np.median(v)), 3),
                         "min_us": round(float(np.min(v)), 3),
                         "mean_us": round(float(np.mean(v)), 3)}
                     for k, v in d.items() if v}
    return out


def c1_cold_kernel(kernel_prefix: str, b_alg_bytes: int, peak_gbs) -> dict | None:
    """Kernel-only duration of C1's cold launches, from the committed
    rocprofv3 kernel trace summary (profiles/c1_poisson512_cold_warm_kernels.json,
    written by --summarize-trace), as us and as a fraction of the stream rate."""
    f = ROOT / "profiles" / "c1_poisson512_cold_warm_kernels.json"
    if not f.exists():
        return None
    summ = json.loads(f.read_text())["kernels"]
    for name, rec in summ.items():
        if name.split("<")[0].endswith(kernel_prefix) and "cold" in rec:
            us = rec["cold"]["median_us"]
            gbs = b_alg_bytes / us / 1e3
            return {"cold_kernel_us": us, "cold_kernel_GBs": round(gbs, 1),
                    "cold_kernel_frac_of_measured_stream":
                        round(gbs / peak_gbs, 3) if peak_gbs else None,
                    "warm_kernel_us": rec.get("warm", {}).get("median_us"),
                    "source": f"rocprofv3 kernel trace, {f.name}"}
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--n", type=int, default=4_000_000)
    ap.add_argument("--sigmas", default="1,64,4096,65536,N",
                    help="comma list of SELL sigma values (N = n_rows); 1 and 64 are "
                         "17x / 9-11x padded (13 GB each)")
    ap.add_argument("--cs", default="64,32", help="comma list of SELL chunk sizes")
    ap.add_argument("--mtx", default="",
                    help="also time FP64 SpMV on this Matrix Market file (CRS row "
                         "blocks and SELL-64 at sigma 1 and 4096), reported under extra")
    ap.add_argument("--cpu-config-seconds", type=float, default=2.0,
                    help="doubling-loop batch length of each per-config CPU baseline")
    ap.add_argument("--summarize-trace", default="",
                    help="print per-kernel durations from a rocprofv3 --kernel-trace "
                         "CSV directory (cold launches: those right after a copy_kernel) "
                         "and exit; runs no GPU work")
    ap.add_argument("--only", default="",
                    help="comma list of sections to run: "
                         "stream,sell,gather,vendor,lowp,crs,c1,cg,c4,c5,cpu")
    args = ap.parse_args()
    if args.summarize_trace:
        print(json.dumps(summarize_trace(args.summarize_trace), indent=1))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU only; there is no multi-GPU path (rows "
              "would shard as contiguous row blocks with one x halo exchange per "
              "product).", file=sys.stderr)
        return 2
    only = set(filter(None, args.only.split(",")))
    want = (lambda s: not only or s in only)
    host = host_description()
    extra: dict = {}
    ctx = capi.Context(0)

    peak = None
    if want("stream"):
        log("stream peak (4 GiB arrays)")
        peak = stream_peak(ctx)
        extra["stream"] = peak
    peak_gbs = peak["copy_GBs"] if peak else None

    log(f"generating Lomax band matrix N={args.n}")
    A = matgen.lomax_band(args.n)
    x = matgen.uniform_f64(matgen.SEED_X, A.n_cols)
    balg = b_alg(A.n_rows, A.n_cols, A.nnz)
    rp, col, val = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val)
    dx = ctx.upload(x)
    dy = ctx.zeros(A.n_rows + 64, np.float64)

    best = None
    sell = {}
    if want("sell"):
        sigmas = [A.n_rows if t == "N" else int(t) for t in args.sigmas.split(",")]
        for C_ in (int(t) for t in args.cs.split(",")):
            for sigma in sigmas:
                t0 = time.perf_counter()
                s = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, val, 8,
                                     C_, sigma)
                conv_s = time.perf_counter() - t0
                yy = ctx.zeros(s.info.n_rows_padded, np.float64)
                ms = time_launches(ctx, lambda: s.apply(capi.F64, dx, yy),
                                   args.steps, args.warmup)
                gbs = balg / ms / 1e6
                fmt = (s.info.n_elements * 12 + (s.info.n_chunks + 1) * 4 +
                       s.info.n_chunks * 4 + A.n_cols * 8 +
                       s.info.n_rows_padded * 8)
                key = f"C{C_}_s{'N' if sigma == A.n_rows else sigma}"
                sell[key] = {"ms": round(ms, 4), "GBs_alg": round(gbs, 1),
                             "GFLOPs": round(2 * A.nnz / ms / 1e6, 1),
                             "GBs_format": round(fmt / ms / 1e6, 1),
                             "fill": round(s.info.n_elements / A.nnz, 4),
                             "conversion_s": round(conv_s, 3),
                             "tail_waves": dict(zip(("slots", "waves", "entries"),
                                                    s.tails()))}
                log(f"SELL {key}: {ms:.4f} ms {gbs:.1f} GB/s")
                sell[key]["sigma_listed"] = sigma in (1, 64, 4096, A.n_rows)
                if sell[key]["sigma_listed"] and (best is None or ms < best[1]):
                    best = (f"SELL-{C_}-{'N' if sigma == A.n_rows else sigma}", ms,
                            "scs_spmv_w64" if C_ == 64 else "scs_spmv_generic")
                yy.free()
                s.destroy()
        extra["sell_f64"] = sell

    gather = None
    if want("gather"):
        log("random-gather ceiling over the band window")
        gather = gather_ceiling(ctx, A, 50, 5)
        extra["gather_ceiling"] = gather

    if want("vendor"):
        extra["rocsparse_csr_f64"] = vendor_section(ctx, A, rp, col, val, dx, args.steps,
                                                    args.warmup)

    if want("lowp"):
        extra["sell_lowp"] = lowp_section(ctx, A, rp, col, val, args.steps, args.warmup)

    crs = {}
    if want("crs"):
        ms = time_launches(ctx, lambda: ctx.crs_spmv(
            capi.F64, A.n_rows, A.n_cols, A.nnz, rp, col, val, dx, dy, None),
            args.steps, args.warmup)
        crs["lane_per_row"] = {"ms": round(ms, 4), "GBs_alg": round(balg / ms / 1e6, 1)}
        plan = ctx.crs_plan(A.n_rows, A.nnz, rp)
        ms = time_launches(ctx, lambda: ctx.crs_spmv(
            capi.F64, A.n_rows, A.n_cols, A.nnz, rp, col, val, dx, dy, plan),
            args.steps, args.warmup)
        crs["row_blocks"] = {"ms": round(ms, 4), "GBs_alg": round(balg / ms / 1e6, 1)}
        if best is None or ms < best[1]:
            best = ("CRS row blocks", ms, "crs_spmv_blocked")
        # host buffers handed over: copy A and x in, one SpMV, y back (pageable
        # numpy memory, synchronous copies), as a caller without device-resident
        # operands pays it
        t0 = time.perf_counter()
        hrp, hcol, hval, hx = (ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val),
                               ctx.upload(x))
        hy = ctx.zeros(A.n_rows, np.float64)
        ctx.crs_spmv(capi.F64, A.n_rows, A.n_cols, A.nnz, hrp, hcol, hval, hx, hy, plan)
        ctx.download(hy)
        pcie_ms = (time.perf_counter() - t0) * 1e3
        crs["pcie_inclusive"] = {"ms": round(pcie_ms, 2), "GBs_alg": round(balg / pcie_ms / 1e6, 1),
                                 "what": "upload row_ptr, col, val, x; CRS row-block SpMV; "
                                         "download y (pageable host memory)"}
        for d in (hrp, hcol, hval, hx, hy):
            d.free()
        ctx.crs_plan_destroy(plan)
        log(f"CRS {crs}")
        extra["crs_f64"] = crs

    if want("c1"):
        P = matgen.poisson2d(512)
        px = matgen.uniform_f64(matgen.SEED_X, P.n_cols)
        pb = b_alg(P.n_rows, P.n_cols, P.nnz)
        prp, pcol, pval, pdx = (ctx.upload(P.row_ptr), ctx.upload(P.col),
                                ctx.upload(P.val), ctx.upload(px))
        ps = ctx.scs_from_crs(P.n_rows, P.n_cols, P.nnz, prp, pcol, pval, 8, 64, 1)
        pplan = ctx.crs_plan(P.n_rows, P.nnz, prp)
        pdy = ctx.zeros(ps.info.n_rows_padded, np.float64)
        scrub_n = 1 << 27  # 1 GiB read + 1 GiB written between launches
        s1, s2 = ctx.zeros(scrub_n, np.float64), ctx.zeros(scrub_n, np.float64)

        def c1_times(fn) -> dict:
            warm_py = time_launches(ctx, fn, 2000, 100)
            warm = graph_ms(ctx, fn, check=lambda: ctx.download(pdy))
            cold = []
            for _ in range(20):
                # 1 GiB of plain stores (they allocate in L2 and the MALL, so no
                # operand survives there), then a 1 GiB non-temporal copy
                ctx._check(ctx.lib.smax_hip_zero(ctx.handle, s1.ptr, s1.nbytes))
                ctx.stream_copy(s1, s2, scrub_n)
                ctx.timer_start()
                fn()
                cold.append(ctx.timer_stop())
            cold_ms = float(np.median(cold))
            return {
                "B_alg": pb, "warm_ms": round(warm, 5),
                "warm_GBs": round(pb / warm / 1e6, 1),
                "warm_method": "HIP graph of 200 back-to-back SpMVs, 10 replays",
                "warm_frac_of_measured_stream": (round(pb / warm / 1e6 / peak_gbs, 3)
                                                 if peak_gbs else None),
                "warm_host_launch_ms": round(warm_py, 5),
                "cold_ms": round(cold_ms, 5), "cold_GBs": round(pb / cold_ms / 1e6, 1),
                "cold_method": "1 GiB of plain stores and a 1 GiB copy between launches, "
                               "median of 20 event-timed"}

        extra["c1_poisson512_f64_sell64_1"] = c1_times(lambda: ps.apply(capi.F64, pdx, pdy))
        extra["c1_poisson512_f64_sell64_1"]["kernel_trace"] = c1_cold_kernel(
            "scs_spmv_w64", pb, peak_gbs)
        # config 1 is the reference's plain CRS SpMV: both CRS schedules too
        extra["c1_poisson512_f64_crs"] = {
            "lane_per_row": c1_times(lambda: ctx.crs_spmv(capi.F64, P.n_rows, P.n_cols, P.nnz,
                                                          prp, pcol, pval, pdx, pdy, None)),
            "row_blocks": c1_times(lambda: ctx.crs_spmv(capi.F64, P.n_rows, P.n_cols, P.nnz,
                                                        prp, pcol, pval, pdx, pdy, pplan))}
        for key, kern in (("lane_per_row", "crs_spmv_lane"), ("row_blocks", "crs_spmv_blocked")):
            extra["c1_poisson512_f64_crs"][key]["kernel_trace"] = c1_cold_kernel(kern, pb, peak_gbs)
        # the dispatch floor: an 8-byte fill kernel timed the same two ways
        # (both C1 timings are within a few microseconds of it)
        floor = c1_times(lambda: ctx._check(ctx.lib.smax_hip_zero(ctx.handle, pdy.ptr, 8)))
        extra["c1_dispatch_floor"] = {"warm_ms": floor["warm_ms"], "cold_ms": floor["cold_ms"],
                                      "kernel": "zero_kernel on 8 bytes"}
        ctx.crs_plan_destroy(pplan)
        for d in (s1, s2, pdy, prp, pcol, pval, pdx):
            d.free()
        ps.destroy()
        log(f"C1 {extra['c1_poisson512_f64_sell64_1']} {extra['c1_poisson512_f64_crs']}")

    if args.mtx:
        extra["mtx"] = mtx_section(ctx, args.mtx, args.steps, args.warmup)

    if want("cg"):
        extra["c3_cg"] = cg_section(ctx)
    if want("c4"):
        extra["c4_spmm_adjoint_f32"] = c4_section(ctx, 20, 3)
    if want("c5"):
        extra["c5_bcrs_spmm_bf16"] = c5_section(ctx, 20, 3)

    base = None
    if want("cpu"):
        log(f"CPU baseline ({args.cpu_seconds} s sample)")
        base = cpu_baseline(A, x, args.cpu_seconds)
        extra["cpu"] = cpu_configs(ctx, A, x, args.cpu_config_seconds, host)
        base["host"] = host

    if best is None:  # a partial run (--only without sell/crs): extras only
        print(json.dumps({"metric": METRIC, "value": None, "unit": "GB/s",
                          "n_gpus": 1, "extra": extra, "cpu_baseline": base}),
              flush=True)
        ctx.close()
        return 0
    fmt, ms, kernel = best
    value = balg / ms / 1e6
    roof = {"bound": "hbm", "achieved": round(value, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(value / HBM_PEAK_GBS, 4),
            "traffic": hbm_traffic(kernel, A.n_rows, fmt), "kernel": kernel}
    if peak_gbs:
        roof["frac_of_measured_stream"] = round(value / peak_gbs, 4)
        roof["measured_stream_GBs"] = peak_gbs
    if gather:
        rate = A.nnz / ms / 1e6
        roof["gather"] = {"spmv_G_gathers_per_s": round(rate, 1),
                          "ceiling_G_gathers_per_s": gather["G_gathers_per_s"],
                          "frac": round(rate / gather["G_gathers_per_s"], 4),
                          "window_MB": gather["window_MB"]}
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "GB/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"lomax_band N={A.n_rows} nnz={A.nnz} FP64 SpMV, "
                               f"best of CRS and SELL-C-sigma (C in {{32,64}}, "
                               f"sigma in {{1,64,4096,N}}): {fmt}",
                   "format": fmt, "n_rows": A.n_rows, "nnz": A.nnz,
                   "B_alg_bytes": balg},
        "roofline": roof,
        "cpu_baseline": base,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
