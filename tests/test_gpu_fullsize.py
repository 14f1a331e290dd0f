"""The benchmark workloads at their full sizes, checked through properties
that do not need a CPU run of the whole workload where that would be slow:
cross-format bit equality (every SpMV is the reference's fma chain, so CRS,
SELL and BCRS must agree bit for bit), the oracle on sampled rows, the FP64
residual of the solvers, and HIP-graph replays of the SELL SpMV."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from smaxkernels_amd import capi, matgen
from tests.helpers import assert_bitwise, fd_gradient_check, sddmm_tree

pytestmark = pytest.mark.gpu

# at collection, before the library binds a HIP runtime (torch brings its own)
torch = pytest.importorskip("torch")


def sampled_rows_crs(A, rows):
    """CRS sub-matrix of the given rows (for oracle checks on a sample)."""
    lens = np.diff(A.row_ptr)[rows]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.arange(A.row_ptr[r], A.row_ptr[r + 1]) for r in rows])
    return rp, A.col[idx], A.val[idx]


@pytest.fixture(scope="module")
def lomax4m():
    return matgen.lomax_band(4_000_000)


def test_config2_all_formats_agree_bitwise(ctx, lomax4m):
    """N = 4M Lomax band, FP64: CRS (both schedules) and SELL-64/32 at every
    listed sigma (1, 64, 4096, N) and at 65536 give the same y, bit for bit,
    and match the oracle on 2000 sampled rows (incl. the longest)."""
    A = lomax4m
    x = matgen.uniform_f64(matgen.SEED_X, A.n_cols)
    rp, col, val, dx = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val), ctx.upload(x)
    y = ctx.zeros(A.n_rows, np.float64)
    ctx.crs_spmv(capi.F64, A.n_rows, A.n_cols, A.nnz, rp, col, val, dx, y, None)
    ref = ctx.download(y)
    lens = np.diff(A.row_ptr)
    rng = np.random.default_rng(8)
    rows = np.unique(np.concatenate([rng.choice(A.n_rows, 1990, replace=False),
                                     np.argsort(-lens)[:10]]))
    srp, scol, sval = sampled_rows_crs(A, rows)
    assert_bitwise(ref[rows], oracle.crs_spmv(srp, scol, sval, x), "sampled rows vs oracle")
    plan = ctx.crs_plan(A.n_rows, A.nnz, rp)
    ctx.crs_spmv(capi.F64, A.n_rows, A.n_cols, A.nnz, rp, col, val, dx, y, plan)
    assert_bitwise(ctx.download(y), ref, "CRS row blocks")
    ctx.crs_plan_destroy(plan)
    for C_ in (64, 32):
        for sigma in (1, 64, 4096, 65536, A.n_rows):
            s = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, val, 8, C_, sigma)
            y.free()
            y = ctx.zeros(A.n_rows, np.float64)
            s.apply(capi.F64, dx, y, unpermuted=True)
            assert_bitwise(ctx.download(y), ref, f"SELL-{C_}-{sigma}")
            s.destroy()
    for d in (rp, col, val, dx, y):
        d.free()


def test_config1_graph_replays_match_eager(ctx):
    """The SELL SpMV captured in a HIP graph (200 launches), replayed with
    eager work and synchronisation in between: every replay equals the eager
    result (the launch records no memset and no host work)."""
    P = matgen.poisson2d(512)
    x = matgen.uniform_f64(matgen.SEED_X, P.n_cols)
    rp, col, val, dx = ctx.upload(P.row_ptr), ctx.upload(P.col), ctx.upload(P.val), ctx.upload(x)
    s = ctx.scs_from_crs(P.n_rows, P.n_cols, P.nnz, rp, col, val, 8, 64, 1)
    y = ctx.zeros(s.info.n_rows_padded, np.float64)
    s.apply(capi.F64, dx, y)
    want = ctx.download(y)
    assert_bitwise(want[s.arrays(np.float64)["perm"]],
                   oracle.crs_spmv(P.row_ptr, P.col, P.val, x), "eager vs oracle")
    st = torch.cuda.Stream()
    ctx.set_stream(st.cuda_stream)
    try:
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=st):
            for _ in range(200):
                s.apply(capi.F64, dx, y)
        for rep in range(4):
            # cleared on the graph's stream: the replay must rewrite all of y
            ctx._check(ctx.lib.smax_hip_zero(ctx.handle, y.ptr, y.nbytes))
            scratch = torch.ones(1 << 20, device="cuda") * rep  # eager work between replays
            with torch.cuda.stream(st):
                g.replay()
            torch.cuda.synchronize()
            assert float(scratch.sum()) == float(rep * (1 << 20))
            assert_bitwise(ctx.download(y), want, f"replay {rep}")
    finally:
        ctx.set_stream(None)
    for d in (rp, col, val, dx, y):
        d.free()
    s.destroy()


def test_config3_solvers_full_size(ctx):
    """160^3 27-point stencil: FP64 CG and the BF16- and FP32-value
    reliable-update IR-CG reach ||b - Ax|| / ||b|| <= 1e-10, re-checked with
    the oracle's FP64 SpMV.  The refined solutions agree with the FP64 CG
    solution to within what the two residuals allow:
    ||x - x_64|| / ||x_64|| <= kappa (rho + rho_64) / (1 - rho_64), with
    kappa = lambda_max / lambda_min from the stencil's closed-form spectrum
    (eigenvalues 27.01 - prod_d (1 + 2 cos(pi j_d / 161)), kappa ~ 1.8e3)."""
    A = matgen.stencil27(160)
    b = matgen.uniform_f64(matgen.SEED_B, A.n_rows)
    rp, col, val = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val)
    S = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, val, 8, 64, 1)
    db = ctx.upload(b)
    c = np.cos(np.pi * np.array([1, 160]) / 161)
    lam_min = 27.01 - (1 + 2 * c[0]) ** 3
    lam_max = 27.01 - (1 + 2 * c[0]) ** 2 * (1 + 2 * c[1])
    kappa = lam_max / lam_min
    x64 = rho64 = None
    for mixed, lo in ((False, capi.BF16), (True, capi.BF16), (True, capi.F32)):
        dx = ctx.zeros(A.n_rows, np.float64)
        prm = capi.CgParams(rel_tol=1e-10, max_iters=20000, check_every=25,
                            inner_rel_tol=0.1, max_inner=5000,
                            inner_value_type=lo, keep_diag_f32=1)
        rep = ctx.cg(S, db, dx, prm, mixed)
        assert rep.status == capi.SUCCESS, (mixed, lo)
        xh = ctx.download(dx)
        r = b - oracle.crs_spmv(A.row_ptr, A.col, A.val, xh)
        rho = np.linalg.norm(r) / np.linalg.norm(b)
        assert rho <= 1e-10, (mixed, lo)
        if x64 is None:
            x64, rho64 = xh, rho
        else:
            err = np.linalg.norm(xh - x64) / np.linalg.norm(x64)
            assert err <= kappa * (rho + rho64) / (1 - rho64), (mixed, lo, err)
        dx.free()
    for d in (rp, col, val, db):
        d.free()
    S.destroy()


@pytest.mark.parametrize("k", [8, 64])
def test_config5_bcrs_spmm_full_size(ctx, k):
    """96^3 7-point block adjacency, 4x4 BF16 blocks, k in {8, 64}: BCRS SpMM
    equals k BCRS SpMVs column by column (the reference's definition), bit
    for bit, forward and through the transpose; sampled block rows match the
    oracle."""
    A = matgen.block7(96, 4)
    v16 = matgen.to_bf16(A.val)
    rp, col, val = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(v16)
    B = ctx.bcrs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, val, 2, 4, 4, 4, 4, False)
    for d in (rp, col, val):
        d.free()
    X = matgen.uniform_f32(matgen.SEED_X, A.n_rows * k)
    dX, dY = ctx.upload(X), ctx.zeros(A.n_rows * k, np.float32)
    B.spmm(capi.BF16, k, dX, dY)
    Y = ctx.download(dY).reshape(A.n_rows, k)
    dx, dy = ctx.zeros(A.n_rows, np.float32), ctx.zeros(A.n_rows, np.float32)
    for v in (0, k // 2 - 1, k - 1):
        dx.free()
        dx = ctx.upload(np.ascontiguousarray(X.reshape(A.n_rows, k)[:, v]))
        B.spmv(capi.BF16, dx, dy)
        assert_bitwise(Y[:, v], ctx.download(dy), f"column {v}")
    rows = np.arange(0, A.n_rows, 997)
    srp, scol, _ = sampled_rows_crs(A, rows)
    sval = np.concatenate([v16[A.row_ptr[r]:A.row_ptr[r + 1]] for r in rows])
    want = oracle.crs_spmm(srp, scol, sval, k, True, X)
    assert_bitwise(Y[rows].ravel(), want, "sampled rows vs oracle CRS SpMM")
    T = B.transpose()
    dY.free()
    dY = ctx.zeros(A.n_rows * k, np.float32)
    T.spmm(capi.BF16, k, dX, dY)
    YT = ctx.download(dY).reshape(A.n_rows, k)
    # the 7-point block adjacency is structurally symmetric: A^T X on sampled
    # rows through the oracle on an explicit transpose
    trp, tcol, tval = oracle.crs_transpose(A.n_rows, A.n_cols, A.row_ptr, A.col, v16)
    srp, scol, sval = sampled_rows_crs(matgen.CRS(A.n_cols, A.n_rows, trp, tcol, tval), rows)
    assert_bitwise(YT[rows].ravel(), oracle.crs_spmm(srp, scol, sval, k, True, X),
                   "A^T sampled rows vs oracle")
    for d in (dX, dY, dx, dy):
        d.free()
    B.destroy()
    T.destroy()


def test_config4_spmm_adjoint_sddmm_full_size(ctx):
    """1M-row Lomax band (nnz ~ 16M), FP32 values, k = 32, loss = sum(Y o W):
    - forward Y = A X: CRS row-major without and with a plan, CRS
      column-major (the reference's default layout, ld = n_rows) with a
      plan, and SELL-64-65536 SpMM all give the same Y bit for bit, and
      2000 sampled rows (incl. the longest) match the oracle's CRS SpMM;
    - adjoint A^T W through the SELL layout of A^T matches the oracle's CRS
      SpMM of an explicit transpose on 2000 sampled rows, bit for bit;
    - SDDMM dvals_ij = sum_v W_iv X_jv lies within its tree's rounding bound
      of the exact sum on 200k sampled entries;
    - both gradients are within 1e-6 of FP64 central differences on 64 value
      and 64 X entries."""
    k = 32
    A = matgen.lomax_band(1_000_000)
    v32 = A.val.astype(np.float32)
    X = matgen.uniform_f32(matgen.SEED_X, A.n_cols * k)
    W = matgen.uniform_f32(matgen.SEED_W, A.n_rows * k)
    rp, col, val = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(v32)
    dX, dW = ctx.upload(X), ctx.upload(W)
    dY = ctx.zeros(A.n_rows * k, np.float32)
    plan = ctx.crs_plan(A.n_rows, A.nnz, rp)
    ctx.crs_spmm(capi.F32, A.n_rows, A.n_cols, A.nnz, rp, col, val, k, True, dX, k, dY, k)
    Y = ctx.download(dY)
    lens = np.diff(A.row_ptr)
    rng = np.random.default_rng(4)
    rows = np.unique(np.concatenate([rng.choice(A.n_rows, 1990, replace=False),
                                     np.argsort(-lens)[:10]]))
    srp, scol, sval = sampled_rows_crs(A, rows)
    Ys = Y.reshape(A.n_rows, k)[rows].ravel()
    assert_bitwise(Ys, oracle.crs_spmm(srp, scol, sval.astype(np.float32), k, True, X),
                   "sampled rows vs oracle")
    ctx.crs_spmm(capi.F32, A.n_rows, A.n_cols, A.nnz, rp, col, val, k, True, dX, k, dY, k,
                 plan)
    assert_bitwise(ctx.download(dY), Y, "row-major with plan")
    Xc = ctx.upload(np.ascontiguousarray(X.reshape(A.n_cols, k).T).ravel())
    Yc = ctx.zeros(A.n_rows * k, np.float32)
    ctx.crs_spmm(capi.F32, A.n_rows, A.n_cols, A.nnz, rp, col, val, k, False, Xc, A.n_rows,
                 Yc, A.n_rows, plan)
    assert_bitwise(ctx.download(Yc).reshape(k, A.n_rows).T.ravel(), Y, "column-major")
    for d in (Xc, Yc):
        d.free()
    ctx.crs_plan_destroy(plan)
    S = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, val, 4, 64, 65536)
    S.apply_spmm(capi.F32, k, dX, dY)
    assert_bitwise(ctx.download(dY), Y, "SELL-64-65536 SpMM")

    T = ctx.scs_transpose_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, val, 4, 64, A.n_cols)
    dG = ctx.zeros(A.n_cols * k, np.float32)
    T.apply_spmm(capi.F32, k, dW, dG)
    G = ctx.download(dG)
    trp, tcol, tval = oracle.crs_transpose(A.n_rows, A.n_cols, A.row_ptr, A.col, v32)
    At = matgen.CRS(A.n_cols, A.n_rows, trp, tcol, tval)
    tl = np.diff(trp)
    trows = np.unique(np.concatenate([rng.choice(A.n_cols, 1990, replace=False),
                                      np.argsort(-tl)[:10]]))
    srp, scol, sval = sampled_rows_crs(At, trows)
    assert_bitwise(G.reshape(A.n_cols, k)[trows].ravel(),
                   oracle.crs_spmm(srp, scol, sval, k, True, W), "A^T W sampled rows")

    ddv = ctx.zeros(A.nnz, np.float32)
    S.sddmm(capi.F32, k, dW, dX, ddv)
    dv = ctx.download(ddv)
    ent = rng.choice(A.nnz, 200_000, replace=False)
    rows_of = np.repeat(np.arange(A.n_rows), lens)
    Wl = W.reshape(-1, k).astype(np.longdouble)[rows_of[ent]]
    Xl = X.reshape(-1, k).astype(np.longdouble)[A.col[ent]]
    exact = (Wl * Xl).sum(axis=1)
    scale = np.abs(Wl * Xl).sum(axis=1)
    g, q = sddmm_tree(k)
    u = 2.0 ** -24
    assert (np.abs(dv[ent].astype(np.longdouble) - exact)
            <= (q + np.log2(g)) * u * scale * (1 + 4 * u) + 1e-300).all()

    fd = fd_gradient_check(A, v32, X, W, k, dv.astype(np.float64),
                           G.astype(np.float64).reshape(A.n_cols, k))
    assert fd["max_rel_err"] <= 1e-6, fd
    for d in (rp, col, val, dX, dW, dY, dG, ddv):
        d.free()
    S.destroy()
    T.destroy()
