"""FP64 CG and mixed-precision IR-CG on the GPU.  Every reported residual is
re-checked on the host with the oracle's FP64 CRS SpMV."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from smaxkernels_amd import capi, matgen

pytestmark = pytest.mark.gpu


def true_rel_residual(A, b, x):
    r = b - oracle.crs_spmv(A.row_ptr, A.col, A.val, x)
    return np.linalg.norm(r) / np.linalg.norm(b)


def solve(ctx, A, C_, sigma, mixed, lo=capi.BF16, diag=True, tol=1e-10,
          inner=1e-4, max_iters=20000, restart=0, coding=0):
    rp, col, v = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val)
    S = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, 8, C_, sigma)
    b = matgen.uniform_f64(matgen.SEED_B, A.n_rows)
    db, dx = ctx.upload(b), ctx.zeros(A.n_rows, np.float64)
    prm = capi.CgParams(rel_tol=tol, max_iters=max_iters, check_every=10,
                        inner_rel_tol=inner, max_inner=5000,
                        inner_value_type=lo, keep_diag_f32=int(diag),
                        restart_inner=restart, column_coding=coding)
    rep = ctx.cg(S, db, dx, prm, mixed)
    x = ctx.download(dx)
    S.destroy()
    return rep, b, x


@pytest.mark.parametrize("C_,sigma", [(64, 1), (64, 4096), (32, 1 << 30)])
def test_fp64_cg_reaches_1e10(ctx, C_, sigma):
    A = matgen.stencil27(20)
    rep, b, x = solve(ctx, A, C_, sigma, mixed=False)
    assert rep.status == capi.SUCCESS
    assert rep.rel_residual <= 1e-10
    assert true_rel_residual(A, b, x) <= 1.5e-10
    assert 10 < rep.iterations < 500


@pytest.mark.parametrize("sigma", [1, 1 << 30])
def test_cg_with_and_without_16bit_columns(ctx, sigma):
    """40^3 stencil: at sigma = 1 every slot-space column lies within 1641 of
    its slot, so the solvers read 16-bit column offsets; a global sort
    (sigma >= N) scatters rows, offsets exceed int16, and they keep int32
    columns.  Both reach 1e-10 in FP64 and BF16-refined CG."""
    A = matgen.stencil27(40)
    S = oracle.crs_to_scs(A.row_ptr, A.col, A.val, 64, sigma)
    rows = np.repeat(np.arange(A.n_rows), np.diff(A.row_ptr))
    spread = np.abs(S["perm"][A.col].astype(np.int64) - S["perm"][rows]).max()
    assert (spread <= 32767) == (sigma == 1)
    for mixed in (False, True):
        rep, b, x = solve(ctx, A, 64, sigma, mixed=mixed, inner=1e-1)
        assert rep.status == capi.SUCCESS
        assert true_rel_residual(A, b, x) <= 1e-10


@pytest.mark.parametrize("sigma", [1, 64])
def test_column_codings_give_identical_iterates(ctx, sigma):
    """One-byte table indices (a 27-point stencil chunk uses at most 27 slot
    offsets at sigma = 1), 16-bit offsets and 32-bit columns name the same
    column for every element: FP64 CG and BF16-refined CG give bit-identical
    solutions, iteration counts and residuals with each.  At sigma = 64 rows
    are permuted inside chunks, which may leave a chunk with too many
    distinct offsets for a table; the automatic choice then falls back."""
    A = matgen.stencil27(40)
    for mixed in (False, True):
        runs = [solve(ctx, A, 64, sigma, mixed=mixed, inner=1e-1, coding=c)
                for c in (capi.COLUMNS_AUTO, capi.COLUMNS_16, capi.COLUMNS_32)]
        for rep, _, x in runs:
            assert rep.status == capi.SUCCESS
        # the coding that ran: tables at sigma = 1, 16-bit and 32-bit as asked
        used = [rep.column_coding for rep, _, _ in runs]
        assert used[1:] == [capi.COLUMNS_16, capi.COLUMNS_32]
        assert used[0] == capi.COLUMNS_8 if sigma == 1 else used[0] in (capi.COLUMNS_8,
                                                                        capi.COLUMNS_16)
        rep0, b, x0 = runs[0]
        assert true_rel_residual(A, b, x0) <= 1e-10
        for rep, _, x in runs[1:]:
            assert rep.iterations == rep0.iterations
            assert rep.outer_iterations == rep0.outer_iterations
            assert rep.rel_residual == rep0.rel_residual
            assert np.array_equal(x.view(np.uint64), x0.view(np.uint64))


def random_band_spd(n, half, per_row, seed):
    """Symmetric, strictly diagonally dominant (so SPD) matrix with per_row
    random off-diagonal pairs per row inside +-half of the diagonal."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    i = np.repeat(np.arange(n), per_row)
    j = i + rng.integers(-half, half + 1, i.size)
    keep = (i != j) & (j >= 0) & (j < n)
    j = np.where(keep, j, 0)
    v = -rng.uniform(0.1, 1.0, keep.sum())
    M = sp.coo_matrix((v, (i[keep], j[keep])), shape=(n, n))
    M = (M + M.T).tocsr()
    M.sum_duplicates()
    M = M + sp.diags(np.asarray(abs(M).sum(axis=1)).ravel() + 1.0)
    M = M.tocsr()
    M.sort_indices()
    return matgen.CRS(n, n, M.indptr.astype(np.int32), M.indices.astype(np.int32),
                      M.data.astype(np.float64))


def test_column_tables_fall_back_when_chunks_use_many_offsets(ctx):
    """Random band rows: a 64-row chunk uses far more than 32 distinct slot
    offsets (all within int16), so the solvers read 16-bit offsets; forcing
    32-bit columns gives the same iterates."""
    A = random_band_spd(20000, 1000, 8, 5)
    for mixed in (False, True):
        runs = [solve(ctx, A, 64, 1, mixed=mixed, inner=1e-1, coding=c)
                for c in (capi.COLUMNS_AUTO, capi.COLUMNS_32)]
        assert [rep.column_coding for rep, _, _ in runs] == [capi.COLUMNS_16,
                                                              capi.COLUMNS_32]
        (rep0, b, x0), (rep1, _, x1) = runs
        assert rep0.status == capi.SUCCESS and rep1.status == capi.SUCCESS
        assert rep0.iterations == rep1.iterations
        assert np.array_equal(x0.view(np.uint64), x1.view(np.uint64))
        assert true_rel_residual(A, b, x0) <= 1e-10


def test_solvers_follow_refreshed_values(ctx):
    """The solvers keep rounded (FP32/BF16) copies of the values; refreshing
    the handle's values drops them, so a second solve runs on the new
    matrix (2A here: same pattern, still SPD)."""
    A = matgen.stencil27(24)
    rp, col, v = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val)
    S = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, v, 8, 64, 1)
    b = matgen.uniform_f64(matgen.SEED_B, A.n_rows)
    db = ctx.upload(b)
    prm = capi.CgParams(rel_tol=1e-10, max_iters=20000, check_every=10,
                        inner_rel_tol=1e-1, max_inner=5000,
                        inner_value_type=capi.BF16, keep_diag_f32=1)
    for scale in (1.0, 2.0):
        if scale != 1.0:
            S.refresh_values(ctx.upload(A.val * scale))
        for mixed in (True, False):
            dx = ctx.zeros(A.n_rows, np.float64)
            rep = ctx.cg(S, db, dx, prm, mixed)
            assert rep.status == capi.SUCCESS
            x = ctx.download(dx)
            r = b - scale * oracle.crs_spmv(A.row_ptr, A.col, A.val, x)
            assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.5e-10
    S.destroy()


@pytest.mark.parametrize("restart", [0, 1])
@pytest.mark.parametrize("inner", [1e-1, 1e-4])
@pytest.mark.parametrize("lo,diag", [(capi.F32, False), (capi.BF16, True),
                                     (capi.BF16, False)])
def test_ircg_reaches_1e10(ctx, lo, diag, inner, restart):
    """BF16 without the FP32 diagonal rounds 26.01 to 26: the low-precision
    operator is then far enough from A that reliable updates must fall back
    to restarted refinement, and still reach the target."""
    A = matgen.stencil27(20)
    rep, b, x = solve(ctx, A, 64, 1, mixed=True, lo=lo, diag=diag, inner=inner,
                      restart=restart)
    assert rep.status == capi.SUCCESS, rep.rel_residual
    assert rep.rel_residual <= 1e-10
    assert true_rel_residual(A, b, x) <= 1.5e-10
    assert rep.outer_iterations >= 2


def test_ircg_permuted_layout(ctx):
    """sigma > 1 reorders rows: the solver works on P A P^T in slot space."""
    A = matgen.stencil27(16)
    rep, b, x = solve(ctx, A, 32, 1 << 30, mixed=True, lo=capi.F32, diag=True)
    assert rep.rel_residual <= 1e-10
    assert true_rel_residual(A, b, x) <= 1.5e-10


def test_solvers_agree(ctx):
    A = matgen.stencil27(14)
    _, b, x64 = solve(ctx, A, 64, 1, mixed=False, tol=1e-12)
    _, _, xir = solve(ctx, A, 64, 1, mixed=True, tol=1e-12)
    assert np.max(np.abs(x64 - xir)) <= 1e-9 * np.max(np.abs(x64))


def test_not_converged_is_reported(ctx):
    A = matgen.stencil27(20)
    rep, _, _ = solve(ctx, A, 64, 1, mixed=False, max_iters=5)
    assert rep.status == capi.ERROR_NOT_CONVERGED
    assert rep.rel_residual > 1e-10
