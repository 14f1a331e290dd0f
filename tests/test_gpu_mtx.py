"""A Matrix Market file through the GPU path: the CRS the reference's
benchmarks build from it (mirrored symmetric entries, file order within
rows) goes through the CRS kernels and the GPU SELL conversion, and every
output equals the oracle's CRS SpMV on the same arrays, bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from smaxkernels_amd import capi, matgen
from tests.helpers import assert_bitwise

pytestmark = pytest.mark.gpu


def write_lower_symmetric(path, A):
    """Lower triangle of a symmetric CRS matrix, column by column (the usual
    SuiteSparse order), as a symmetric Matrix Market file."""
    rows = np.repeat(np.arange(A.n_rows), np.diff(A.row_ptr))
    keep = rows >= A.col
    r, c, v = rows[keep], A.col[keep], A.val[keep]
    o = np.lexsort((r, c))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        f.write(f"{A.n_rows} {A.n_cols} {len(o)}\n")
        for i in o:
            f.write(f"{r[i] + 1} {c[i] + 1} {float(v[i])!r}\n")


@pytest.mark.parametrize("C_,sigma", [(64, 1), (64, 4096), (32, 1 << 20)])
def test_mtx_symmetric_file_spmv(ctx, tmp_path, C_, sigma):
    S = matgen.stencil27(14)
    S.val = matgen.uniform_f64(3, S.nnz)
    # symmetrise the values: v(i,j) = v(j,i) = the lower triangle's value
    rows = np.repeat(np.arange(S.n_rows), np.diff(S.row_ptr))
    dense = {}
    for i, j, v in zip(rows, S.col, S.val):
        dense[(max(i, j), min(i, j))] = v
    S.val = np.array([dense[(max(i, j), min(i, j))] for i, j in zip(rows, S.col)])
    p = tmp_path / "s.mtx"
    write_lower_symmetric(p, S)
    A = matgen.read_mtx(p)
    assert A.nnz == S.nnz
    # rows are the stencil's rows, in a different (file) order within each row
    assert_bitwise(A.row_ptr, S.row_ptr)
    x = matgen.uniform_f64(matgen.SEED_X, A.n_cols)
    want = oracle.crs_spmv(A.row_ptr, A.col, A.val, x)
    rp, col, val, dx = ctx.upload(A.row_ptr), ctx.upload(A.col), ctx.upload(A.val), ctx.upload(x)
    y = ctx.zeros(A.n_rows, np.float64)
    plan = ctx.crs_plan(A.n_rows, A.nnz, rp)
    ctx.crs_spmv(capi.F64, A.n_rows, A.n_cols, A.nnz, rp, col, val, dx, y, plan)
    assert_bitwise(ctx.download(y), want, "CRS row blocks")
    ctx.crs_plan_destroy(plan)
    s = ctx.scs_from_crs(A.n_rows, A.n_cols, A.nnz, rp, col, val, 8, C_, sigma)
    s.apply(capi.F64, dx, y, unpermuted=True)
    assert_bitwise(ctx.download(y), want, f"SELL-{C_}-{sigma}")
    s.destroy()
    for d in (rp, col, val, dx, y):
        d.free()
