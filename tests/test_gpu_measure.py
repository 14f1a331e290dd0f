"""Measurement helpers of the C ABI: the stream copy/triad kernels behind the
measured bandwidth peak and the random-gather kernel behind the gather
ceiling.  Their results are checked exactly (they only move or sum values
that are exact in binary64)."""
from __future__ import annotations

import numpy as np
import pytest

from smaxkernels_amd import capi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 1001, 65536])
def test_stream_copy_and_triad(ctx, n):
    b = np.arange(n, dtype=np.float64)
    c = (np.arange(n, dtype=np.float64) % 7) - 3.0
    db, dc, da = ctx.upload(b), ctx.upload(c), ctx.zeros(n, np.float64)
    ctx.stream_copy(db, da, n)
    assert np.array_equal(ctx.download(da), b)
    ctx.stream_triad(da, db, dc, 0.5, n)
    assert np.array_equal(ctx.download(da), b + 0.5 * c)  # exact: halves of small integers
    for d in (da, db, dc):
        d.free()


@pytest.mark.parametrize("loads", [1, 2, 4, 8])
@pytest.mark.parametrize("n", [1, 2, 3, 1001, 65537, 3_000_001])
def test_stream_copy_wide(ctx, loads, n):
    """The copy with several 16-byte loads in flight per lane moves every
    element, including the odd last one and partial workgroup spans; an
    unsupported load count is refused."""
    src = np.arange(n, dtype=np.float64) * 1.5 - 7.0
    ds, dd = ctx.upload(src), ctx.upload(np.full(n, -1.0))
    ctx.stream_copy_wide(ds, dd, n, loads)
    assert np.array_equal(ctx.download(dd), src)
    st = ctx.lib.smax_hip_stream_copy_wide(ctx.handle, ds.ptr, dd.ptr, n, 3)
    assert st == capi.ERROR_INVALID_VALUE
    for d in (ds, dd):
        d.free()


@pytest.mark.parametrize("per_lane", [8, 16, 32])
@pytest.mark.parametrize("n,window", [(100_003, 1), (100_003, 1000), (100_003, 100_003),
                                      (4096, 4000)])
def test_gather_rate_reads_inside_the_window(ctx, per_lane, n, window):
    """x = index: every gathered value lies in lane i's window
    [lo, lo + window) with lo = clamp(i - window/2, 0, n - window), so
    y[i] / per_lane lies in [lo, lo + window - 1]; with x = 1, y = per_lane."""
    dx, dy = ctx.upload(np.arange(n, dtype=np.float64)), ctx.zeros(n, np.float64)
    ctx.gather_rate(dx, dy, n, window, per_lane)
    y = ctx.download(dy)
    i = np.arange(n)
    lo = np.clip(i - window // 2, 0, n - window)
    assert np.all(y >= per_lane * lo) and np.all(y <= per_lane * (lo + window - 1))
    if window == 1:
        assert np.array_equal(y, per_lane * lo.astype(np.float64))
    dx.free()
    dx = ctx.upload(np.ones(n))
    ctx.gather_rate(dx, dy, n, window, per_lane)
    assert np.all(ctx.download(dy) == per_lane)
    dx.free()
    dy.free()


def test_gather_rate_rejects_bad_arguments(ctx):
    dx, dy = ctx.zeros(16, np.float64), ctx.zeros(16, np.float64)
    for n, window, k in ((16, 17, 16), (16, 0, 16), (0, 1, 16), (16, 4, 7)):
        with pytest.raises(capi.SmaxHipError):
            ctx.gather_rate(dx, dy, n, window, k)
    dx.free()
    dy.free()
