"""tamx_tile_norms_f64_noalloc, the tile norms without any device
allocation (no counterpart in the reference; for callers that cannot
allocate), through the C ABI against tamx_tile_norms_f64 with a plan of the
same tiles: bit for bit, on two calls in a row, the second with fewer and
smaller tiles, so the shared static scratch must not carry anything over.
Also: the size limits are refused on the host before any launch."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from tiledarray_amd import _lib
from tiledarray_amd import engine as E

pytestmark = pytest.mark.gpu


def _tiles(rng, ntiles, lo, hi, absent=0.2):
    """A pool holding ntiles tiles of random volumes in [lo, hi] (some
    absent), in a random order inside the pool."""
    vols = rng.integers(lo, hi + 1, ntiles).astype(np.int64)
    present = rng.random(ntiles) >= absent
    order = rng.permutation(ntiles)
    offs = np.full(ntiles, -1, np.int64)
    pos = 0
    for t in order:
        if present[t]:
            offs[t] = pos
            pos += int(vols[t])
    pool = torch.empty(max(1, pos), dtype=torch.float64, device="cuda")
    pool.uniform_(-1.0, 1.0)
    return offs, vols, pool


def _planned(ctx, offs, vols, pool):
    plan = E.NormPlan(ctx, offs, vols)
    out = torch.empty(offs.size, dtype=torch.float32, device="cuda")
    E.call("tamx_tile_norms_f64", ctx.ptr, plan.ptr, pool.data_ptr(), out.data_ptr(),
           ctx.stream())
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _noalloc(ctx, offs, vols, pool):
    d_off = torch.as_tensor(offs, device="cuda")
    d_vol = torch.as_tensor(vols, device="cuda")
    out = torch.full((offs.size,), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    E.call("tamx_tile_norms_f64_noalloc", ctx.ptr, d_off.data_ptr(), d_vol.data_ptr(),
           offs.size, pool.data_ptr(), pool.numel(), out.data_ptr(), ctx.stream())
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_norms_equal_planned_norms_twice(ctx):
    rng = np.random.default_rng(77)
    # first call: 3000 tiles up to 300000 elements (many span several
    # 65536-element chunks); second: 500 tiles up to 9000 elements
    for ntiles, lo, hi in ((3000, 0, 300000), (500, 1, 9000)):
        offs, vols, pool = _tiles(rng, ntiles, lo, hi)
        ref = _planned(ctx, offs, vols, pool)
        got = _noalloc(ctx, offs, vols, pool)
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
        assert (got[offs < 0] == 0).all()
        # and both equal the oracle's sequential norms to float rounding
        h = pool.cpu().numpy()
        orc = oracle.tile_norms(h, offs, vols)
        np.testing.assert_allclose(got, orc, rtol=1e-6, atol=0)


def test_noalloc_limits_refused_before_launch(ctx):
    L = E.lib()
    pool = torch.zeros(16, dtype=torch.float64, device="cuda")
    big = _lib.NOALLOC_MAX_TILES + 1
    d = torch.zeros(big, dtype=torch.int64, device="cuda")
    out = torch.zeros(big, dtype=torch.float32, device="cuda")
    rc = L.tamx_tile_norms_f64_noalloc(ctx.ptr, d.data_ptr(), d.data_ptr(), big,
                                       pool.data_ptr(), pool.numel(), out.data_ptr(),
                                       ctx.stream())
    assert rc == _lib.TAMX_ERR_CAPACITY and b"tiles exceed" in L.tamx_last_error()
    rc = L.tamx_tile_norms_f64_noalloc(ctx.ptr, d.data_ptr(), d.data_ptr(), 4,
                                       pool.data_ptr(), _lib.NOALLOC_MAX_CHUNKS * 65536,
                                       out.data_ptr(), ctx.stream())
    assert rc == _lib.TAMX_ERR_CAPACITY and b"chunks" in L.tamx_last_error()
    torch.cuda.synchronize()
    assert bool((out == 0).all())
