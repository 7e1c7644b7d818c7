"""Arena tensors of tensors on the device (outer tiles of several cells, one
slab each) contracted through the grouped launch, against the oracle's arena
contraction (oracle.arena_contract, whose closed forms tests/test_arena.py
pins).  Both regimes: the inner GEMM C(i,j;a,b) = A(i,k;a,c) B(k,j;c,b),
batched over result outer tiles with the operands' cells read where they
lie, and the inner outer product C(i,j;p,q) = A(i,k;p) B(k,j;q) with
complete k-runs as one GEMM (arena_einsum.h:1170).

Bars: FP64 integer-valued cells bit-exact; FP32 relative Frobenius error
<= 1e-5; outer-tile shapes (scaled norms, result estimate) bit-exact for
FP64 cells, within 1e-5 for FP32 ones; the result cells packed slab by
slab."""
import numpy as np
import pytest
import torch

import oracle
from tiledarray_amd.arena import ArenaArray, ArenaContraction

pytestmark = pytest.mark.gpu

EPS = float(np.finfo(np.float32).eps)

# (i bounds, k bounds, j bounds, fill, empty-tile probability)
CASES = [
    ([0, 3, 7, 8], [0, 4, 6, 11], [0, 2, 5, 9], 0.8, 0.0),
    ([0, 4, 8, 12, 16], [0, 4, 8, 12, 16], [0, 4, 8, 12, 16], 0.5, 0.2),
    ([0, 1, 2], [0, 5], [0, 3], 1.0, 0.0),
]
IDS = ["ragged", "sparse", "long_runs"]


def check_result(c, A, B, ea, ec, eb, alpha, f64):
    torch.cuda.synchronize()
    R = c.result()
    ib, kb, jb = A.rows.boundaries, A.cols.boundaries, B.cols.boundaries
    Cref, est, sa, sb = oracle.arena_contract(
        ib, kb, jb, ea, ec, eb, A.pool.cpu().numpy(), A.cell_off, B.pool.cpu().numpy(),
        B.cell_off, R.cell_off, R.pool.numel(), alpha=alpha, threshold=c.threshold)
    NI, NK, NJ = c.NI, c.NK, c.NJ
    # shapes: FP64 cells bit-exact; FP32 cells sum their squares in float,
    # in another order than the oracle, so within float rounding
    same = np.testing.assert_array_equal if f64 else (
        lambda x, y: np.testing.assert_allclose(x, y, rtol=1e-5, atol=0))
    same(c.a_tn.cpu().numpy()[:NI * NK], sa.reshape(-1))
    same(c.b_tn.cpu().numpy()[:NK * NJ], sb.reshape(-1))
    same(c.c_est.cpu().numpy()[:NI * NJ], est.reshape(-1))
    # result cells: present exactly in the non-zero result tiles
    ti = np.repeat(np.arange(NI), np.diff(ib))
    tj = np.repeat(np.arange(NJ), np.diff(jb))
    np.testing.assert_array_equal(R.present, est[ti[:, None], tj[None, :]] >= c.threshold)
    got = R.pool.cpu().numpy()
    used = np.zeros(got.size, bool)
    for i in range(R.shape[0]):
        for j in range(R.shape[1]):
            o = R.cell_off[i, j]
            if o >= 0:
                used[o:o + ea[i] * eb[j]] = True
    if f64:
        np.testing.assert_array_equal(got[used], Cref[used])
    else:
        d = np.linalg.norm((got[used] - Cref[used]).astype(np.float64))
        assert d <= 1e-5 * max(np.linalg.norm(Cref[used].astype(np.float64)), 1e-30)
    # slab by slab: cells of result tile (I, J) in row-major order, tiles in order
    order = [R.cell_off[i, j] for I in range(NI) for J in range(NJ)
             for i in range(ib[I], ib[I + 1]) for j in range(jb[J], jb[J + 1])
             if R.cell_off[i, j] >= 0]
    assert order == sorted(order)
    return R


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_arena_inner_gemm(ctx, dtype, case):
    ib, kb, jb, fill, empty = case
    rng = np.random.default_rng(len(ib) * 31 + len(kb))
    ea, ec, eb = (rng.integers(3, 41, b[-1]) for b in (ib, kb, jb))
    f64 = dtype == torch.float64
    A = ArenaArray.random(ib, kb, [(0, ea), (1, ec)], fill=fill, empty_tiles=empty, dtype=dtype,
                          integer=f64, seed=5)
    B = ArenaArray.random(kb, jb, [(0, ec), (1, eb)], fill=fill, empty_tiles=empty, dtype=dtype,
                          integer=f64, seed=6)
    c = ArenaContraction(ctx, A, B, alpha=1.5)
    assert c.regime == "inner_gemm"
    c.run()
    R = check_result(c, A, B, ea, ec, eb, 1.5, f64)
    # a second run over the same operands leaves the same result
    first = R.pool.clone()
    c.run()
    torch.cuda.synchronize()
    assert torch.equal(c.result().pool, first)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_arena_outer_product(ctx, dtype, case):
    ib, kb, jb, fill, empty = case
    rng = np.random.default_rng(len(ib) * 17 + len(jb))
    ep, eq = rng.integers(1, 70, ib[-1]), rng.integers(1, 70, jb[-1])
    f64 = dtype == torch.float64
    A = ArenaArray.random(ib, kb, [(0, ep)], fill=fill, empty_tiles=empty, dtype=dtype,
                          integer=f64, seed=8)
    B = ArenaArray.random(kb, jb, [(1, eq)], fill=fill, empty_tiles=empty, dtype=dtype,
                          integer=f64, seed=9)
    c = ArenaContraction(ctx, A, B, alpha=-2.0)
    assert c.regime == "outer_product"
    c.run()
    check_result(c, A, B, ep, np.ones(kb[-1], np.int64), eq, -2.0, f64)
    if fill == 1.0:
        assert c.nmerged > 0 and c.counters()["npairs"] == c.nmerged


def test_arena_tile_norms_match_oracle(ctx):
    # outer-tile norms over the cells where they lie, FP64: bit-exact
    ib, kb = [0, 3, 7, 8], [0, 4, 6, 11]
    rng = np.random.default_rng(3)
    ea, ec = rng.integers(1, 30, 8), rng.integers(1, 30, 11)
    A = ArenaArray.random(ib, kb, [(0, ea), (1, ec)], fill=0.7, dtype=torch.float64, seed=2)
    n = A.tile_norms(ctx)
    torch.cuda.synchronize()
    ref = oracle.tot_tile_norms(A.pool.cpu().numpy(), A.tile_cells,
                                A.cell_off.reshape(-1)[A.slab_cells],
                                A.cell_vol.reshape(-1)[A.slab_cells])
    np.testing.assert_array_equal(n.cpu().numpy()[:ref.size], ref)
