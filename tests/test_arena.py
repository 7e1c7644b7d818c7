"""Arena tensors of tensors on the host side (no GPU): the oracle's arena
contraction against closed forms (the dense block product of the cell
matrices; tiles screened out by their scaled norms), the slab layout, and
the host plan of the inner-outer-product (ce+e) regime executed in numpy
against the oracle.  The reference's own ToT tests pin no values
(tests/tot_expressions.cpp:4246-4269) and its arena tests compare two
implementations (tests/einsum.cpp:1270-1378), so the pins here are the
mathematics."""
import numpy as np
import pytest
import torch

import oracle
from tiledarray_amd.arena import ArenaArray, plan_ce_e, slab_layout
from tiledarray_amd.tiled import TiledRange1

EPS = float(np.finfo(np.float32).eps)


def host_arena(rows, cols, inner, fill, seed, scale_tile=None):
    rows, cols = TiledRange1(rows), TiledRange1(cols)
    rng = np.random.default_rng(seed)
    present = rng.random((rows.extent, cols.extent)) < fill
    vol = np.ones(present.shape, np.int64)
    for m, e in inner:
        vol *= np.asarray(e)[:, None] if m == 0 else np.asarray(e)[None, :]
    off, _, _, _ = slab_layout(rows, cols, present, vol)
    n = int(np.where(present, vol, 0).sum())
    pool = rng.integers(-8, 9, max(1, n)).astype(np.float64)
    if scale_tile is not None:  # shrink one outer tile's cells
        (tr, tc), f = scale_tile
        for r in range(rows.boundaries[tr], rows.boundaries[tr + 1]):
            for c in range(cols.boundaries[tc], cols.boundaries[tc + 1]):
                if off[r, c] >= 0:
                    pool[off[r, c]:off[r, c] + vol[r, c]] *= f
    return ArenaArray(rows, cols, inner, present, torch.as_tensor(pool))


def dense(X, er, ec):
    """The cell matrix as one dense matrix (empty cells zero)."""
    ra = np.concatenate([[0], np.cumsum(er)])
    rc = np.concatenate([[0], np.cumsum(ec)])
    D = np.zeros((ra[-1], rc[-1]))
    pool = X.pool.numpy()
    for r in range(er.size):
        for c in range(ec.size):
            o = X.cell_off[r, c]
            if o >= 0:
                D[ra[r]:ra[r + 1], rc[c]:rc[c + 1]] = pool[o:o + er[r] * ec[c]].reshape(er[r], ec[c])
    return D, ra, rc


def packed_result(ea, eb):
    co = np.zeros((ea.size, eb.size), np.int64)
    at = 0
    for i in range(ea.size):
        for j in range(eb.size):
            co[i, j] = at
            at += ea[i] * eb[j]
    return co, at


SETUP = [
    # (i bounds, k bounds, j bounds, fill)
    ([0, 3, 7, 8], [0, 4, 6], [0, 2, 5], 0.8),
    ([0, 1], [0, 1], [0, 1], 1.0),
    ([0, 5, 10], [0, 3, 6, 9], [0, 4, 7], 0.5),
]


@pytest.mark.parametrize("setup", SETUP, ids=["ragged", "one_cell", "half"])
def test_oracle_arena_contract_is_dense_block_product(setup):
    ib, kb, jb, fill = setup
    rng = np.random.default_rng(len(ib) * 7 + len(kb))
    ea = rng.integers(1, 6, ib[-1])
    ec = rng.integers(1, 6, kb[-1])
    eb = rng.integers(1, 6, jb[-1])
    A = host_arena(ib, kb, [(0, ea), (1, ec)], fill, 1)
    B = host_arena(kb, jb, [(0, ec), (1, eb)], fill, 2)
    co, n = packed_result(ea, eb)
    C, est, sa, sb = oracle.arena_contract(ib, kb, jb, ea, ec, eb, A.pool.numpy(), A.cell_off,
                                           B.pool.numpy(), B.cell_off, co, n, alpha=1.5,
                                           threshold=1e-30)
    Ad, ra, _ = dense(A, ea, ec)
    Bd, _, rbb = dense(B, ec, eb)
    Cd = 1.5 * (Ad @ Bd)
    for i in range(ea.size):
        for j in range(eb.size):
            got = C[co[i, j]:co[i, j] + ea[i] * eb[j]].reshape(ea[i], eb[j])
            np.testing.assert_array_equal(got, Cd[ra[i]:ra[i + 1], rbb[j]:rbb[j + 1]])


def test_oracle_arena_screening_drops_small_tiles():
    # A's outer tile (0, 1) is scaled to 1e-20: its scaled norm falls below
    # the float epsilon threshold, so its cells take part in no product and
    # result tiles fed only by it are zero
    ib, kb, jb = [0, 3, 6], [0, 2, 5], [0, 3, 4]
    rng = np.random.default_rng(9)
    ea, ec, eb = rng.integers(1, 5, 6), rng.integers(1, 5, 5), rng.integers(1, 5, 4)
    A = host_arena(ib, kb, [(0, ea), (1, ec)], 1.0, 3, scale_tile=((0, 1), 1e-20))
    B = host_arena(kb, jb, [(0, ec), (1, eb)], 1.0, 4)
    co, n = packed_result(ea, eb)
    C, est, sa, sb = oracle.arena_contract(ib, kb, jb, ea, ec, eb, A.pool.numpy(), A.cell_off,
                                           B.pool.numpy(), B.cell_off, co, n, threshold=EPS)
    assert sa[0, 1] == 0.0 and sa[0, 0] > 0 and sa[1, 1] > 0
    pool = A.pool.numpy().copy()
    for r in range(0, 3):
        for c in range(2, 5):
            o = A.cell_off[r, c]
            pool[o:o + ea[r] * ec[c]] = 0.0
    Az = ArenaArray(A.rows, A.cols, A.inner, A.present, torch.as_tensor(pool))
    Ad, ra, _ = dense(Az, ea, ec)
    Bd, _, rbb = dense(B, ec, eb)
    Cd = Ad @ Bd
    for i in range(ea.size):
        for j in range(eb.size):
            got = C[co[i, j]:co[i, j] + ea[i] * eb[j]].reshape(ea[i], eb[j])
            np.testing.assert_array_equal(got, Cd[ra[i]:ra[i + 1], rbb[j]:rbb[j + 1]])


def test_slab_layout_tiles_are_contiguous_row_major_runs():
    rows, cols = TiledRange1([0, 3, 4, 9]), TiledRange1([0, 2, 7])
    rng = np.random.default_rng(5)
    present = rng.random((9, 7)) < 0.6
    vol = rng.integers(1, 9, (9, 7))
    off, order, pre, slab = slab_layout(rows, cols, present, vol)
    t = 0
    for tr in range(rows.ntiles):
        for tc in range(cols.ntiles):
            at = slab[t]
            ids = order[pre[t]:pre[t + 1]]
            rr = range(rows.boundaries[tr], rows.boundaries[tr + 1])
            cc = range(cols.boundaries[tc], cols.boundaries[tc + 1])
            assert list(ids) == [r * 7 + c for r in rr for c in cc]
            for r in rr:
                for c in cc:
                    if present[r, c]:
                        assert off[r, c] == at
                        at += vol[r, c]
                    else:
                        assert off[r, c] == -1
            assert at == slab[t + 1]
            t += 1
    assert slab[-1] == np.where(present, vol, 0).sum()


def run_plan_on_host(A, B, tasks, pairs, n):
    """The grouped GEMM (op_a = TRANS, op_b = NO_TRANS) of a host work list,
    in numpy, with beta 0."""
    a, b = A.pool.numpy(), B.pool.numpy()
    C = np.zeros(n)
    for (c_off, m, nn, ldc, pb, pc, _flags) in tasks:
        acc = np.zeros((m, nn))
        for (ao, bo, k, lda, ldb, _t1, _t2, _r) in pairs[pb:pb + pc]:
            Am = np.stack([a[ao + kk * lda:ao + kk * lda + m] for kk in range(k)])  # k x m
            Bm = np.stack([b[bo + kk * ldb:bo + kk * ldb + nn] for kk in range(k)])  # k x n
            acc += Am.T @ Bm
        C[c_off:c_off + m * nn] = acc.reshape(-1)
    return C


@pytest.mark.parametrize("fill", [1.0, 0.85, 0.4])
def test_ce_e_plan_executed_on_host_equals_oracle(fill):
    ib, kb, jb = [0, 3, 7], [0, 4, 5, 9], [0, 2, 6]
    rng = np.random.default_rng(int(fill * 100))
    ep = rng.integers(1, 7, ib[-1])
    eq = rng.integers(1, 7, jb[-1])
    A = host_arena(ib, kb, [(0, ep)], fill, 11)
    B = host_arena(kb, jb, [(1, eq)], fill, 12)
    ones = np.ones(kb[-1], np.int64)
    # shapes of the oracle, then the plan from them
    co0, n0 = packed_result(ep, eq)
    _, est, sa, sb = oracle.arena_contract(ib, kb, jb, ep, ones, eq, A.pool.numpy(), A.cell_off,
                                           B.pool.numpy(), B.cell_off, co0, n0, threshold=EPS)
    tasks, pairs, c_off, n, merged = plan_ce_e(A, B, sa, sb, est, EPS)
    if fill == 1.0:  # every run complete: one GEMM per task and K tile
        assert merged == len(pairs) > 0 and any(p[2] > 1 for p in pairs)
    # each (i, j, k) with both cells present in non-zero tiles is covered once
    Cref, _, _, _ = oracle.arena_contract(ib, kb, jb, ep, ones, eq, A.pool.numpy(), A.cell_off,
                                          B.pool.numpy(), B.cell_off, c_off, n, threshold=EPS)
    C = run_plan_on_host(A, B, tasks, pairs, n)
    np.testing.assert_array_equal(C, Cref)
    # result cells are packed slab by slab
    order = [c_off[i, j] for I in range(len(ib) - 1) for J in range(len(jb) - 1)
             for i in range(ib[I], ib[I + 1]) for j in range(jb[J], jb[J + 1]) if c_off[i, j] >= 0]
    assert order == sorted(order)
