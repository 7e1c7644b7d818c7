"""Pins the CPU oracle against closed forms, brute force and hand-derived
cases (the reference ships no golden vectors for this path; its own tests pin
results by invariants, which these follow).

A mistake planted in any oracle function (a dropped scaling term, a swapped
Trans flag, a wrong permutation direction, a missing beta) fails one of these.
"""
import itertools
import math

import numpy as np
import pytest

import oracle
from oracle import permopt


def brute_gemm(op_a, op_b, alpha, A, B, beta=0.0, C=None):
    Am = A.T if op_a else A
    Bm = B.T if op_b else B
    m, k = Am.shape
    n = Bm.shape[1]
    out = np.zeros((m, n)) if C is None else C.astype(np.float64).copy() * beta
    for i in range(m):
        for j in range(n):
            s = 0.0
            for p in range(k):
                s += float(Am[i, p]) * float(Bm[p, j])
            out[i, j] += alpha * s
    return out


# ----------------------------------------------------------------- norms --

def test_norm_all_ones_tile():
    # a tile of ones has norm sqrt(volume) (examples/gemm/ta_sparse.cpp:146-166)
    for shape in [(1,), (7,), (16, 16), (64, 3), (5, 6, 7)]:
        x = np.ones(shape)
        assert oracle.tile_norm(x) == np.float32(math.sqrt(x.size))


def test_norm_sequential_double_order():
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, 1000)
    s = 0.0
    for v in x:  # reduce_op_serial order (vector_op.h:562-577)
        s += v * v
    assert oracle.tile_norm(x) == np.float32(math.sqrt(s))


def test_norm_float_tiles_accumulate_in_float():
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, 777).astype(np.float32)
    s = np.float32(0)
    for v in x:
        s = np.float32(s + np.float32(v * v))
    assert oracle.tile_norm(x) == np.float32(np.sqrt(s))


def test_tile_norms_pool_and_absent_tiles():
    rng = np.random.default_rng(3)
    pool = rng.uniform(-2, 2, 100)
    offs = np.array([0, -1, 30, 60], np.int64)
    vols = np.array([30, 10, 30, 40], np.int64)
    n = oracle.tile_norms(pool, offs, vols)
    assert n[1] == 0
    for t in (0, 2, 3):
        assert n[t] == np.float32(np.sqrt(np.sum(pool[offs[t]:offs[t] + vols[t]] ** 2)))


# ------------------------------------------------------- shape scaling ---

def test_scale_rank2_closed_form_and_order():
    rng = np.random.default_rng(4)
    e0 = np.array([3, 5, 7], np.float32)
    e1 = np.array([2, 11], np.float32)
    norms = rng.uniform(0, 10, (3, 2)).astype(np.float32)
    out, z = oracle.scale_tile_norms(norms, [e0, e1], 0.0)
    assert z == 0
    for i in range(3):
        for j in range(2):
            x = np.float32(1) / e0[i]
            y = np.float32(1) / e1[j]
            assert out[i, j] == np.float32(norms[i, j] * np.float32(x * y))
            assert abs(out[i, j] - norms[i, j] / (e0[i] * e1[j])) <= 1e-6 * out[i, j]


def test_scale_rank1_divides():
    e0 = np.array([3, 7], np.float32)
    norms = np.array([1.0, 2.0], np.float32)
    out, _ = oracle.scale_tile_norms(norms, [e0], 0.0)
    assert out[0] == np.float32(np.float32(1.0) / np.float32(3))
    assert out[1] == np.float32(np.float32(2.0) / np.float32(7))


@pytest.mark.parametrize("rank", [3, 4, 5])
def test_scale_higher_rank_equals_inverse_volume(rank):
    rng = np.random.default_rng(rank)
    exts = [rng.integers(1, 9, rng.integers(1, 4)).astype(np.float32) for _ in range(rank)]
    shape = tuple(len(e) for e in exts)
    norms = rng.uniform(0.5, 3, shape).astype(np.float32)
    out, z = oracle.scale_tile_norms(norms, exts, 0.0)
    vol = np.ones(shape)
    for d, e in enumerate(exts):
        vol = vol * e.reshape([-1 if q == d else 1 for q in range(rank)])
    np.testing.assert_allclose(out, norms / vol, rtol=1e-6)


def test_scale_screening_counts_zeros():
    e = [np.array([1, 1], np.float32), np.array([1, 1], np.float32)]
    norms = np.array([[1e-9, 1.0], [0.5, 1e-3]], np.float32)
    out, z = oracle.scale_tile_norms(norms, e, 1e-2)
    assert z == 2
    assert out[0, 0] == 0 and out[1, 1] == 0 and out[0, 1] == 1.0


# ------------------------------------------------------------ shape gemm --

def test_shape_gemm_closed_form_notrans():
    # result == |alpha| * sum_k ||A_ik|| ||B_kj|| / (M_i N_j)
    # (tests/sparse_shape.cpp:1478-1541 pattern, tolerance as there)
    rng = np.random.default_rng(5)
    em, ek, en = [rng.integers(2, 30, n).astype(np.float32) for n in (4, 5, 3)]
    rawA = rng.uniform(0, 5, (4, 5)).astype(np.float32)
    rawB = rng.uniform(0, 5, (5, 3)).astype(np.float32)
    sA, _ = oracle.scale_tile_norms(rawA, [em, ek], 0.0)
    sB, _ = oracle.scale_tile_norms(rawB, [ek, en], 0.0)
    est, z = oracle.shape_gemm(sA, sB, 0, 0, 4, 3, 5, oracle.k_sizes([ek]), 7.2, 0.0)
    expect = 7.2 * (rawA @ rawB.astype(np.float64)) / np.outer(em, en)
    np.testing.assert_allclose(est, expect, rtol=1e-5)
    assert z == 0


def test_shape_gemm_threshold_and_outer_product():
    sA = np.array([[1.0, 0.0]], np.float32)
    sB = np.array([[0.0], [1.0]], np.float32)
    est, z = oracle.shape_gemm(sA, sB, 0, 0, 1, 1, 2, np.array([1, 1], np.float32), 1.0, 0.5)
    assert est[0, 0] == 0 and z == 1
    est, z = oracle.shape_gemm(np.array([2.0], np.float32), np.array([3.0, 0.1], np.float32),
                               0, 0, 1, 2, 0, np.zeros(0, np.float32), 2.0, 1.0)
    assert est[0, 0] == np.float32(12.0) and est[0, 1] == 0 and z == 1


def test_shape_gemm_trans_uniform_k_matches_notrans():
    # with equal k extents the contiguous-run scaling of the reference
    # (sparse_shape.h:1632-1647) is layout independent
    rng = np.random.default_rng(6)
    L = rng.uniform(0, 1, (6, 4)).astype(np.float32)  # M x K
    R = rng.uniform(0, 1, (4, 5)).astype(np.float32)  # K x N
    ks = np.full(4, 16, np.float32)
    a, _ = oracle.shape_gemm(L, R, 0, 0, 6, 5, 4, ks, 1.0, 0.0)
    b, _ = oracle.shape_gemm(L.T.copy(), R, 1, 0, 6, 5, 4, ks, 1.0, 0.0)
    c, _ = oracle.shape_gemm(L, R.T.copy(), 0, 1, 6, 5, 4, ks, 1.0, 0.0)
    np.testing.assert_allclose(a, b, rtol=1e-6)
    np.testing.assert_allclose(a, c, rtol=1e-6)


def test_shape_gemm_trans_uneven_k_follows_storage_runs():
    # hand-computed: left stored K x M = [[1, 2], [3, 4]] (K=2, M=2), Trans;
    # the reference multiplies storage element s by k_sizes[s % K]:
    # [[1*ka, 2*kb], [3*ka, 4*kb]], then the Trans-aware GEMM.
    ka, kb = 2.0, 5.0
    L = np.array([[1, 2], [3, 4]], np.float32)
    R = np.array([[1], [1]], np.float32)  # K x N, NoTrans: rows scaled by k
    est, _ = oracle.shape_gemm(L, R, 1, 0, 2, 1, 2, np.array([ka, kb], np.float32), 1.0, 0.0)
    Ls = np.array([[1 * ka, 2 * kb], [3 * ka, 4 * kb]])  # K x M storage
    Rs = np.array([[1 * ka], [1 * kb]])
    np.testing.assert_allclose(est, Ls.T @ Rs, rtol=1e-6)


# ------------------------------------------------------------------ gemm --

@pytest.mark.parametrize("op_a,op_b", list(itertools.product([0, 1], [0, 1])))
def test_gemm_brute_force_all_ops(op_a, op_b):
    rng = np.random.default_rng(10 + 2 * op_a + op_b)
    m, n, k = 5, 4, 6
    A = rng.integers(-9, 10, (k, m) if op_a else (m, k)).astype(np.float64)
    B = rng.integers(-9, 10, (n, k) if op_b else (k, n)).astype(np.float64)
    C0 = rng.integers(-9, 10, (m, n)).astype(np.float64)
    np.testing.assert_array_equal(oracle.gemm(op_a, op_b, 3.0, A, B), brute_gemm(op_a, op_b, 3.0, A, B))
    np.testing.assert_array_equal(oracle.gemm(op_a, op_b, 3.0, A, B, 1.0, C0),
                                  brute_gemm(op_a, op_b, 3.0, A, B, 1.0, C0))
    Af, Bf = A.astype(np.float32), B.astype(np.float32)
    np.testing.assert_array_equal(oracle.gemm(op_a, op_b, 3.0, Af, Bf),
                                  brute_gemm(op_a, op_b, 3.0, A, B).astype(np.float32))


# --------------------------------------------------------------- permute --

@pytest.mark.parametrize("rank", [2, 3, 4, 5, 6])
def test_permute_to_convention_and_round_trip(rank):
    rng = np.random.default_rng(rank)
    shape = tuple(int(x) for x in rng.integers(1, 6, rank))
    x = rng.uniform(-1, 1, shape)
    for _ in range(4):
        perm = rng.permutation(rank)
        y = oracle.permute(x, perm)
        for d in range(rank):
            assert y.shape[perm[d]] == x.shape[d]
        idx = tuple(int(rng.integers(0, s)) for s in shape)
        tidx = [0] * rank
        for d in range(rank):
            tidx[perm[d]] = idx[d]
        assert y[tuple(tidx)] == x[idx]
        inv = np.argsort(perm)
        np.testing.assert_array_equal(oracle.permute(y, inv), x)


def test_permute_float32():
    x = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    np.testing.assert_array_equal(oracle.permute(x, [2, 0, 1]), np.transpose(x, [1, 2, 0]))


# -------------------------------------------------------------- contract --

def _blocksparse(rng, ext_r, ext_c, mask, integer=True, dtype=np.float64):
    offs = np.full(mask.size, -1, np.int64)
    vols = np.multiply.outer(ext_r, ext_c).reshape(-1)
    pos = 0
    for t in range(mask.size):
        if mask.reshape(-1)[t]:
            offs[t] = pos
            pos += vols[t]
    pool = (rng.integers(-20, 21, pos) if integer else rng.uniform(-1, 1, pos)).astype(dtype)
    return pool, offs


def _dense_of(pool, offs, ext_r, ext_c, trans=False):
    R, C = (ext_c, ext_r) if trans else (ext_r, ext_c)
    br, bc = np.concatenate([[0], np.cumsum(R)]), np.concatenate([[0], np.cumsum(C)])
    out = np.zeros((br[-1], bc[-1]))
    for i in range(len(R)):
        for j in range(len(C)):
            o = offs[i * len(C) + j]
            if o >= 0:
                out[br[i]:br[i + 1], bc[j]:bc[j + 1]] = pool[o:o + R[i] * C[j]].reshape(R[i], C[j])
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("op_a,op_b", list(itertools.product([0, 1], [0, 1])))
def test_contract_integer_exact_vs_dense_product(op_a, op_b, dtype):
    # integer-valued inputs make every partial sum exact, so the block-sparse
    # one-rank Summa must equal the dense product exactly on the non-zero
    # result tiles (tests/dist_eval_contraction_eval.cpp:375-472 pattern);
    # FP32 pools run the float restatement (ora_contract_f32), whose sums of
    # at most 40 products of integers below 2^6 are exact in float too
    rng = np.random.default_rng(20 + op_a * 2 + op_b)
    em, ek, en = rng.integers(1, 9, 4), rng.integers(1, 9, 5), rng.integers(1, 9, 3)
    am = rng.random((4, 5)) < 0.6
    bm = rng.random((5, 3)) < 0.6
    if op_a:
        A, aoff = _blocksparse(rng, ek, em, am.T)
        Ad = _dense_of(A, aoff, ek, em).T
    else:
        A, aoff = _blocksparse(rng, em, ek, am)
        Ad = _dense_of(A, aoff, em, ek)
    if op_b:
        B, boff = _blocksparse(rng, en, ek, bm.T)
        Bd = _dense_of(B, boff, en, ek).T
    else:
        B, boff = _blocksparse(rng, ek, en, bm)
        Bd = _dense_of(B, boff, ek, en)
    cmask = rng.random((4, 3)) < 0.8
    coff = np.full(12, -1, np.int64)
    pos = 0
    for t in range(12):
        if cmask.reshape(-1)[t]:
            coff[t] = pos
            pos += em[t // 3] * en[t % 3]
    A, B = A.astype(dtype), B.astype(dtype)
    C, npairs = oracle.contract(op_a, op_b, em, en, ek, A, aoff, B, boff, coff, pos, alpha=2.0)
    assert C.dtype == dtype
    ref = 2.0 * Ad @ Bd
    Cd = _dense_of(C, coff, em, en)
    bm_r, bn_r = np.concatenate([[0], np.cumsum(em)]), np.concatenate([[0], np.cumsum(en)])
    expect_pairs = 0
    for i in range(4):
        for j in range(3):
            blk = ref[bm_r[i]:bm_r[i + 1], bn_r[j]:bn_r[j + 1]]
            got = Cd[bm_r[i]:bm_r[i + 1], bn_r[j]:bn_r[j + 1]]
            if cmask[i, j]:
                np.testing.assert_array_equal(got, blk)
                expect_pairs += int(np.sum(am[i, :] & bm[:, j]))
            else:
                assert not got.any()
    assert npairs == expect_pairs


def test_contract_pair_filter_skips_small_products():
    rng = np.random.default_rng(30)
    em, ek, en = np.array([2]), np.array([2, 2]), np.array([2])
    A, aoff = _blocksparse(rng, em, ek, np.ones((1, 2), bool))
    B, boff = _blocksparse(rng, ek, en, np.ones((2, 1), bool))
    an = np.array([1.0, 1e-6], np.float32)
    bn = np.array([1.0, 1.0], np.float32)
    coff = np.array([0], np.int64)
    C, npairs = oracle.contract(0, 0, em, en, ek, A, aoff, B, boff, coff, 4, a_norms=an,
                                b_norms=bn, pair_filter_threshold=1e-3)
    assert npairs == 1  # 1e-6*1 < 1e-3/2: the second pair is dropped
    expect = A[:4].reshape(2, 2) @ B[:4].reshape(2, 2)
    np.testing.assert_array_equal(C.reshape(2, 2), expect)


# --------------------------------------------------------------- permopt --

@pytest.mark.parametrize("left,right,expect", [
    ("mk", "kn", (list("mk"), list("kn"), list("mn"), permopt.IDENTITY, permopt.IDENTITY)),
    ("km", "kn", (list("mk"), list("kn"), list("mn"), permopt.MATRIX_TRANSPOSE, permopt.IDENTITY)),
    ("mk", "nk", (list("mk"), list("kn"), list("mn"), permopt.IDENTITY, permopt.MATRIX_TRANSPOSE)),
    ("cdij", "abcd", (list("ijcd"), list("cdab"), list("ijab"), permopt.MATRIX_TRANSPOSE,
                      permopt.MATRIX_TRANSPOSE)),
    ("cdij", "acbd", (list("ijcd"), list("cdab"), list("ijab"), permopt.MATRIX_TRANSPOSE,
                      permopt.GENERAL)),
    ("ikj", "kl", (list("ijk"), list("kl"), list("ijl"), permopt.GENERAL, permopt.IDENTITY)),
])
def test_permopt_hand_derived(left, right, expect):
    # derived by hand from permopt.h:254-381; the argument index lists come
    # back in logical GEMM order (outer, inner) / (inner, outer) whatever the
    # storage order, the permtypes say how storage maps onto them
    assert permopt.compute_index_list_contraction(list(left), list(right)) == expect


def test_permopt_inner_order_mismatch_respects_preference():
    # left "i,k,l" vs right "l,k,j": inner orders differ and the ranks are
    # equal, so prefer_to_permute_left decides which side is permuted
    l, r, res, lt, rt = permopt.compute_index_list_contraction(list("ikl"), list("lkj"), True)
    assert (l, r, res, lt, rt) == (list("ilk"), list("lkj"), list("ij"), permopt.GENERAL,
                                   permopt.IDENTITY)
    l, r, res, lt, rt = permopt.compute_index_list_contraction(list("ikl"), list("lkj"), False)
    assert (l, r, res, lt, rt) == (list("ikl"), list("klj"), list("ij"), permopt.IDENTITY,
                                   permopt.GENERAL)
    # the lower-rank argument is permuted whatever the preference
    l, r, res, lt, rt = permopt.compute_index_list_contraction(list("ikl"), list("lk"), True)
    assert lt == permopt.IDENTITY and rt == permopt.GENERAL
    l, r, res, lt, rt = permopt.compute_index_list_contraction(list("kl"), list("lkj"), False)
    assert lt == permopt.GENERAL and rt == permopt.IDENTITY


def test_tot_tile_norms_closed_forms_and_cell_order():
    # outer-tile norms of a tensor of tensors (Tensor<Tensor>::squared_norm,
    # tensor.h:3768-3776): all-ones cells give sqrt(total elements), empty
    # cells are skipped, a one-cell tile equals the plain tile norm
    pool = np.ones(100)
    tc = np.array([0, 3, 5, 5])
    off = np.array([0, -1, 10, 40, 70])
    vol = np.array([10, 99, 30, 30, 30])
    got = oracle.tot_tile_norms(pool, tc, off, vol)
    np.testing.assert_allclose(got, np.float32([np.sqrt(40), np.sqrt(60), 0.0]), rtol=1e-7)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, 77)
    assert oracle.tot_tile_norms(x, [0, 1], [0], [77])[0] == oracle.tile_norms(x, [0], [77])[0]
    # each cell is summed on its own and then added: in float, a 4096 cell
    # followed by a cell of three ones gives 2^24 + 3 (nested), not 2^24
    # (element by element, each 1 lost below the 2^24 spacing)
    pool32 = np.array([4096.0, 1.0, 1.0, 1.0], np.float32)
    n = oracle.tot_tile_norms(pool32, [0, 2], [0, 1], [1, 3])[0]
    nested = np.float32(np.float32(4096.0) * np.float32(4096.0)) + np.float32(3.0)
    assert n == np.sqrt(nested, dtype=np.float32) and n != np.float32(4096.0)
