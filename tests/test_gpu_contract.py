"""GPU parity of the contraction path against the CPU oracle.

Bars (stated per test): bit-exact for integer-valued data, norms, shape
estimates and sparsity patterns; relative Frobenius error <= 1e-13 for the
FP64 MFMA path on random data; <= 1e-5 for FP32 data.
"""
import itertools

import numpy as np
import pytest
import torch

import oracle
from tiledarray_amd import DistArray, TiledRange, TiledRange1
from tiledarray_amd import engine as E
from tiledarray_amd import workloads as W

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-13


def rel_fro(x, ref):
    d = np.linalg.norm((np.asarray(x, np.float64) - ref).ravel())
    r = np.linalg.norm(np.asarray(ref, np.float64).ravel())
    return d / r if r > 0 else d


def host_pool(arr):
    return arr.pool.detach().cpu().numpy()


def matrix_ops_arrays(rng, trm, trk, trn, op_a, op_b, dtype, integer, fill=1.0):
    """A (m x k, stored k x m if op_a) and B (k x n, stored n x k if op_b)."""
    ga = (trk, trm) if op_a else (trm, trk)
    gb = (trn, trk) if op_b else (trk, trn)
    ma = rng.random((ga[0].ntiles, ga[1].ntiles)) < fill
    mb = rng.random((gb[0].ntiles, gb[1].ntiles)) < fill
    A = DistArray.allocate(TiledRange(list(ga)), ma, dtype=dtype)
    B = DistArray.allocate(TiledRange(list(gb)), mb, dtype=dtype)
    for X in (A, B):
        n = X.pool.numel()
        h = rng.integers(-16, 17, n) if integer else rng.uniform(-1, 1, n)
        X.pool.copy_(torch.as_tensor(h, dtype=dtype))
    return A, B


def oracle_contract(c, A, B, a_present=None, b_present=None):
    """Runs the oracle on the same fused structure the product used and
    returns its C pool laid out at the product's result offsets."""
    a_off = A.offsets.copy()
    b_off = B.offsets.copy()
    if a_present is not None:
        a_off[~a_present] = -1
    if b_present is not None:
        b_off[~b_present] = -1
    c_off = c.c_off.cpu().numpy()
    alpha = c.alpha
    Cref, npairs = oracle.contract(c.op_a, c.op_b, c.mext_h, c.next_h, c.kext_h, host_pool(A),
                                   a_off, host_pool(B), b_off, c_off, c.c_capacity, alpha=alpha)
    return Cref, npairs, c_off


@pytest.mark.parametrize("op_a,op_b", list(itertools.product([0, 1], [0, 1])))
@pytest.mark.parametrize("tile", [(64, 64, 16), (64, 64, 32), (128, 128, 16), (128, 64, 16),
                                  (64, 128, 16), (32, 32, 16), (32, 32, 32), (128, 64, 32),
                                  (64, 128, 32)])
def test_grouped_gemm_integer_exact_all_ops(ctx, op_a, op_b, tile):
    rng = np.random.default_rng(100 + op_a * 2 + op_b)
    trm = TiledRange1([0, 40, 130, 200, 333])
    trk = TiledRange1([0, 17, 64, 150, 151, 300])
    trn = TiledRange1([0, 96, 97, 250])
    A, B = matrix_ops_arrays(rng, trm, trk, trn, op_a, op_b, torch.float64, True, fill=0.8)
    la = "k,m" if op_a else "m,k"
    lb = "n,k" if op_b else "k,n"
    c = E.Contraction(ctx, "m,n", A, la, B, lb, alpha=3.0, tile=tile)
    assert (c.op_a, c.op_b) == (op_a, op_b)
    c.run()
    torch.cuda.synchronize()
    Cref, npairs, c_off = oracle_contract(c, A, B)
    got = c.C.cpu().numpy()
    np.testing.assert_array_equal(got[:Cref.size][c_off_mask(c_off, c)], Cref[c_off_mask(c_off, c)])
    assert c.counters()["npairs"] == npairs


def c_off_mask(c_off, c):
    m = np.zeros(c.c_capacity, bool)
    for t, o in enumerate(c_off):
        if o >= 0:
            m[o:o + c.mext_h[t // c.Nt] * c.next_h[t % c.Nt]] = True
    return m


def test_dense_uniform_random_fp64_tolerance(ctx):
    # the dense 128-tile matrix product (as the 2048^3 workload, scaled down)
    A, B, _ = W.dense_gemm(n=512, tile=128)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n")
    c.run()
    res = c.result()
    ref = A.to_dense() @ B.to_dense()
    err = rel_fro(res.to_dense(), ref)
    assert err <= FP64_TOL, err
    cnt = c.counters()
    assert cnt["npairs"] == 4 ** 3 and cnt["ntasks"] == 16
    assert cnt["flops"] == 2 * 512 ** 3


def test_beta_accumulate_and_empty_worklist(ctx):
    # a task with beta = 1 adds to the C tile; a list with zero tasks is a no-op
    import ctypes
    from tiledarray_amd import _lib
    hip = ctypes.CDLL("libamdhip64.so")
    m, n, k = 48, 40, 24
    A = torch.randint(-5, 6, (m * k,), dtype=torch.float64, device="cuda")
    B = torch.randint(-5, 6, (k * n,), dtype=torch.float64, device="cuda")
    C0 = torch.randint(-5, 6, (m * n,), dtype=torch.float64, device="cuda")
    C = C0.clone()
    wl = E.WorkList(ctx, 4, 4, 4)
    tasks, pairs, counters = wl.buffers()
    task = _lib.GemmTask(0, m, n, n, 0, 2, _lib.TAMX_TASK_BETA1)
    pr = (_lib.GemmPair * 2)(_lib.GemmPair(0, 0, k, k, n, 0, 0, 0),
                             _lib.GemmPair(0, 0, k, k, n, 0, 0, 0))
    cnt = (ctypes.c_int64 * 8)(1, 2, m * n, 0, 0, 0, 0, 0)
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert hip.hipMemcpy(tasks, ctypes.byref(task), ctypes.sizeof(task), 1) == 0
    assert hip.hipMemcpy(pairs, pr, ctypes.sizeof(pr), 1) == 0
    assert hip.hipMemcpy(counters, cnt, ctypes.sizeof(cnt), 1) == 0
    E.call("tamx_grouped_gemm_f64", ctx.ptr, wl.ptr, 0, 0, ctypes.c_double(0.5), A.data_ptr(),
           B.data_ptr(), C.data_ptr(), ctx.stream())
    torch.cuda.synchronize()
    ref = C0.view(m, n) + 0.5 * 2 * (A.view(m, k) @ B.view(k, n))
    assert torch.equal(C.view(m, n), ref)
    # zero tasks: C unchanged
    cnt0 = (ctypes.c_int64 * 8)(0, 0, 0, 0, 0, 0, 0, 0)
    assert hip.hipMemcpy(counters, cnt0, ctypes.sizeof(cnt0), 1) == 0
    before = C.clone()
    E.call("tamx_grouped_gemm_f64", ctx.ptr, wl.ptr, 0, 0, ctypes.c_double(1.0), A.data_ptr(),
           B.data_ptr(), C.data_ptr(), ctx.stream())
    torch.cuda.synchronize()
    assert torch.equal(C, before)


def test_blocksparse_screening_bit_exact_and_values(ctx):
    # the banded block-sparse workload at a reduced size: norms, scaled
    # shapes, the result estimate and the result pattern must equal the
    # oracle's bit for bit; values within the FP64 tolerance
    thr = 1e-8
    A, B, info = W.blocksparse_gemm(n=3000, wlo=24, whi=96, seed=11)
    ra = E.tile_norms(ctx, A).cpu().numpy()
    rb = E.tile_norms(ctx, B).cpu().numpy()
    vols = A.trange.tile_volumes()
    oa = oracle.tile_norms(host_pool(A), A.offsets, vols)
    ob = oracle.tile_norms(host_pool(B), B.offsets, vols)
    np.testing.assert_array_equal(ra, oa)
    np.testing.assert_array_equal(rb, ob)
    E.make_sparse(ctx, A, thr)
    E.make_sparse(ctx, B, thr)
    ext = [r.extents for r in A.trange.ranges]
    sa, za = oracle.scale_tile_norms(oa.reshape(A.trange.tiles_shape), ext, thr)
    sb, zb = oracle.scale_tile_norms(ob.reshape(B.trange.tiles_shape), ext, thr)
    np.testing.assert_array_equal(A.shape_norms.cpu().numpy(), sa.reshape(-1))
    np.testing.assert_array_equal(B.shape_norms.cpu().numpy(), sb.reshape(-1))
    # the scaling on its own, with the count of screened-out tiles
    for o, s_ref, z_ref in ((oa, sa, za), (ob, sb, zb)):
        n = torch.as_tensor(o, device="cuda").clone()
        zc = torch.zeros(1, dtype=torch.int64, device="cuda")
        E.scale_norms(ctx, n, A.trange, thr, zero_count=zc)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(n.cpu().numpy(), s_ref.reshape(-1))
        assert 0 < z_ref < n.numel() and int(zc.item()) == z_ref
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=thr)
    c.run()
    torch.cuda.synchronize()
    nt = A.trange.tiles_shape[0]
    est, zc = oracle.shape_gemm(sa, sb, 0, 0, nt, nt, nt, oracle.k_sizes([ext[0]]), 1.0, thr)
    np.testing.assert_array_equal(c.c_norms.cpu().numpy(), est.reshape(-1))
    c_off = c.c_off.cpu().numpy()
    np.testing.assert_array_equal(c_off >= 0, est.reshape(-1) >= thr)
    a_present = (A.offsets >= 0) & (sa.reshape(-1) >= thr)
    b_present = (B.offsets >= 0) & (sb.reshape(-1) >= thr)
    Cref, npairs, _ = oracle_contract(c, A, B, a_present, b_present)
    assert c.counters()["npairs"] == npairs
    mask = c_off_mask(c_off, c)
    err = rel_fro(c.C.cpu().numpy()[mask], Cref[mask])
    assert err <= FP64_TOL, err


@pytest.mark.parametrize("op_a,op_b", list(itertools.product([0, 1], [0, 1])))
@pytest.mark.parametrize("M,N,K", [(13, 5, 9), (4, 11, 3), (1, 1, 1), (7, 6, 0)])
def test_shape_gemm_direct_bit_exact(ctx, op_a, op_b, M, N, K):
    # the shape estimate through the C ABI, two problems in one batch, against
    # the oracle: estimate bit for bit and the count of screened-out tiles
    import ctypes
    rng = np.random.default_rng(1000 + 17 * M + 5 * N + K + 2 * op_a + op_b)
    nb = 2
    kk = max(K, 1)
    left = (rng.random((nb, M * kk)) * 2.0 ** rng.integers(-3, 2, (nb, M * kk))).astype(np.float32)
    right = (rng.random((nb, kk * N)) * 2.0 ** rng.integers(-3, 2, (nb, kk * N))).astype(np.float32)
    ksz = rng.integers(1, 300, (nb, kk)).astype(np.float32)
    # a threshold that screens out about half of the first problem's tiles
    est0, _ = oracle.shape_gemm(left[0], right[0], op_a, op_b, M, N, K, ksz[0], 1.5, 0.0)
    thr = np.float32(np.median(est0))
    dl, dr, dk = (torch.as_tensor(x.reshape(-1), device="cuda") for x in (left, right, ksz))
    out = torch.empty(nb * M * N, dtype=torch.float32, device="cuda")
    zc = torch.zeros(1, dtype=torch.int64, device="cuda")
    E.call("tamx_shape_gemm_f32", ctx.ptr, dl.data_ptr(), dr.data_ptr(), op_a, op_b, M, N, K, nb,
           dk.data_ptr(), ctypes.c_float(1.5), ctypes.c_float(thr), out.data_ptr(), zc.data_ptr(),
           ctx.stream())
    torch.cuda.synchronize()
    zsum = 0
    for b in range(nb):
        est, z = oracle.shape_gemm(left[b], right[b], op_a, op_b, M, N, K, ksz[b], 1.5, thr)
        np.testing.assert_array_equal(out.cpu().numpy()[b * M * N:(b + 1) * M * N], est.reshape(-1))
        zsum += z
    assert 0 < zsum < nb * M * N or M * N == 1, zsum
    assert int(zc.item()) == zsum


def test_pair_filter_matches_oracle(ctx):
    # the per-pair rule (|A_ik| |B_kj| < threshold / Kt drops the pair, the
    # rule TA compiles out) against the oracle: the same pairs, values within
    # the FP64 tolerance; the threshold drops about half of the pairs
    thr = 1e-8
    A, B, _ = W.blocksparse_gemm(n=2000, wlo=30, whi=90, seed=14, integer=True)
    E.make_sparse(ctx, A, thr)
    E.make_sparse(ctx, B, thr)
    sa, sb = A.shape_norms.cpu().numpy(), B.shape_norms.cpu().numpy()
    a_present = (A.offsets >= 0) & (sa >= thr)
    b_present = (B.offsets >= 0) & (sb >= thr)
    nt = A.trange.tiles_shape[0]
    pa, pb = a_present.reshape(nt, nt), b_present.reshape(nt, nt)
    na, nb = sa.reshape(nt, nt), sb.reshape(nt, nt)
    prods = [na[i, k] * nb[k, j] for i in range(nt) for k in range(nt) for j in range(nt)
             if pa[i, k] and pb[k, j]]
    pthr = float(np.float32(np.median(prods)) * np.float32(nt))
    c0 = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=thr)
    c0.run()
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=thr, pair_threshold=pthr)
    c.run()
    torch.cuda.synchronize()
    a_off, b_off = A.offsets.copy(), B.offsets.copy()
    a_off[~a_present] = -1
    b_off[~b_present] = -1
    c_off = c.c_off.cpu().numpy()
    Cref, npairs = oracle.contract(c.op_a, c.op_b, c.mext_h, c.next_h, c.kext_h, host_pool(A),
                                   a_off, host_pool(B), b_off, c_off, c.c_capacity,
                                   alpha=c.alpha, a_norms=sa, b_norms=sb,
                                   pair_filter_threshold=pthr)
    n0 = c0.counters()["npairs"]
    assert c.counters()["npairs"] == npairs and 0 < npairs < n0
    mask = c_off_mask(c_off, c)
    np.testing.assert_array_equal(c.C.cpu().numpy()[mask], Cref[mask])


def test_blocksparse_integer_bit_exact(ctx):
    A, B, _ = W.blocksparse_gemm(n=2000, wlo=30, whi=90, seed=12, integer=True)
    E.make_sparse(ctx, A, 1e-8)
    E.make_sparse(ctx, B, 1e-8)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=1e-8)
    c.run()
    torch.cuda.synchronize()
    sa = A.shape_norms.cpu().numpy()
    sb = B.shape_norms.cpu().numpy()
    Cref, npairs, c_off = oracle_contract(c, A, B, (A.offsets >= 0) & (sa >= A.shape_threshold),
                                          (B.offsets >= 0) & (sb >= B.shape_threshold))
    mask = c_off_mask(c_off, c)
    np.testing.assert_array_equal(c.C.cpu().numpy()[mask], Cref[mask])


@pytest.mark.parametrize("rank", [2, 3, 4, 5, 6])
def test_permute_bit_exact_vs_oracle(ctx, rank):
    rng = np.random.default_rng(200 + rank)
    ranges = [TiledRange1(np.concatenate([[0], np.cumsum(rng.integers(1, 7 if rank > 3 else 40,
                                                                     rng.integers(1, 4)))]))
              for _ in range(rank)]
    tr = TiledRange(ranges)
    X = DistArray.allocate(tr, rng.random(tr.ntiles) < 0.8)
    X.pool.copy_(torch.as_tensor(rng.uniform(-1, 1, X.pool.numel())))
    for _ in range(3):
        perm = [int(p) for p in rng.permutation(rank)]
        Y = E.permute_array(ctx, X, perm)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(Y.to_dense(), oracle.permute(X.to_dense(), perm))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_permute_large_and_mixed_planes(ctx, dtype):
    # planes above 96 x 96 elements move in blocks (one side or both split
    # evenly), smaller ones whole; one plan holds both kinds, with ragged
    # edges on every tile
    rng = np.random.default_rng(77)
    tr = TiledRange([TiledRange1([0, 130, 170, 301]), TiledRange1([0, 97, 100, 260]),
                     TiledRange1([0, 3, 5])])
    X = DistArray.allocate(tr, rng.random(tr.ntiles) < 0.9, dtype=dtype)
    X.pool.copy_(torch.as_tensor(rng.uniform(-1, 1, X.pool.numel()), dtype=dtype))
    for perm in ([1, 0, 2], [2, 1, 0], [0, 2, 1], [1, 2, 0]):
        Y = E.permute_array(ctx, X, perm)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(Y.to_dense(), oracle.permute(X.to_dense(), perm))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("bt, dt", [([0, 13, 24], [0, 11, 22]), ([0, 45, 90], [0, 19, 40]),
                                    ([0, 3, 150], [0, 7, 9])],
                         ids=["short_rows", "long_rows", "skewed"])
def test_permute_joined_rows(ctx, dtype, bt, dt):
    # (a,c,b,d) -> (c,d,a,b), the ladder's permutation: consecutive a are
    # consecutive items of one workgroup (tens of thousands of planes, so a
    # workgroup takes several), and the cache line where the rows of a and
    # a + 1 meet is written by both items together; odd row lengths put
    # that line anywhere, and a row can lie inside one line
    rng = np.random.default_rng(91)
    tr = TiledRange([TiledRange1([0, 37, 100]), TiledRange1([0, 55, 100]),
                     TiledRange1(bt), TiledRange1(dt)])
    X = DistArray.allocate(tr, rng.random(tr.ntiles) < 0.85, dtype=dtype)
    X.pool.copy_(torch.as_tensor(rng.uniform(-1, 1, X.pool.numel()), dtype=dtype))
    Y = E.permute_array(ctx, X, [2, 0, 3, 1])
    torch.cuda.synchronize()
    np.testing.assert_array_equal(Y.to_dense(), oracle.permute(X.to_dense(), [2, 0, 3, 1]))
    # a second execution into the same result leaves it unchanged
    first = Y.pool.clone()
    Y._perm_plan.run(X.pool, Y.pool)
    torch.cuda.synchronize()
    assert torch.equal(Y.pool, first)


def test_ladder_permutes_and_result_transpose(ctx):
    # R(a,b,i,j) = T(c,d,i,j) * V(a,c,b,d): V is permuted (general), T folds
    # into Trans, and the result permutation is fused into the GEMM store
    T, V, info = W.ladder(o=8, v=40, otiles=2, wlo=9, whi=17, seed=3)
    c = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d")
    assert c.op_a == 1 and c.op_b == 0 and c.store_trans
    c.run()
    R = c.result()
    ref = np.einsum("cdij,acbd->abij", T.to_dense(), V.to_dense())
    err = rel_fro(R.to_dense(), ref)
    assert err <= FP64_TOL, err


@pytest.mark.parametrize("target", ["i,j,a,b", "a,i,b,j", "j,b,a,i"])
def test_ladder_result_permutations(ctx, target):
    T, V, _ = W.ladder(o=6, v=24, otiles=2, wlo=7, whi=12, seed=4, v_order="abcd")
    c = E.Contraction(ctx, target, T, "c,d,i,j", V, "a,b,c,d")
    c.run()
    R = c.result()
    ref = np.einsum("cdij,abcd->" + target.replace(",", ""), T.to_dense(), V.to_dense())
    assert rel_fro(R.to_dense(), ref) <= FP64_TOL


def test_fp32_grouped_gemm_tolerance(ctx):
    rng = np.random.default_rng(7)
    trm = TiledRange1([0, 16, 60, 96])
    trk = TiledRange1([0, 33, 80])
    trn = TiledRange1([0, 50, 51, 96])
    A, B = matrix_ops_arrays(rng, trm, trk, trn, 0, 0, torch.float32, False)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", precision="f32")
    c.run()
    R = c.result()
    ref = A.to_dense().astype(np.float64) @ B.to_dense().astype(np.float64)
    assert rel_fro(R.to_dense(), ref) <= 1e-5


def test_incongruent_tilings_raise(ctx):
    A = DistArray.allocate(TiledRange([[0, 4, 8], [0, 3, 8]]))
    B = DistArray.allocate(TiledRange([[0, 4, 8], [0, 8]]))
    with pytest.raises(ValueError):
        E.Contraction(ctx, "m,n", A, "m,k", B, "k,n")


@pytest.mark.parametrize("short", ["tasks", "pairs", "c_elements"])
def test_worklist_capacity_overflow(ctx, short):
    # A work list (or result pool) too small for the schedule: the builder
    # marks TAMX_ERR_CAPACITY on the device, hands the GEMM an empty list and
    # every result tile offset -1, so C is left as it was and every kernel
    # exits normally (TA throws on the same resource limits,
    # contraction_eval.h:1855-1898); the host reads the mark as an error.
    from tiledarray_amd import _lib
    rng = np.random.default_rng(21)
    trm = TiledRange1([0, 40, 130, 200])
    trk = TiledRange1([0, 17, 64, 150])
    trn = TiledRange1([0, 96, 97, 250])
    A, B = matrix_ops_arrays(rng, trm, trk, trn, 0, 0, torch.float64, True)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n")
    grid = c.Mt * c.Nt
    if short == "tasks":
        c.wl = E.WorkList(ctx, grid - 1, grid * c.Kt, grid)
    elif short == "pairs":
        c.wl = E.WorkList(ctx, grid, grid * c.Kt - 1, grid)
    else:
        c._desc.c_capacity = c.c_capacity - 1
    c.C.fill_(7.0)
    c.run()
    torch.cuda.synchronize()
    cnt = c.wl.counters()
    assert cnt[_lib.WL_STATUS] == _lib.TAMX_ERR_CAPACITY
    assert cnt[_lib.WL_NTASKS] == 0 and cnt[_lib.WL_NPAIRS] == 0
    # the totals that were asked for
    assert cnt[5] == grid and cnt[6] == grid * c.Kt
    assert bool((c.C == 7.0).all())
    assert bool((c.c_off == -1).all())
    with pytest.raises(_lib.TamxError) as ei:
        c.counters()
    assert ei.value.status == _lib.TAMX_ERR_CAPACITY
    # the same context and kernels then run a list that fits
    c.wl = E.WorkList(ctx, grid, grid * c.Kt, grid)
    c._desc.c_capacity = c.c_capacity
    c.run()
    torch.cuda.synchronize()
    Cref, npairs, c_off = oracle_contract(c, A, B)
    mask = c_off_mask(c_off, c)
    np.testing.assert_array_equal(c.C.cpu().numpy()[mask], Cref[mask])
    assert c.counters()["npairs"] == npairs


def test_grid_capacity_and_out_of_memory_status(ctx):
    # a result grid beyond the work list's grid capacity is refused on the
    # host before any launch; a work list beyond device memory returns
    # TAMX_ERR_OOM with a message, and allocations work again after it
    import ctypes
    from tiledarray_amd import _lib
    lib = E.lib()
    small = E.WorkList(ctx, 1, 1, 1)
    rng = np.random.default_rng(3)
    trm = TiledRange1([0, 8, 16])
    A, B = matrix_ops_arrays(rng, trm, trm, trm, 0, 0, torch.float64, True)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n")
    c.shape_step()
    rc = lib.tamx_contraction_worklist(ctx.ptr, ctypes.byref(c._desc), small.ptr, ctx.stream())
    assert rc == _lib.TAMX_ERR_CAPACITY
    assert b"grid capacity" in lib.tamx_last_error()
    # about 215 GB per list at the largest capacities: the second or third
    # cannot fit in one GPU's memory
    cap = (1 << 31) - 2
    made, status = [], _lib.TAMX_OK
    for _ in range(3):
        p = ctypes.c_void_p()
        status = lib.tamx_worklist_create(ctx.ptr, cap, cap, cap, ctypes.byref(p))
        if status != _lib.TAMX_OK:
            break
        made.append(p)
    msg = lib.tamx_last_error()
    for p in made:
        assert lib.tamx_worklist_destroy(p) == _lib.TAMX_OK
    assert status == _lib.TAMX_ERR_OOM, (status, msg)
    assert b"allocation" in msg
    c.run()
    torch.cuda.synchronize()
    Cref, npairs, c_off = oracle_contract(c, A, B)
    mask = c_off_mask(c_off, c)
    np.testing.assert_array_equal(c.C.cpu().numpy()[mask], Cref[mask])
