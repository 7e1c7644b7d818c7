"""Split-precision (BF16-limb) contraction against the FP64 oracle.

Bars: integer-valued tiles whose entries fit the first limb are represented
exactly, so every limb count must reproduce the oracle bit for bit; for
random data the relative Frobenius error must stay below the bound of the
limb expansion (measured separately and stated per limb count below), and 5
or more limbs must meet the 1e-10 target.
"""
import itertools

import numpy as np
import pytest
import torch

import oracle
from tiledarray_amd import DistArray, TiledRange, TiledRange1
from tiledarray_amd import engine as E
from tiledarray_amd import workloads as W

from test_gpu_contract import c_off_mask, matrix_ops_arrays, oracle_contract, rel_fro

pytestmark = pytest.mark.gpu

# relative Frobenius error bounds per limb count (the first limb carries 8
# bits, every later one 9; the bounds leave ~30x headroom over a numpy
# simulation of the same expansion: U[-1,1] tiles 1.5e-8 / 3.3e-11 / 4.9e-13
# at 3 / 4 / 5 limbs, ladder tiles 3.6e-7 / 1.2e-9 / 6.0e-12).  From five
# limbs on, the FP32 rounding of the lower-level accumulator (summed over
# 128 k before each FP64 fold) sets a floor near 2e-12.
SPLIT_TOL = {2: 1e-3, 3: 1e-6, 4: 1e-9, 5: 1e-11, 6: 1e-11}
LADDER_TOL = {3: 1e-5, 4: 3e-8, 5: 2e-10}


@pytest.mark.parametrize("nslices", [2, 3, 5, 6])
@pytest.mark.parametrize("op_a,op_b", list(itertools.product([0, 1], [0, 1])))
def test_split_integer_exact(ctx, nslices, op_a, op_b):
    rng = np.random.default_rng(300 + nslices * 4 + op_a * 2 + op_b)
    trm = TiledRange1([0, 40, 130, 200])
    trk = TiledRange1([0, 17, 64, 150, 151, 300])
    trn = TiledRange1([0, 96, 97, 180])
    A, B = matrix_ops_arrays(rng, trm, trk, trn, op_a, op_b, torch.float64, True, fill=0.8)
    la = "k,m" if op_a else "m,k"
    lb = "n,k" if op_b else "k,n"
    c = E.Contraction(ctx, "m,n", A, la, B, lb, alpha=2.0, precision=f"bf16x{nslices}")
    c.run()
    torch.cuda.synchronize()
    Cref, npairs, c_off = oracle_contract(c, A, B)
    mask = c_off_mask(c_off, c)
    np.testing.assert_array_equal(c.C.cpu().numpy()[mask], Cref[mask])


def test_norms_with_max_and_premax_split(ctx):
    # the fused norm + max pass gives the same norms bit for bit and each
    # tile's exact max |a|; the split fed with those maxima writes the same
    # limbs as the two-pass split, and the contraction is unchanged
    A, B, _ = W.blocksparse_gemm(n=1200, wlo=24, whi=96, seed=8)
    E.make_sparse(ctx, A, 1e-8)
    E.make_sparse(ctx, B, 1e-8)
    for X in (A, B):
        plain = E.tile_norms(ctx, X)
        mx = torch.full((X.trange.ntiles,), -1.0, dtype=torch.float64, device="cuda")
        fused = E.tile_norms(ctx, X, maxabs=mx)
        torch.cuda.synchronize()
        assert torch.equal(plain, fused)
        ref = np.zeros(X.trange.ntiles)
        for t in range(X.trange.ntiles):
            if X.offsets[t] >= 0:
                ref[t] = np.abs(X.tile(t).cpu().numpy()).max()
        np.testing.assert_array_equal(mx.cpu().numpy(), ref)
        X._mx = mx
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=1e-8, precision="bf16x4")
    c.run()
    torch.cuda.synchronize()
    limbs = [c.a_split.limbs.clone(), c.b_split.limbs.clone()]
    C0 = c.C.clone()
    for sp in (c.a_split, c.b_split):
        sp.limbs.zero_()
    c.split_step(maxabs=(A._mx, B._mx))
    c.gemm_step()
    torch.cuda.synchronize()
    assert torch.equal(c.a_split.limbs, limbs[0]) and torch.equal(c.b_split.limbs, limbs[1])
    mask = torch.as_tensor(c_off_mask(c.c_off.cpu().numpy(), c), device="cuda")
    assert torch.equal(c.C[:mask.numel()][mask], C0[:mask.numel()][mask])


@pytest.mark.parametrize("precision", ["bf16x4", "i8x5"])
def test_split_per_operand_overlap_equals_sequential(ctx, precision):
    # each operand's split enqueued on the side stream right after its own
    # norm pass (the other operand's norms, the shape estimate and the work
    # list run beside it) gives the limbs and the result of the sequential
    # step bit for bit, step after step
    A, B, _ = W.blocksparse_gemm(n=1500, wlo=24, whi=96, seed=21)
    E.make_sparse(ctx, A, 1e-8)
    E.make_sparse(ctx, B, 1e-8)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=1e-8, precision=precision)
    c.run()
    torch.cuda.synchronize()
    limbs = [c.a_split.limbs.clone(), c.b_split.limbs.clone()]
    C0 = c.C.clone()
    mask = torch.as_tensor(c_off_mask(c.c_off.cpu().numpy(), c), device="cuda")
    mx = {id(X): torch.zeros(X.trange.ntiles, dtype=torch.float64, device="cuda") for X in (A, B)}
    for _ in range(3):
        for sp in (c.a_split, c.b_split):
            sp.limbs.zero_()
        c.C.zero_()
        for w, X in enumerate((A, B)):
            E.tile_norms(ctx, X, out=X.shape_norms, maxabs=mx[id(X)])
            E.scale_norms(ctx, X.shape_norms, X.trange, 1e-8)
            c.split_operand_async(w, mx[id(X)])
        c.plan_and_join()
        c.gemm_step()
        torch.cuda.synchronize()
        assert torch.equal(c.a_split.limbs, limbs[0]) and torch.equal(c.b_split.limbs, limbs[1])
        assert torch.equal(c.C[:mask.numel()][mask], C0[:mask.numel()][mask])


@pytest.mark.parametrize("tile", [(128, 64), (64, 128)])
@pytest.mark.parametrize("precision", ["bf16x2", "bf16x3", "bf16x4", "i8x3", "i8x4"])
@pytest.mark.parametrize("op_a,op_b", list(itertools.product([0, 1], [0, 1])))
def test_split_wide_tiles_integer_exact(ctx, tile, precision, op_a, op_b):
    # the wide macro tiles (one wave per SIMD) on tiles wider and narrower
    # than one macro tile, ragged in both directions
    nslices = int(precision.split("x")[-1])
    rng = np.random.default_rng(900 + nslices * 4 + op_a * 2 + op_b + 50 * ("i8" in precision))
    trm = TiledRange1([0, 40, 170, 300, 301])
    trk = TiledRange1([0, 17, 64, 150, 151, 300])
    trn = TiledRange1([0, 96, 97, 250])
    A, B = matrix_ops_arrays(rng, trm, trk, trn, op_a, op_b, torch.float64, True, fill=0.8)
    la = "k,m" if op_a else "m,k"
    lb = "n,k" if op_b else "k,n"
    c = E.Contraction(ctx, "m,n", A, la, B, lb, alpha=2.0, precision=precision, tile=tile)
    c.run()
    torch.cuda.synchronize()
    Cref, npairs, c_off = oracle_contract(c, A, B)
    mask = c_off_mask(c_off, c)
    np.testing.assert_array_equal(c.C.cpu().numpy()[mask], Cref[mask])


@pytest.mark.parametrize("tile", [(128, 64), (64, 128)])
@pytest.mark.parametrize("precision", ["bf16x4", "i8x4"])
def test_split_wide_tiles_match_default(ctx, tile, precision):
    # same limbs, same pair order: the wide kernel's result equals the
    # default kernel's bit for bit on random data
    A, B, _ = W.blocksparse_gemm(n=1500, wlo=24, whi=96, seed=21)
    E.make_sparse(ctx, A, 1e-8)
    E.make_sparse(ctx, B, 1e-8)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=1e-8, precision=precision)
    c.run()
    torch.cuda.synchronize()
    C0 = c.C.clone()
    c.set_tile(tile)
    c.C.zero_()
    c.gemm_step()
    torch.cuda.synchronize()
    mask = torch.as_tensor(c_off_mask(c.c_off.cpu().numpy(), c), device="cuda")
    assert torch.equal(c.C[:mask.numel()][mask], C0[:mask.numel()][mask])


@pytest.mark.parametrize("precision", ["i8x5", "i8x6"])
@pytest.mark.parametrize("op_a,op_b", list(itertools.product([0, 1], [0, 1])))
def test_split_i8_eight_wave_integer_exact(ctx, precision, op_a, op_b):
    # 128 x 64 macro tiles from five INT8 limbs on (eight waves, B staged by
    # four of them) on tiles wider and narrower than one macro tile, ragged
    # in both directions, and taller than 128 rows
    nslices = int(precision.split("x")[-1])
    rng = np.random.default_rng(1300 + nslices * 4 + op_a * 2 + op_b)
    trm = TiledRange1([0, 40, 170, 300, 301])
    trk = TiledRange1([0, 17, 64, 150, 151, 300])
    trn = TiledRange1([0, 96, 97, 250])
    A, B = matrix_ops_arrays(rng, trm, trk, trn, op_a, op_b, torch.float64, True, fill=0.8)
    la = "k,m" if op_a else "m,k"
    lb = "n,k" if op_b else "k,n"
    c = E.Contraction(ctx, "m,n", A, la, B, lb, alpha=2.0, precision=precision,
                      tile=(128, 64))
    c.run()
    torch.cuda.synchronize()
    Cref, npairs, c_off = oracle_contract(c, A, B)
    mask = c_off_mask(c_off, c)
    np.testing.assert_array_equal(c.C.cpu().numpy()[mask], Cref[mask])


@pytest.mark.parametrize("precision", ["i8x5", "i8x6"])
def test_split_i8_eight_wave_matches_default(ctx, precision):
    # same limbs, same pair order, same levels and folds: the eight-wave
    # kernel's result equals the default kernel's bit for bit on random data
    A, B, _ = W.blocksparse_gemm(n=1500, wlo=24, whi=96, seed=23)
    E.make_sparse(ctx, A, 1e-8)
    E.make_sparse(ctx, B, 1e-8)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=1e-8, precision=precision)
    c.run()
    torch.cuda.synchronize()
    C0 = c.C.clone()
    for _ in range(2):
        c.set_tile((128, 64))
        c.C.zero_()
        c.gemm_step()
        torch.cuda.synchronize()
        mask = torch.as_tensor(c_off_mask(c.c_off.cpu().numpy(), c), device="cuda")
        assert torch.equal(c.C[:mask.numel()][mask], C0[:mask.numel()][mask])


@pytest.mark.parametrize("bk", [32, 64, 128])
@pytest.mark.parametrize("precision", ["i8x3", "i8x4", "i8x5"])
@pytest.mark.parametrize("op_a,op_b", list(itertools.product([0, 1], [0, 1])))
def test_split_i8_two_wave_integer_exact(ctx, precision, op_a, op_b, bk):
    # the INT8 kernels on the 32x32x32 MFMA (64 x 64 macro tiles; bk 32: four
    # waves with a four-stage ring of 32-k chunks, bk 64: four waves with two
    # 64-k buffers, bk 128: two waves) on tiles ragged in both directions,
    # narrower and wider than a macro tile, and k extents that are not
    # multiples of a chunk (17, 47, 86, 1, 149)
    nslices = int(precision.split("x")[-1])
    if bk == 128 and nslices not in (4, 5):
        pytest.skip("the two-wave kernel takes 4 or 5 limbs")
    rng = np.random.default_rng(1300 + nslices * 4 + op_a * 2 + op_b)
    trm = TiledRange1([0, 40, 170, 300, 301])
    trk = TiledRange1([0, 17, 64, 150, 151, 300])
    trn = TiledRange1([0, 96, 97, 250])
    A, B = matrix_ops_arrays(rng, trm, trk, trn, op_a, op_b, torch.float64, True, fill=0.8)
    la = "k,m" if op_a else "m,k"
    lb = "n,k" if op_b else "k,n"
    c = E.Contraction(ctx, "m,n", A, la, B, lb, alpha=2.0, precision=precision,
                      tile=(64, 64, bk))
    c.run()
    torch.cuda.synchronize()
    Cref, npairs, c_off = oracle_contract(c, A, B)
    mask = c_off_mask(c_off, c)
    np.testing.assert_array_equal(c.C.cpu().numpy()[mask], Cref[mask])


@pytest.mark.parametrize("bk", [32, 64, 128])
@pytest.mark.parametrize("precision", ["i8x2", "i8x4", "i8x5", "i8x6"])
def test_split_i8_two_wave_matches_default(ctx, precision, bk):
    # same limbs, same pair order, same folds: the 32x32x32 kernels' results
    # equal the 16x16x64 kernel's bit for bit on random block-sparse data,
    # and on a second run over a smaller problem after it
    if bk == 128 and precision not in ("i8x4", "i8x5"):
        pytest.skip("the two-wave kernel takes 4 or 5 limbs")
    for n, seed in ((1500, 31), (700, 32)):
        A, B, _ = W.blocksparse_gemm(n=n, wlo=24, whi=150, seed=seed)
        E.make_sparse(ctx, A, 1e-8)
        E.make_sparse(ctx, B, 1e-8)
        c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=1e-8,
                          precision=precision, alpha=-1.5)
        c.run()
        torch.cuda.synchronize()
        C0 = c.C.clone()
        c.set_tile((64, 64, bk))
        c.C.zero_()
        c.gemm_step()
        torch.cuda.synchronize()
        mask = torch.as_tensor(c_off_mask(c.c_off.cpu().numpy(), c), device="cuda")
        assert torch.equal(c.C[:mask.numel()][mask], C0[:mask.numel()][mask])


def test_split_sigma_is_tile_exponent(ctx):
    rng = np.random.default_rng(5)
    tr = TiledRange1([0, 20, 45])
    A, B = matrix_ops_arrays(rng, tr, tr, tr, 0, 0, torch.float64, False)
    A.pool.mul_(1e-3)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", precision="bf16x3")
    c.run()
    torch.cuda.synchronize()
    sa, sb = c.sigmas()
    for X, sig in ((A, sa), (B, sb)):
        for t in range(X.trange.ntiles):
            tile = X.tile(t).cpu().numpy()
            m = np.abs(tile).max()
            assert 2.0 ** (sig[t] - 1) <= m < 2.0 ** sig[t]


@pytest.mark.parametrize("nslices", [2, 3, 4, 5, 6])
def test_split_blocksparse_error_bound(ctx, nslices):
    thr = 1e-8
    A, B, _ = W.blocksparse_gemm(n=1500, wlo=24, whi=96, seed=21)
    E.make_sparse(ctx, A, thr)
    E.make_sparse(ctx, B, thr)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=thr, precision=f"bf16x{nslices}")
    c.run()
    torch.cuda.synchronize()
    sa, sb = A.shape_norms.cpu().numpy(), B.shape_norms.cpu().numpy()
    Cref, npairs, c_off = oracle_contract(c, A, B, (A.offsets >= 0) & (sa >= thr),
                                          (B.offsets >= 0) & (sb >= thr))
    mask = c_off_mask(c_off, c)
    err = rel_fro(c.C.cpu().numpy()[mask], Cref[mask])
    assert err <= SPLIT_TOL[nslices], (nslices, err)


@pytest.mark.parametrize("nslices", [3, 4, 5])
def test_split_ladder_error_bound(ctx, nslices):
    T, V, _ = W.ladder(o=8, v=40, otiles=2, wlo=9, whi=17, seed=3)
    c = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", precision=f"bf16x{nslices}")
    c.run()
    R = c.result()
    ref = np.einsum("cdij,acbd->abij", T.to_dense(), V.to_dense())
    err = rel_fro(R.to_dense(), ref)
    assert err <= LADDER_TOL[nslices], (nslices, err)


@pytest.mark.parametrize("nslices", [2, 3, 4, 5])
@pytest.mark.parametrize("is_right,op", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_split_limbs_closed_form(ctx, nslices, is_right, op):
    # the limbs of one tile against the expansion's definition:
    #   L_s 2^(9s) is an integer l_s, |l_0| <= 256, |l_s| <= 256,
    #   |a - 2^(sigma-8) sum_s L_s| <= 2^(sigma - 9S), padding is zero
    import ctypes
    from tiledarray_amd import _lib
    rows, cols = 37, 45  # stored (row-major) tile
    rng = np.random.default_rng(700 + nslices)
    a = rng.uniform(-1, 1, (rows, cols)) * 3.7 * 2.0 ** rng.integers(-3, 4, (rows, cols))
    pool = torch.as_tensor(a.reshape(-1), device="cuda")
    plan = ctypes.c_void_p()
    off = np.zeros(1, np.int64)
    r32 = np.array([rows], np.int32)
    c32 = np.array([cols], np.int32)
    E.call("tamx_split_plan_create", ctx.ptr, nslices, is_right, op, 1, off.ctypes.data,
           r32.ctypes.data, c32.ctypes.data, ctypes.byref(plan))
    try:
        n = _lib.lib().tamx_split_plan_limb_elems(plan)
        limbs = torch.zeros(n, dtype=torch.bfloat16, device="cuda")
        E.call("tamx_split_f64", ctx.ptr, plan, pool.data_ptr(), limbs.data_ptr(), ctx.stream())
        torch.cuda.synchronize()
        sig = np.zeros(1, np.int32)
        E.call("tamx_split_plan_sigmas", plan, sig.ctypes.data)
    finally:
        _lib.lib().tamx_split_plan_destroy(plan)
    sigma = int(sig[0])
    assert 2.0 ** (sigma - 1) <= np.abs(a).max() < 2.0 ** sigma
    # logical limb layout: op(A) as m x K (A) or op(B)^T as n x K (B)
    trans = (op == 1) if not is_right else (op == 0)
    logical = a.T if trans else a
    m, k = logical.shape
    mp, kp = -(-m // 16) * 16, -(-k // 32) * 32
    L = limbs.float().cpu().numpy().astype(np.float64)[:nslices * mp * kp]
    L = L.reshape(nslices, mp, kp)
    for s_ in range(nslices):
        l = L[s_] * 2.0 ** (9 * s_)
        assert np.array_equal(l, np.rint(l)), s_
        assert np.abs(l).max() <= 256, s_
    assert not L[:, m:, :].any() and not L[:, :, k:].any()
    rec = 2.0 ** (sigma - 8) * L[:, :m, :k].sum(axis=0)
    assert np.abs(logical - rec).max() <= 2.0 ** (sigma - 9 * nslices)


def test_split_error_drops_nine_bits_per_limb(ctx):
    # one more limb divides the error by about 2^9 (between 2^7 and 2^11)
    A, B, _ = W.dense_gemm(n=384, tile=128)
    errs = {}
    ref = A.to_dense() @ B.to_dense()
    for S in (2, 3, 4):
        c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", precision=f"bf16x{S}")
        c.run()
        errs[S] = rel_fro(c.result().to_dense(), ref)
    for S in (2, 3):
        ratio = errs[S] / errs[S + 1]
        assert 2.0 ** 7 <= ratio <= 2.0 ** 11, (S, errs)


@pytest.mark.parametrize("nslices", [3, 4])
def test_split_fused_permute_matches_permuted_copy(ctx, nslices):
    # V stored (a,c,b,d) is split straight from its stored layout through an
    # index map; the limbs and the result equal those from a permuted copy
    T, V, _ = W.ladder(o=8, v=40, otiles=2, wlo=9, whi=17, seed=4)
    cf = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d",
                       precision=f"bf16x{nslices}")
    cp = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d",
                       precision=f"bf16x{nslices}", fuse_permute=False)
    assert hasattr(cf.B, "_view_src") and hasattr(cp.B, "_perm_plan")
    cf.run()
    cp.run()
    torch.cuda.synchronize()
    assert torch.equal(cf.b_split.limbs, cp.b_split.limbs)
    assert torch.equal(cf.a_split.limbs, cp.a_split.limbs)
    np.testing.assert_array_equal(cf.result().to_dense(), cp.result().to_dense())


@pytest.mark.parametrize("annot", ["c,a,d,b", "d,c,b,a", "b,d,a,c"])
def test_split_fused_permute_other_orders(ctx, annot):
    # other general permutations of the right operand (the stride-1 source
    # mode lands in the limb rows or in k), integer data: exact
    rng = np.random.default_rng(11)
    tv = TiledRange1([0, 5, 12, 20])
    to = TiledRange1([0, 3, 6])
    T = DistArray.allocate(TiledRange([tv, tv, to, to]), rng.random(36) < 0.9)
    T.pool.copy_(torch.as_tensor(rng.integers(-8, 9, T.pool.numel()), dtype=torch.float64))
    V = DistArray.allocate(TiledRange([tv, tv, tv, tv]), rng.random(81) < 0.9)
    V.pool.copy_(torch.as_tensor(rng.integers(-8, 9, V.pool.numel()), dtype=torch.float64))
    c = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, annot, precision="bf16x3")
    c.run()
    R = c.result().to_dense()
    ref = np.einsum("cdij," + annot.replace(",", "") + "->abij", T.to_dense(), V.to_dense())
    np.testing.assert_array_equal(R, ref)
    # the same operand on the left
    c = E.Contraction(ctx, "i,j,a,b", V, annot, T, "c,d,i,j", precision="bf16x3")
    c.run()
    np.testing.assert_array_equal(c.result().to_dense(), ref.transpose(2, 3, 0, 1))


@pytest.mark.parametrize("op_a,op_b", list(itertools.product([0, 1], [0, 1])))
def test_f32x2_integer_exact(ctx, op_a, op_b):
    # small integers split exactly into hi (lo = 0) and their products and
    # 16-term FP32 sums are exact
    rng = np.random.default_rng(400 + op_a * 2 + op_b)
    trm = TiledRange1([0, 40, 130])
    trk = TiledRange1([0, 17, 64, 150])
    trn = TiledRange1([0, 96, 97, 180])
    A, B = matrix_ops_arrays(rng, trm, trk, trn, op_a, op_b, torch.float64, True)
    c = E.Contraction(ctx, "m,n", A, "k,m" if op_a else "m,k", B, "n,k" if op_b else "k,n",
                      precision="f32x2")
    c.run()
    torch.cuda.synchronize()
    Cref, _, c_off = oracle_contract(c, A, B)
    mask = c_off_mask(c_off, c)
    np.testing.assert_array_equal(c.C.cpu().numpy()[mask], Cref[mask])


def test_f32x2_error_bound(ctx):
    # FP32 accumulation rounds: ~24 bits, bound 1e-6 relative (Frobenius)
    T, V, _ = W.ladder(o=8, v=40, otiles=2, wlo=9, whi=17, seed=3)
    c = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", precision="f32x2")
    c.run()
    R = c.result()
    ref = np.einsum("cdij,acbd->abij", T.to_dense(), V.to_dense())
    err = rel_fro(R.to_dense(), ref)
    assert err <= 1e-6, err


# ------------------------------------------------------------- INT8 limbs --
# relative Frobenius error bounds for INT8 limbs of 7 bits: a numpy
# simulation of the same expansion on U[-1,1] 512^3 products (64-wide tiles)
# gives 9.5e-7 / 8.3e-9 / 7.1e-11 / 6.0e-13 at 3 / 4 / 5 / 6 limbs (the
# INT32 levels are exact, so no accumulator floor); the bounds leave ~30x
I8_TOL = {3: 3e-5, 4: 3e-7, 5: 2e-9, 6: 2e-11}


def _i8_split_tile(ctx, a, nslices, is_right, op):
    """INT8 limbs of the one stored tile `a` (rows x cols, row-major): the
    limb array as (nslices, m_pad, k_pad), sigma, and op(a) as m x k."""
    import ctypes
    from tiledarray_amd import _lib
    rows, cols = a.shape
    pool = torch.as_tensor(np.ascontiguousarray(a).reshape(-1), device="cuda")
    plan = ctypes.c_void_p()
    off = np.zeros(1, np.int64)
    r32 = np.array([rows], np.int32)
    c32 = np.array([cols], np.int32)
    E.call("tamx_split_plan_create_typed", ctx.ptr, _lib.LIMB_I8, nslices, is_right, op, 1,
           off.ctypes.data, r32.ctypes.data, c32.ctypes.data, ctypes.byref(plan))
    try:
        assert _lib.lib().tamx_split_plan_limb_type(plan) == _lib.LIMB_I8
        n = _lib.lib().tamx_split_plan_limb_elems(plan)
        limbs = torch.full((n,), 99, dtype=torch.uint8, device="cuda")
        E.call("tamx_split_f64", ctx.ptr, plan, pool.data_ptr(), limbs.data_ptr(), ctx.stream())
        torch.cuda.synchronize()
        sig = np.zeros(1, np.int32)
        E.call("tamx_split_plan_sigmas", plan, sig.ctypes.data)
    finally:
        _lib.lib().tamx_split_plan_destroy(plan)
    trans = (op == 1) if not is_right else (op == 0)
    logical = a.T if trans else a
    m, k = logical.shape
    mp, kp = -(-m // 16) * 16, -(-k // 64) * 64
    L = limbs.cpu().numpy().view(np.int8).astype(np.float64)[:nslices * mp * kp]
    # each plane in blocks of 32 k: block b holds rows 0..mp-1, 32 limbs each
    L = L.reshape(nslices, kp // 32, mp, 32).transpose(0, 2, 1, 3).reshape(nslices, mp, kp)
    return L, int(sig[0]), logical


def _i8_limbs_by_definition(logical, sigma, nslices):
    # l_s = rint(x_s) (ties to even), x_0 = a 2^(7-sigma),
    # x_{s+1} = 128 (x_s - l_s); every step is exact in FP64
    xs = np.ldexp(logical, 7 - sigma)
    out = []
    for _ in range(nslices):
        ls = np.rint(xs)
        out.append(ls)
        xs = (xs - ls) * 128.0
    return np.stack(out)


@pytest.mark.parametrize("nslices", [2, 3, 4, 5, 6])
@pytest.mark.parametrize("is_right,op", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_split_i8_limbs_closed_form(ctx, nslices, is_right, op):
    # the INT8 limbs of one tile against their definition, bit for bit, and:
    #   |l_0| <= 127, |l_s| <= 64 after it,
    #   |a - 2^(sigma-7) sum_s l_s 2^-7s| <= 2^(sigma - 7S - 1),
    #   K padded to 64, padding zero
    rng = np.random.default_rng(800 + nslices)
    a = rng.uniform(-1, 1, (37, 45)) * 3.7 * 2.0 ** rng.integers(-3, 4, (37, 45))
    L, sigma, logical = _i8_split_tile(ctx, a, nslices, is_right, op)
    m, k = logical.shape
    assert np.abs(L[0]).max() <= 127 and np.abs(L[1:]).max() <= 64
    assert not L[:, m:, :].any() and not L[:, :, k:].any()
    assert np.array_equal(L[:, :m, :k], _i8_limbs_by_definition(logical, sigma, nslices))
    w = 2.0 ** (-7 * np.arange(nslices))
    rec = 2.0 ** (sigma - 7) * np.tensordot(w, L[:, :m, :k], axes=1)
    assert np.abs(logical - rec).max() <= 2.0 ** (sigma - 7 * nslices - 1)
    # sigma: the tile max in [2^(sigma-1), 2^sigma), or just below 2^(sigma-1)
    # when it would round to 128 in the first limb
    mx = np.abs(a).max()
    assert mx < 2.0 ** sigma and (mx >= 2.0 ** (sigma - 1) or mx * 2.0 ** (8 - sigma) > 127)


@pytest.mark.parametrize("is_right,op", [(0, 0), (1, 1)])
def test_split_i8_limbs_ties(ctx, is_right, op):
    # values halfway between limb steps: the first limb of (j + 1/2) / 128 and
    # the second of j / 128 + (i + 1/2) / 2^14 are ties, which round to even
    # (tile max 125.5 / 128, so sigma = 0 and x_0 = 128 a)
    j = np.arange(-126, 126, dtype=np.float64)
    first = (j + 0.5) / 128.0
    second = np.clip(j, -125, 125) / 128.0 + (np.arange(j.size) % 128 - 63.5) / 2.0 ** 14
    a = np.concatenate([first, second, -first, -second]).reshape(16, -1)
    L, sigma, logical = _i8_split_tile(ctx, a, 3, is_right, op)
    m, k = logical.shape
    assert sigma == 0
    exp = _i8_limbs_by_definition(logical, sigma, 3)
    assert np.array_equal(L[:, :m, :k], exp)
    # the ties did round to even
    t0 = np.abs(np.ldexp(logical, 7) - np.trunc(np.ldexp(logical, 7))) == 0.5
    assert t0.sum() >= 2 * j.size and np.all(L[0, :m, :k][t0] % 2 == 0)


@pytest.mark.parametrize("precision", ["f64", "bf16x3", "i8x4"])
@pytest.mark.parametrize("empty", ["A", "B", "both"])
@pytest.mark.parametrize("sparse", [True, False])
def test_contraction_with_empty_operand(ctx, precision, empty, sparse):
    # an operand that stores no tile.  Sparse (screened shapes): the result
    # shape is zero, so no task, no pair, every result tile absent.  Dense
    # (no screening): every result tile is a task with no pair, written as
    # zeros (beta = 0, ContractReduce with nothing to reduce)
    rng = np.random.default_rng(1400)
    trm = TiledRange1([0, 40, 130])
    trk = TiledRange1([0, 17, 64, 150])
    trn = TiledRange1([0, 96, 97])
    A, B = matrix_ops_arrays(rng, trm, trk, trn, 0, 0, torch.float64, True, fill=0.8)
    if empty in ("A", "both"):
        A = DistArray.allocate(TiledRange([trm, trk]), np.zeros((2, 3), bool), dtype=torch.float64)
    if empty in ("B", "both"):
        B = DistArray.allocate(TiledRange([trk, trn]), np.zeros((3, 2), bool), dtype=torch.float64)
    kw = {}
    if sparse:
        E.make_sparse(ctx, A, 1e-8)
        E.make_sparse(ctx, B, 1e-8)
        kw["threshold"] = 1e-8
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", precision=precision, **kw)
    c.C.fill_(7.0)
    c.run()
    torch.cuda.synchronize()
    cnt = c.counters()
    assert cnt["npairs"] == 0 and cnt["flops"] == 0
    c_off = c.c_off.cpu().numpy()
    if sparse:
        assert cnt["ntasks"] == 0 and (c_off < 0).all()
    else:
        assert cnt["ntasks"] == 4 and (c_off >= 0).all()
        mask = c_off_mask(c_off, c)
        assert not c.C.cpu().numpy()[mask].any()


@pytest.mark.parametrize("e", [-1070, -700, 700, 1000])
def test_split_i8_limbs_extreme_scales(ctx, e):
    # tiles from the subnormal range to near the top of the FP64 range: the
    # scale exponent follows the tile and the limbs equal their definition
    # bit for bit (x_0 = a 2^(7-sigma) is formed exactly by ldexp)
    rng = np.random.default_rng(1200 + e)
    a = np.ldexp(rng.uniform(-1, 1, (23, 70)), e)
    assert np.all(np.isfinite(a)) and np.abs(a).max() > 0
    L, sigma, logical = _i8_split_tile(ctx, a, 4, 0, 0)
    m, k = logical.shape
    assert np.array_equal(L[:, :m, :k], _i8_limbs_by_definition(logical, sigma, 4))
    mx = np.abs(a).max()
    assert mx < np.ldexp(1.0, sigma) and mx >= np.ldexp(1.0, sigma - 2)


@pytest.mark.parametrize("precision", ["i8x3", "bf16x3"])
def test_split_power_of_two_scaled_operands_exact(ctx, precision):
    # A scaled by 2^600 and B by 2^-600: the per-tile scale exponents move
    # by 600 each, the limbs stay the same, and the product of integer-valued
    # tiles equals the oracle's unscaled product bit for bit
    rng = np.random.default_rng(1300)
    trm = TiledRange1([0, 40, 130, 200])
    trk = TiledRange1([0, 17, 64, 150, 151, 300])
    trn = TiledRange1([0, 96, 97, 180])
    A, B = matrix_ops_arrays(rng, trm, trk, trn, 0, 0, torch.float64, True, fill=0.8)
    Ah, Bh = A.pool.clone(), B.pool.clone()
    A.pool.mul_(2.0 ** 600)
    B.pool.mul_(2.0 ** -600)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", precision=precision)
    c.run()
    torch.cuda.synchronize()
    A.pool.copy_(Ah)
    B.pool.copy_(Bh)
    Cref, npairs, c_off = oracle_contract(c, A, B)
    mask = c_off_mask(c_off, c)
    np.testing.assert_array_equal(c.C.cpu().numpy()[mask], Cref[mask])


@pytest.mark.parametrize("nslices", [2, 3, 5, 6])
@pytest.mark.parametrize("op_a,op_b", list(itertools.product([0, 1], [0, 1])))
def test_split_i8_integer_exact(ctx, nslices, op_a, op_b):
    rng = np.random.default_rng(400 + nslices * 4 + op_a * 2 + op_b)
    trm = TiledRange1([0, 40, 130, 200])
    trk = TiledRange1([0, 17, 64, 150, 151, 300])
    trn = TiledRange1([0, 96, 97, 180])
    A, B = matrix_ops_arrays(rng, trm, trk, trn, op_a, op_b, torch.float64, True, fill=0.8)
    la = "k,m" if op_a else "m,k"
    lb = "n,k" if op_b else "k,n"
    c = E.Contraction(ctx, "m,n", A, la, B, lb, alpha=2.0, precision=f"i8x{nslices}")
    c.run()
    torch.cuda.synchronize()
    Cref, npairs, c_off = oracle_contract(c, A, B)
    mask = c_off_mask(c_off, c)
    np.testing.assert_array_equal(c.C.cpu().numpy()[mask], Cref[mask])


@pytest.mark.parametrize("nslices", [3, 4, 5, 6])
def test_split_i8_blocksparse_error_bound(ctx, nslices):
    A, B, _ = W.blocksparse_gemm(n=1500, wlo=24, whi=96, seed=33)
    E.make_sparse(ctx, A, 1e-8)
    E.make_sparse(ctx, B, 1e-8)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=1e-8,
                      precision=f"i8x{nslices}")
    c.run()
    torch.cuda.synchronize()
    Cref, npairs, c_off = oracle_contract(c, A, B)
    mask = c_off_mask(c_off, c)
    err = rel_fro(c.C.cpu().numpy()[mask], Cref[mask])
    assert err <= I8_TOL[nslices], (nslices, err)


def test_split_i8_long_pairs_fold(ctx):
    # pairs longer than the 4096-k fold interval of the INT32 levels
    rng = np.random.default_rng(12)
    trm = TiledRange1([0, 70])
    trk = TiledRange1([0, 5000, 9000])
    trn = TiledRange1([0, 33])
    A, B = matrix_ops_arrays(rng, trm, trk, trn, 0, 0, torch.float64, True, fill=1.0)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", precision="i8x3")
    c.run()
    torch.cuda.synchronize()
    Cref, npairs, c_off = oracle_contract(c, A, B)
    mask = c_off_mask(c_off, c)
    np.testing.assert_array_equal(c.C.cpu().numpy()[mask], Cref[mask])


def test_split_i8_ladder_index_map(ctx):
    # the general operand read through the index map, INT8 limbs: equal to
    # the permuted copy's result bit for bit, and within the error bound
    T, V, _ = W.ladder(o=8, v=40, otiles=2, wlo=9, whi=17, seed=4)
    cf = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", precision="i8x5")
    cp = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", precision="i8x5",
                       fuse_permute=False)
    assert hasattr(cf.B, "_view_src") and hasattr(cp.B, "_perm_plan")
    cf.run()
    cp.run()
    torch.cuda.synchronize()
    assert torch.equal(cf.b_split.limbs, cp.b_split.limbs)
    R = cf.result().to_dense()
    np.testing.assert_array_equal(R, cp.result().to_dense())
    ref = np.einsum("cdij,acbd->abij", T.to_dense(), V.to_dense())
    assert rel_fro(R, ref) <= LADDER_TOL[4]  # five INT8 limbs ~ four BF16 ones


def test_split_i8_autotune_and_wide_tile_rejected(ctx):
    from tiledarray_amd import _lib
    A, B, _ = W.blocksparse_gemm(n=800, wlo=24, whi=96, seed=6)
    E.make_sparse(ctx, A, 1e-8)
    E.make_sparse(ctx, B, 1e-8)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=1e-8, precision="i8x4")
    c.run()
    torch.cuda.synchronize()
    tile, ms = c.tune(reps=1, force=True)
    assert tile[:2] in ((64, 64), (128, 64), (64, 128)) and ms > 0.0
    # five INT8 limbs have 64 x 64 and the eight-wave 128 x 64, not 64 x 128
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=1e-8, precision="i8x5")
    c.run()
    torch.cuda.synchronize()
    tile, ms = c.tune(reps=1, force=True)
    assert tile[:2] in ((64, 64), (128, 64)) and ms > 0.0
    c.set_tile((64, 128))
    with pytest.raises(_lib.TamxError) as err:
        c.gemm_step()
    assert err.value.status == _lib.TAMX_ERR_UNSUPPORTED
