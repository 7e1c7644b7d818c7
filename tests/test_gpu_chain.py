"""Result orders written by the GEMM itself, and contractions chained with
the intermediate kept in HBM.

A general result permutation (contract_reduce.h:370-378) is fused into the
store of every GEMM kernel (FP64, FP32, BF16 and INT8 limbs, two FP32
limbs): each result tile is written through its index map straight into
the target order.  Bars: the fused store equals the permute pass over the
finished tiles bit for bit for every precision; FP64 integer-valued data
equal the oracle (and einsum) exactly; result() reads nothing back from the
device.

Chains D = (T * V) * E in the style of TA expressions whose result feeds the
next one (expressions/expr.h:412-421): the first result stays a device
DistArray (pool, offsets and, when screened, the estimated shape
SparseShape::gemm gives it, sparse_shape.h:1589-1681) and is the next
contraction's operand without any host copy of its tiles or offsets.
Integer-valued data: exact against einsum (dense) and against the oracle
run step by step on the same screened structure (sparse), patterns
bit-exact.
"""
import numpy as np
import pytest
import torch

import oracle
from tiledarray_amd import DistArray, TiledRange, TiledRange1
from tiledarray_amd import engine as E
from tiledarray_amd import workloads as W

from test_gpu_contract import c_off_mask, oracle_contract, rel_fro

pytestmark = pytest.mark.gpu


def _integer_ladder(seed, o=6, v=30, otiles=2, wlo=6, whi=11):
    T, V, _ = W.ladder(o=o, v=v, otiles=otiles, wlo=wlo, whi=whi, seed=seed, v_order="abcd")
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    for X in (T, V):
        X.pool.copy_(torch.randint(-6, 7, X.pool.shape, generator=g, device="cuda",
                                   dtype=torch.int64).to(X.pool.dtype))
    return T, V


def _no_sync_result(c):
    """c.result() with torch's synchronising calls turned into errors."""
    torch.cuda.synchronize()
    torch.cuda.set_sync_debug_mode("error")
    try:
        R = c.result()
    finally:
        torch.cuda.set_sync_debug_mode("default")
    return R


@pytest.mark.parametrize("target", ["b,a,i,j", "a,i,b,j", "j,b,a,i", "i,b,a,j"])
@pytest.mark.parametrize("precision", ["f64", "i8x5", "bf16x3", "f32x2"])
def test_fused_result_permute_equals_permute_pass(ctx, target, precision):
    T, V = _integer_ladder(11)
    kw = dict(precision=precision)
    cf = E.Contraction(ctx, target, T, "c,d,i,j", V, "a,b,c,d", **kw)
    assert cf.c_scatter is not None
    cf.run()
    Rf = _no_sync_result(cf)
    cp = E.Contraction(ctx, target, T, "c,d,i,j", V, "a,b,c,d", fuse_result_permute=False, **kw)
    assert cp.c_scatter is None
    cp.run()
    Rp = cp.result()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(Rf.offsets, Rp.offsets)
    m = torch.as_tensor(c_off_mask(cf.c_off.cpu().numpy(), cf), device="cuda")
    n = m.numel()
    assert torch.equal(Rf.pool[:n][m], Rp.pool[:n][m])
    ref = np.einsum("cdij,abcd->" + target.replace(",", ""), T.to_dense(), V.to_dense())
    if precision == "f64":
        # integer data: every product and sum exact
        np.testing.assert_array_equal(Rf.to_dense(), ref)
    else:
        assert rel_fro(Rf.to_dense(), ref) <= (1e-6 if precision != "bf16x3" else 1e-5)


@pytest.mark.parametrize("target", ["b,a,i,j", "j,b,a,i"])
def test_fused_result_permute_eight_wave_split(ctx, target):
    # the eight-wave INT8 kernel (128 x 64 macro tiles, five limbs) writes
    # through the same scatter map as the 64 x 64 kernel: equal results
    T, V = _integer_ladder(13)
    out = []
    for tile in (None, (128, 64)):
        c = E.Contraction(ctx, target, T, "c,d,i,j", V, "a,b,c,d", precision="i8x5", tile=tile)
        assert c.c_scatter is not None
        c.run()
        out.append(_no_sync_result(c))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out[0].offsets, out[1].offsets)
    assert torch.equal(out[0].pool, out[1].pool)


def test_fused_result_permute_fp32_and_alpha(ctx):
    T, V = _integer_ladder(12)
    T32 = DistArray(T.trange, T.pool.float(), T.offsets)
    V32 = DistArray(V.trange, V.pool.float(), V.offsets)
    c = E.Contraction(ctx, "j,a,i,b", T32, "c,d,i,j", V32, "a,b,c,d", precision="f32", alpha=-2.0)
    assert c.c_scatter is not None
    c.run()
    R = _no_sync_result(c)
    ref = -2.0 * np.einsum("cdij,abcd->jaib", T.to_dense(), V.to_dense())
    np.testing.assert_array_equal(R.to_dense(), ref.astype(np.float32))


def test_fused_result_permute_against_oracle(ctx):
    # the natural result tiles from the oracle, permuted tile by tile as TA's
    # post-op does, equal the fused store bit for bit
    T, V = _integer_ladder(13, v=44, wlo=9, whi=16)
    target = "i,b,a,j"
    c = E.Contraction(ctx, target, T, "c,d,i,j", V, "a,b,c,d", alpha=3.0)
    c.run()
    R = _no_sync_result(c)
    Cref, npairs, c_off = oracle_contract(c, c.A, c.B)
    rp = [int(c.plan.result_perm[d]) for d in range(c.plan.result_rank)]
    assert c.counters()["npairs"] == npairs
    nat = TiledRange(c.natural_ranges)
    ext = nat.tile_extents_table()
    got = R.pool.cpu().numpy()
    for t, o in enumerate(c_off):
        if o < 0:
            continue
        e = [int(x) for x in ext[t]]
        vol = int(np.prod(e))
        want = oracle.permute(Cref[o:o + vol].reshape(e), rp).reshape(-1)
        np.testing.assert_array_equal(got[o:o + vol], want)


def test_chain_dense_intermediate_stays_on_device(ctx):
    # R1(b,a,i,j) = T(c,d,i,j) * V(a,b,c,d): a general result order, fused
    # into the store; R2(i,j,x) = R1(b,a,i,j) * E(b,a,x) reads R1 as the
    # GEMM's A operand in place (its contracted modes lead, so op_a = Trans)
    T, V = _integer_ladder(21)
    trv = T.trange.ranges[0]
    trx = TiledRange1([0, 5, 13, 20])
    Eop = DistArray.allocate(TiledRange([trv, trv, trx]))
    Eop.pool.copy_(torch.randint(-4, 5, Eop.pool.shape, device="cuda").double())
    c1 = E.Contraction(ctx, "b,a,i,j", T, "c,d,i,j", V, "a,b,c,d")
    c1.run()
    R1 = _no_sync_result(c1)
    c2 = E.Contraction(ctx, "i,j,x", R1, "b,a,i,j", Eop, "b,a,x")
    assert c2.A is R1 and c2.op_a == 1
    c2.run()
    R2 = c2.result()
    torch.cuda.synchronize()
    # R1's offsets were never read back to the host
    assert R1._offsets is None
    ref = np.einsum("cdij,abcd,bax->ijx", T.to_dense(), V.to_dense(), Eop.to_dense())
    np.testing.assert_array_equal(R2.to_dense(), ref)


def test_chain_sparse_against_oracle(ctx):
    # the screened chain: R1's shape is the estimate of the first product,
    # and the second product screens with it, as TA's expression does
    thr = 1e-3
    T, V = _integer_ladder(22, v=40, wlo=8, whi=14)
    # banded magnitudes so that screening removes tiles and pairs
    for X, modes in ((T, (0, 2)), (V, (0, 2))):
        ext = X.trange.tile_extents_table()
        shp = X.trange.tiles_shape
        for t in range(X.trange.ntiles):
            idx = np.unravel_index(t, shp)
            d = abs(int(idx[modes[0]]) - int(idx[modes[1]]))
            if X.offsets[t] >= 0 and d >= 2:
                X.tile(t).mul_(2.0 ** (-12 * d))
    trv = T.trange.ranges[0]
    trx = TiledRange1([0, 7, 16])
    Eop = DistArray.allocate(TiledRange([trv, trv, trx]))
    Eop.pool.copy_(torch.randint(-4, 5, Eop.pool.shape, device="cuda").double())
    for X in (T, V, Eop):
        E.make_sparse(ctx, X, thr)
    c1 = E.Contraction(ctx, "b,a,i,j", T, "c,d,i,j", V, "a,b,c,d", threshold=thr)
    c1.run()
    R1 = _no_sync_result(c1)
    assert R1.shape_norms is not None and R1.shape_threshold == c1.threshold
    c2 = E.Contraction(ctx, "i,j,x", R1, "b,a,i,j", Eop, "b,a,x", threshold=thr)
    c2.run()
    torch.cuda.synchronize()
    assert R1._offsets is None
    # step 1 on the oracle: its estimate equals the device's bit for bit,
    # its tiles (permuted into the target order) the device's R1 exactly
    sa = c1.A.shape_norms.cpu().numpy()
    sb = c1.B.shape_norms.cpu().numpy()
    est1, _ = oracle.shape_gemm(np.where(sa >= thr, sa, 0), np.where(sb >= thr, sb, 0),
                                c1.op_a, c1.op_b, c1.Mt, c1.Nt, c1.Kt,
                                c1.k_sizes.cpu().numpy(), 1.0, thr)
    np.testing.assert_array_equal(c1.c_norms.cpu().numpy(), est1.reshape(-1))
    a_pres = sa >= thr
    b_pres = sb >= thr
    Cref1, np1, c_off1 = oracle_contract(c1, c1.A, c1.B, a_pres, b_pres)
    assert c1.counters()["npairs"] == np1
    rp = [int(c1.plan.result_perm[d]) for d in range(c1.plan.result_rank)]
    ext = TiledRange(c1.natural_ranges).tile_extents_table()
    got1 = R1.pool.cpu().numpy()
    for t, o in enumerate(c_off1):
        if o >= 0:
            e = [int(x) for x in ext[t]]
            vol = int(np.prod(e))
            np.testing.assert_array_equal(got1[o:o + vol],
                                          oracle.permute(Cref1[o:o + vol].reshape(e), rp).reshape(-1))
    # step 2 on the oracle from R1's shape and tiles
    s1 = R1.shape_norms.cpu().numpy()
    se = Eop.shape_norms.cpu().numpy()
    est2, _ = oracle.shape_gemm(np.where(s1 >= thr, s1, 0), np.where(se >= thr, se, 0),
                                c2.op_a, c2.op_b, c2.Mt, c2.Nt, c2.Kt,
                                c2.k_sizes.cpu().numpy(), 1.0, thr)
    np.testing.assert_array_equal(c2.c_norms.cpu().numpy(), est2.reshape(-1))
    Cref2, np2, c_off2 = oracle_contract(c2, c2.A, c2.B, s1 >= thr, se >= thr)
    assert c2.counters()["npairs"] == np2 > 0
    np.testing.assert_array_equal((c_off2 >= 0), est2.reshape(-1) >= thr)
    mask = c_off_mask(c_off2, c2)
    np.testing.assert_array_equal(c2.C.cpu().numpy()[mask], Cref2[mask])
