"""Batched tensor-of-tensor contractions (one persistent grouped launch for
all instances) and their hipGraph capture, against the CPU oracle.

Bars: FP64 integer-valued cells bit-exact; FP32 relative Frobenius error
<= 1e-5 per instance; shapes (cell norms, scaled norms, result estimate,
result pattern) bit-exact; every graph replay equal to the eager result.
"""
import numpy as np
import pytest
import torch

import oracle
from tiledarray_amd import engine as E
from tiledarray_amd import workloads as W

pytestmark = pytest.mark.gpu


def make(ctx, nbatch, dtype, integer=False, seed=7, fill=0.5):
    w = W.tot_batch(nbatch=nbatch, dtype=dtype, integer=integer, seed=seed, fill=fill)
    g = w["grid"]
    c = E.BatchedContraction(ctx, nbatch, g, g, g, w["a_dims"], w["b_dims"], w["c_dims"],
                             w["A"], w["a_off"], w["B"], w["b_off"])
    return w, c


def oracle_batch(w, c):
    """Per instance: shapes and the one-rank contraction, at the product's
    result offsets; returns (C pool, c_norms, pattern)."""
    g = w["grid"]
    Ah, Bh = w["A"].cpu().numpy(), w["B"].cpu().numpy()
    c_off = c.c_off.cpu().numpy()
    Cref = np.zeros(c.c_capacity, Ah.dtype)
    est_all = []
    thr = c.threshold
    for b in range(w["nbatch"]):
        ad, cd, bd = w["a_dims"][b], w["c_dims"][b], w["b_dims"][b]
        sl = slice(b * g * g, (b + 1) * g * g)
        a_off, b_off = w["a_off"][sl], w["b_off"][sl]
        na = oracle.tile_norms(Ah, a_off, np.multiply.outer(ad, cd).reshape(-1)).reshape(g, g)
        nb = oracle.tile_norms(Bh, b_off, np.multiply.outer(cd, bd).reshape(-1)).reshape(g, g)
        sa, _ = oracle.scale_tile_norms(na, [ad, cd], thr)
        sb, _ = oracle.scale_tile_norms(nb, [cd, bd], thr)
        est, _ = oracle.shape_gemm(sa, sb, 0, 0, g, g, g, oracle.k_sizes([cd]), 1.0, thr)
        est_all.append(est.reshape(-1))
        ao = np.where(sa.reshape(-1) >= thr, a_off, -1)
        bo = np.where(sb.reshape(-1) >= thr, b_off, -1)
        part, _ = oracle.contract(0, 0, ad, bd, cd, Ah, ao, Bh, bo, c_off[sl], c.c_capacity)
        for t in range(g * g):
            o = c_off[sl][t]
            if o >= 0:
                n = ad[t // g] * bd[t % g]
                Cref[o:o + n] = part[o:o + n]
    return Cref, np.concatenate(est_all), c_off


def used_mask(w, c, c_off):
    g = w["grid"]
    m = np.zeros(c.c_capacity, bool)
    for b in range(w["nbatch"]):
        for t in range(g * g):
            o = c_off[b * g * g + t]
            if o >= 0:
                m[o:o + w["a_dims"][b][t // g] * w["b_dims"][b][t % g]] = True
    return m


def test_tot_fp64_integer_exact_and_shapes(ctx):
    w, c = make(ctx, 24, torch.float64, integer=True)
    c.run()
    torch.cuda.synchronize()
    Cref, est, c_off = oracle_batch(w, c)
    np.testing.assert_array_equal(c.c_norms.cpu().numpy(), est)
    np.testing.assert_array_equal(c_off >= 0, est >= c.threshold)
    m = used_mask(w, c, c_off)
    np.testing.assert_array_equal(c.C.cpu().numpy()[m], Cref[m])


def test_tot_fp32_tolerance(ctx):
    w, c = make(ctx, 40, torch.float32, seed=8)
    c.run()
    torch.cuda.synchronize()
    Cref, est, c_off = oracle_batch(w, c)
    np.testing.assert_array_equal(c_off >= 0, est >= c.threshold)
    got = c.C.cpu().numpy().astype(np.float64)
    g = w["grid"]
    for b in range(w["nbatch"]):
        idx = []
        for t in range(g * g):
            o = c_off[b * g * g + t]
            if o >= 0:
                idx.append(np.arange(o, o + w["a_dims"][b][t // g] * w["b_dims"][b][t % g]))
        if not idx:
            continue
        idx = np.concatenate(idx)
        r = Cref[idx].astype(np.float64)
        assert np.linalg.norm(got[idx] - r) <= 1e-5 * np.linalg.norm(r), b


def test_tot_graph_replays_match_eager(ctx):
    w, c = make(ctx, 64, torch.float32, seed=9)
    graph = c.capture()
    for rep in range(3):
        # eager work between replays: new operand data, then the eager result
        w["A"].mul_(1.5 if rep % 2 == 0 else -0.5)
        w["B"].add_(0.25)
        torch.cuda.synchronize()
        c.run()
        torch.cuda.synchronize()
        eager = c.C.clone()
        eager_off = c.c_off.clone()
        c.C.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(c.c_off, eager_off)
        m = torch.as_tensor(used_mask(w, c, eager_off.cpu().numpy()), device=c.C.device)
        assert torch.equal(c.C[m], eager[m]), rep


def test_tot_empty_instances(ctx):
    # instances with no cells at all produce no tasks
    w, c = make(ctx, 8, torch.float64, integer=True, fill=0.0)
    c.run()
    torch.cuda.synchronize()
    cnt = c.counters()
    assert cnt["ntasks"] == 0 and cnt["npairs"] == 0
    assert (c.c_off.cpu().numpy() < 0).all()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fused_mode_product_integer_exact(ctx, dtype):
    # C("n,i,j") = 2 A("n,i,k") * B("n,k,j") with n a Hadamard mode (TA's
    # batched Summa): integer tiles, some absent, ragged tilings of every
    # mode; the result equals the per-element product (numpy, exact on
    # integers) everywhere, zero where no pair contributes
    from tiledarray_amd import DistArray, TiledRange, TiledRange1
    rng = np.random.default_rng(1500)
    trn = TiledRange1([0, 3, 7, 8])
    tri = TiledRange1([0, 20, 45])
    trk = TiledRange1([0, 17, 40, 41])
    trj = TiledRange1([0, 33, 50])
    ma = rng.random((3, 2, 3)) < 0.7
    mb = rng.random((3, 3, 2)) < 0.7
    A = DistArray.allocate(TiledRange([trn, tri, trk]), ma, dtype=dtype)
    B = DistArray.allocate(TiledRange([trn, trk, trj]), mb, dtype=dtype)
    for X in (A, B):
        X.pool.copy_(torch.as_tensor(rng.integers(-9, 10, X.pool.numel()), dtype=dtype))
    c = E.FusedContraction(ctx, A, B, alpha=2.0)
    c.run()
    R = c.result()
    torch.cuda.synchronize()
    ref = 2.0 * np.einsum("nik,nkj->nij", A.to_dense().astype(np.float64),
                          B.to_dense().astype(np.float64))
    got = R.to_dense().astype(np.float64)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got, ref)
    # pairs: per element, the k tiles where both slices are stored
    t = np.repeat(np.arange(3), trn.extents)
    expect = int(sum((ma[t[n]][:, :, None] & mb[t[n]][None, :, :]).sum() for n in range(8)))
    assert c.counters()["npairs"] == expect


@pytest.mark.parametrize("target,left,right", [
    ("n,i,j", "n,i,k", "n,k,j"),
    ("i,n,j", "i,k,n", "j,n,k"),
    ("j,i,n", "k,n,i", "n,j,k"),
    ("n,m,i,j", "n,i,k,m", "m,k,n,j"),     # two fused modes
    ("n,i,j", "n,p,i,q", "q,n,p,j"),       # two contracted modes
    ("i,n", "n,i,k", "k,n"),               # no free mode in B
    ("n,i,j", "i,n", "n,j"),               # no contracted mode (Hadamard outer product)
])
def test_fused_contract_any_order(ctx, target, left, right):
    # fused, contracted and free indices anywhere in the operands and the
    # result, one or more of each (none for the contracted or a free group):
    # operands and result permuted by the permute kernel around the fused
    # product; exact against numpy on integers
    from tiledarray_amd import DistArray, TiledRange, TiledRange1
    rng = np.random.default_rng(1600)
    tr = {"n": TiledRange1([0, 2, 5]), "i": TiledRange1([0, 20, 45]),
          "k": TiledRange1([0, 17, 40, 41]), "j": TiledRange1([0, 33, 50]),
          "m": TiledRange1([0, 1, 3]), "p": TiledRange1([0, 6, 13]),
          "q": TiledRange1([0, 4, 5, 11])}
    L, R_, T = left.split(","), right.split(","), target.split(",")
    A = DistArray.allocate(TiledRange([tr[x] for x in L]),
                           rng.random(tuple(tr[x].ntiles for x in L)) < 0.8)
    B = DistArray.allocate(TiledRange([tr[x] for x in R_]),
                           rng.random(tuple(tr[x].ntiles for x in R_)) < 0.8)
    for X in (A, B):
        X.pool.copy_(torch.as_tensor(rng.integers(-9, 10, X.pool.numel()), dtype=torch.float64))
    R, cnt = E.fused_contract(ctx, target, A, left, B, right, alpha=-1.0)
    torch.cuda.synchronize()
    ref = -np.einsum(f"{''.join(L)},{''.join(R_)}->{''.join(T)}", A.to_dense(), B.to_dense())
    got = R.to_dense()
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got, ref)


def test_fused_contract_sparse_tile_screening(ctx):
    # the fused product with screened shapes: the result estimate is the
    # shape GEMM of each fused tile-slab (SparseShape::gemm_batched), so a
    # result tile exists exactly when the oracle's per-slab estimate reaches
    # the threshold, and its values are the product over the operand tiles
    # the shapes keep (exact on integers scaled by powers of two)
    from tiledarray_amd import DistArray, TiledRange, TiledRange1
    rng = np.random.default_rng(1700)
    thr = 1e-6
    trn, tri = TiledRange1([0, 2, 5, 6]), TiledRange1([0, 20, 45, 51])
    trk, trj = TiledRange1([0, 17, 40, 41]), TiledRange1([0, 33, 50])
    A = DistArray.allocate(TiledRange([trn, tri, trk]), rng.random((3, 3, 3)) < 0.9)
    B = DistArray.allocate(TiledRange([trn, trk, trj]), rng.random((3, 3, 2)) < 0.9)
    for X in (A, B):
        X.pool.copy_(torch.as_tensor(rng.integers(-9, 10, X.pool.numel()), dtype=torch.float64))
        # some tiles scaled far down: their shapes fall below the threshold
        vols = X.trange.tile_volumes()
        for t in np.nonzero(X.offsets >= 0)[0]:
            if rng.random() < 0.35:
                o = X.offsets[t]
                X.pool[o:o + vols[t]] *= 2.0 ** -40
    E.make_sparse(ctx, A, thr)
    E.make_sparse(ctx, B, thr)
    f = E.FusedContraction(ctx, A, B, threshold=thr)
    f.run()
    R = f.result()
    torch.cuda.synchronize()
    sa = A.shape_norms.cpu().numpy().reshape(3, 3, 3)
    sb = B.shape_norms.cpu().numpy().reshape(3, 3, 2)
    kz = oracle.k_sizes([trk.extents])
    present = np.zeros((3, 3, 2), bool)
    for h in range(3):
        est, _ = oracle.shape_gemm(sa[h], sb[h], 0, 0, 3, 2, 3, kz, 1.0, thr)
        present[h] = est >= thr
    assert 0 < present.sum() < present.size
    assert np.array_equal(R.offsets.reshape(3, 3, 2) >= 0, present)
    # reference over the operand tiles the shapes keep
    Ad, Bd = A.to_dense(), B.to_dense()
    for X, D, s, tr3 in ((A, Ad, sa, (trn, tri, trk)), (B, Bd, sb, (trn, trk, trj))):
        for idx in zip(*np.nonzero(s < thr)):
            sl = tuple(slice(r.boundaries[a], r.boundaries[a + 1]) for r, a in zip(tr3, idx))
            D[sl] = 0.0
    ref = np.einsum("nik,nkj->nij", Ad, Bd)
    got = R.to_dense()
    for h, i, j in zip(*np.nonzero(present)):
        sl = (slice(trn.boundaries[h], trn.boundaries[h + 1]),
              slice(tri.boundaries[i], tri.boundaries[i + 1]),
              slice(trj.boundaries[j], trj.boundaries[j + 1]))
        np.testing.assert_array_equal(got[sl], ref[sl])


@pytest.mark.parametrize("precision", ["i8x3", "bf16x3"])
def test_tot_split_precision_integer_exact(ctx, precision):
    # the batched launch with split-precision GEMMs: every cell is one tile of
    # the limb split (ordinals as in the work list's pairs); integer cells
    # whose entries fit the first limb reproduce the oracle bit for bit
    w = W.tot_batch(nbatch=24, dtype=torch.float64, integer=True, seed=9, fill=0.5)
    g = w["grid"]
    c = E.BatchedContraction(ctx, 24, g, g, g, w["a_dims"], w["b_dims"], w["c_dims"],
                             w["A"], w["a_off"], w["B"], w["b_off"], precision=precision)
    c.run()
    torch.cuda.synchronize()
    Cref, est, c_off = oracle_batch(w, c)
    np.testing.assert_array_equal(c_off >= 0, est >= c.threshold)
    m = used_mask(w, c, c_off)
    assert m.any()
    np.testing.assert_array_equal(c.C.cpu().numpy()[m], Cref[m])


def test_fused_contract_split_precision(ctx):
    # the fused product through the limb split (5 INT8 limbs) on random data,
    # against the FP64 fused product: within the five-limb error bound
    from tiledarray_amd import DistArray, TiledRange, TiledRange1
    rng = np.random.default_rng(1800)
    tr = {"n": TiledRange1([0, 2, 5]), "i": TiledRange1([0, 40, 90]),
          "k": TiledRange1([0, 70, 150]), "j": TiledRange1([0, 64, 100])}
    A = DistArray.allocate(TiledRange([tr[x] for x in "nik"]))
    B = DistArray.allocate(TiledRange([tr[x] for x in "nkj"]))
    for X in (A, B):
        X.pool.copy_(torch.as_tensor(rng.uniform(-1, 1, X.pool.numel())))
    R64, _ = E.fused_contract(ctx, "n,i,j", A, "n,i,k", B, "n,k,j")
    R8, _ = E.fused_contract(ctx, "n,i,j", A, "n,i,k", B, "n,k,j", precision="i8x5")
    R4, _ = E.fused_contract(ctx, "n,i,j", A, "n,i,k", B, "n,k,j", precision="i8x4")
    torch.cuda.synchronize()
    ref = R64.to_dense()
    err = np.linalg.norm(R8.to_dense() - ref) / np.linalg.norm(ref)
    err4 = np.linalg.norm(R4.to_dense() - ref) / np.linalg.norm(ref)
    # five limbs meet the 1e-10 target; one limb fewer drops 7 bits, so four
    # limbs must be measurably worse (the test tells limb counts apart)
    assert 0 < err <= 1e-10, err
    assert err4 >= 16 * err, (err4, err)


def test_fused_contract_screened_split_matches_fp64_pattern(ctx):
    # per-tile screening with the limb split: the same result tiles as the
    # screened FP64 product, values within the five-limb bound
    from tiledarray_amd import DistArray, TiledRange, TiledRange1
    rng = np.random.default_rng(1900)
    thr = 1e-6
    tr = {"n": TiledRange1([0, 2, 5]), "i": TiledRange1([0, 40, 90, 100]),
          "k": TiledRange1([0, 70, 150]), "j": TiledRange1([0, 64, 100])}
    A = DistArray.allocate(TiledRange([tr[x] for x in "kni"]))
    B = DistArray.allocate(TiledRange([tr[x] for x in "njk"]))
    for X in (A, B):
        X.pool.copy_(torch.as_tensor(rng.uniform(-1, 1, X.pool.numel())))
        vols = X.trange.tile_volumes()
        for t in range(X.trange.ntiles):
            if rng.random() < 0.3:
                o = X.offsets[t]
                X.pool[o:o + vols[t]] *= 1e-9
    E.make_sparse(ctx, A, thr)
    E.make_sparse(ctx, B, thr)
    R64, c64 = E.fused_contract(ctx, "i,n,j", A, "k,n,i", B, "n,j,k", threshold=thr)
    R8, c8 = E.fused_contract(ctx, "i,n,j", A, "k,n,i", B, "n,j,k", threshold=thr,
                              precision="i8x5")
    R4, _ = E.fused_contract(ctx, "i,n,j", A, "k,n,i", B, "n,j,k", threshold=thr,
                             precision="i8x4")
    torch.cuda.synchronize()
    assert c64["npairs"] == c8["npairs"] and c64["ntasks"] == c8["ntasks"]
    assert np.array_equal(R64.offsets >= 0, R8.offsets >= 0)
    assert 0 < (R64.offsets >= 0).sum() < R64.offsets.size
    # the screened results carry their shapes, equal to the FP64 product's
    assert R8.shape_norms is not None and R8.shape_threshold == R64.shape_threshold
    assert torch.equal(R8.shape_norms, R64.shape_norms)
    ref = R64.to_dense()
    err = np.linalg.norm(R8.to_dense() - ref) / np.linalg.norm(ref)
    err4 = np.linalg.norm(R4.to_dense() - ref) / np.linalg.norm(ref)
    assert 0 < err <= 1e-10, err
    assert err4 >= 16 * err, (err4, err)
