"""The contraction path at the benchmark workloads' full sizes.

Block-sparse 16384^3 (tile widths U[64,256]): the oracle runs the whole
product on the host (BLAS per tile pair), so the FP64 result is checked to
1e-13 and the split precisions to their stated bounds, with the tile-pair
count equal.  Ladder o=32, v=256 (V = 34 GB): too large for the host oracle,
so size-independent properties are checked instead, each bit for bit:
- the GEMM reading V through the index map equals the GEMM on a permuted copy;
- the limb split reading V through the index map equals the split of a copy;
- scaling V by 2 doubles R exactly (FP64, and the split, whose per-tile scale
  moves by one binade while the limbs stay the same).
"""
import pytest
import torch

from tiledarray_amd import engine as E
from tiledarray_amd import workloads as W

from test_gpu_contract import c_off_mask, oracle_contract, rel_fro

pytestmark = pytest.mark.gpu


def present(X):
    s = X.shape_norms.cpu().numpy()
    return (X.offsets >= 0) & (s >= X.shape_threshold)


@pytest.fixture(scope="module")
def blocksparse_full(ctx):
    thr = 1e-8
    A, B, _ = W.blocksparse_gemm(n=16384, wlo=64, whi=256)
    E.make_sparse(ctx, A, thr)
    E.make_sparse(ctx, B, thr)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=thr)
    c.run()
    torch.cuda.synchronize()
    Cref, npairs, c_off = oracle_contract(c, A, B, present(A), present(B))
    mask = c_off_mask(c_off, c)
    yield A, B, c, Cref, npairs, mask
    del A, B, c
    torch.cuda.empty_cache()


def test_blocksparse_full_fp64(blocksparse_full):
    A, B, c, Cref, npairs, mask = blocksparse_full
    assert c.counters()["npairs"] == npairs
    err = rel_fro(c.C.cpu().numpy()[mask], Cref[mask])
    assert err <= 1e-13, err


@pytest.mark.parametrize("precision,tol", [("bf16x4", 2e-10), ("i8x5", 2e-10), ("i8x4", 3e-7)])
def test_blocksparse_full_split(ctx, blocksparse_full, precision, tol):
    A, B, c, Cref, npairs, mask = blocksparse_full
    cs = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=c.threshold,
                       precision=precision)
    cs.run()
    torch.cuda.synchronize()
    assert cs.counters()["npairs"] == npairs
    assert torch.equal(cs.c_off, c.c_off)
    err = rel_fro(cs.C.cpu().numpy()[mask], Cref[mask])
    assert err <= tol, (precision, err)
    del cs
    torch.cuda.empty_cache()


def _run(ctx, T, V, **kw):
    c = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", threshold=1e-10, **kw)
    c.run()
    torch.cuda.synchronize()
    return c


def _same_result(c1, c2):
    assert torch.equal(c1.c_off, c2.c_off)
    mask = torch.as_tensor(c_off_mask(c1.c_off.cpu().numpy(), c1), device=c1.C.device)
    return torch.equal(c1.C[:mask.numel()][mask], c2.C[:mask.numel()][mask])


def test_ladder_full_index_maps_and_linearity(ctx):
    T, V, _ = W.ladder(o=32, v=256, otiles=2, wlo=48, whi=96)
    E.make_sparse(ctx, T, 1e-10)
    E.make_sparse(ctx, V, 1e-10)
    # FP64: gathered GEMM = GEMM on the permuted copy
    cc = _run(ctx, T, V, fuse_permute=False)
    cg = _run(ctx, T, V, fuse_permute=True)
    assert cg.b_gather is not None and hasattr(cc.B, "_perm_plan")
    assert _same_result(cc, cg)
    del cc
    torch.cuda.empty_cache()
    # the limb split through the index map = the split of the copy
    sm = _run(ctx, T, V, precision="i8x4")
    sc = _run(ctx, T, V, precision="i8x4", fuse_permute=False)
    assert hasattr(sm.B, "_view_src") and hasattr(sc.B, "_perm_plan")
    assert torch.equal(sm.b_split.limbs, sc.b_split.limbs)
    assert _same_result(sm, sc)
    del sc
    torch.cuda.empty_cache()
    # linearity: V -> 2V doubles R exactly, FP64 and split (the shapes are
    # kept, so the tile pattern is the same)
    R1, S1 = cg.C.clone(), sm.C.clone()
    V.pool.mul_(2.0)
    cg.run()
    sm.run()
    torch.cuda.synchronize()
    mask = torch.as_tensor(c_off_mask(cg.c_off.cpu().numpy(), cg), device=cg.C.device)
    assert torch.equal(cg.C[:mask.numel()][mask], 2.0 * R1[:mask.numel()][mask])
    assert torch.equal(sm.C[:mask.numel()][mask], 2.0 * S1[:mask.numel()][mask])
    del cg, sm, T, V
    torch.cuda.empty_cache()
