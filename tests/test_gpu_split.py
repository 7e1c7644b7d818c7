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

# relative Frobenius error bounds per limb count on U[-1,1] / Gaussian data
# (a limb carries ~8 bits; the bound leaves ~30x headroom over what the
# expansion's truncation gives)
SPLIT_TOL = {2: 3e-3, 3: 3e-5, 4: 1e-7, 5: 1e-10, 6: 1e-12}


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


@pytest.mark.parametrize("nslices", [3, 5])
def test_split_ladder_error_bound(ctx, nslices):
    T, V, _ = W.ladder(o=8, v=40, otiles=2, wlo=9, whi=17, seed=3)
    c = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", precision=f"bf16x{nslices}")
    c.run()
    R = c.result()
    ref = np.einsum("cdij,acbd->abij", T.to_dense(), V.to_dense())
    err = rel_fro(R.to_dense(), ref)
    assert err <= SPLIT_TOL[nslices], (nslices, err)


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
