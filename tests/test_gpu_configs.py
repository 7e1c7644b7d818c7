"""The dense and batched benchmark workloads at their stated sizes, against
the oracle.

Dense C("m,n") = A("m,k") * B("k,n"), FP64, M = N = K = 2048 in 128-wide
tiles (16 per mode, 4096 tile pairs): integer-valued inputs reproduce the
oracle bit for bit; U[-1,1] inputs reach relative Frobenius error <= 1e-13.

1024 independent tensor-of-tensor products C_n(i,j;a,b) = A_n(i,k;a,c) *
B_n(k,j;c,b) (8x8 outer grids, half the cells present, inner extents
U[16,96]), FP32 and FP64, in one persistent grouped launch captured in a
hipGraph: patterns bit-exact per instance, shapes bit-exact for FP64 cells
and within float rounding for FP32 cells; FP64 integer cells
bit-exact, FP64 U[-1,1] cells <= 1e-13 and FP32 cells <= 1e-5 relative
Frobenius error per instance; every replay of the graph equals the eager
result bit for bit.
"""
import numpy as np
import pytest
import torch

from tiledarray_amd import engine as E
from tiledarray_amd import workloads as W

from test_gpu_contract import c_off_mask, oracle_contract, rel_fro
from test_gpu_tot import oracle_batch, used_mask

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("integer", [True, False])
def test_dense_2048_against_oracle(ctx, integer):
    A, B, info = W.dense_gemm(n=2048, tile=128, integer=integer)
    assert A.trange.tiles_shape == (16, 16)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n")
    c.run()
    torch.cuda.synchronize()
    Cref, npairs, c_off = oracle_contract(c, A, B)
    cnt = c.counters()
    assert npairs == cnt["npairs"] == 4096
    assert cnt["ntasks"] == 256 and cnt["flops"] == 2 * 2048 ** 3
    mask = c_off_mask(c_off, c)
    assert mask.sum() == 2048 * 2048
    got = c.C.cpu().numpy()
    if integer:
        np.testing.assert_array_equal(got[mask], Cref[mask])
    else:
        err = rel_fro(got[mask], Cref[mask])
        assert err <= 1e-13, err
    # the result as a tiled array in target order equals the dense product
    R = c.result()
    ref = A.to_dense() @ B.to_dense()
    if integer:
        np.testing.assert_array_equal(R.to_dense(), ref)
    else:
        assert rel_fro(R.to_dense(), ref) <= 1e-13


def _per_instance_errors(w, c, Cref, c_off):
    """Relative Frobenius error of every instance's result cells."""
    got = c.C.cpu().numpy().astype(np.float64)
    g = w["grid"]
    errs = []
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
        errs.append(np.linalg.norm(got[idx] - r) / np.linalg.norm(r))
    return np.asarray(errs)


@pytest.mark.parametrize("dtype,integer,tol", [(torch.float64, True, 0.0),
                                               (torch.float64, False, 1e-13),
                                               (torch.float32, False, 1e-5)])
def test_tot_1024_graph_against_oracle(ctx, dtype, integer, tol):
    w = W.tot_batch(nbatch=1024, dtype=dtype, integer=integer)
    g = w["grid"]
    c = E.BatchedContraction(ctx, 1024, g, g, g, w["a_dims"], w["b_dims"], w["c_dims"],
                             w["A"], w["a_off"], w["B"], w["b_off"])
    graph = c.capture()
    c.C.zero_()
    c.c_norms.zero_()
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    Cref, est, c_off = oracle_batch(w, c)
    # the result pattern bit for bit; the shape estimate bit for bit for
    # FP64 cells (norms summed in double, then cast to float), within float
    # rounding for FP32 cells (TA sums their squares in float, sequentially,
    # tensor/tensor.h:3768-3776; the device sums in float in another order)
    if dtype == torch.float64:
        np.testing.assert_array_equal(c.c_norms.cpu().numpy(), est)
    else:
        np.testing.assert_allclose(c.c_norms.cpu().numpy(), est, rtol=1e-5, atol=0)
    np.testing.assert_array_equal(c_off >= 0, est >= c.threshold)
    cnt = c.counters()
    assert cnt["ntasks"] == int((c_off >= 0).sum()) > 1024 * 8
    m = used_mask(w, c, c_off)
    if integer:
        np.testing.assert_array_equal(c.C.cpu().numpy()[m], Cref[m])
    else:
        errs = _per_instance_errors(w, c, Cref, c_off)
        assert errs.size > 1000 and errs.max() <= tol, float(errs.max())
    # further replays, with eager work and synchronisations between them,
    # reproduce the replayed result bit for bit
    first = c.C.clone()
    for rep in range(2):
        c.C.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(c.C, first), rep
