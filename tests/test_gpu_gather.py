"""Gathered grouped GEMM (tamx_grouped_gemm_gather_f64 / _f32): an operand
that needs a general permutation is read through a per-tile index map of its
stored tiles instead of a permuted copy.

Bars: the MFMA order does not depend on how the operands are staged, so a
gathered contraction equals the one on a permuted copy bit for bit (every
compiled tile shape); integer-valued data reproduces the oracle exactly for
general orders on either side and on both sides at once, in FP64 and FP32.
"""
import ctypes

import numpy as np
import pytest
import torch

from tiledarray_amd import DistArray, TiledRange, TiledRange1, _lib
from tiledarray_amd import engine as E
from tiledarray_amd import workloads as W

pytestmark = pytest.mark.gpu


def integer_array(rng, ranges, dtype=torch.float64, fill=0.9, lo=-8, hi=9):
    tr = TiledRange(ranges)
    X = DistArray.allocate(tr, rng.random(tr.ntiles) < fill, dtype=dtype)
    X.pool.copy_(torch.as_tensor(rng.integers(lo, hi, X.pool.numel()), dtype=dtype))
    return X


def einsum_ref(target, A, left, B, right):
    spec = f"{left.replace(',', '')},{right.replace(',', '')}->{target.replace(',', '')}"
    return np.einsum(spec, A.to_dense().astype(np.float64), B.to_dense().astype(np.float64))


@pytest.mark.parametrize("annot", ["a,c,b,d", "c,a,d,b", "b,d,a,c", "a,d,c,b"])
@pytest.mark.parametrize("side", ["right", "left"])
def test_gather_integer_exact_general_orders(ctx, annot, side):
    rng = np.random.default_rng(sum(map(ord, annot + side)))
    tv = TiledRange1([0, 5, 12, 20, 21])
    to = TiledRange1([0, 3, 7])
    T = integer_array(rng, [tv, tv, to, to])
    V = integer_array(rng, [tv, tv, tv, tv])
    if side == "right":
        args = ("a,b,i,j", T, "c,d,i,j", V, annot)
    else:
        args = ("i,j,a,b", V, annot, T, "c,d,i,j")
    c = E.Contraction(ctx, *args, fuse_permute=True)
    gathered = c.b_gather if side == "right" else c.a_gather
    assert gathered is not None
    c.run()
    np.testing.assert_array_equal(c.result().to_dense(), einsum_ref(*args))


def test_gather_both_operands_general(ctx):
    rng = np.random.default_rng(5)
    tv = TiledRange1([0, 6, 13, 19])
    to = TiledRange1([0, 4, 9])
    T = integer_array(rng, [tv, to, tv, to])  # stored (c,i,d,j)
    V = integer_array(rng, [tv, tv, tv, tv])  # stored (a,c,b,d)
    args = ("a,b,i,j", T, "c,i,d,j", V, "a,c,b,d")
    c = E.Contraction(ctx, *args, fuse_permute=True)
    assert c.a_gather is not None and c.b_gather is not None
    c.run()
    np.testing.assert_array_equal(c.result().to_dense(), einsum_ref(*args))


def test_gather_fp32_integer_exact(ctx):
    rng = np.random.default_rng(8)
    tv = TiledRange1([0, 7, 16, 24])
    to = TiledRange1([0, 5, 8])
    T = integer_array(rng, [tv, tv, to, to], dtype=torch.float32, lo=-4, hi=5)
    V = integer_array(rng, [tv, tv, tv, tv], dtype=torch.float32, lo=-4, hi=5)
    args = ("a,b,i,j", T, "c,d,i,j", V, "a,c,b,d")
    c = E.Contraction(ctx, *args, precision="f32", fuse_permute=True)
    assert c.b_gather is not None
    c.run()
    np.testing.assert_array_equal(c.result().to_dense(), einsum_ref(*args))


GATHER_TILES = [(64, 64, 16), (128, 64, 16), (64, 128, 16)]


@pytest.mark.parametrize("tile", GATHER_TILES)
def test_gather_equals_permuted_copy_every_tile(ctx, tile):
    # random FP64 data: the gathered GEMM and the GEMM on a permuted copy
    # run the same MFMAs in the same order
    T, V, _ = W.ladder(o=8, v=40, otiles=2, wlo=9, whi=17, seed=6)
    cg = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", tile=tile,
                       fuse_permute=True)
    cp = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", tile=tile,
                       fuse_permute=False)
    assert cg.b_gather is not None and cp.b_gather is None and hasattr(cp.B, "_perm_plan")
    cg.run()
    cp.run()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(cg.result().to_dense(), cp.result().to_dense())


def test_gather_other_tiles_unsupported(ctx):
    T, V, _ = W.ladder(o=6, v=24, otiles=2, wlo=7, whi=12, seed=2)
    c = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", tile=(32, 32, 16),
                      fuse_permute=True)
    c.permute_step()
    c.shape_step()
    c.worklist_step()
    with pytest.raises(_lib.TamxError) as err:
        c.gemm_step()
    assert err.value.status == _lib.TAMX_ERR_UNSUPPORTED


def test_default_layout_choice(ctx):
    # the FP64 GEMM permutes a general operand into a copy by default, the
    # limb split reads it through the index map
    T, V, _ = W.ladder(o=6, v=24, otiles=2, wlo=7, whi=12, seed=2)
    c = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d")
    assert c.b_gather is None and hasattr(c.B, "_perm_plan")
    c = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", precision="bf16x3")
    assert hasattr(c.B, "_view_src")


def test_gather_ladder_against_oracle_tolerance(ctx):
    T, V, _ = W.ladder(o=8, v=48, otiles=2, wlo=10, whi=20, seed=9)
    c = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", fuse_permute=True)
    c.run()
    R = c.result().to_dense()
    # V(a,c,b,d) T(c,d,i,j) summed over c, d gives (a,b,i,j)
    ref = np.tensordot(V.to_dense(), T.to_dense(), axes=([1, 3], [0, 1]))
    err = np.linalg.norm(R - ref) / np.linalg.norm(ref)
    assert err <= 1e-14, err


def test_gather_autotune_keeps_c_and_matches(ctx):
    T, V, _ = W.ladder(o=8, v=40, otiles=2, wlo=9, whi=17, seed=12)
    c = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", fuse_permute=True)
    c.run()
    torch.cuda.synchronize()
    before = c.C.clone()
    tile, ms = c.tune(reps=1, force=True)
    torch.cuda.synchronize()
    assert torch.equal(c.C, before) and ms > 0.0
    assert tile[:3] in GATHER_TILES
    c.gemm_step()
    torch.cuda.synchronize()
    assert torch.equal(c.C, before)


def test_gather_plan_rejects_bad_maps(ctx):
    p = ctypes.c_void_p()
    ext = np.array([4, 4], np.int32)
    st = np.array([4, 1], np.int64)
    L = _lib.lib()
    # five modes
    assert L.tamx_gather_plan_create(ctx.ptr, 1, 5, 1, ext.ctypes.data, st.ctypes.data,
                                     ext.ctypes.data, st.ctypes.data,
                                     ctypes.byref(p)) == _lib.TAMX_ERR_BAD_ARG
    # negative stride
    bad = np.array([4, -1], np.int64)
    assert L.tamx_gather_plan_create(ctx.ptr, 1, 2, 2, ext.ctypes.data, bad.ctypes.data,
                                     ext.ctypes.data, st.ctypes.data,
                                     ctypes.byref(p)) == _lib.TAMX_ERR_BAD_ARG
    # offsets past 31 bits
    big = np.array([1 << 20, 1 << 12], np.int32)
    bst = np.array([1 << 12, 1], np.int64)
    assert L.tamx_gather_plan_create(ctx.ptr, 1, 2, 2, big.ctypes.data, bst.ctypes.data,
                                     ext.ctypes.data, st.ctypes.data,
                                     ctypes.byref(p)) == _lib.TAMX_ERR_BAD_ARG
    assert L.tamx_gather_plan_create(ctx.ptr, 1, 2, 2, ext.ctypes.data, st.ctypes.data,
                                     ext.ctypes.data, st.ctypes.data, ctypes.byref(p)) == 0
    assert L.tamx_gather_plan_destroy(p) == 0
