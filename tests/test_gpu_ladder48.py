"""The ladder contraction R("a,b,i,j") = T("c,d,i,j") * V("a,c,b,d") at its
larger stated size, o = 48 (2 tiles), v = 320 (tiles U[48,96]): V holds
320^4 doubles (84 GB), beyond the host oracle as a whole.

Checked against the oracle on whole result tile columns R(a,b,.,.) of a
sample of (a,b) tiles: each needs only T and the V tiles (a,.,b,.), which
are copied to the host, permuted per tile as TA permutes on read, and
contracted by the oracle.  Per sampled column:
- T's tile norms (all of them) and the column's V tile norms bit-exact, so
  the column's shape estimate, recomputed by the oracle from them, must
  equal the GPU's estimate and pattern bit for bit;
- FP64: relative Frobenius error <= 1e-13;
- split precisions (3 BF16 limbs, 2 FP32 limbs, 4 and 5 INT8 limbs): max and
  Frobenius relative errors within the bounds of their limb expansions, the
  same result pattern and tile pairs as FP64.
Over the whole result, bit for bit: the GEMM reading V through the index map
equals the GEMM on a permuted copy, and V -> 2V doubles R.
"""
import numpy as np
import pytest
import torch

import oracle
from tiledarray_amd import engine as E
from tiledarray_amd import workloads as W

from test_gpu_contract import c_off_mask

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

THR = 1e-10
# max and Frobenius relative error bounds on the sampled columns (ladder
# tiles, K = w_c w_d up to 9216): 3 BF16 limbs carry 8 + 2 x 9 bits, 2 FP32
# limbs round their FP32 accumulation (about 24 bits), INT8 limbs 7 bits each
SPLIT_BOUND = {"bf16x3": 5e-6, "f32x2": 1e-6, "i8x4": 1e-6, "i8x5": 1e-8}


def _run(ctx, T, V, **kw):
    c = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", threshold=THR, **kw)
    c.run()
    torch.cuda.synchronize()
    return c


def _column_oracle(c, T, V, cols):
    """For result tile columns `cols` (indices into the natural result grid's
    Nt columns, i.e. (a,b) tiles): the oracle's shape estimate and result
    tiles, after checking the GPU's tile norms it depends on."""
    plan = c.plan
    rp = [int(x) for x in plan.right_perm[:4]]
    Mt, Nt, Kt = c.Mt, c.Nt, c.Kt
    assert c.op_b == 0  # V in GEMM layout is k x n: (c,d) x (a,b)
    # T: all tile norms bit-exact, scaled norms from them
    Th = T.pool.cpu().numpy()
    tvol = T.trange.tile_volumes()
    tn = oracle.tile_norms(Th, T.offsets, tvol)
    np.testing.assert_array_equal(tn, E.tile_norms(c.ctx, T).cpu().numpy())
    st, _ = oracle.scale_tile_norms(tn.reshape(T.trange.tiles_shape),
                                    [r.extents for r in T.trange.ranges], THR)
    np.testing.assert_array_equal(st.reshape(-1), T.shape_norms.cpu().numpy())
    a_off = np.where(st.reshape(-1) >= THR, T.offsets, -1)
    v_raw = E.tile_norms(c.ctx, V).cpu().numpy()
    v_scaled = V.shape_norms.cpu().numpy()
    vext = V.trange.tile_extents_table()
    bshape = c.B.trange.tiles_shape
    # contracted tile extents (c, d): the first two modes of V in GEMM layout
    ksz = oracle.k_sizes([c.B.trange.ranges[0].extents, c.B.trange.ranges[1].extents])
    b_pool, b_off, pos = [], np.full(Kt * Nt, -1, np.int64), 0
    sv = np.zeros((Kt, Nt), np.float32)
    for j in cols:
        col_tiles = []
        for k in range(Kt):
            bt = k * Nt + j
            idx_p = np.unravel_index(bt, bshape)
            src = int(np.ravel_multi_index([idx_p[rp[d]] for d in range(4)], V.trange.tiles_shape))
            off = int(V.offsets[src])
            e = [int(x) for x in vext[src]]
            tile = V.pool[off:off + int(np.prod(e))].cpu().numpy().reshape(e)
            n = oracle.tile_norm(tile.reshape(-1))
            assert n == v_raw[src], (j, k)
            col_tiles.append((bt, tile, src))
        # the column's scaled norms from the oracle's scaling of its tiles
        # (each tile's inverse-volume factor depends on its own extents only)
        for bt, tile, src in col_tiles:
            s1, _ = oracle.scale_tile_norms(
                np.array([[[[v_raw[src]]]]], np.float32),
                [np.array([x], np.float32) for x in vext[src]], THR)
            assert s1.reshape(-1)[0] == v_scaled[src]
            sv[bt // Nt, j] = v_scaled[src] if v_scaled[src] >= THR else 0.0
            if v_scaled[src] >= THR:
                pt = oracle.permute(tile, rp).reshape(-1)
                b_off[bt] = pos
                b_pool.append(pt)
                pos += pt.size
    Bh = np.concatenate(b_pool)
    # the shape estimate of the sampled columns, from the checked norms
    sT = np.where(st.reshape(-1) >= THR, st.reshape(-1), 0).reshape(Kt, Mt)
    est = {}
    for j in cols:
        e, _ = oracle.shape_gemm(sT, sv[:, j:j + 1].copy(), int(c.op_a), 0, Mt, 1, Kt, ksz,
                                 1.0, THR)
        est[j] = e.reshape(-1)
    c_off = np.full(Mt * Nt, -1, np.int64)
    vol = 0
    for j in cols:
        for i in range(Mt):
            if est[j][i] >= THR:
                c_off[i * Nt + j] = vol
                vol += int(c.mext_h[i] * c.next_h[j])
    oracle.set_num_threads(max(1, min(16, len(cols) * Mt)))
    ref, npairs = oracle.contract(c.op_a, c.op_b, c.mext_h, c.next_h, c.kext_h, Th, a_off, Bh,
                                  b_off, c_off, max(1, vol))
    tiles = {}
    for j in cols:
        for i in range(Mt):
            o = c_off[i * Nt + j]
            if o >= 0:
                m, n = int(c.mext_h[i]), int(c.next_h[j])
                tiles[(i, j)] = ref[o:o + m * n].reshape(m, n)
    return est, tiles, npairs


def _gpu_tiles(c, keys):
    """The GPU's result tiles (i, j) of the natural grid as (m x n) arrays
    (the store is transposed into the (a,b,i,j) target order)."""
    off = c.c_off.cpu().numpy()
    out = {}
    for i, j in keys:
        o = int(off[i * c.Nt + j])
        m, n = int(c.mext_h[i]), int(c.next_h[j])
        out[(i, j)] = c.C[o:o + m * n].cpu().numpy().reshape(n, m).T
    return out


def _errors(got, ref):
    num = sum(float(np.sum((got[k] - ref[k]) ** 2)) for k in ref)
    den = sum(float(np.sum(ref[k] ** 2)) for k in ref)
    mx = max(float(np.abs(got[k] - ref[k]).max()) for k in ref)
    rmax = max(float(np.abs(ref[k]).max()) for k in ref)
    return (num / den) ** 0.5, mx / rmax


@pytest.fixture(scope="module")
def ladder48(ctx):
    T, V, info = W.ladder(o=48, v=320, otiles=2, wlo=48, whi=96)
    E.make_sparse(ctx, T, THR)
    E.make_sparse(ctx, V, THR)
    c = _run(ctx, T, V, fuse_permute=False)
    assert hasattr(c.B, "_perm_plan")
    # sampled (a,b) columns: the two with the smallest w_a w_b
    cols = [int(x) for x in np.argsort(c.next_h, kind="stable")[:2]]
    print(f"ladder o=48 v=320: v tiles {info['v_widths']}, sampled columns {cols}", flush=True)
    est, ref, npairs = _column_oracle(c, T, V, cols)
    cn = c.c_norms.cpu().numpy()
    off = c.c_off.cpu().numpy()
    for j in cols:
        for i in range(c.Mt):
            assert cn[i * c.Nt + j] == est[j][i], (i, j)
            assert (off[i * c.Nt + j] >= 0) == (est[j][i] >= THR), (i, j)
    data = {"T": T, "V": V, "cols": cols, "ref": ref, "npairs_sample": npairs,
            "f64": _gpu_tiles(c, ref.keys()), "c_off": c.c_off.clone(), "C": c.C.clone(),
            "counters": c.counters(), "mask": c_off_mask(off, c)}
    del c
    torch.cuda.empty_cache()
    yield data
    data.clear()
    torch.cuda.empty_cache()


def test_ladder48_fp64_against_oracle(ladder48):
    fro, mx = _errors(ladder48["f64"], ladder48["ref"])
    print(f"FP64: rel. Frobenius {fro:.3e}, rel. max {mx:.3e} over "
          f"{len(ladder48['ref'])} result tiles", flush=True)
    assert len(ladder48["ref"]) >= 2
    assert fro <= 1e-13, fro


def test_ladder48_index_map_equals_copy_and_linearity(ctx, ladder48):
    T, V = ladder48["T"], ladder48["V"]
    cg = _run(ctx, T, V, fuse_permute=True)
    assert cg.b_gather is not None
    assert torch.equal(cg.c_off, ladder48["c_off"])
    mask = torch.as_tensor(ladder48["mask"], device=cg.C.device)
    n = mask.numel()
    assert torch.equal(cg.C[:n][mask], ladder48["C"][:n][mask])
    # linearity: V -> 2V doubles R exactly (the shape estimate moves by a
    # power of two and keeps the pattern)
    V.pool.mul_(2.0)
    try:
        cg.run()
        torch.cuda.synchronize()
        assert torch.equal(cg.c_off, ladder48["c_off"])
        assert torch.equal(cg.C[:n][mask], 2.0 * ladder48["C"][:n][mask])
    finally:
        V.pool.mul_(0.5)
        torch.cuda.synchronize()
    del cg
    torch.cuda.empty_cache()


@pytest.mark.parametrize("precision", ["bf16x3", "f32x2", "i8x4", "i8x5"])
def test_ladder48_split_precision(ctx, ladder48, precision):
    T, V = ladder48["T"], ladder48["V"]
    cs = _run(ctx, T, V, precision=precision)
    assert torch.equal(cs.c_off, ladder48["c_off"])
    assert cs.counters()["npairs"] == ladder48["counters"]["npairs"]
    got = _gpu_tiles(cs, ladder48["ref"].keys())
    fro, mx = _errors(got, ladder48["ref"])
    print(f"{precision}: rel. Frobenius {fro:.3e}, rel. max {mx:.3e}", flush=True)
    assert 0 < fro <= SPLIT_BOUND[precision], (precision, fro)
    assert mx <= 10 * SPLIT_BOUND[precision], (precision, mx)
    del cs
    torch.cuda.empty_cache()
