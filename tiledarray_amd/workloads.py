"""Synthetic inputs of the contraction benchmarks, generated on the device.

Tilings and sparsity masks come from numpy's PCG64 seeded per workload (so
they are identical on every machine); tile data comes from torch's device
generator seeded the same way.  Each builder returns DistArrays resident in
HBM plus a dict describing what was generated.

Workloads (the benchmark drivers of examples/gemm/ta_dense.cpp,
ta_sparse.cpp and ta_cc_abcd.cpp set the pattern):
  dense_gemm        C("m,n") = A("m,k") * B("k,n"), uniform tiles, U[-1,1]
  blocksparse_gemm  banded block-sparse FP64 matrices with uneven tiles,
                    tile (i,j) kept w.p. exp(-|i-j|/8), entries
                    U[-1,1] * 10^(-|i-j|/4)
  ladder            R("a,b,i,j") = T("c,d,i,j") * V("a,c,b,d"), o modes in 2
                    tiles, v modes tiled unevenly, T ~ N(0, 1e-2),
                    V ~ N(0,1) * exp(-|a-c|/64)
"""
import math

import numpy as np
import torch

from .tiled import DistArray, TiledRange, TiledRange1

BASE_SEED = 0x5EED2026


def random_tiling(extent, lo, hi, rng):
    """Tile widths drawn uniformly from [lo, hi]; the last tile absorbs the
    remainder (clipped so the boundaries end exactly at `extent`)."""
    b = [0]
    while b[-1] < extent:
        w = int(rng.integers(lo, hi + 1))
        if extent - (b[-1] + w) < lo:  # fold a short remainder into this tile
            w = extent - b[-1]
        b.append(min(extent, b[-1] + w))
    return TiledRange1(b)


def _fill_uniform(pool, gen):
    pool.uniform_(-1.0, 1.0, generator=gen)


def _tile_scale_(arr, scales):
    """Multiplies every present tile t by scales[t] (host array)."""
    offs = arr.offsets
    vols = arr.trange.tile_volumes()
    present = offs >= 0
    if not present.any():
        return
    per = torch.as_tensor(np.asarray(scales, np.float64)[present], device=arr.pool.device,
                          dtype=arr.pool.dtype)
    reps = torch.as_tensor(vols[present], device=arr.pool.device)
    order = np.argsort(offs[present])  # pool order
    per = per[torch.as_tensor(order, device=arr.pool.device)]
    reps = reps[torch.as_tensor(order, device=arr.pool.device)]
    arr.pool.mul_(torch.repeat_interleave(per, reps))


def dense_gemm(n=2048, tile=128, seed=BASE_SEED, dtype=torch.float64, device="cuda",
               integer=False):
    tr = TiledRange1.uniform(n, tile)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    A = DistArray.allocate(TiledRange([tr, tr]), dtype=dtype, device=device)
    B = DistArray.allocate(TiledRange([tr, tr]), dtype=dtype, device=device)
    for X in (A, B):
        if integer:
            X.pool.copy_(torch.randint(-8, 9, X.pool.shape, generator=gen, device=device))
        else:
            _fill_uniform(X.pool, gen)
    return A, B, {"M": n, "N": n, "K": n, "tile": tile, "seed": seed}


def blocksparse_gemm(n=16384, wlo=64, whi=256, band=8.0, decay=4.0, seed=BASE_SEED,
                     dtype=torch.float64, device="cuda", integer=False):
    rng = np.random.default_rng(seed)
    tr = random_tiling(n, wlo, whi, rng)
    nt = tr.ntiles
    i, j = np.meshgrid(np.arange(nt), np.arange(nt), indexing="ij")
    dist = np.abs(i - j)
    prob = np.exp(-dist / band)
    mask_a = rng.random((nt, nt)) < prob
    mask_b = rng.random((nt, nt)) < prob
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    A = DistArray.allocate(TiledRange([tr, tr]), mask_a, dtype=dtype, device=device)
    B = DistArray.allocate(TiledRange([tr, tr]), mask_b, dtype=dtype, device=device)
    scale = 10.0 ** (-dist.reshape(-1) / decay)
    for X in (A, B):
        if integer:
            X.pool.copy_(torch.randint(-8, 9, X.pool.shape, generator=gen, device=device))
        else:
            _fill_uniform(X.pool, gen)
            _tile_scale_(X, scale)
    info = {"M": n, "N": n, "K": n, "tiles_per_mode": nt, "widths": [wlo, whi],
            "fill_a": float(mask_a.mean()), "fill_b": float(mask_b.mean()), "seed": seed}
    return A, B, info


def ladder(o=32, v=256, otiles=2, wlo=48, whi=96, seed=BASE_SEED, device="cuda",
           dtype=torch.float64, v_order="acbd"):
    """T("c,d,i,j") and V stored in `v_order` ("acbd" forces a general
    permutation of V before the GEMM)."""
    rng = np.random.default_rng(seed)
    trv = random_tiling(v, wlo, whi, rng)
    tro = TiledRange1.uniform(o, (o + otiles - 1) // otiles)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    T = DistArray.allocate(TiledRange([trv, trv, tro, tro]), dtype=dtype, device=device)
    T.pool.normal_(0.0, math.sqrt(1e-2), generator=gen)
    modes = {"a": trv, "b": trv, "c": trv, "d": trv}
    V = DistArray.allocate(TiledRange([modes[x] for x in v_order]), dtype=dtype, device=device)
    V.pool.normal_(0.0, 1.0, generator=gen)
    # exp(-|a-c|/64) on every element: per tile, a broadcast over the other modes
    pa, pc = v_order.index("a"), v_order.index("c")
    ext = V.trange.tile_extents_table()
    shp = V.trange.tiles_shape
    fac_cache = {}
    for t in range(V.trange.ntiles):
        idx = np.unravel_index(t, shp)
        a0 = int(trv.boundaries[idx[pa]])
        c0 = int(trv.boundaries[idx[pc]])
        wa, wc = int(ext[t][pa]), int(ext[t][pc])
        key = (a0 - c0, wa, wc)
        f = fac_cache.get(key)
        if f is None:
            ga = a0 + np.arange(wa)[:, None]
            gc = c0 + np.arange(wc)[None, :]
            f = torch.as_tensor(np.exp(-np.abs(ga - gc) / 64.0), device=device, dtype=dtype)
            fac_cache[key] = f
        tile = V.tile(t)
        shape = [1] * 4
        shape[pa], shape[pc] = wa, wc
        fv = f if pa < pc else f.t()
        tile.mul_(fv.reshape(shape))
    info = {"o": o, "v": v, "v_tiles": trv.ntiles, "v_widths": trv.extents.tolist(),
            "T_sigma2": 1e-2, "v_order": v_order, "seed": seed}
    return T, V, info


def tot_batch(nbatch=1024, grid=8, fill=0.5, lo=16, hi=96, seed=BASE_SEED, dtype=torch.float32,
              device="cuda", integer=False):
    """nbatch independent tensor-of-tensor products C_n(i,j;a,b) =
    A_n(i,k;a,c) * B_n(k,j;c,b) on grid x grid outer cells, each outer cell
    present with probability `fill`; inner extents a_i, c_k, b_j drawn from
    [lo, hi] per outer index per instance.  Returns a dict of host metadata
    and device pools: cells of A stored (a_i x c_k) row-major, of B (c_k x b_j)."""
    rng = np.random.default_rng(seed)
    g = grid
    ad = rng.integers(lo, hi + 1, (nbatch, g))
    cd = rng.integers(lo, hi + 1, (nbatch, g))
    bd = rng.integers(lo, hi + 1, (nbatch, g))
    ma = rng.random((nbatch, g, g)) < fill
    mb = rng.random((nbatch, g, g)) < fill
    va = (ad[:, :, None] * cd[:, None, :]).reshape(-1)
    vb = (cd[:, :, None] * bd[:, None, :]).reshape(-1)

    def offsets(mask, vol):
        m = mask.reshape(-1)
        off = np.full(m.size, -1, np.int64)
        cs = np.concatenate([[0], np.cumsum(np.where(m, vol, 0))])
        off[m] = cs[:-1][m]
        return off, int(cs[-1])

    a_off, na = offsets(ma, va)
    b_off, nb = offsets(mb, vb)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    A = torch.empty(max(1, na), dtype=dtype, device=device)
    B = torch.empty(max(1, nb), dtype=dtype, device=device)
    for X in (A, B):
        if integer:
            X.copy_(torch.randint(-8, 9, X.shape, generator=gen, device=device))
        else:
            X.uniform_(-1.0, 1.0, generator=gen)
    return {"nbatch": nbatch, "grid": g, "a_dims": ad, "c_dims": cd, "b_dims": bd,
            "a_off": a_off, "b_off": b_off, "a_vol": va, "b_vol": vb, "A": A, "B": B,
            "fill_a": float(ma.mean()), "fill_b": float(mb.mean()), "seed": seed}
