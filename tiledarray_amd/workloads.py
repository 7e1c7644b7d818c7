"""Synthetic inputs of the contraction benchmarks, generated on the device.

Tilings and sparsity masks come from numpy's PCG64 seeded per workload (so
they are identical on every machine).  Tile data comes from a counter-based
generator keyed by (base seed, workload, tensor, tile ordinal) and counted
by the element's ordinal inside its tile (SplitMix64; `keyed_fill`), so a
tile's values do not depend on where it lies in the pool, on the other
tiles, or on the machine.  Each builder returns DistArrays resident in HBM
plus a dict describing what was generated.

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


# workload ids in the generator key
CFG_DENSE, CFG_BLOCKSPARSE, CFG_LADDER, CFG_TOT = 1, 2, 3, 4
GENERATOR = ("SplitMix64 (Steele, Lea, Flood 2014) keyed by (seed, workload, tensor, "
             "tile ordinal), counter = element ordinal in the tile")

_M64 = (1 << 64) - 1
_GAMMA = 0x9E3779B97F4A7C15


def _s64(c):
    """An unsigned 64-bit constant as the int64 with the same bits."""
    c &= _M64
    return c - (1 << 64) if c >= (1 << 63) else c


def _mix_int(z):
    """SplitMix64's output function on a Python int (mod 2^64)."""
    z &= _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def _srl(x, k):
    """Logical right shift of int64 tensor bits."""
    return (x >> k) & ((1 << (64 - k)) - 1)


def mix64(z):
    """SplitMix64's output function on an int64 tensor (two's-complement
    wraparound gives the unsigned arithmetic)."""
    z = (z ^ _srl(z, 30)) * _s64(0xBF58476D1CE4E5B9)
    z = (z ^ _srl(z, 27)) * _s64(0x94D049BB133111EB)
    return z ^ _srl(z, 31)


def tile_key(seed, config, tensor, tile):
    """The generator state of one tile, as an unsigned int."""
    k = _mix_int(_mix_int(_mix_int(seed) ^ config) ^ tensor)
    return _mix_int((k + (tile + 1) * _GAMMA) & _M64)


def keyed_fill(pool, offsets, volumes, config, tensor, seed=BASE_SEED, kind="uniform",
               batch=1 << 26):
    """Fills the present tiles of `pool` (offsets < 0: absent): element e of
    tile t gets draw number e of the stream keyed by (seed, config, tensor,
    t), draw i being mix64(key + (i + 1) * golden gamma).  kind: "uniform"
    U[-1, 1) from the top 53 bits, "normal" N(0, 1) by Box-Muller from draws
    2e and 2e + 1, "int" integers in [-8, 8]."""
    offsets = np.asarray(offsets, np.int64)
    volumes = np.asarray(volumes, np.int64)
    present = np.nonzero((offsets >= 0) & (volumes > 0))[0]
    order = present[np.argsort(offsets[present], kind="stable")]
    dev = pool.device
    g = _s64(_GAMMA)

    def run(ts):
        t = np.asarray(ts, np.int64)
        o0 = int(offsets[t[0]])
        vol = torch.as_tensor(volumes[t], device=dev)
        n = int(volumes[t].sum())
        keys = torch.as_tensor([_s64(tile_key(seed, config, tensor, int(x))) for x in t],
                               dtype=torch.int64, device=dev)
        starts = torch.as_tensor(offsets[t] - o0, device=dev)
        e = torch.arange(n, dtype=torch.int64, device=dev) - torch.repeat_interleave(starts, vol)
        k = torch.repeat_interleave(keys, vol)
        out = pool[o0:o0 + n]
        if kind == "normal":
            z1 = mix64(k + (2 * e + 1) * g)
            z2 = mix64(k + (2 * e + 2) * g)
            u1 = (_srl(z1, 11) + 1).to(torch.float64) * 2.0 ** -53  # (0, 1]
            u2 = _srl(z2, 11).to(torch.float64) * 2.0 ** -53
            out.copy_(torch.sqrt(-2.0 * torch.log(u1)) * torch.cos(2.0 * math.pi * u2))
        else:
            z = mix64(k + (e + 1) * g)
            if kind == "int":
                out.copy_((_srl(z, 11) % 17 - 8).to(pool.dtype))
            else:
                out.copy_(_srl(z, 11).to(torch.float64) * 2.0 ** -52 - 1.0)

    # runs of tiles that lie back to back in the pool, a batch of elements
    # at a time
    run_t, run_n = [], 0
    for t in order:
        contiguous = run_t and offsets[run_t[-1]] + volumes[run_t[-1]] == offsets[t]
        if run_t and (not contiguous or run_n + volumes[t] > batch):
            run(run_t)
            run_t, run_n = [], 0
        run_t.append(int(t))
        run_n += int(volumes[t])
    if run_t:
        run(run_t)


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
    A = DistArray.allocate(TiledRange([tr, tr]), dtype=dtype, device=device)
    B = DistArray.allocate(TiledRange([tr, tr]), dtype=dtype, device=device)
    for w, X in enumerate((A, B)):
        keyed_fill(X.pool, X.offsets, X.trange.tile_volumes(), CFG_DENSE, w, seed,
                   "int" if integer else "uniform")
    return A, B, {"M": n, "N": n, "K": n, "tile": tile, "seed": seed, "generator": GENERATOR}


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
    A = DistArray.allocate(TiledRange([tr, tr]), mask_a, dtype=dtype, device=device)
    B = DistArray.allocate(TiledRange([tr, tr]), mask_b, dtype=dtype, device=device)
    scale = 10.0 ** (-dist.reshape(-1) / decay)
    for w, X in enumerate((A, B)):
        keyed_fill(X.pool, X.offsets, X.trange.tile_volumes(), CFG_BLOCKSPARSE, w, seed,
                   "int" if integer else "uniform")
        if not integer:
            _tile_scale_(X, scale)
    info = {"M": n, "N": n, "K": n, "tiles_per_mode": nt, "widths": [wlo, whi],
            "fill_a": float(mask_a.mean()), "fill_b": float(mask_b.mean()), "seed": seed,
            "generator": GENERATOR}
    return A, B, info


def ladder(o=32, v=256, otiles=2, wlo=48, whi=96, seed=BASE_SEED, device="cuda",
           dtype=torch.float64, v_order="acbd"):
    """T("c,d,i,j") and V stored in `v_order` ("acbd" forces a general
    permutation of V before the GEMM)."""
    rng = np.random.default_rng(seed)
    trv = random_tiling(v, wlo, whi, rng)
    tro = TiledRange1.uniform(o, (o + otiles - 1) // otiles)
    T = DistArray.allocate(TiledRange([trv, trv, tro, tro]), dtype=dtype, device=device)
    keyed_fill(T.pool, T.offsets, T.trange.tile_volumes(), CFG_LADDER, 0, seed, "normal")
    T.pool.mul_(math.sqrt(1e-2))
    modes = {"a": trv, "b": trv, "c": trv, "d": trv}
    V = DistArray.allocate(TiledRange([modes[x] for x in v_order]), dtype=dtype, device=device)
    keyed_fill(V.pool, V.offsets, V.trange.tile_volumes(), CFG_LADDER, 1, seed, "normal")
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
            "T_sigma2": 1e-2, "v_order": v_order, "seed": seed, "generator": GENERATOR}
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
    A = torch.zeros(max(1, na), dtype=dtype, device=device)
    B = torch.zeros(max(1, nb), dtype=dtype, device=device)
    for w, (X, off, vol) in enumerate(((A, a_off, va), (B, b_off, vb))):
        keyed_fill(X, off, vol, CFG_TOT, w, seed, "int" if integer else "uniform")
    return {"nbatch": nbatch, "grid": g, "a_dims": ad, "c_dims": cd, "b_dims": bd,
            "a_off": a_off, "b_off": b_off, "a_vol": va, "b_vol": vb, "A": A, "B": B,
            "fill_a": float(ma.mean()), "fill_b": float(mb.mean()), "seed": seed,
            "generator": GENERATOR}
