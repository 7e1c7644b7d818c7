"""Device-resident tiled tensors on one rank: the host-side data model the
contraction path runs on.

Mirrors TiledArray's layout conventions:
  - TiledRange1: tile boundaries of one mode (tiled_range1.h);
  - TiledRange: one TiledRange1 per mode; tiles are enumerated row-major
    over the tile grid (tiled_range.h, range.h:80-150);
  - every tile is a dense row-major block (TA::Tensor, tensor/tensor.h:454-458);
  - DistArray: a block-sparse tensor whose non-zero tiles live back to back
    in one HBM pool, tile t at element offset offsets[t] (-1 = zero tile),
    with its SparseShape (per-element scaled Frobenius norms, float) when it
    was made sparse (sparse_shape.h:77, conversions/make_array.h:152-205).
"""
import numpy as np
import torch


class TiledRange1:
    """Tile boundaries of one mode, e.g. TiledRange1([0, 128, 256])."""

    def __init__(self, boundaries):
        b = np.asarray(boundaries, dtype=np.int64)
        if b.ndim != 1 or b.size < 1 or np.any(np.diff(b) < 0):
            raise ValueError("TiledRange1 boundaries must be a non-decreasing sequence")
        self.boundaries = b

    @classmethod
    def uniform(cls, extent, tile):
        b = list(range(0, extent, tile)) + [extent]
        return cls(b)

    @property
    def ntiles(self):
        return int(self.boundaries.size - 1)

    @property
    def extents(self):
        return np.diff(self.boundaries)

    @property
    def extent(self):
        return int(self.boundaries[-1] - self.boundaries[0])

    def __eq__(self, other):
        return isinstance(other, TiledRange1) and np.array_equal(self.boundaries, other.boundaries)

    def __repr__(self):
        return f"TiledRange1({self.boundaries.tolist()})"


class TiledRange:
    def __init__(self, ranges):
        self.ranges = [r if isinstance(r, TiledRange1) else TiledRange1(r) for r in ranges]

    @property
    def rank(self):
        return len(self.ranges)

    @property
    def tiles_shape(self):
        return tuple(r.ntiles for r in self.ranges)

    @property
    def ntiles(self):
        return int(np.prod(self.tiles_shape, dtype=np.int64))

    @property
    def elements_shape(self):
        return tuple(r.extent for r in self.ranges)

    def tile_extents_table(self):
        """(ntiles, rank) int64 table of every tile's extents, row-major."""
        grids = np.meshgrid(*[r.extents for r in self.ranges], indexing="ij")
        return np.stack([g.reshape(-1) for g in grids], axis=1).astype(np.int64)

    def tile_volumes(self):
        return np.prod(self.tile_extents_table(), axis=1).astype(np.int64)

    def permuted(self, perm):
        """Tiled range of the TO-permuted tensor (result[perm[d]] = self[d])."""
        out = [None] * self.rank
        for d, p in enumerate(perm):
            out[p] = self.ranges[d]
        return TiledRange(out)

    def __eq__(self, other):
        return isinstance(other, TiledRange) and all(a == b for a, b in zip(self.ranges, other.ranges)) \
            and self.rank == other.rank

    def __repr__(self):
        return f"TiledRange({self.ranges})"


class DistArray:
    """A block-sparse tiled tensor resident on one GPU.

    pool     1-D torch tensor on the device holding the non-zero tiles;
    offsets  host int64 array (or None until fetched), -1 for zero tiles;
    d_offsets the same on the device;
    shape_norms optional device float32 tile-grid tensor of scaled norms and
    shape_threshold the threshold they were screened with (None = dense).
    """

    def __init__(self, trange, pool, offsets=None, d_offsets=None, shape_norms=None,
                 shape_threshold=None):
        self.trange = trange if isinstance(trange, TiledRange) else TiledRange(trange)
        self.pool = pool
        self._offsets = None if offsets is None else np.asarray(offsets, np.int64)
        if d_offsets is None:
            d_offsets = torch.as_tensor(self._offsets, dtype=torch.int64, device=pool.device)
        self.d_offsets = d_offsets
        self.shape_norms = shape_norms
        self.shape_threshold = shape_threshold

    # -- construction -------------------------------------------------------
    @classmethod
    def allocate(cls, trange, mask=None, dtype=torch.float64, device="cuda"):
        trange = trange if isinstance(trange, TiledRange) else TiledRange(trange)
        vols = trange.tile_volumes()
        present = np.ones(trange.ntiles, bool) if mask is None else np.asarray(mask, bool).reshape(-1)
        offsets = np.full(trange.ntiles, -1, np.int64)
        csum = np.concatenate([[0], np.cumsum(np.where(present, vols, 0))])
        offsets[present] = csum[:-1][present]
        pool = torch.empty(int(csum[-1]), dtype=dtype, device=device)
        return cls(trange, pool, offsets)

    @classmethod
    def from_dense(cls, dense, trange, mask=None, device="cuda"):
        """Tiles cut from a dense numpy/torch array (zero tiles dropped by mask)."""
        dense_t = torch.as_tensor(dense)
        arr = cls.allocate(trange, mask, dtype=dense_t.dtype, device=device)
        host = torch.empty(arr.pool.numel(), dtype=dense_t.dtype)
        for t, lo, hi in arr.tile_bounds():
            off = arr.offsets[t]
            if off < 0:
                continue
            blk = dense_t[tuple(slice(int(a), int(b)) for a, b in zip(lo, hi))]
            host[off:off + blk.numel()] = blk.reshape(-1)
        arr.pool.copy_(host.to(arr.pool.device))
        return arr

    # -- accessors ----------------------------------------------------------
    @property
    def offsets(self):
        if self._offsets is None:
            self._offsets = self.d_offsets.cpu().numpy().astype(np.int64)
        return self._offsets

    @property
    def dtype(self):
        return self.pool.dtype

    def tile_bounds(self):
        """Yields (ordinal, lobound, upbound) of every tile, row-major."""
        shp = self.trange.tiles_shape
        for t in range(self.trange.ntiles):
            idx = np.unravel_index(t, shp)
            lo = [r.boundaries[i] for r, i in zip(self.trange.ranges, idx)]
            hi = [r.boundaries[i + 1] for r, i in zip(self.trange.ranges, idx)]
            yield t, lo, hi

    def is_zero(self, t):
        return self.offsets[t] < 0

    def tile(self, t):
        """Device view of tile t reshaped to its extents (None if zero)."""
        off = int(self.offsets[t])
        if off < 0:
            return None
        ext = self.trange.tile_extents_table()[t]
        return self.pool[off:off + int(np.prod(ext))].view(*[int(e) for e in ext])

    def to_dense(self):
        """Host numpy copy with zero tiles filled with zeros."""
        host = self.pool.detach().cpu().numpy()
        out = np.zeros(self.trange.elements_shape, dtype=host.dtype)
        ext = self.trange.tile_extents_table()
        for t, lo, hi in self.tile_bounds():
            off = self.offsets[t]
            if off < 0:
                continue
            n = int(np.prod(ext[t]))
            out[tuple(slice(int(a), int(b)) for a, b in zip(lo, hi))] = \
                host[off:off + n].reshape([int(e) for e in ext[t]])
        return out

    def nonzero_tiles(self):
        return int(np.count_nonzero(self.offsets >= 0))
