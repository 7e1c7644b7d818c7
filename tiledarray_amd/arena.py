"""Tensor-of-tensor arrays in an arena in HBM, and their contraction through
the grouped launch.

Reference: TiledArray keeps the cells of a tensor-of-tensors (ToT) tile in an
arena (tensor/arena.h:66, bump-allocated pages; tensor/arena_tensor.h:67,
ArenaTensor, a view onto one cell) and contracts such tiles either cell by
cell (make_fused_contraction_lambda, tensor/arena_einsum.h:1803) or, where
the cells of a k-run lie at one stride, as one strided GEMM per run (the
ce+e regime arena_strided_dgemm_ce_e, arena_einsum.h:1170; the ce+ce
regimes, :1339 and :1585), choosing at expressions/cont_engine.h:1490-1560
with the plan of make_contraction_arena_plan (arena_einsum.h:957).

Layout.  An ArenaArray is a matrix of cells (outer modes r, c) tiled by two
TiledRange1s.  Every outer tile owns one slab of a single device pool: its
present cells in row-major order of the tile's cell grid, each cell's inner
tensor contiguous and row-major; slabs follow each other in row-major
outer-tile order.  The cell table (element offset of cell (r, c), -1 for an
empty cell) lives on the device beside the pool, so the kernels address
cells where they lie: nothing is repacked.  Inner extents are separable:
each inner mode takes its extent from one outer index (cell (i, k) of
A(i,k;a,c) is a_i x c_k), which is the shape every consistent contraction
of such arrays has (the contracted inner extent c must agree for every i
and j that meet at k).

Shapes are TiledArray's: outer-tile norms (the Frobenius norm over all cells
of the tile, Tensor<Tensor<T>>::squared_norm, tensor/tensor.h:3768-3776),
scaled by the inverse number of cells in the tile and screened
(sparse_shape.h:150-217), and the result estimate over outer tiles with the
contracted tiles' cell counts as k sizes (sparse_shape.h:1589-1681).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import call, lib
from .engine import FLOAT_EPS, BatchedContraction, WorkList
from .tiled import TiledRange1

_TASK = np.dtype(_lib.GemmTask)
_PAIR = np.dtype(_lib.GemmPair)


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                 ctypes.c_int, ctypes.c_void_p]
    return h


def slab_layout(rows, cols, present, vol):
    """Arena layout of a cell matrix: (cell_off [nr, nc] with -1 for empty
    cells, slab_cells (cell ordinals r * nc + c in slab order, empty cells
    included), tile_cells (prefix of slab_cells per outer tile, row-major
    outer tiles), slab_off (element offset of each tile's slab, plus the
    total at the end))."""
    nr, nc = rows.extent, cols.extent
    present = np.asarray(present, bool).reshape(nr, nc)
    vol = np.asarray(vol, np.int64).reshape(nr, nc)
    cell_off = np.full((nr, nc), -1, np.int64)
    order, tile_cells, slab_off = [], [0], [0]
    at = 0
    for tr in range(rows.ntiles):
        r0, r1 = int(rows.boundaries[tr]), int(rows.boundaries[tr + 1])
        for tc in range(cols.ntiles):
            c0, c1 = int(cols.boundaries[tc]), int(cols.boundaries[tc + 1])
            ids = (np.arange(r0, r1)[:, None] * nc + np.arange(c0, c1)[None, :]).reshape(-1)
            v = np.where(present.reshape(-1)[ids], vol.reshape(-1)[ids], 0)
            start = at + np.concatenate([[0], np.cumsum(v)[:-1]])
            pres = present.reshape(-1)[ids]
            cell_off.reshape(-1)[ids[pres]] = start[pres]
            at += int(v.sum())
            order.append(ids)
            tile_cells.append(tile_cells[-1] + ids.size)
            slab_off.append(at)
    slab_cells = np.concatenate(order) if order else np.zeros(0, np.int64)
    return cell_off, slab_cells, np.asarray(tile_cells, np.int64), np.asarray(slab_off, np.int64)


class ArenaArray:
    """A ToT matrix of cells in an arena: rows, cols (TiledRange1 over cell
    indices), inner (one (outer mode, extents per outer index) pair per
    inner mode), present (bool, nr x nc), pool (device tensor holding the
    slabs).  cell_off (optional, host int64 nr x nc) gives the cells'
    offsets when the pool was written elsewhere in slab order (a contraction
    result); otherwise the slab layout is computed here."""

    def __init__(self, rows, cols, inner, present, pool, cell_off=None):
        self.rows = rows if isinstance(rows, TiledRange1) else TiledRange1(rows)
        self.cols = cols if isinstance(cols, TiledRange1) else TiledRange1(cols)
        nr, nc = self.rows.extent, self.cols.extent
        self.inner = [(int(m), np.asarray(e, np.int64)) for m, e in inner]
        for m, e in self.inner:
            if m not in (0, 1) or e.size != (nr if m == 0 else nc):
                raise ValueError("ArenaArray: each inner mode needs one extent per "
                                 "index of its outer mode")
        self.present = np.asarray(present, bool).reshape(nr, nc)
        self.pool = pool
        self.dtype = pool.dtype
        vol = np.ones((nr, nc), np.int64)
        for m, e in self.inner:
            vol *= e[:, None] if m == 0 else e[None, :]
        self.cell_vol = vol
        off, self.slab_cells, self.tile_cells, self.slab_off = slab_layout(
            self.rows, self.cols, self.present, vol)
        self.cell_off = off if cell_off is None else np.asarray(cell_off, np.int64).reshape(nr, nc)
        self.d_cell_off = torch.as_tensor(self.cell_off.reshape(-1), device=pool.device)
        self._norm_plan = None

    @property
    def shape(self):
        return self.rows.extent, self.cols.extent

    @property
    def tiles_shape(self):
        return self.rows.ntiles, self.cols.ntiles

    def inner_shape(self, r, c):
        return tuple(int(e[r if m == 0 else c]) for m, e in self.inner)

    def cell(self, r, c):
        """The cell's inner tensor as a device view (None when empty)."""
        o = int(self.cell_off[r, c])
        if o < 0:
            return None
        return self.pool[o:o + int(self.cell_vol[r, c])].view(self.inner_shape(r, c))

    @classmethod
    def random(cls, rows, cols, inner, fill=0.7, empty_tiles=0.0, dtype=torch.float64,
               integer=False, seed=0, device="cuda"):
        """Random cells (integers in [-8, 8] with `integer`, else U[-1,1]);
        each cell present with probability `fill`, each outer tile emptied
        whole with probability `empty_tiles`."""
        rows = rows if isinstance(rows, TiledRange1) else TiledRange1(rows)
        cols = cols if isinstance(cols, TiledRange1) else TiledRange1(cols)
        rng = np.random.default_rng(seed)
        present = rng.random((rows.extent, cols.extent)) < fill
        if empty_tiles > 0:
            for tr in range(rows.ntiles):
                for tc in range(cols.ntiles):
                    if rng.random() < empty_tiles:
                        present[rows.boundaries[tr]:rows.boundaries[tr + 1],
                                cols.boundaries[tc]:cols.boundaries[tc + 1]] = False
        vol = np.ones(present.shape, np.int64)
        for m, e in inner:
            e = np.asarray(e, np.int64)
            vol *= e[:, None] if m == 0 else e[None, :]
        n = int(np.where(present, vol, 0).sum())
        gen = torch.Generator(device=device)
        gen.manual_seed(seed)
        pool = torch.empty(max(1, n), dtype=dtype, device=device)
        if integer:
            pool.copy_(torch.randint(-8, 9, pool.shape, generator=gen, device=device))
        else:
            pool.uniform_(-1.0, 1.0, generator=gen)
        return cls(rows, cols, inner, present, pool)

    def tile_norms(self, ctx, out=None):
        """Outer-tile Frobenius norms (float32, device, row-major tile grid)
        over the cells where they lie (tamx_norm_plan_create_cells)."""
        if self._norm_plan is None or self._norm_plan[0] is not ctx:
            ids = self.slab_cells
            off = np.ascontiguousarray(self.cell_off.reshape(-1)[ids], np.int64)
            vol = np.ascontiguousarray(self.cell_vol.reshape(-1)[ids], np.int64)
            tc = np.ascontiguousarray(self.tile_cells, np.int64)
            p = ctypes.c_void_p()
            call("tamx_norm_plan_create_cells", ctx.ptr, self.rows.ntiles * self.cols.ntiles,
                 tc.ctypes.data, off.ctypes.data, vol.ctypes.data, ctypes.byref(p))
            self._norm_plan = (ctx, p)
        nt = self.rows.ntiles * self.cols.ntiles
        if out is None:
            out = torch.empty(max(1, nt), dtype=torch.float32, device=self.pool.device)
        f = "tamx_tile_norms_f64" if self.dtype == torch.float64 else "tamx_tile_norms_f32"
        call(f, ctx.ptr, self._norm_plan[1], self.pool.data_ptr(), out.data_ptr(), ctx.stream())
        return out

    def __del__(self):
        try:
            if self._norm_plan is not None:
                lib().tamx_norm_plan_destroy(self._norm_plan[1])
                self._norm_plan = None
        except Exception:
            pass


def plan_ce_e(A, B, sa, sb, est, threshold):
    """Work list of the inner outer product C(i,j;p,q) = sum_k A(i,k;p)
    B(k,j;q) (the ce+e regime, arena_einsum.h:1170-1330) from the arenas'
    presence and the screened outer-tile shapes sa (A tiles), sb (B tiles),
    est (result tiles).  Tasks: the cells of the non-zero result tiles,
    placed slab by slab (row-major result tiles, row-major cells), P x Q
    each.  Pairs, per task and per K tile where both operand tiles are
    non-zero, ascending K: one GEMM over the whole k-run when row i of A's
    tile (I, K) and B's whole tile (K, J) are complete (A as the K x P slab
    read transposed, lda = P; B as the K x Q run at the tile's row length),
    else one K = 1 GEMM per k with both cells present (the reference's
    per-cell fallback).  Every pair is op_a = TRANS, op_b = NO_TRANS.
    Returns (tasks, pairs, c_off [nI x nJ, -1 empty], elements, merged)."""
    ea, eb = A.inner[0][1], B.inner[0][1]
    rb, kb, cb = A.rows.boundaries, A.cols.boundaries, B.cols.boundaries
    NI, NK, NJ = A.rows.ntiles, A.cols.ntiles, B.cols.ntiles
    tasks, pairs = [], []
    c_off = np.full((A.rows.extent, B.cols.extent), -1, np.int64)
    at = 0
    nmerged = 0
    a_run = np.zeros((A.rows.extent, NK), bool)
    for K in range(NK):
        a_run[:, K] = A.present[:, kb[K]:kb[K + 1]].all(axis=1)
    b_tile = np.zeros((NK, NJ), bool)
    for K in range(NK):
        for J in range(NJ):
            b_tile[K, J] = B.present[kb[K]:kb[K + 1], cb[J]:cb[J + 1]].all()
    for I in range(NI):
        for J in range(NJ):
            if not est[I, J] >= threshold:
                continue
            Ks = [K for K in range(NK) if sa[I, K] >= threshold and sb[K, J] >= threshold]
            rowlen = int(eb[cb[J]:cb[J + 1]].sum())
            for i in range(int(rb[I]), int(rb[I + 1])):
                P = int(ea[i])
                for j in range(int(cb[J]), int(cb[J + 1])):
                    Q = int(eb[j])
                    c_off[i, j] = at
                    pb = len(pairs)
                    for K in Ks:
                        k0, k1 = int(kb[K]), int(kb[K + 1])
                        if a_run[i, K] and b_tile[K, J]:
                            pairs.append((A.cell_off[i, k0], B.cell_off[k0, j], k1 - k0, P,
                                          rowlen, 0, 0, 0))
                            nmerged += 1
                            continue
                        for k in range(k0, k1):
                            if A.present[i, k] and B.present[k, j]:
                                pairs.append((A.cell_off[i, k], B.cell_off[k, j], 1, P, Q,
                                              0, 0, 0))
                    tasks.append((at, P, Q, Q, pb, len(pairs) - pb, 0))
                    at += P * Q
    return tasks, pairs, c_off, at, nmerged


class ArenaContraction:
    """C = alpha * A * B over the outer cell matrices of two ArenaArrays
    (C(i,j) += A(i,k) . B(k,j) for every k), on the device, with the result
    an ArenaArray tiled by A's rows and B's columns.  Two regimes:

    - "inner_gemm" (C(i,j;a,b) = A(i,k;a,c) B(k,j;c,b)): every cell pair is
      one inner GEMM.  The launch is a batched grouped GEMM with one batch
      per result outer tile, whose operand cell tables point at the arenas'
      cells (shared between batches, nothing copied); outer-tile shapes are
      replicated onto cells, so the screening decisions are the outer
      tiles', and the result cells come out packed slab by slab.  Fully on
      the device (no host synchronisation per step).
    - "outer_product" (C(i,j;p,q) = A(i,k;p) B(k,j;q)), the ce+e regime
      (arena_einsum.h:1170): a k-run of A cells (one outer tile's k cells of
      row i) lies at one stride in its slab, and so does the run of B cells
      (k, j) of a whole B tile; a complete run pair is one GEMM with K the
      run length (A as the K x P slab, transposed; B as the K x Q slab at
      the tile's row stride), and every other present cell pair is a K = 1
      GEMM (the reference's per-cell fallback).  All in one grouped launch;
      the work list is planned on the host from the arenas' presence and the
      device screening (one synchronisation per plan).

    precision: the cells' type (FP64 or FP32 MFMA)."""

    def __init__(self, ctx, A, B, alpha=1.0, threshold=FLOAT_EPS, tile=None):
        if not (A.cols == B.rows):
            raise ValueError("ArenaContraction: A's column tiling must equal B's row tiling")
        if A.dtype != B.dtype:
            raise ValueError("ArenaContraction: operands of different types")
        self.ctx, self.A, self.B = ctx, A, B
        self.alpha = float(alpha)
        self.threshold = float(np.float32(threshold))
        self.dtype = A.dtype
        dev = A.pool.device
        ia, ib = A.inner, B.inner
        if (len(ia) == 2 and len(ib) == 2 and [m for m, _ in ia] == [0, 1]
                and [m for m, _ in ib] == [0, 1] and np.array_equal(ia[1][1], ib[0][1])):
            self.regime = "inner_gemm"
            self.ea, self.ec, self.eb = ia[0][1], ia[1][1], ib[1][1]
        elif len(ia) == 1 and len(ib) == 1 and ia[0][0] == 0 and ib[0][0] == 1:
            self.regime = "outer_product"
            self.ea, self.eb = ia[0][1], ib[0][1]
            self.ec = np.ones(A.cols.extent, np.int64)
        else:
            raise ValueError("ArenaContraction: inner modes must be (a_i, c_k) x (c_k, b_j) "
                             "or (p_i) x (q_j)")
        self.NI, self.NK, self.NJ = A.rows.ntiles, A.cols.ntiles, B.cols.ntiles
        self.tile = tile
        # outer-tile shape buffers
        self.a_tn = torch.zeros(max(1, self.NI * self.NK), dtype=torch.float32, device=dev)
        self.b_tn = torch.zeros(max(1, self.NK * self.NJ), dtype=torch.float32, device=dev)
        self.c_est = torch.zeros(max(1, self.NI * self.NJ), dtype=torch.float32, device=dev)
        f32 = np.float32
        self.a_sizes = torch.as_tensor(np.concatenate(
            [A.rows.extents, A.cols.extents]).astype(f32), device=dev)
        self.b_sizes = torch.as_tensor(np.concatenate(
            [B.rows.extents, B.cols.extents]).astype(f32), device=dev)
        self.k_sizes = torch.as_tensor(A.cols.extents.astype(f32), device=dev)
        if self.regime == "inner_gemm":
            self._build_batched(dev)
        else:
            self.C = None
            self.wl = None

    # -- shapes ---------------------------------------------------------------
    def shape_step(self):
        """Outer-tile norms, inverse-volume scaling, result estimate."""
        s = self.ctx.stream()
        self.A.tile_norms(self.ctx, out=self.a_tn)
        self.B.tile_norms(self.ctx, out=self.b_tn)
        for norms, sizes, grid in ((self.a_tn, self.a_sizes, (self.NI, self.NK)),
                                   (self.b_tn, self.b_sizes, (self.NK, self.NJ))):
            nt = np.ascontiguousarray(grid, np.int64)
            call("tamx_shape_scale", self.ctx.ptr, norms.data_ptr(), 2, nt.ctypes.data, 1,
                 sizes.data_ptr(), 0, ctypes.c_float(self.threshold), None, s)
        call("tamx_shape_gemm_f32", self.ctx.ptr, self.a_tn.data_ptr(), self.b_tn.data_ptr(),
             0, 0, self.NI, self.NJ, self.NK, 1, self.k_sizes.data_ptr(),
             ctypes.c_float(abs(self.alpha)), ctypes.c_float(self.threshold),
             self.c_est.data_ptr(), None, s)

    # -- inner GEMM regime --------------------------------------------------------
    def _build_batched(self, dev):
        A, B = self.A, self.B
        rb, cb = A.rows.boundaries, B.cols.boundaries
        TI = int(A.rows.extents.max()) if A.rows.ntiles else 1
        TJ = int(B.cols.extents.max()) if B.cols.ntiles else 1
        KT = A.cols.extent
        nb = self.NI * self.NJ
        self.TI, self.TJ = TI, TJ
        # batch b = (I, J): rows of I (padded to TI), all k cells, columns of J
        mdims = np.ones((nb, TI), np.int64)
        ndims = np.ones((nb, TJ), np.int64)
        a_off = np.full((nb, TI, KT), -1, np.int64)
        b_off = np.full((nb, KT, TJ), -1, np.int64)
        # gather indices of the replicated outer-tile shapes (the last slot of
        # each padded shape vector is zero: padding rows, columns, cells)
        ktile = np.repeat(np.arange(self.NK), A.cols.extents)
        a_idx = np.full((nb, TI, KT), self.NI * self.NK, np.int64)
        b_idx = np.full((nb, KT, TJ), self.NK * self.NJ, np.int64)
        c_idx = np.full((nb, TI, TJ), self.NI * self.NJ, np.int64)
        self.valid = np.zeros((nb, TI, TJ), bool)
        for I in range(self.NI):
            r0, r1 = int(rb[I]), int(rb[I + 1])
            for J in range(self.NJ):
                c0, c1 = int(cb[J]), int(cb[J + 1])
                b = I * self.NJ + J
                mdims[b, :r1 - r0] = self.ea[r0:r1]
                ndims[b, :c1 - c0] = self.eb[c0:c1]
                a_off[b, :r1 - r0, :] = A.cell_off[r0:r1, :]
                b_off[b, :, :c1 - c0] = B.cell_off[:, c0:c1]
                a_idx[b, :r1 - r0, :] = I * self.NK + ktile[None, :]
                b_idx[b, :, :c1 - c0] = ktile[:, None] * self.NJ + J
                c_idx[b, :r1 - r0, :c1 - c0] = I * self.NJ + J
                self.valid[b, :r1 - r0, :c1 - c0] = True
        kdims = np.tile(self.ec[None, :], (nb, 1))
        self.batch = BatchedContraction(
            self.ctx, nb, TI, TJ, KT, mdims, ndims, kdims, A.pool, a_off.reshape(-1), B.pool,
            b_off.reshape(-1), threshold=self.threshold, alpha=self.alpha, tile=self.tile,
            cell_shapes=False)
        self.d_a_idx = torch.as_tensor(a_idx.reshape(-1), device=dev)
        self.d_b_idx = torch.as_tensor(b_idx.reshape(-1), device=dev)
        self.d_c_idx = torch.as_tensor(c_idx.reshape(-1), device=dev)
        self._pad = torch.zeros(1, dtype=torch.float32, device=dev)
        self.C = self.batch.C
        self.wl = self.batch.wl

    def _replicate_shapes(self):
        """Every cell takes its outer tile's scaled norm (operands) or
        estimate (result); padding cells take the zero slot past the end."""
        bt = self.batch
        for dst, src, n_src, idx in ((bt.a_norms, self.a_tn, self.NI * self.NK, self.d_a_idx),
                                     (bt.b_norms, self.b_tn, self.NK * self.NJ, self.d_b_idx),
                                     (bt.c_norms, self.c_est, self.NI * self.NJ, self.d_c_idx)):
            full = torch.cat([src[:n_src], self._pad])
            torch.index_select(full, 0, idx, out=dst[:idx.numel()])

    # -- outer-product regime ---------------------------------------------------------
    def plan_outer_product(self):
        """Host plan of the ce+e work list from the screening on the device
        (one synchronisation), uploaded into the work list."""
        A, B, thr = self.A, self.B, self.threshold
        sa = self.a_tn.cpu().numpy()[:self.NI * self.NK].reshape(self.NI, self.NK)
        sb = self.b_tn.cpu().numpy()[:self.NK * self.NJ].reshape(self.NK, self.NJ)
        est = self.c_est.cpu().numpy()[:self.NI * self.NJ].reshape(self.NI, self.NJ)
        tasks, pairs, c_off, at, nmerged = plan_ce_e(A, B, sa, sb, est, thr)
        self.op_c_off = c_off
        self.nmerged = nmerged
        ntasks, npairs = len(tasks), len(pairs)
        dev = A.pool.device
        self.C = torch.zeros(max(1, at), dtype=self.dtype, device=dev)
        self.wl = WorkList(self.ctx, max(1, ntasks), max(1, npairs), max(1, ntasks))
        t_arr = np.array(tasks, dtype=_TASK) if ntasks else np.zeros(1, _TASK)
        p_arr = np.array(pairs, dtype=_PAIR) if npairs else np.zeros(1, _PAIR)
        cnt = np.zeros(_lib.WL_NCOUNTERS, np.int64)
        cnt[_lib.WL_NTASKS], cnt[_lib.WL_NPAIRS], cnt[_lib.WL_CELEMS] = ntasks, npairs, at
        if npairs:
            # GEMM flops of the pairs: 2 m n k with m, n from the pair's task
            reps = np.repeat(np.arange(ntasks), t_arr["pair_count"])
            cnt[_lib.WL_FLOPS] = int((2 * t_arr["m"][reps].astype(np.int64) *
                                      t_arr["n"][reps] * p_arr["k"]).sum())
        tp, pp, cp = self.wl.buffers()
        h = _hip()
        s = self.ctx.stream()
        for dst, src in ((tp, t_arr), (pp, p_arr), (cp, cnt)):
            if h.hipMemcpyAsync(dst, src.ctypes.data, src.nbytes, 1, s) != 0:
                raise _lib.TamxError(_lib.TAMX_ERR_DEVICE, "hipMemcpyAsync", "work list upload")
        torch.cuda.current_stream(dev).synchronize()  # the host arrays die here
        self._hold = (t_arr, p_arr, cnt)
        return ntasks, npairs

    # -- steps ------------------------------------------------------------------
    def worklist_step(self):
        if self.regime == "inner_gemm":
            self._replicate_shapes()
            self.batch.worklist_step()
        else:
            self.plan_outer_product()

    def gemm_step(self):
        if self.regime == "inner_gemm":
            self.batch.gemm_step()
            return
        s = self.ctx.stream()
        cfg = None if self.tile is None else ctypes.byref(
            _lib.GemmConfig(*(tuple(self.tile) + (16, 0))[:4]))
        f64 = self.dtype == torch.float64
        call("tamx_grouped_gemm_f64_config" if f64 else "tamx_grouped_gemm_f32_config",
             self.ctx.ptr, self.wl.ptr, 1, 0,
             ctypes.c_double(self.alpha) if f64 else ctypes.c_float(self.alpha),
             self.A.pool.data_ptr(), self.B.pool.data_ptr(), self.C.data_ptr(), cfg, s)

    def run(self):
        self.shape_step()
        self.worklist_step()
        self.gemm_step()

    def counters(self):
        c = self.wl.counters()
        return {"ntasks": c[_lib.WL_NTASKS], "npairs": c[_lib.WL_NPAIRS],
                "celems": c[_lib.WL_CELEMS], "flops": c[_lib.WL_FLOPS]}

    def result_cell_offsets(self):
        """Offsets of the result cells (host int64, nr x nc, -1 empty)."""
        if self.regime == "outer_product":
            return self.op_c_off
        A, B = self.A, self.B
        off = self.batch.c_off.cpu().numpy().reshape(self.NI * self.NJ, self.TI, self.TJ)
        out = np.full((A.rows.extent, B.cols.extent), -1, np.int64)
        rb, cb = A.rows.boundaries, B.cols.boundaries
        for I in range(self.NI):
            for J in range(self.NJ):
                r0, r1, c0, c1 = int(rb[I]), int(rb[I + 1]), int(cb[J]), int(cb[J + 1])
                out[r0:r1, c0:c1] = off[I * self.NJ + J, :r1 - r0, :c1 - c0]
        return out

    def result(self):
        """The result ArenaArray (its cells where the launch wrote them:
        slab by slab, row-major cells; synchronises for the cell table)."""
        off = self.result_cell_offsets()
        inner = [(0, self.ea), (1, self.eb)]
        return ArenaArray(self.A.rows, self.B.cols, inner, off >= 0, self.C, cell_off=off)
