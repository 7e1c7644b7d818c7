"""Single-rank contraction engine over the C ABI of libtamx.

`contract(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,b,c,d")` evaluates
C(target) = A(left) * B(right) the way TiledArray's expression layer does on
one rank, with every data-parallel step on the GPU:

  1. plan (tamx_plan_contraction): operand permutation types, GEMM ops, the
     natural result order and its permutation to the target
     (expressions/cont_engine.h:354-470, expressions/permopt.h:254-381);
  2. permute operand tiles whose layout is `general` (permute-on-read,
     dist_eval/array_eval.h:330-365) with one tamx_permute launch;
  3. fuse the operands into block matrices (GemmHelper,
     math/gemm_helper.h:41-278) and estimate the result shape
     (SparseShape::gemm, sparse_shape.h:1589-1681);
  4. build the device work list of surviving tile pairs (Summa's one-rank
     schedule, dist_eval/contraction_eval.h:1094-1384) and run all of them in
     one persistent grouped GEMM (ReducePairTask + ContractReduce,
     reduce_task.h:961-1026, tile_op/contract_reduce.h:409-453);
  5. apply the result permutation, fused into the GEMM store when it is a
     matrix transpose (contract_reduce.h:370-378).

A `Contraction` object does steps 1-3's host work once; `run()` then
enqueues only device work, with no host synchronisation, so it can be
replayed or captured in a hipGraph.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import call, lib
from .tiled import DistArray, TiledRange

FLOAT_EPS = float(np.finfo(np.float32).eps)  # SparseShape<float> default threshold


class Context:
    """Owns a tamx_context for one device."""

    def __init__(self, device=0):
        if not torch.cuda.is_available():
            raise RuntimeError("tiledarray_amd needs a GPU: torch.cuda.is_available() is False")
        self.device = device
        torch.cuda.set_device(device)
        p = ctypes.c_void_p()
        call("tamx_context_create", device, ctypes.byref(p))
        self.ptr = p
        self.num_cus = lib().tamx_context_num_cus(p)

    def stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def close(self):
        if self.ptr:
            lib().tamx_context_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A launch sequence captured through tamx_graph_begin/_end.  `fn` is
    called once on a side stream under capture; replay() relaunches it on
    the current stream."""

    def __init__(self, ctx, fn):
        self.ctx = ctx
        self.ptr = None
        side = torch.cuda.Stream(ctx.device)
        side.wait_stream(torch.cuda.current_stream(ctx.device))
        with torch.cuda.stream(side):
            s = ctypes.c_void_p(side.cuda_stream)
            call("tamx_graph_begin", ctx.ptr, s)
            p = ctypes.c_void_p()
            try:
                fn()
            except BaseException:
                # end the capture so the stream is usable again, drop the graph
                lib().tamx_graph_end(ctx.ptr, s, ctypes.byref(p))
                if p:
                    lib().tamx_graph_destroy(p)
                raise
            call("tamx_graph_end", ctx.ptr, s, ctypes.byref(p))
            self.ptr = p
        torch.cuda.current_stream(ctx.device).wait_stream(side)

    def replay(self):
        call("tamx_graph_launch", self.ptr, self.ctx.stream())

    def __del__(self):
        try:
            if self.ptr:
                lib().tamx_graph_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


def autotune(ctx, wl, dtype, op_a, op_b, A, B, C, reps=3, force=False, warmup=True,
             a_gather=None, b_gather=None):
    """Fastest compiled tile configuration for the built work list `wl`
    (tamx_autotune_grouped_gemm, or its gathered form when a GatherPlan is
    given); returns ((bm, bn, bk, wgs_per_cu), ms)."""
    prec = _lib.PREC_F64 if dtype == torch.float64 else _lib.PREC_F32
    cfg = _lib.GemmConfig()
    ms = ctypes.c_float()
    flags = (_lib.TUNE_FORCE if force else 0) | (0 if warmup else _lib.TUNE_NO_WARMUP)
    if a_gather is None and b_gather is None:
        call("tamx_autotune_grouped_gemm", ctx.ptr, wl.ptr, prec, int(op_a), int(op_b),
             A.data_ptr(), B.data_ptr(), C.data_ptr(), int(reps), flags, ctypes.byref(cfg),
             ctypes.byref(ms), ctx.stream())
    else:
        call("tamx_autotune_grouped_gemm_gather", ctx.ptr, wl.ptr, prec, int(op_a), int(op_b),
             A.data_ptr(), _plan_ptr(a_gather), B.data_ptr(), _plan_ptr(b_gather),
             C.data_ptr(), int(reps), flags, ctypes.byref(cfg), ctypes.byref(ms),
             ctx.stream())
    return (cfg.bm, cfg.bn, cfg.bk, cfg.wgs_per_cu), ms.value


def _plan_ptr(plan):
    return None if plan is None else plan.ptr


class WorkList:
    def __init__(self, ctx, task_capacity, pair_capacity, grid_capacity):
        self.ctx = ctx
        p = ctypes.c_void_p()
        call("tamx_worklist_create", ctx.ptr, int(task_capacity), int(pair_capacity),
             int(grid_capacity), ctypes.byref(p))
        self.ptr = p
        self.task_capacity, self.pair_capacity, self.grid_capacity = \
            int(task_capacity), int(pair_capacity), int(grid_capacity)

    def counters(self):
        buf = (ctypes.c_int64 * _lib.WL_NCOUNTERS)()
        call("tamx_worklist_read_counters", self.ptr, buf, self.ctx.stream())
        return list(buf)

    def buffers(self):
        t, p, c = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        call("tamx_worklist_buffers", self.ptr, ctypes.byref(t), ctypes.byref(p), ctypes.byref(c))
        return t.value, p.value, c.value

    def __del__(self):
        try:
            if self.ptr:
                lib().tamx_worklist_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


class NormPlan:
    def __init__(self, ctx, offsets, volumes):
        self.ctx = ctx
        offsets = np.ascontiguousarray(offsets, np.int64)
        volumes = np.ascontiguousarray(volumes, np.int64)
        p = ctypes.c_void_p()
        call("tamx_norm_plan_create", ctx.ptr, offsets.ctypes.data, volumes.ctypes.data,
             offsets.size, ctypes.byref(p))
        self.ptr = p
        self.ntiles = offsets.size
        self.volumes = volumes

    def __del__(self):
        try:
            if self.ptr:
                lib().tamx_norm_plan_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


class PermutePlan:
    def __init__(self, ctx, perm, extents, src_off, dst_off):
        self.ctx = ctx
        perm = np.ascontiguousarray(perm, np.int32)
        extents = np.ascontiguousarray(extents, np.int64)
        src_off = np.ascontiguousarray(src_off, np.int64)
        dst_off = np.ascontiguousarray(dst_off, np.int64)
        p = ctypes.c_void_p()
        call("tamx_permute_plan_create", ctx.ptr, int(perm.size), perm.ctypes.data,
             int(src_off.size), extents.ctypes.data, src_off.ctypes.data, dst_off.ctypes.data,
             ctypes.byref(p))
        self.ptr = p
        self.volume = int(lib().tamx_permute_plan_volume(p))

    def run(self, src, dst):
        f = "tamx_permute_f64" if src.dtype == torch.float64 else "tamx_permute_f32"
        call(f, self.ctx.ptr, self.ptr, src.data_ptr(), dst.data_ptr(), self.ctx.stream())

    def __del__(self):
        try:
            if self.ptr:
                lib().tamx_permute_plan_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class GatherPlan:
    """Device index map of a permuted view's tiles (tamx_gather_plan): the
    grouped GEMM reads the view's stored tiles through it."""

    def __init__(self, ctx, maps):
        nt, nr, nk, rext, rstr, kext, kstr = maps
        p = ctypes.c_void_p()
        call("tamx_gather_plan_create", ctx.ptr, nt, nr, nk, rext.ctypes.data, rstr.ctypes.data,
             kext.ctypes.data, kstr.ctypes.data, ctypes.byref(p))
        self.ptr = p

    def __del__(self):
        try:
            if self.ptr:
                lib().tamx_gather_plan_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


class _SplitOperand:
    def __init__(self, ptr, limbs, ntiles):
        self.ptr, self.limbs, self.ntiles = ptr, limbs, ntiles

    def __del__(self):
        try:
            if self.ptr:
                lib().tamx_split_plan_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


# ------------------------------------------------------------- shapes ------

def tile_norms(ctx, arr, out=None, maxabs=None):
    """Raw per-tile Frobenius norms (float32, device), Tensor::norm().  With
    `maxabs` (float64, one per tile; FP64 arrays only) the same pass also
    writes each tile's max |a|, the scale input of the limb split."""
    plan = getattr(arr, "_norm_plan", None)
    if plan is None or plan.ctx is not ctx:
        plan = NormPlan(ctx, arr.offsets, arr.trange.tile_volumes())
        arr._norm_plan = plan
    if out is None:
        out = torch.empty(arr.trange.ntiles, dtype=torch.float32, device=arr.pool.device)
    if maxabs is not None:
        if arr.dtype != torch.float64 or maxabs.dtype != torch.float64:
            raise ValueError("tile maxima are produced for FP64 arrays into float64 tensors")
        call("tamx_tile_norms_max_f64", ctx.ptr, plan.ptr, arr.pool.data_ptr(), out.data_ptr(),
             maxabs.data_ptr(), ctx.stream())
        return out
    f = "tamx_tile_norms_f64" if arr.dtype == torch.float64 else "tamx_tile_norms_f32"
    call(f, ctx.ptr, plan.ptr, arr.pool.data_ptr(), out.data_ptr(), ctx.stream())
    return out


def _sizes_vector(trange, device):
    s = np.concatenate([r.extents for r in trange.ranges]).astype(np.float32)
    return torch.as_tensor(s, device=device)


def scale_norms(ctx, norms, trange, threshold, zero_count=None, sizes=None):
    """In place: norms /= tile volume (reference op order), screen < threshold."""
    nt = np.ascontiguousarray(trange.tiles_shape, np.int64)
    if sizes is None:
        sizes = _sizes_vector(trange, norms.device)
    call("tamx_shape_scale", ctx.ptr, norms.data_ptr(), trange.rank, nt.ctypes.data, 1,
         sizes.data_ptr(), 0, ctypes.c_float(threshold), _ptr(zero_count), ctx.stream())
    return norms


def make_sparse(ctx, arr, threshold=FLOAT_EPS):
    """Attaches a SparseShape to `arr` built from its tile norms (make_array
    semantics, conversions/make_array.h:152-205): norms, inverse-volume
    scaling and screening.  Tiles screened to zero stay in the pool but are
    treated as absent by the contraction."""
    n = tile_norms(ctx, arr)
    scale_norms(ctx, n, arr.trange, threshold)
    arr.shape_norms = n
    arr.shape_threshold = float(np.float32(threshold))
    return arr


def k_sizes(inner_ranges):
    """recursive_outer_product of contracted tile extents (float,
    sparse_shape.h:106-131 with the identity op, used at :1626-1630)."""
    def rop(vs):
        if len(vs) == 1:
            return vs[0].astype(np.float32)
        mid = (len(vs) >> 1) + (len(vs) & 1)
        l, r = rop(vs[:mid]), rop(vs[mid:])
        return (l[:, None] * r[None, :]).reshape(-1).astype(np.float32)
    return rop([r.extents.astype(np.float32) for r in inner_ranges])


def _fused_extents(ranges):
    e = np.ones(1, np.int64)
    for r in ranges:
        e = np.multiply.outer(e, r.extents.astype(np.int64)).reshape(-1)
    return e


def plan_contraction(left, right, target, prefer_to_permute_left=True):
    p = _lib.ContractionPlan()
    call("tamx_plan_contraction", left.encode(), right.encode(), target.encode(),
         1 if prefer_to_permute_left else 0, ctypes.byref(p))
    return p


def permuted_view(arr, perm):
    """The TO-permuted array's structure (tiled range, tile offsets, shape)
    over the *unpermuted* source pool, for consumers that read the source
    through an index map (the limb split): no permuted copy is made.  The
    view records its source (`_view_src`, `_view_perm`); its pool must not
    be read as if permuted."""
    perm = [int(x) for x in perm]
    new_tr = arr.trange.permuted(perm)
    grid = np.arange(arr.trange.ntiles).reshape(arr.trange.tiles_shape)
    inv = np.argsort(perm)
    src_of_new = np.transpose(grid, inv).reshape(-1)  # source ordinal of each permuted tile
    new_offs = arr.offsets[src_of_new].copy()
    out = DistArray(new_tr, arr.pool, new_offs)
    if arr.shape_norms is not None:
        out.shape_norms = arr.shape_norms.view(arr.trange.tiles_shape).permute(
            *[int(i) for i in inv]).contiguous().view(-1)
        out.shape_threshold = arr.shape_threshold
    out._view_src, out._view_perm, out._view_src_of = arr, perm, src_of_new
    return out


def permute_array(ctx, arr, perm):
    """TO-permuted copy of `arr` (tile data, tiled range and shape)."""
    perm = [int(x) for x in perm]
    rank = arr.trange.rank
    new_tr = arr.trange.permuted(perm)
    ext = arr.trange.tile_extents_table()
    offs = arr.offsets
    # tile ordinal of every source tile in the permuted grid
    grid = np.arange(arr.trange.ntiles).reshape(arr.trange.tiles_shape)
    inv = np.argsort(perm)
    new_ord_of_src = np.empty(arr.trange.ntiles, np.int64)
    new_ord_of_src[np.transpose(grid, inv).reshape(-1)] = np.arange(arr.trange.ntiles)
    new_offs = np.full(arr.trange.ntiles, -1, np.int64)
    new_offs[new_ord_of_src] = offs  # each tile keeps its place in the pool
    present = offs >= 0
    plan = PermutePlan(ctx, perm, ext[present], offs[present], offs[present])
    pool = torch.empty_like(arr.pool)
    if plan.volume:
        plan.run(arr.pool, pool)
    out = DistArray(new_tr, pool, new_offs)
    if arr.shape_norms is not None:
        out.shape_norms = arr.shape_norms.view(arr.trange.tiles_shape).permute(
            *[int(i) for i in inv]).contiguous().view(-1)
        out.shape_threshold = arr.shape_threshold
    out._perm_plan = plan
    return out


# -------------------------------------------------------- contraction ------

class Contraction:
    """A prepared single-rank contraction C(target) = alpha * A(left) * B(right).

    Host work (planning, operand permutation plans, fusion metadata, result
    allocation) happens here; run() enqueues device work only.

    precision: "f64" / "f32" (the operands' type on FP64 / FP32 MFMA),
    "bf16xS" or "i8xS" (FP64 operands split into S BF16 or INT8 limbs,
    2 <= S <= 6, products on BF16 / INT8 MFMA, sums and C in FP64) or
    "f32x2" (two FP32 limbs).  fuse_permute: how a `general` operand is
    brought into GEMM layout (None: an index map for the limb precisions, a
    permuted copy otherwise; True: an index map; False: a copy).
    fuse_result_permute: a general result order is written by the GEMM
    through each result tile's index map (True) or by a permute pass over
    the finished tiles (False).
    """

    def __init__(self, ctx, target, A, left, B, right, alpha=1.0, threshold=FLOAT_EPS,
                 pair_threshold=0.0, precision="f64", tile=None, permute_operands=True,
                 fuse_permute=None, fuse_result_permute=True):
        self.ctx = ctx
        self.alpha = float(alpha)
        self.precision = precision
        self.set_tile(tile)
        self.plan = plan = plan_contraction(left, right, target)
        self.threshold = float(np.float32(threshold))
        dev = A.pool.device
        # operands in GEMM layout.  A `general` operand is either permuted
        # into a copy (as the reference permutes on read) or read through an
        # index map of its stored tiles, with no copy.  fuse_permute=None
        # picks the map for the limb split (it stages every element once
        # anyway) and the copy for the FP64 / FP32 GEMM, whose gathered
        # loads cost about 10% of its rate on the ladder workloads, more
        # than the permute pass; True forces the map (saving the copy's
        # memory), False the copy.  The FP32 two-limb GEMM always copies.
        if fuse_permute is None:
            fuse = precision.startswith(("bf16x", "i8x"))
        else:
            fuse = bool(fuse_permute) and precision != "f32x2"
        self.A_src, self.B_src = A, B

        def gemm_layout(X, permtype, perm, inner):
            if permtype != 3:
                return X
            n_in = int(sum(inner))
            if fuse and 1 <= n_in <= 4 and 1 <= len(inner) - n_in <= 4:
                return permuted_view(X, perm)
            return permute_array(ctx, X, perm)
        self.A = gemm_layout(A, plan.left_permtype, plan.left_perm[:plan.left_rank],
                             [plan.left_gemm_is_inner[d] for d in range(plan.left_rank)])
        self.B = gemm_layout(B, plan.right_permtype, plan.right_perm[:plan.right_rank],
                             [plan.right_gemm_is_inner[d] for d in range(plan.right_rank)])
        self.permute_operands = permute_operands
        Ag, Bg = self.A, self.B
        l_inner = [bool(plan.left_gemm_is_inner[d]) for d in range(plan.left_rank)]
        r_inner = [bool(plan.right_gemm_is_inner[d]) for d in range(plan.right_rank)]
        l_outer_r = [Ag.trange.ranges[d] for d in range(plan.left_rank) if not l_inner[d]]
        l_inner_r = [Ag.trange.ranges[d] for d in range(plan.left_rank) if l_inner[d]]
        r_outer_r = [Bg.trange.ranges[d] for d in range(plan.right_rank) if not r_inner[d]]
        r_inner_r = [Bg.trange.ranges[d] for d in range(plan.right_rank) if r_inner[d]]
        # contracted modes must be tiled identically (cont_engine.h:614-633)
        if len(l_inner_r) != len(r_inner_r) or any(a != b for a, b in zip(l_inner_r, r_inner_r)):
            raise ValueError("contracted modes of the operands have incongruent tilings")
        self.op_a, self.op_b = int(plan.op_a), int(plan.op_b)
        self.Mt = int(np.prod([r.ntiles for r in l_outer_r], dtype=np.int64))
        self.Nt = int(np.prod([r.ntiles for r in r_outer_r], dtype=np.int64))
        self.Kt = int(np.prod([r.ntiles for r in l_inner_r], dtype=np.int64))
        mext, next_, kext = (_fused_extents(l_outer_r), _fused_extents(r_outer_r),
                             _fused_extents(l_inner_r))
        self.mext_h, self.next_h, self.kext_h = mext, next_, kext
        self.d_mext = torch.as_tensor(mext.astype(np.int32), device=dev)
        self.d_next = torch.as_tensor(next_.astype(np.int32), device=dev)
        self.d_kext = torch.as_tensor(kext.astype(np.int32), device=dev)
        self.k_sizes = torch.as_tensor(k_sizes(l_inner_r), device=dev)
        self.sparse = Ag.shape_norms is not None and Bg.shape_norms is not None
        if (Ag.shape_norms is None) != (Bg.shape_norms is None):
            raise ValueError("operands must both be dense or both be sparse")
        # result structure in natural order (left outer, right outer)
        self.natural_ranges = l_outer_r + r_outer_r
        self.c_norms = torch.ones(self.Mt * self.Nt, dtype=torch.float32, device=dev)
        self.c_off = torch.empty(self.Mt * self.Nt, dtype=torch.int64, device=dev)
        self.c_capacity = int(mext.sum() * next_.sum())
        self.dtype = A.pool.dtype
        # (zeroed once: the part past the result tiles then holds no stale
        # bits, so whole-buffer comparisons are well defined)
        self.C = torch.zeros(self.c_capacity, dtype=self.dtype, device=dev)
        self.store_trans = bool(plan.result_perm_matrix_transpose)
        # a general result permutation (contract_reduce.h:370-378) is fused
        # into the GEMM's store: every result tile is written straight into
        # the target order through its index map, so no permute pass runs
        # and result() needs nothing from the host
        self.c_scatter = None
        if fuse_result_permute and not plan.result_perm_identity and not self.store_trans:
            maps = self._scatter_maps(l_outer_r, r_outer_r,
                                      [int(plan.result_perm[d]) for d in range(plan.result_rank)])
            if maps is not None:
                self.c_scatter = GatherPlan(ctx, maps)
        grid = self.Mt * self.Nt
        self.wl = WorkList(ctx, grid, max(1, grid * self.Kt), grid)
        self.pair_threshold = float(pair_threshold)
        self._side = None  # side stream of split_and_plan, made on first use
        self._desc = self._make_desc()
        self.nslices = 0
        self.limb_type = _lib.LIMB_I8 if precision.startswith("i8x") else _lib.LIMB_BF16
        # gathered operands of the FP64 / FP32 GEMM and the load orientation
        # their maps read best along (the op passed for them)
        self.a_gather = self.b_gather = None
        self.gemm_op_a, self.gemm_op_b = self.op_a, self.op_b
        if precision in ("f64", "f32"):
            if hasattr(self.A, "_view_src"):
                self.a_gather = GatherPlan(ctx, self._gather_maps(self.A, False))
                self.gemm_op_a = 0 if self._k_fastest(self.A, False) else 1
            if hasattr(self.B, "_view_src"):
                self.b_gather = GatherPlan(ctx, self._gather_maps(self.B, True))
                self.gemm_op_b = 1 if self._k_fastest(self.B, True) else 0
        if precision.startswith(("bf16x", "i8x")):
            if self.dtype != torch.float64:
                raise ValueError("split precision takes FP64 operands")
            self.nslices = int(precision.split("x")[-1])
            self.a_split = self._split_plan(self.A, False, self.op_a)
            self.b_split = self._split_plan(self.B, True, self.op_b)
        elif precision not in ("f64", "f32", "f32x2"):
            raise ValueError(f"unknown precision {precision!r}")

    def _split_plan(self, X, is_right, op):
        """Split plan over the operand's fused tile grid (ordinals as in the
        work list's pairs) and its BF16 limb pool."""
        if hasattr(X, "_view_src"):
            return self._split_plan_general(X, is_right)
        Mt, Nt, Kt = self.Mt, self.Nt, self.Kt
        if not is_right:
            outer, rows_first = self.mext_h, not op  # NoTrans: (m, k)
            go, gi = Mt, Kt
        else:
            outer, rows_first = self.next_h, bool(op)  # Trans: (n, k)
            go, gi = Nt, Kt
        if rows_first:  # grid (outer x inner)
            rows = np.repeat(outer, gi)
            cols = np.tile(self.kext_h, go)
        else:  # grid (inner x outer)
            rows = np.repeat(self.kext_h, go)
            cols = np.tile(outer, gi)
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        offs = np.ascontiguousarray(X.offsets, np.int64)
        p = ctypes.c_void_p()
        call("tamx_split_plan_create_typed", self.ctx.ptr, self.limb_type, self.nslices,
             1 if is_right else 0, int(op), offs.size, offs.ctypes.data, rows.ctypes.data,
             cols.ctypes.data, ctypes.byref(p))
        return _SplitOperand(p, self._limb_pool(p, X), offs.size)

    def _limb_pool(self, plan, X):
        n = int(lib().tamx_split_plan_limb_elems(plan))
        dtype = torch.uint8 if self.limb_type == _lib.LIMB_I8 else torch.bfloat16
        return torch.zeros(max(1, n), dtype=dtype, device=X.pool.device)

    def _split_plan_general(self, X, is_right):
        """Split plan of a permuted view: limb rows are the operand's outer
        (non-contracted) modes, limb k its contracted modes (_gather_maps)."""
        nt, nr, nk, rext, rstr, kext, kstr = self._gather_maps(X, is_right)
        offs = np.ascontiguousarray(X.offsets, np.int64)
        p = ctypes.c_void_p()
        call("tamx_split_plan_create_general_typed", self.ctx.ptr, self.limb_type, self.nslices,
             nt, offs.ctypes.data, nr, nk, rext.ctypes.data, rstr.ctypes.data,
             kext.ctypes.data, kstr.ctypes.data, ctypes.byref(p))
        return _SplitOperand(p, self._limb_pool(p, X), nt)

    @staticmethod
    def _scatter_maps(l_outer, r_outer, rp):
        """Index maps of the natural result tiles (left-outer modes as rows,
        right-outer modes as columns) into the stored target-order tiles:
        natural mode d lands at target position rp[d] ("TO" convention), so
        its stride is the product of the target extents after rp[d].  None
        when a side has more than four modes (the result is then permuted
        after the GEMM)."""
        nl, nr = len(l_outer), len(r_outer)
        if not (1 <= nl <= 4 and 1 <= nr <= 4):
            return None
        ranges = list(l_outer) + list(r_outer)
        rank = nl + nr
        Mt = int(np.prod([r.ntiles for r in l_outer], dtype=np.int64))
        Nt = int(np.prod([r.ntiles for r in r_outer], dtype=np.int64))
        nt = Mt * Nt
        rext = np.ones((nt, nl), np.int32)
        rstr = np.zeros((nt, nl), np.int64)
        kext = np.ones((nt, nr), np.int32)
        kstr = np.zeros((nt, nr), np.int64)
        lidx = np.array(np.unravel_index(np.arange(Mt), [r.ntiles for r in l_outer])).T
        ridx = np.array(np.unravel_index(np.arange(Nt), [r.ntiles for r in r_outer])).T
        for i in range(Mt):
            for j in range(Nt):
                t = i * Nt + j
                e = ([int(l_outer[d].extents[lidx[i, d]]) for d in range(nl)] +
                     [int(r_outer[d].extents[ridx[j, d]]) for d in range(nr)])
                tgt = [0] * rank
                for d in range(rank):
                    tgt[rp[d]] = e[d]
                tstride = [int(np.prod(tgt[q + 1:], dtype=np.int64)) for q in range(rank)]
                rext[t] = e[:nl]
                kext[t] = e[nl:]
                rstr[t] = [tstride[rp[d]] for d in range(nl)]
                kstr[t] = [tstride[rp[d]] for d in range(nl, rank)]
        return nt, nl, nr, rext, rstr, kext, kstr

    def _gemm_modes(self, is_right):
        plan = self.plan
        rank = plan.right_rank if is_right else plan.left_rank
        inner = [bool((plan.right_gemm_is_inner if is_right else plan.left_gemm_is_inner)[d])
                 for d in range(rank)]
        return ([d for d in range(rank) if not inner[d]], [d for d in range(rank) if inner[d]])

    def _k_fastest(self, X, is_right):
        """Whether the stored stride-1 mode of a permuted view is its last
        contracted mode (else loads run along the rows)."""
        row_modes, k_modes = self._gemm_modes(is_right)
        rank = len(row_modes) + len(k_modes)
        inv = np.argsort(X._view_perm)
        last = [dp for dp in range(rank) if inv[dp] == rank - 1][0]
        return last != row_modes[-1]

    def _gather_maps(self, X, is_right):
        """Per-tile index maps of a permuted view: the extents and stored
        strides of its outer modes (rows of op(X), GEMM order) and of its
        contracted modes (k), by the view's tile ordinals."""
        plan = self.plan
        rank = plan.right_rank if is_right else plan.left_rank
        inner = [bool((plan.right_gemm_is_inner if is_right else plan.left_gemm_is_inner)[d])
                 for d in range(rank)]
        src, perm, src_of = X._view_src, X._view_perm, X._view_src_of
        inv = np.argsort(perm)  # permuted mode d' is source mode inv[d']
        sext = src.trange.tile_extents_table().astype(np.int64)  # by source ordinal
        row_modes = [d for d in range(rank) if not inner[d]]
        k_modes = [d for d in range(rank) if inner[d]]
        if not (1 <= len(row_modes) <= 4 and 1 <= len(k_modes) <= 4):
            raise ValueError("an index map covers 1 to 4 outer and 1 to 4 inner modes")
        nt = X.trange.ntiles
        rext = np.ones((nt, len(row_modes)), np.int32)
        rstr = np.zeros((nt, len(row_modes)), np.int64)
        kext = np.ones((nt, len(k_modes)), np.int32)
        kstr = np.zeros((nt, len(k_modes)), np.int64)
        for t in range(nt):
            e = sext[src_of[t]]
            strides = np.ones(rank, np.int64)  # row-major strides of the source tile
            for d in range(rank - 2, -1, -1):
                strides[d] = strides[d + 1] * e[d + 1]
            for w, dp in enumerate(row_modes):
                rext[t, w], rstr[t, w] = e[inv[dp]], strides[inv[dp]]
            for w, dp in enumerate(k_modes):
                kext[t, w], kstr[t, w] = e[inv[dp]], strides[inv[dp]]
        return nt, len(row_modes), len(k_modes), rext, rstr, kext, kstr

    def split_step(self, maxabs=None):
        """Split both operands into BF16 limbs (part of the split path).
        `maxabs`: optional (A, B) per-tile max |a| tensors (tile_norms with
        maxabs, by the operands' tile ordinals) that spare the max pass."""
        for w, (X, sp) in enumerate(((self.A, self.a_split), (self.B, self.b_split))):
            m = None if maxabs is None else maxabs[w]
            if m is None:
                call("tamx_split_f64", self.ctx.ptr, sp.ptr, X.pool.data_ptr(),
                     sp.limbs.data_ptr(), self.ctx.stream())
            else:
                call("tamx_split_f64_premax", self.ctx.ptr, sp.ptr, X.pool.data_ptr(),
                     m.data_ptr(), sp.limbs.data_ptr(), self.ctx.stream())

    def _side_stream(self):
        if self._side is None:
            self._side = torch.cuda.Stream(self.ctx.device)  # default priority
        return self._side

    def split_operand_async(self, which, maxabs=None):
        """Limb split of one operand (0: A, 1: B) on a side stream, after
        the work already enqueued on the current stream (typically that
        operand's norm pass, which gives `maxabs`), so that one operand's
        split runs while the other's norms are computed.  plan_and_join()
        runs the planning steps and makes the current stream wait for the
        splits."""
        X, sp = ((self.A, self.a_split), (self.B, self.b_split))[which]
        main = torch.cuda.current_stream(self.ctx.device)
        side = self._side_stream()
        side.wait_stream(main)
        with torch.cuda.stream(side):
            if maxabs is None:
                call("tamx_split_f64", self.ctx.ptr, sp.ptr, X.pool.data_ptr(),
                     sp.limbs.data_ptr(), self.ctx.stream())
            else:
                call("tamx_split_f64_premax", self.ctx.ptr, sp.ptr, X.pool.data_ptr(),
                     maxabs.data_ptr(), sp.limbs.data_ptr(), self.ctx.stream())
        if maxabs is not None:
            maxabs.record_stream(side)

    def plan_and_join(self):
        """Shape estimate and work list on the current stream, then wait for
        the limb splits enqueued by split_operand_async."""
        self.shape_step()
        self.worklist_step()
        torch.cuda.current_stream(self.ctx.device).wait_stream(self._side_stream())

    def split_and_plan(self, maxabs=None):
        """The limb split of both operands and the planning steps (shape
        estimate, work list), which do not depend on each other: the split
        runs on a side stream, the planning on the current one, and the
        current stream waits for both before it returns, so a gemm_step
        enqueued after it sees the limbs."""
        dev = self.ctx.device
        main = torch.cuda.current_stream(dev)
        side = self._side_stream()
        side.wait_stream(main)
        with torch.cuda.stream(side):
            self.split_step(maxabs)
        if maxabs is not None:
            for m in maxabs:
                if m is not None:
                    m.record_stream(side)
        self.shape_step()
        self.worklist_step()
        main.wait_stream(side)

    def sigmas(self):
        out = []
        for sp in (self.a_split, self.b_split):
            buf = np.zeros(sp.ntiles, np.int32)
            call("tamx_split_plan_sigmas", sp.ptr, buf.ctypes.data)
            out.append(buf)
        return out

    def _make_desc(self):
        d = _lib.ContractionDesc()
        d.op_a, d.op_b = self.op_a, self.op_b
        d.Mt, d.Nt, d.Kt, d.nbatch = self.Mt, self.Nt, self.Kt, 1
        d.mext, d.next, d.kext = (self.d_mext.data_ptr(), self.d_next.data_ptr(),
                                  self.d_kext.data_ptr())
        d.a_off, d.b_off = self.A.d_offsets.data_ptr(), self.B.d_offsets.data_ptr()
        if self.sparse:
            d.a_norms, d.b_norms = self.A.shape_norms.data_ptr(), self.B.shape_norms.data_ptr()
            d.a_threshold, d.b_threshold = self.A.shape_threshold, self.B.shape_threshold
            d.c_threshold = self.threshold
        else:
            d.a_norms = d.b_norms = None
            d.a_threshold = d.b_threshold = 0.0
            d.c_threshold = 0.0
        d.c_norms = self.c_norms.data_ptr()
        d.pair_threshold = self.pair_threshold
        d.store_trans = 1 if self.store_trans else 0
        d.c_off = self.c_off.data_ptr()
        d.c_capacity = self.c_capacity
        return d

    def set_tile(self, tile):
        """tile: None (library default) or (bm, bn[, bk[, wgs_per_cu]])."""
        self.tile = None if tile is None else tuple(int(x) for x in tile)
        if self.tile is not None:
            t = self.tile + (16, 0)[len(self.tile) - 2:] if len(self.tile) < 4 else self.tile
            self.tile_cfg = _lib.GemmConfig(*t[:4])

    def tune(self, reps=3, force=False, warmup=True):
        """Autotunes the tile configuration on the current work list (built
        by worklist_step, with the limbs split for a split precision) and
        keeps it; FP64, FP32, BF16-limb and INT8-limb precisions."""
        if self.nslices:
            cfg = _lib.GemmConfig()
            ms = ctypes.c_float()
            flags = (_lib.TUNE_FORCE if force else 0) | (0 if warmup else _lib.TUNE_NO_WARMUP)
            call("tamx_autotune_grouped_gemm_split", self.ctx.ptr, self.wl.ptr, self.a_split.ptr,
                 self.a_split.limbs.data_ptr(), self.b_split.ptr, self.b_split.limbs.data_ptr(),
                 self.C.data_ptr(), int(reps), flags, ctypes.byref(cfg), ctypes.byref(ms),
                 self.ctx.stream())
            tile = (cfg.bm, cfg.bn, cfg.bk, cfg.wgs_per_cu)
            self.set_tile(tile)
            return tile, ms.value
        if self.precision not in ("f64", "f32"):
            raise ValueError("autotuning covers the f64, f32 and limb-split grouped GEMMs")
        tile, ms = autotune(self.ctx, self.wl, self.C.dtype, self.gemm_op_a, self.gemm_op_b,
                            self.A.pool, self.B.pool, self.C, reps=reps, force=force,
                            warmup=warmup, a_gather=self.a_gather, b_gather=self.b_gather)
        self.set_tile(tile)
        return tile, ms

    # -- device steps -------------------------------------------------------
    def permute_step(self):
        """Re-permutes `general` operands from their source arrays."""
        for src, dst in ((self.A_src, self.A), (self.B_src, self.B)):
            if dst is not src and hasattr(dst, "_perm_plan"):
                dst._perm_plan.run(src.pool, dst.pool)

    def shape_step(self):
        """Result shape estimate (SparseShape::gemm) into c_norms."""
        if not self.sparse:
            return
        call("tamx_shape_gemm_f32", self.ctx.ptr, self.A.shape_norms.data_ptr(),
             self.B.shape_norms.data_ptr(), self.op_a, self.op_b, self.Mt, self.Nt, self.Kt, 1,
             self.k_sizes.data_ptr(), ctypes.c_float(abs(self.alpha)),
             ctypes.c_float(self.threshold), self.c_norms.data_ptr(), None, self.ctx.stream())

    def worklist_step(self):
        call("tamx_contraction_worklist", self.ctx.ptr, ctypes.byref(self._desc), self.wl.ptr,
             self.ctx.stream())

    def gemm_step(self):
        s = self.ctx.stream()
        A, B, C = self.A.pool.data_ptr(), self.B.pool.data_ptr(), self.C.data_ptr()
        cfg = None if self.tile is None else ctypes.byref(self.tile_cfg)
        if self.c_scatter is not None:
            sc = self.c_scatter.ptr
            if self.precision in ("f64", "f32"):
                f64 = self.precision == "f64"
                call("tamx_grouped_gemm_scatter_f64" if f64 else "tamx_grouped_gemm_scatter_f32",
                     self.ctx.ptr, self.wl.ptr, self.gemm_op_a, self.gemm_op_b,
                     ctypes.c_double(self.alpha) if f64 else ctypes.c_float(self.alpha),
                     A, _plan_ptr(self.a_gather), B, _plan_ptr(self.b_gather), C, sc, cfg, s)
            elif self.precision == "f32x2":
                call("tamx_grouped_gemm_f32x2_scatter", self.ctx.ptr, self.wl.ptr, self.op_a,
                     self.op_b, ctypes.c_double(self.alpha), A, B, C, sc, s)
            else:
                call("tamx_grouped_gemm_split_scatter", self.ctx.ptr, self.wl.ptr,
                     ctypes.c_double(self.alpha), self.a_split.ptr,
                     self.a_split.limbs.data_ptr(), self.b_split.ptr,
                     self.b_split.limbs.data_ptr(), C, sc, cfg, s)
        elif self.a_gather is not None or self.b_gather is not None:
            f64 = self.precision == "f64"
            call("tamx_grouped_gemm_gather_f64" if f64 else "tamx_grouped_gemm_gather_f32",
                 self.ctx.ptr, self.wl.ptr, self.gemm_op_a, self.gemm_op_b,
                 ctypes.c_double(self.alpha) if f64 else ctypes.c_float(self.alpha),
                 A, _plan_ptr(self.a_gather), B, _plan_ptr(self.b_gather), C, cfg, s)
        elif self.precision == "f64":
            call("tamx_grouped_gemm_f64_config", self.ctx.ptr, self.wl.ptr, self.op_a, self.op_b,
                 ctypes.c_double(self.alpha), A, B, C, cfg, s)
        elif self.precision == "f32":
            call("tamx_grouped_gemm_f32_config", self.ctx.ptr, self.wl.ptr, self.op_a, self.op_b,
                 ctypes.c_float(self.alpha), A, B, C, cfg, s)
        elif self.precision == "f32x2":
            call("tamx_grouped_gemm_f32x2", self.ctx.ptr, self.wl.ptr, self.op_a, self.op_b,
                 ctypes.c_double(self.alpha), A, B, C, s)
        else:
            call("tamx_grouped_gemm_split_config", self.ctx.ptr, self.wl.ptr,
                 ctypes.c_double(self.alpha), self.a_split.ptr, self.a_split.limbs.data_ptr(),
                 self.b_split.ptr, self.b_split.limbs.data_ptr(), C, cfg, s)

    def run(self):
        if self.permute_operands:
            self.permute_step()
        if self.nslices:
            self.split_and_plan()
        else:
            self.shape_step()
            self.worklist_step()
        self.gemm_step()

    # -- results ------------------------------------------------------------
    def counters(self):
        c = self.wl.counters()
        if c[_lib.WL_STATUS] != 0:
            raise _lib.TamxError(c[_lib.WL_STATUS], "tamx_contraction_worklist",
                                 f"capacity exceeded (tasks {c[5]}, pairs {c[6]})")
        return {"ntasks": c[_lib.WL_NTASKS], "npairs": c[_lib.WL_NPAIRS],
                "celems": c[_lib.WL_CELEMS], "flops": c[_lib.WL_FLOPS]}

    def result(self):
        """The result DistArray in target order.  Its tiles stay in HBM and
        nothing is read back: identity, transposed and (up to four modes a
        side) general result orders are written by the GEMM itself; only a
        general order with more modes a side runs a permute pass, whose
        plan needs the result offsets on the host."""
        plan = self.plan
        nat = TiledRange(self.natural_ranges)
        rp = [int(plan.result_perm[d]) for d in range(plan.result_rank)]
        norms = self.c_norms if self.sparse else None
        thr = self.threshold if self.sparse else None
        if plan.result_perm_identity:
            return DistArray(nat, self.C, d_offsets=self.c_off, shape_norms=norms,
                             shape_threshold=thr)
        tgt = nat.permuted(rp)
        inv = [int(i) for i in np.argsort(rp)]
        grid_perm = lambda t: t.view(nat.tiles_shape).permute(*inv).contiguous().view(-1)
        tnorms = grid_perm(norms) if norms is not None else None
        if self.store_trans or self.c_scatter is not None:
            return DistArray(tgt, self.C, d_offsets=grid_perm(self.c_off), shape_norms=tnorms,
                             shape_threshold=thr)
        # general result permutation: permute every non-zero tile in place
        # of its pool slot (contract_reduce.h:370-378)
        offs = self.c_off.cpu().numpy()
        present = offs >= 0
        ext = nat.tile_extents_table()
        pplan = PermutePlan(self.ctx, rp, ext[present], offs[present], offs[present])
        out = torch.empty_like(self.C)
        if pplan.volume:
            pplan.run(self.C, out)
        return DistArray(tgt, out, d_offsets=grid_perm(self.c_off), shape_norms=tnorms,
                         shape_threshold=thr)


def contract(ctx, target, A, left, B, right, **kw):
    """C(target) = alpha * A(left) * B(right); returns (result, counters)."""
    c = Contraction(ctx, target, A, left, B, right, **kw)
    c.run()
    return c.result(), c.counters()


class BatchedContraction:
    """nbatch independent block-sparse products C_n = A_n * B_n in one
    persistent grouped launch: the batched (fused-index) product of
    tile_op/batched_contract_reduce.h:48-240 and the tensor-of-tensor
    element-op GEMM of tensor/tensor.h:3226-3680 (outer cells are tiles, the
    inner cell matrices are their data).

    Every instance has its own Mt x Kt and Kt x Nt cell grids (the same
    grid sizes) and its own inner extents; shapes are built from cell norms
    (float, inverse-volume scaled) and screened with `threshold`, exactly as
    for a single contraction, batched.  precision: None (the cells' type,
    FP64 or FP32 MFMA) or "i8xS" / "bf16xS" (FP64 cells split into S INT8 or
    BF16 limbs, each cell one tile of the split plan)."""

    def __init__(self, ctx, nbatch, Mt, Nt, Kt, mdims, ndims, kdims, A, a_off, B, b_off,
                 threshold=FLOAT_EPS, alpha=1.0, tile=None, precision=None, cell_shapes=True):
        self.ctx = ctx
        dev = A.device
        self.nbatch, self.Mt, self.Nt, self.Kt = int(nbatch), int(Mt), int(Nt), int(Kt)
        self.dtype = A.dtype
        self.A, self.B = A, B
        self.alpha = float(alpha)
        self.threshold = float(np.float32(threshold))
        md = np.asarray(mdims, np.int64).reshape(nbatch, Mt)
        nd = np.asarray(ndims, np.int64).reshape(nbatch, Nt)
        kd = np.asarray(kdims, np.int64).reshape(nbatch, Kt)
        self.mdims, self.ndims, self.kdims = md, nd, kd
        self.a_off_h = np.ascontiguousarray(a_off, np.int64)
        self.b_off_h = np.ascontiguousarray(b_off, np.int64)
        self.d_a_off = torch.as_tensor(self.a_off_h, device=dev)
        self.d_b_off = torch.as_tensor(self.b_off_h, device=dev)
        self.a_norms = torch.zeros(nbatch * Mt * Kt, dtype=torch.float32, device=dev)
        self.b_norms = torch.zeros(nbatch * Kt * Nt, dtype=torch.float32, device=dev)
        self.c_norms = torch.zeros(nbatch * Mt * Nt, dtype=torch.float32, device=dev)
        # cell_shapes=False: the caller writes a_norms, b_norms and c_norms
        # itself (FusedContraction's per-tile screening), so the cell norm
        # plans and size vectors of shape_step are not built
        self.cell_shapes = bool(cell_shapes)
        if self.cell_shapes:
            a_vol = (md[:, :, None] * kd[:, None, :]).reshape(-1)
            b_vol = (kd[:, :, None] * nd[:, None, :]).reshape(-1)
            self.a_plan = NormPlan(ctx, self.a_off_h, a_vol)
            self.b_plan = NormPlan(ctx, self.b_off_h, b_vol)
            f32 = np.float32
            self.a_sizes = torch.as_tensor(
                np.concatenate([md, kd], axis=1).astype(f32).reshape(-1), device=dev)
            self.b_sizes = torch.as_tensor(
                np.concatenate([kd, nd], axis=1).astype(f32).reshape(-1), device=dev)
            self.k_sizes = torch.as_tensor(kd.astype(f32).reshape(-1), device=dev)
        self.d_mext = torch.as_tensor(md.astype(np.int32).reshape(-1), device=dev)
        self.d_next = torch.as_tensor(nd.astype(np.int32).reshape(-1), device=dev)
        self.d_kext = torch.as_tensor(kd.astype(np.int32).reshape(-1), device=dev)
        self.c_capacity = int(sum(int(md[b].sum()) * int(nd[b].sum()) for b in range(nbatch)))
        self.C = torch.zeros(max(1, self.c_capacity), dtype=self.dtype, device=dev)
        self.c_off = torch.empty(nbatch * Mt * Nt, dtype=torch.int64, device=dev)
        grid = nbatch * Mt * Nt
        self.wl = WorkList(ctx, grid, grid * Kt, grid)
        d = _lib.ContractionDesc()
        d.op_a, d.op_b = 0, 0
        d.Mt, d.Nt, d.Kt, d.nbatch = self.Mt, self.Nt, self.Kt, self.nbatch
        d.mext, d.next, d.kext = (self.d_mext.data_ptr(), self.d_next.data_ptr(),
                                  self.d_kext.data_ptr())
        d.a_off, d.b_off = self.d_a_off.data_ptr(), self.d_b_off.data_ptr()
        d.a_norms, d.b_norms = self.a_norms.data_ptr(), self.b_norms.data_ptr()
        d.a_threshold = d.b_threshold = d.c_threshold = self.threshold
        d.c_norms = self.c_norms.data_ptr()
        d.pair_threshold = 0.0
        d.store_trans = 0
        d.c_off = self.c_off.data_ptr()
        d.c_capacity = self.c_capacity
        self._desc = d
        self.tile_cfg = None if tile is None else _lib.GemmConfig(*(tuple(tile) + (16, 0))[:4])
        # split precision (FP64 cells): every cell is one tile of the limb
        # split, with the ordinal the work list's pairs carry (b, i, k) and
        # (b, k, j)
        self.precision = precision or ("f64" if self.dtype == torch.float64 else "f32")
        self.nslices = 0
        if self.precision.startswith(("bf16x", "i8x")):
            if self.dtype != torch.float64:
                raise ValueError("BatchedContraction: split precisions need FP64 cells")
            self.nslices = int(self.precision.split("x")[-1])
            self.limb_type = _lib.LIMB_I8 if self.precision.startswith("i8x") else _lib.LIMB_BF16
            a_rows = np.repeat(md, Kt, axis=1).reshape(-1)
            a_cols = np.tile(kd, (1, Mt)).reshape(-1)
            b_rows = np.repeat(kd, Nt, axis=1).reshape(-1)
            b_cols = np.tile(nd, (1, Kt)).reshape(-1)
            self.a_split = self._split_plan(self.a_off_h, a_rows, a_cols, False)
            self.b_split = self._split_plan(self.b_off_h, b_rows, b_cols, True)
        elif self.precision not in ("f64", "f32"):
            raise ValueError(f"BatchedContraction: unknown precision {self.precision!r}")

    def _split_plan(self, offs, rows, cols, is_right):
        offs = np.ascontiguousarray(offs, np.int64)
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        p = ctypes.c_void_p()
        call("tamx_split_plan_create_typed", self.ctx.ptr, self.limb_type, self.nslices,
             1 if is_right else 0, 0, offs.size, offs.ctypes.data, rows.ctypes.data,
             cols.ctypes.data, ctypes.byref(p))
        n = int(lib().tamx_split_plan_limb_elems(p))
        dtype = torch.uint8 if self.limb_type == _lib.LIMB_I8 else torch.bfloat16
        return _SplitOperand(p, torch.zeros(max(1, n), dtype=dtype, device=self.A.device),
                             offs.size)

    def shape_step(self):
        """Cell norms, inverse-volume scaling, batched result estimate."""
        if not self.cell_shapes:
            raise RuntimeError("BatchedContraction: built without cell shapes")
        s = self.ctx.stream()
        f = "tamx_tile_norms_f64" if self.dtype == torch.float64 else "tamx_tile_norms_f32"
        call(f, self.ctx.ptr, self.a_plan.ptr, self.A.data_ptr(), self.a_norms.data_ptr(), s)
        call(f, self.ctx.ptr, self.b_plan.ptr, self.B.data_ptr(), self.b_norms.data_ptr(), s)
        for norms, sizes, grid in ((self.a_norms, self.a_sizes, (self.Mt, self.Kt)),
                                   (self.b_norms, self.b_sizes, (self.Kt, self.Nt))):
            nt = np.ascontiguousarray(grid, np.int64)
            call("tamx_shape_scale", self.ctx.ptr, norms.data_ptr(), 2, nt.ctypes.data,
                 self.nbatch, sizes.data_ptr(), 1, ctypes.c_float(self.threshold), None, s)
        call("tamx_shape_gemm_f32", self.ctx.ptr, self.a_norms.data_ptr(), self.b_norms.data_ptr(),
             0, 0, self.Mt, self.Nt, self.Kt, self.nbatch, self.k_sizes.data_ptr(),
             ctypes.c_float(abs(self.alpha)), ctypes.c_float(self.threshold),
             self.c_norms.data_ptr(), None, s)

    def worklist_step(self):
        call("tamx_contraction_worklist", self.ctx.ptr, ctypes.byref(self._desc), self.wl.ptr,
             self.ctx.stream())

    def gemm_step(self):
        s = self.ctx.stream()
        cfg = None if self.tile_cfg is None else ctypes.byref(self.tile_cfg)
        if self.nslices:
            for X, sp in ((self.A, self.a_split), (self.B, self.b_split)):
                call("tamx_split_f64", self.ctx.ptr, sp.ptr, X.data_ptr(), sp.limbs.data_ptr(), s)
            call("tamx_grouped_gemm_split_config", self.ctx.ptr, self.wl.ptr,
                 ctypes.c_double(self.alpha), self.a_split.ptr, self.a_split.limbs.data_ptr(),
                 self.b_split.ptr, self.b_split.limbs.data_ptr(), self.C.data_ptr(), cfg, s)
        elif self.dtype == torch.float64:
            call("tamx_grouped_gemm_f64_config", self.ctx.ptr, self.wl.ptr, 0, 0,
                 ctypes.c_double(self.alpha), self.A.data_ptr(), self.B.data_ptr(),
                 self.C.data_ptr(), cfg, s)
        else:
            call("tamx_grouped_gemm_f32_config", self.ctx.ptr, self.wl.ptr, 0, 0,
                 ctypes.c_float(self.alpha), self.A.data_ptr(), self.B.data_ptr(),
                 self.C.data_ptr(), cfg, s)

    def run(self):
        self.shape_step()
        self.worklist_step()
        self.gemm_step()

    def counters(self):
        c = self.wl.counters()
        if c[_lib.WL_STATUS] != 0:
            raise _lib.TamxError(c[_lib.WL_STATUS], "tamx_contraction_worklist",
                                 "capacity exceeded")
        return {"ntasks": c[_lib.WL_NTASKS], "npairs": c[_lib.WL_NPAIRS],
                "celems": c[_lib.WL_CELEMS], "flops": c[_lib.WL_FLOPS]}

    def tune(self, reps=3, force=False, warmup=True):
        """Autotunes the tile configuration on the current work list (built
        by worklist_step) and keeps it."""
        tile, ms = autotune(self.ctx, self.wl, self.dtype, 0, 0, self.A, self.B, self.C,
                            reps=reps, force=force, warmup=warmup)
        self.tile_cfg = _lib.GemmConfig(*tile)
        return tile, ms

    def capture(self):
        """Captures run() into a graph (tamx_graph_begin/_end); returns a
        Graph whose replay() relaunches it.  Every step is device work with
        no host synchronisation, no allocation and no memset, so the graph
        replays the whole contraction."""
        side = torch.cuda.Stream(self.ctx.device)
        side.wait_stream(torch.cuda.current_stream(self.ctx.device))
        with torch.cuda.stream(side):
            self.run()  # warm caches (occupancy queries) outside the capture
        torch.cuda.current_stream(self.ctx.device).wait_stream(side)
        torch.cuda.synchronize()
        return Graph(self.ctx, self.run)


class FusedContraction:
    """C(F, I, J) = alpha * A(F, I, K) * B(F, K, J) with F Hadamard (fused)
    modes, leading in all three, K contracted and I, J free mode groups (the
    groups hold nf, ni, nk, nj modes; ni, nk or nj may be 0): TA's batched
    Summa with BatchedContractReduce (expressions/cont_engine.h:684-1100,
    tile_op/batched_contract_reduce.h:48-240, dist_eval/contraction_eval.h:
    96-187), where every element of the fused modes is an independent
    product C_f = A_f * B_f (tensor.h:3211-3215).

    Each group's modes are fused into one (their tile combinations in
    row-major order, as GemmHelper fuses modes).  Every fused element is one
    instance of the batched launch (BatchedContraction): its cells are slices
    of the stored tiles, A tile (f, i, k) giving the slice at offset
    + p * m_i * k_k for element p of fused tile f, so nothing is copied in.
    result() copies the instances' result cells into (F, I, J) tiles with the
    permute kernel (identity order).

    threshold None: no screening, every result tile is formed (dense arrays).
    threshold t: A and B carry their shapes (make_sparse), and the screening
    is TA's per tile, SparseShape::gemm_batched (sparse_shape.h:1695-1880):
    the result estimate is the shape GEMM of each fused tile-slab (batched
    over the fused tiles, contracted rows scaled by the contracted tile
    volumes), tiles below t are zero, and the operand tile norms and the
    estimate are replicated over every element of their fused tile, so a
    pair or a result cell is kept exactly when its tile is."""

    def __init__(self, ctx, A, B, groups=(1, 1, 1, 1), threshold=None, alpha=1.0,
                 precision=None):
        nf, ni, nk, nj = (int(g) for g in groups)
        ta, tb = A.trange.ranges, B.trange.ranges
        if nf < 1 or len(ta) != nf + ni + nk or len(tb) != nf + nk + nj:
            raise ValueError("FusedContraction: A is (F, I, K) and B is (F, K, J), F not empty")
        if not (all(x == y for x, y in zip(ta[:nf], tb[:nf]))
                and all(x == y for x, y in zip(ta[nf + ni:], tb[nf:nf + nk]))):
            raise ValueError("FusedContraction: the fused or contracted tilings of A and B differ")
        if A.pool.dtype != B.pool.dtype:
            raise ValueError("FusedContraction: A and B differ in type")
        self.ctx, self.A, self.B = ctx, A, B
        self.ranges = list(ta[:nf]) + list(ta[nf:nf + ni]) + list(tb[nf + nk:])
        fe = _fused_extents(ta[:nf])
        me = _fused_extents(ta[nf:nf + ni])
        ke = _fused_extents(ta[nf + ni:])
        je = _fused_extents(tb[nf + nk:])
        nt, mt, kt, jt = fe.size, me.size, ke.size, je.size
        # element p of fused tile t: global element index, tile and position
        self.elem_tile = np.repeat(np.arange(nt), fe)
        self.elem_pos = (np.concatenate([np.arange(w) for w in fe]) if fe.sum()
                         else np.zeros(0, np.int64))
        N = int(self.elem_tile.size)
        self.N, self.Mt, self.Kt, self.Jt = N, mt, kt, jt
        self.me, self.je = me, je
        ao = A.offsets.reshape(nt, mt, kt)
        bo = B.offsets.reshape(nt, kt, jt)
        t, p = self.elem_tile, self.elem_pos
        a_slice = (me[:, None] * ke[None, :]).astype(np.int64)   # mt x kt
        b_slice = (ke[:, None] * je[None, :]).astype(np.int64)   # kt x jt
        a_off = np.where(ao[t] >= 0, ao[t] + p[:, None, None] * a_slice[None], -1)
        b_off = np.where(bo[t] >= 0, bo[t] + p[:, None, None] * b_slice[None], -1)
        thr = 0.0 if threshold is None else float(np.float32(threshold))
        nb = max(N, 1)
        self.tile_shapes = threshold is not None
        self.bc = BatchedContraction(ctx, nb, mt, jt, kt, np.tile(me, nb), np.tile(je, nb),
                                     np.tile(ke, nb), A.pool,
                                     a_off.reshape(-1) if N else np.full(mt * kt, -1),
                                     B.pool, b_off.reshape(-1) if N else np.full(kt * jt, -1),
                                     threshold=thr, alpha=alpha, precision=precision,
                                     cell_shapes=not self.tile_shapes)
        if self.tile_shapes:
            if A.shape_norms is None or B.shape_norms is None:
                raise ValueError("FusedContraction: a threshold needs A and B with shapes "
                                 "(make_sparse)")
            dev = A.pool.device
            self.nt, self.alpha, self.threshold = nt, float(alpha), thr
            self.est = torch.zeros(nt * mt * jt, dtype=torch.float32, device=dev)
            kz = k_sizes(ta[nf + ni:]) if nk else np.ones(1, np.float32)
            self.ksz = torch.as_tensor(np.tile(kz, nt), device=dev)
            self.d_elem_tile = torch.as_tensor(self.elem_tile, device=dev)
            # operand tiles are screened by their own shapes' thresholds
            self.bc._desc.a_threshold = float(A.shape_threshold)
            self.bc._desc.b_threshold = float(B.shape_threshold)

    def _tile_shape_step(self):
        """The per-tile result estimate (batched over the fused tiles) and
        the tile norms and estimate replicated over the fused elements."""
        mt, kt, jt, nt = self.Mt, self.Kt, self.Jt, self.nt
        call("tamx_shape_gemm_f32", self.ctx.ptr, self.A.shape_norms.data_ptr(),
             self.B.shape_norms.data_ptr(), 0, 0, mt, jt, kt, nt, self.ksz.data_ptr(),
             ctypes.c_float(abs(self.alpha)), ctypes.c_float(self.threshold),
             self.est.data_ptr(), None, self.ctx.stream())
        if self.N:
            # gathers into the preallocated per-element buffers (no
            # temporaries, so the step can be captured in a graph)
            t, N = self.d_elem_tile, self.N
            torch.index_select(self.A.shape_norms.view(nt, mt * kt), 0, t,
                               out=self.bc.a_norms.view(N, mt * kt))
            torch.index_select(self.B.shape_norms.view(nt, kt * jt), 0, t,
                               out=self.bc.b_norms.view(N, kt * jt))
            torch.index_select(self.est.view(nt, mt * jt), 0, t,
                               out=self.bc.c_norms.view(N, mt * jt))

    def run(self):
        if self.tile_shapes:
            self._tile_shape_step()
            self.bc.worklist_step()
            self.bc.gemm_step()
        else:
            self.bc.run()

    def counters(self):
        return self.bc.counters()

    def result(self):
        """The result DistArray over (F, I, J) (synchronises)."""
        N, mt, jt = self.N, self.Mt, self.Jt
        trange = TiledRange(self.ranges)
        ntiles_f = trange.ntiles // (mt * jt)
        c_off = self.bc.c_off.cpu().numpy()[:N * mt * jt].reshape(N, mt, jt)
        present = np.zeros((ntiles_f, mt, jt), bool)
        for t in range(ntiles_f):
            sel = self.elem_tile == t
            if sel.any():
                present[t] = (c_off[sel] >= 0).any(axis=0)
        R = DistArray.allocate(trange, present.reshape(-1), dtype=self.A.pool.dtype,
                               device=self.A.pool.device)
        R.pool.zero_()
        if self.tile_shapes:
            # the screened result carries its shape: the per-tile estimate,
            # (F, I, J) tiles in row-major order as the result's tile grid
            # (SparseShape::gemm_batched, sparse_shape.h:1695-1880)
            R.shape_norms = self.est.clone()
            R.shape_threshold = self.threshold
        ro = R.offsets.reshape(ntiles_f, mt, jt)
        me, je = self.me, self.je
        n, i, j = np.nonzero(c_off >= 0)
        if n.size:
            t, p = self.elem_tile[n], self.elem_pos[n]
            ext = np.stack([me[i], je[j]], axis=1)
            dst = ro[t, i, j] + p * me[i] * je[j]
            plan = PermutePlan(self.ctx, [0, 1], ext, c_off[n, i, j], dst)
            if plan.volume:
                plan.run(self.bc.C, R.pool)
        return R


def fused_contract(ctx, target, A, left, B, right, alpha=1.0, threshold=None, precision=None):
    """target = alpha * A(left) * B(right) where some indices are Hadamard
    (fused: in left, right and target), some contracted (in left and right
    only) and the rest free, in any order: the operands are brought to
    (F, I, K) and (F, K, J) by the permute kernel where their order differs
    (TA permutes `general` operands the same way, tensor/permute.h:121-207),
    with F, I and J in target order and K in left order; the product runs as
    FusedContraction, and the result is permuted from (F, I, J) to the
    target order.  Returns (result, counters)."""
    L = [x.strip() for x in left.split(",")]
    Rr = [x.strip() for x in right.split(",")]
    T = [x.strip() for x in target.split(",")]
    if len(set(L)) != len(L) or len(set(Rr)) != len(Rr) or len(set(T)) != len(T):
        raise ValueError("fused_contract: an index repeats within an operand or the result")
    F = [x for x in T if x in L and x in Rr]
    I = [x for x in T if x in L and x not in Rr]
    J = [x for x in T if x in Rr and x not in L]
    K = [x for x in L if x in Rr and x not in T]
    if not F or set(T) != set(F + I + J) or set(L) != set(F + I + K) or set(Rr) != set(F + K + J):
        raise ValueError("fused_contract: needs a fused index, and every index of the result "
                         "in an operand")

    def in_order(X, idx, want):
        if idx == want:
            return X
        return permute_array(ctx, X, [want.index(x) for x in idx])  # TO convention

    f = FusedContraction(ctx, in_order(A, L, F + I + K), in_order(B, Rr, F + K + J),
                         groups=(len(F), len(I), len(K), len(J)), threshold=threshold,
                         alpha=alpha, precision=precision)
    f.run()
    R = f.result()
    if T != F + I + J:
        R = permute_array(ctx, R, [T.index(x) for x in F + I + J])
    return R, f.counters()
