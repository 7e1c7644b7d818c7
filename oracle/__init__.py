"""CPU oracle for TiledArray's one-rank contraction path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py may import it.
It restates the reference's behaviour (each function in ta_oracle.c cites the
TiledArray file:line it follows) and is never the thing measured or shipped.
"""
import ctypes
import os

import numpy as np

from . import permopt  # noqa: F401  (restated GEMMPermutationOptimizer)

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libta_oracle.so")

_lib = None

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            from .build import build
            build()
        L = ctypes.CDLL(LIBPATH)
        L.ora_blas_name.restype = ctypes.c_char_p
        L.ora_num_threads.restype = ctypes.c_int
        L.ora_set_num_threads.argtypes = [ctypes.c_int]
        L.ora_norm_f64.restype = ctypes.c_float
        L.ora_norm_f64.argtypes = [_vp, ctypes.c_int64]
        L.ora_norm_f32.restype = ctypes.c_float
        L.ora_norm_f32.argtypes = [_vp, ctypes.c_int64]
        for n in ("ora_tile_norms_f64", "ora_tile_norms_f32"):
            getattr(L, n).argtypes = [_vp, _vp, _vp, ctypes.c_int64, _vp]
        for n in ("ora_tot_tile_norms_f64", "ora_tot_tile_norms_f32"):
            getattr(L, n).argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int64, _vp]
        L.ora_scale_tile_norms.restype = ctypes.c_int64
        L.ora_scale_tile_norms.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_float]
        L.ora_shape_gemm.restype = ctypes.c_int64
        L.ora_shape_gemm.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                     ctypes.c_int64, ctypes.c_int64, _vp, ctypes.c_float,
                                     ctypes.c_float, _vp]
        L.ora_dgemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                ctypes.c_int64, ctypes.c_double, _vp, ctypes.c_int64, _vp,
                                ctypes.c_int64, ctypes.c_double, _vp, ctypes.c_int64]
        L.ora_sgemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                ctypes.c_int64, ctypes.c_float, _vp, ctypes.c_int64, _vp,
                                ctypes.c_int64, ctypes.c_float, _vp, ctypes.c_int64]
        for n in ("ora_permute_f64", "ora_permute_f32"):
            getattr(L, n).argtypes = [_vp, _vp, ctypes.c_int, _vp, _vp]
        for n, sc in (("ora_contract_f64", ctypes.c_double), ("ora_contract_f32", ctypes.c_float)):
            f = getattr(L, n)
            f.restype = ctypes.c_int64
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                          ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, sc,
                          _vp, _vp, ctypes.c_float]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def blas_name():
    return lib().ora_blas_name().decode()


def num_threads():
    return lib().ora_num_threads()


def set_num_threads(n):
    lib().ora_set_num_threads(int(n))


def tile_norm(x):
    x = np.ascontiguousarray(x)
    if x.dtype == np.float64:
        return np.float32(lib().ora_norm_f64(_p(x), x.size))
    x = _c(x, np.float32)
    return np.float32(lib().ora_norm_f32(_p(x), x.size))


def tile_norms(pool, offsets, volumes):
    pool = np.ascontiguousarray(pool)
    offsets = _c(offsets, np.int64)
    volumes = _c(volumes, np.int64)
    out = np.zeros(offsets.size, np.float32)
    f = lib().ora_tile_norms_f64 if pool.dtype == np.float64 else lib().ora_tile_norms_f32
    f(_p(pool), _p(offsets), _p(volumes), offsets.size, _p(out))
    return out


def tot_tile_norms(pool, tile_cells, cell_off, cell_vol):
    """Outer-tile norms of a tensor of tensors (cells reduced in order,
    each cell's own sequential sum added; tensor.h:3768-3776)."""
    pool = np.ascontiguousarray(pool)
    tile_cells = _c(tile_cells, np.int64)
    cell_off = _c(cell_off, np.int64)
    cell_vol = _c(cell_vol, np.int64)
    nt = tile_cells.size - 1
    out = np.zeros(max(nt, 0), np.float32)
    f = lib().ora_tot_tile_norms_f64 if pool.dtype == np.float64 else lib().ora_tot_tile_norms_f32
    f(_p(pool), _p(tile_cells), _p(cell_off), _p(cell_vol), nt, _p(out))
    return out


def scale_tile_norms(norms, tile_extents, threshold):
    """SparseShape ctor scaling: norms (tile-grid shaped) are scaled in place
    by inverse tile volume and screened; returns (scaled, zero_count)."""
    norms = np.array(norms, dtype=np.float32, copy=True, order="C")
    rank = len(tile_extents)
    ntiles = np.array([len(e) for e in tile_extents], np.int64)
    sizes = np.concatenate([np.asarray(e, np.float32) for e in tile_extents]).astype(np.float32)
    z = lib().ora_scale_tile_norms(_p(norms), rank, _p(ntiles), _p(sizes), np.float32(threshold))
    return norms, int(z)


def k_sizes(inner_extents):
    """recursive_outer_product over the contracted modes' tile extents with
    the identity op (sparse_shape.h:1626-1630), float products."""
    def rop(vs):
        if len(vs) == 1:
            return np.asarray(vs[0], np.float32)
        mid = (len(vs) >> 1) + (len(vs) & 1)
        l, r = rop(vs[:mid]), rop(vs[mid:])
        return (l[:, None] * r[None, :]).reshape(-1).astype(np.float32)
    return rop([np.asarray(e, np.float32) for e in inner_extents])


def shape_gemm(left, right, op_a, op_b, M, N, K, ksz, abs_factor, threshold):
    left = _c(left, np.float32)
    right = _c(right, np.float32)
    ksz = _c(ksz, np.float32)
    out = np.zeros(M * N, np.float32)
    z = lib().ora_shape_gemm(_p(left), _p(right), int(op_a), int(op_b), M, N, K, _p(ksz),
                             np.float32(abs_factor), np.float32(threshold), _p(out))
    return out.reshape(M, N), int(z)


def gemm(op_a, op_b, alpha, A, B, beta=0.0, C=None):
    """Row-major TA gemm on 2-D numpy matrices (A stored k x m if op_a)."""
    dt = np.float64 if A.dtype == np.float64 else np.float32
    A = _c(A, dt)
    B = _c(B, dt)
    m = A.shape[1] if op_a else A.shape[0]
    k = A.shape[0] if op_a else A.shape[1]
    n = B.shape[0] if op_b else B.shape[1]
    C = np.zeros((m, n), dt) if C is None else _c(C, dt).copy()
    f = lib().ora_dgemm if dt == np.float64 else lib().ora_sgemm
    f(int(op_a), int(op_b), m, n, k, alpha, _p(A), A.shape[1], _p(B), B.shape[1], beta, _p(C), n)
    return C


def permute(src, perm):
    """TA 'TO' convention: result.shape[perm[d]] = src.shape[d]."""
    src = np.ascontiguousarray(src)
    perm = _c(perm, np.int32)
    ext = np.array(src.shape, np.int64)
    rext = np.empty_like(ext)
    rext[perm] = ext
    dst = np.empty(tuple(rext), src.dtype)
    f = lib().ora_permute_f64 if src.dtype == np.float64 else lib().ora_permute_f32
    f(_p(src), _p(dst), src.ndim, _p(ext), _p(perm))
    return dst


def contract(op_a, op_b, mext, next_, kext, A, a_off, B, b_off, c_off, c_size, alpha=1.0,
             a_norms=None, b_norms=None, pair_filter_threshold=0.0):
    """One-rank Summa over fused block-sparse matrices; returns (C pool, npairs)."""
    dt = np.float64 if A.dtype == np.float64 else np.float32
    Mt, Nt, Kt = len(mext), len(next_), len(kext)
    C = np.zeros(int(c_size), dt)
    args = [int(op_a), int(op_b), Mt, Nt, Kt, _p(_c(mext, np.int32)), _p(_c(next_, np.int32)),
            _p(_c(kext, np.int32))]
    mext_, next__, kext_ = _c(mext, np.int32), _c(next_, np.int32), _c(kext, np.int32)
    A = _c(A, dt)
    B = _c(B, dt)
    a_off = _c(a_off, np.int64)
    b_off = _c(b_off, np.int64)
    c_off = _c(c_off, np.int64)
    an = _c(a_norms, np.float32) if a_norms is not None else None
    bn = _c(b_norms, np.float32) if b_norms is not None else None
    f = lib().ora_contract_f64 if dt == np.float64 else lib().ora_contract_f32
    npairs = f(int(op_a), int(op_b), Mt, Nt, Kt, _p(mext_), _p(next__), _p(kext_), _p(A),
               _p(a_off), _p(B), _p(b_off), _p(C), _p(c_off), alpha, _p(an), _p(bn),
               np.float32(pair_filter_threshold))
    del args
    return C, int(npairs)


def arena_slab_cells(row_bounds, col_bounds):
    """Cell ordinals (r * nc + c) in arena slab order: outer tiles row-major,
    cells row-major inside each tile; and the prefix of cells per tile."""
    rb, cb = np.asarray(row_bounds, np.int64), np.asarray(col_bounds, np.int64)
    nc = int(cb[-1] - cb[0])
    ids, pre = [], [0]
    for tr in range(rb.size - 1):
        for tc in range(cb.size - 1):
            t = (np.arange(rb[tr], rb[tr + 1])[:, None] * nc +
                 np.arange(cb[tc], cb[tc + 1])[None, :]).reshape(-1)
            ids.append(t)
            pre.append(pre[-1] + t.size)
    return (np.concatenate(ids) if ids else np.zeros(0, np.int64)), np.asarray(pre, np.int64)


def arena_contract(i_bounds, k_bounds, j_bounds, ea, ec, eb, A, a_cell_off, B, b_cell_off,
                   c_cell_off, c_size, alpha=1.0, threshold=float(np.finfo(np.float32).eps)):
    """Contraction of two arena tensors of tensors, C(i,j) = alpha sum_k
    A(i,k) . B(k,j) over cells, cell (i,k) of A being ea[i] x ec[k] (ec = 1:
    the inner outer product) and (k,j) of B ec[k] x eb[j], as TiledArray
    runs it: outer-tile norms over the tiles' cells (tensor/tensor.h:3768-3776),
    scaled by inverse cells per tile and screened (sparse_shape.h:150-217),
    the outer-tile result estimate with cells per contracted tile as k sizes
    (sparse_shape.h:1589-1681), then per non-zero result tile and per
    non-zero operand tile pair the cell products in ascending k (the per-cell
    contraction of arena_einsum.h:1803, which the strided regimes of
    arena_einsum.h:1170-1800 compute in other groupings).  Result cells are
    written at c_cell_off (nI x nJ, < 0 skipped).  Returns (C pool,
    result-tile estimate, scaled A tile norms, scaled B tile norms)."""
    ib, kb, jb = (np.asarray(x, np.int64) for x in (i_bounds, k_bounds, j_bounds))
    ea, ec, eb = (np.asarray(x, np.int64) for x in (ea, ec, eb))
    nI, nK, nJ = ib.size - 1, kb.size - 1, jb.size - 1
    a_cell_off = np.asarray(a_cell_off, np.int64).reshape(ea.size, ec.size)
    b_cell_off = np.asarray(b_cell_off, np.int64).reshape(ec.size, eb.size)
    c_cell_off = np.asarray(c_cell_off, np.int64).reshape(ea.size, eb.size)
    a_vol = (ea[:, None] * ec[None, :])
    b_vol = (ec[:, None] * eb[None, :])
    ida, pa = arena_slab_cells(ib, kb)
    idb, pb = arena_slab_cells(kb, jb)
    na = tot_tile_norms(A, pa, a_cell_off.reshape(-1)[ida], a_vol.reshape(-1)[ida])
    nb = tot_tile_norms(B, pb, b_cell_off.reshape(-1)[idb], b_vol.reshape(-1)[idb])
    sa, _ = scale_tile_norms(na.reshape(nI, nK), [np.diff(ib), np.diff(kb)], threshold)
    sb, _ = scale_tile_norms(nb.reshape(nK, nJ), [np.diff(kb), np.diff(jb)], threshold)
    est, _ = shape_gemm(sa, sb, 0, 0, nI, nJ, nK, np.diff(kb).astype(np.float32), abs(alpha),
                        threshold)
    # tile decisions replicated onto the cells
    ti = np.repeat(np.arange(nI), np.diff(ib))
    tk = np.repeat(np.arange(nK), np.diff(kb))
    tj = np.repeat(np.arange(nJ), np.diff(jb))
    a_on = sa[ti[:, None], tk[None, :]] >= threshold
    b_on = sb[tk[:, None], tj[None, :]] >= threshold
    c_on = est[ti[:, None], tj[None, :]] >= threshold
    ao = np.where(a_on, a_cell_off, -1).reshape(-1)
    bo = np.where(b_on, b_cell_off, -1).reshape(-1)
    co = np.where(c_on, c_cell_off, -1).reshape(-1)
    C, _ = contract(0, 0, ea, eb, ec, A, ao, B, bo, co, c_size, alpha=alpha)
    return C, est, sa, sb
