"""GPU tests of the runtime services: the tile-configuration autotuner and
graph capture through the C ABI (tamx_autotune_grouped_gemm,
tamx_autotune_grouped_gemm_split, tamx_graph_begin/_end/_launch).

Bars: the autotuner returns a compiled configuration and leaves the
caller's C bit-for-bit unchanged; a contraction run with the tuned
configuration is bit-exact against the oracle on integer-valued data; every
graph replay equals the eager result bit-for-bit.
"""
import ctypes

import numpy as np
import pytest
import torch

from tiledarray_amd import _lib
from tiledarray_amd import engine as E
from tiledarray_amd import workloads as W

from test_gpu_contract import c_off_mask, oracle_contract

pytestmark = pytest.mark.gpu


def blocksparse_integer(ctx, n=1500, seed=21):
    A, B, _ = W.blocksparse_gemm(n=n, wlo=30, whi=90, seed=seed, integer=True)
    E.make_sparse(ctx, A, 1e-8)
    E.make_sparse(ctx, B, 1e-8)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=1e-8)
    return A, B, c


def present(X):
    s = X.shape_norms.cpu().numpy()
    return (X.offsets >= 0) & (s >= X.shape_threshold)


def test_autotune_returns_compiled_config_and_keeps_c(ctx):
    A, B, c = blocksparse_integer(ctx)
    c.run()
    torch.cuda.synchronize()
    before = c.C.clone()
    tile, ms = c.tune(reps=2, force=True)
    torch.cuda.synchronize()
    compiled = [(g.bm, g.bn, g.bk) for g in _lib.gemm_configs()]
    assert tile[:3] in compiled
    assert ms > 0.0
    assert torch.equal(c.C, before)
    # the cached class returns the same configuration without timing
    tile2, ms2 = c.tune()
    assert tile2 == tile and ms2 == pytest.approx(ms)
    # a run with the tuned configuration is bit-exact against the oracle
    c.C.fill_(np.nan)
    c.gemm_step()
    torch.cuda.synchronize()
    Cref, npairs, c_off = oracle_contract(c, A, B, present(A), present(B))
    mask = c_off_mask(c_off, c)
    np.testing.assert_array_equal(c.C.cpu().numpy()[mask], Cref[mask])
    assert c.counters()["npairs"] == npairs


def test_autotune_accumulating_tasks_leave_c_unchanged(ctx):
    # beta = 1 tasks: the trial launches must not accumulate into the caller's C
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    m, n, k = 96, 80, 40
    A = torch.randint(-5, 6, (m * k,), dtype=torch.float64, device="cuda")
    B = torch.randint(-5, 6, (k * n,), dtype=torch.float64, device="cuda")
    C0 = torch.randint(-5, 6, (m * n,), dtype=torch.float64, device="cuda")
    C = C0.clone()
    wl = E.WorkList(ctx, 4, 4, 4)
    tasks, pairs, counters = wl.buffers()
    task = _lib.GemmTask(0, m, n, n, 0, 1, _lib.TAMX_TASK_BETA1)
    pr = _lib.GemmPair(0, 0, k, k, n, 0, 0, 0)
    cnt = (ctypes.c_int64 * 8)(1, 1, m * n, 0, 0, 0, 0, 0)
    assert hip.hipMemcpy(tasks, ctypes.byref(task), ctypes.sizeof(task), 1) == 0
    assert hip.hipMemcpy(pairs, ctypes.byref(pr), ctypes.sizeof(pr), 1) == 0
    assert hip.hipMemcpy(counters, cnt, ctypes.sizeof(cnt), 1) == 0
    tile, ms = E.autotune(ctx, wl, torch.float64, 0, 0, A, B, C, reps=1, force=True)
    torch.cuda.synchronize()
    assert torch.equal(C, C0)
    cfg = _lib.GemmConfig(*tile)
    E.call("tamx_grouped_gemm_f64_config", ctx.ptr, wl.ptr, 0, 0, ctypes.c_double(2.0),
           A.data_ptr(), B.data_ptr(), C.data_ptr(), ctypes.byref(cfg), ctx.stream())
    torch.cuda.synchronize()
    ref = C0.view(m, n) + 2.0 * (A.view(m, k) @ B.view(k, n))
    assert torch.equal(C.view(m, n), ref)


def test_autotune_empty_worklist_returns_default(ctx):
    wl = E.WorkList(ctx, 4, 4, 4)
    x = torch.zeros(16, dtype=torch.float32, device="cuda")
    tile, ms = E.autotune(ctx, wl, torch.float32, 0, 0, x, x, x, reps=1, force=True)
    g0 = _lib.gemm_configs()[0]
    assert tile[:3] == (g0.bm, g0.bn, g0.bk) and ms == 0.0


def test_autotune_split_picks_a_split_tile_with_identical_results(ctx):
    # the split GEMM's tuner times its macro tiles; every one gives the same
    # result bit for bit (same limbs, same pair and MFMA order)
    A, B, _ = blocksparse_integer(ctx, n=1500, seed=31)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=1e-8, precision="bf16x4")
    c.run()
    torch.cuda.synchronize()
    ref = c.C.clone()
    tile, ms = c.tune(reps=1, force=True)
    assert tile[:2] in ((64, 64), (128, 64), (64, 128)) and ms > 0.0
    c.C.zero_()
    c.gemm_step()
    torch.cuda.synchronize()
    mask = torch.as_tensor(c_off_mask(c.c_off.cpu().numpy(), c), device=c.C.device)
    assert torch.equal(c.C[:mask.numel()][mask], ref[:mask.numel()][mask])


def test_autotune_rejects_split_precision(ctx):
    wl = E.WorkList(ctx, 4, 4, 4)
    cfg = _lib.GemmConfig()
    ms = ctypes.c_float()
    rc = _lib.lib().tamx_autotune_grouped_gemm(ctx.ptr, wl.ptr, _lib.PREC_BF16XS, 0, 0, None,
                                               None, None, 1, 0, ctypes.byref(cfg),
                                               ctypes.byref(ms), ctx.stream())
    assert rc == _lib.TAMX_ERR_BAD_ARG


def test_graph_begin_rejects_null_stream(ctx):
    assert _lib.lib().tamx_graph_begin(ctx.ptr, None) == _lib.TAMX_ERR_BAD_ARG


def test_native_graph_replays_match_eager(ctx):
    # the whole single-contraction step (norms, shapes, estimate, work list,
    # GEMM) captured once and replayed with new operand data in between
    A, B, c = blocksparse_integer(ctx, n=1200, seed=5)
    sa = E._sizes_vector(A.trange, A.pool.device)
    sb = E._sizes_vector(B.trange, B.pool.device)

    def step():
        for X, sz in ((A, sa), (B, sb)):
            E.tile_norms(ctx, X, out=X.shape_norms)
            E.scale_norms(ctx, X.shape_norms, X.trange, 1e-8, sizes=sz)
        c.shape_step()
        c.worklist_step()
        c.gemm_step()

    step()
    torch.cuda.synchronize()
    graph = E.Graph(ctx, step)
    for rep in range(3):
        A.pool.mul_(-1.0 if rep % 2 == 0 else 2.0)
        B.pool.add_(1.0)
        torch.cuda.synchronize()
        step()
        torch.cuda.synchronize()
        eager, eager_off = c.C.clone(), c.c_off.clone()
        c.C.zero_()
        c.c_off.fill_(-7)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(c.c_off, eager_off), rep
        mask = torch.as_tensor(c_off_mask(eager_off.cpu().numpy(), c), device=c.C.device)
        assert torch.equal(c.C[:mask.numel()][mask], eager[:mask.numel()][mask]), rep
