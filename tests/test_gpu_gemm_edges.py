"""The grouped GEMM's operand staging at its edges, through the C-ABI on a
hand-built work list: ragged macro tiles, a pair's partial last k chunk,
padded leading dimensions, every transpose combination, leading dimensions at
the 32-bit staging offsets' limit (2^20 elements, staged with buffer loads)
and above it (staged with 64-bit addresses), FP64 and FP32, several macro
tiles.  Integer data make every sum exact, so results are compared with
torch bit for bit.  (The per-pair GEMM this replaces is the tile GEMM of
src/TiledArray/tile_op/contract_reduce.h, whose operands TiledArray passes
with their own leading dimensions.)"""
import ctypes

import pytest
import torch

from tiledarray_amd import _lib
from tiledarray_amd import engine as E

pytestmark = pytest.mark.gpu

LIMIT = 1 << 20  # largest leading dimension staged with 32-bit offsets

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hip


def run_pair(ctx, dtype, op_a, op_b, m, n, k, lda, ldb, tile):
    """C = op(A) op(B) for one task of one pair; returns (C, reference)."""
    a_rows = k if op_a else m
    b_rows = n if op_b else k
    assert lda >= (m if op_a else k) and ldb >= (k if op_b else n)
    g = torch.Generator(device="cuda").manual_seed(m * 7919 + n * 104729 + k + lda + ldb)
    A = torch.randint(-3, 4, (a_rows * lda,), dtype=torch.int32, device="cuda", generator=g).to(dtype)
    B = torch.randint(-3, 4, (b_rows * ldb,), dtype=torch.int32, device="cuda", generator=g).to(dtype)
    C = torch.full((m * n,), 7.0, dtype=dtype, device="cuda")
    wl = E.WorkList(ctx, 4, 4, 4096)
    tasks, pairs, counters = wl.buffers()
    task = _lib.GemmTask(0, m, n, n, 0, 1, 0)
    pr = _lib.GemmPair(0, 0, k, lda, ldb, 0, 0, 0)
    cnt = (ctypes.c_int64 * 8)(1, 1, m * n, 0, 0, 0, 0, 0)
    assert hip().hipMemcpy(tasks, ctypes.byref(task), ctypes.sizeof(task), 1) == 0
    assert hip().hipMemcpy(pairs, ctypes.byref(pr), ctypes.sizeof(pr), 1) == 0
    assert hip().hipMemcpy(counters, cnt, ctypes.sizeof(cnt), 1) == 0
    cfg = _lib.GemmConfig(*tile, 0)
    f64 = dtype == torch.float64
    E.call("tamx_grouped_gemm_f64_config" if f64 else "tamx_grouped_gemm_f32_config",
           ctx.ptr, wl.ptr, op_a, op_b, ctypes.c_double(1.0) if f64 else ctypes.c_float(1.0),
           A.data_ptr(), B.data_ptr(), C.data_ptr(), ctypes.byref(cfg), ctx.stream())
    torch.cuda.synchronize()
    Aop = A.view(a_rows, lda)[:, : (m if op_a else k)]
    Aop = Aop.t() if op_a else Aop
    Bop = B.view(b_rows, ldb)[:, : (k if op_b else n)]
    Bop = Bop.t() if op_b else Bop
    ref = (Aop.double().cpu() @ Bop.double().cpu()).to(dtype)
    return C.view(m, n).cpu(), ref


# (m, n, k, extra lda, extra ldb): ragged in every direction, partial k chunks
SHAPES = [(70, 33, 37, 0, 0), (70, 33, 37, 3, 5), (5, 130, 16, 1, 0), (129, 7, 65, 0, 2)]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("ops", [(0, 0), (1, 0), (0, 1), (1, 1)], ids=["NN", "TN", "NT", "TT"])
@pytest.mark.parametrize("tile", [(64, 64, 16), (128, 64, 16), (32, 32, 16), (64, 64, 32)],
                         ids=lambda t: "x".join(map(str, t)))
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "m%dn%dk%d+%d+%d" % s)
def test_ragged_padded_transposed(ctx, dtype, ops, tile, shape):
    m, n, k, xa, xb = shape
    op_a, op_b = ops
    lda = (m if op_a else k) + xa
    ldb = (k if op_b else n) + xb
    C, ref = run_pair(ctx, dtype, op_a, op_b, m, n, k, lda, ldb, tile)
    assert torch.equal(C, ref)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("case", ["at_limit", "above_limit"])
@pytest.mark.parametrize("ops", [(0, 0), (1, 1)], ids=["NN", "TT"])
def test_leading_dimension_limit(ctx, dtype, case, ops):
    # NN: A is m x k with lda = k; TT: A is k x m and B is n x k with padded
    # leading dimensions.  At the limit the staging uses buffer loads with
    # its largest offsets; above it, 64-bit addresses.
    op_a, op_b = ops
    ld = LIMIT if case == "at_limit" else LIMIT + 5
    if ops == (0, 0):
        m, n, k = 3, 5, ld
        lda, ldb = k, n
    else:
        m, n, k = 7, 6, 9
        lda, ldb = ld, ld
    C, ref = run_pair(ctx, dtype, op_a, op_b, m, n, k, lda, ldb, (64, 64, 16))
    assert torch.equal(C, ref)
