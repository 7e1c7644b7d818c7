"""Benchmark of the contraction hot path on one MI355X.

Default workload: the block-sparse FP64 matrix contraction C("m,n") =
A("m,k") * B("k,n") with M = N = K = 16384, tile widths drawn from [64,256],
tile (i,j) kept with probability exp(-|i-j|/8), entries U[-1,1] *
10^(-|i-j|/4) and SparseShape threshold 1e-8.

One step = one pass of the hot path over inputs already resident in HBM:
tile norms of A and B, inverse-volume scaling and screening, the result shape
estimate, the device work list of surviving tile pairs, and the persistent
grouped FP64 MFMA GEMM.  The metric is nonzero-tile TFLOP/s: the sum over
surviving (i,k,j) of 2*m_i*k_k*n_j, divided by the step time.

Prints one JSON line (rank 0, N = 1).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "nonzero-tile TFLOP/s, FP64 vs split-precision MFMA at stated rel. error"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="blocksparse", choices=["blocksparse", "dense"])
    p.add_argument("--n", type=int, default=None, help="matrix extent (default per workload)")
    p.add_argument("--tile", default=None, help="macro tile BMxBN[xBK[xWGS]] for the grouped GEMM")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--precision", default="f64", help="f64 or bf16xS (S BF16 limbs)")
    return p.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from tiledarray_amd import engine as E
    from tiledarray_amd import workloads as W

    ctx = E.Context(0)
    thr = 1e-8
    if args.workload == "blocksparse":
        n = args.n or 16384
        A, B, info = W.blocksparse_gemm(n=n, wlo=64, whi=256)
        workload = (f"block-sparse FP64 C(m,n)=A(m,k)*B(k,n), M=N=K={n}, tile widths U[64,256], "
                    f"fill exp(-|i-j|/8), entries U[-1,1]*10^(-|i-j|/4), threshold 1e-8")
        sparse = True
    else:
        n = args.n or 2048
        A, B, info = W.dense_gemm(n=n, tile=128)
        workload = f"dense FP64 C(m,n)=A(m,k)*B(k,n), M=N=K={n}, 128-wide tiles, U[-1,1]"
        sparse = False
    tile = tuple(int(x) for x in args.tile.split("x")) if args.tile else None

    sizes = {}
    if sparse:
        E.make_sparse(ctx, A, thr)
        E.make_sparse(ctx, B, thr)
        for X in (A, B):
            sizes[id(X)] = E._sizes_vector(X.trange, X.pool.device)
    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=thr, tile=tile,
                      precision=args.precision)

    def screening():
        for X in (A, B):
            E.tile_norms(ctx, X, out=X.shape_norms)
            E.scale_norms(ctx, X.shape_norms, X.trange, thr, sizes=sizes[id(X)])

    gemm_events = []

    def step(timed=False):
        if sparse:
            screening()
        if c.nslices:
            c.split_step()
        c.shape_step()
        c.worklist_step()
        if timed:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            c.gemm_step()
            e1.record()
            gemm_events.append((e0, e1))
        else:
            c.gemm_step()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    cnt = c.counters()
    flops = cnt["flops"]
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(timed=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    gemm_ms = sum(a.elapsed_time(b) for a, b in gemm_events) / len(gemm_events)
    value = flops / (ms * 1e-3) / 1e12
    achieved = flops / (gemm_ms * 1e-3) / 1e12

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "tiles_per_mode": A.trange.tiles_shape[0],
            "result_tiles": cnt["ntasks"],
            "tile_pairs": cnt["npairs"],
            "flops_per_step": flops,
            "macro_tile": list(tile) if tile else [64, 64, 16],
            "seed": info["seed"],
        },
        "roofline": {
            "bound": "mfma",
            "kernel": "grouped_gemm_kernel<double>",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
            "traffic": None,
            "avg_launch_ms": round(gemm_ms, 4),
        },
    }

    if not args.no_cpu:
        out["cpu_baseline"], out["rel_error"] = cpu_leg(args, c, A, B, thr, sparse, flops, ms)
    print(json.dumps(out), flush=True)


def cpu_leg(args, c, A, B, thr, sparse, gpu_flops, gpu_ms):
    """The oracle (the CPU restatement of TiledArray's one-rank path, MKL
    serial inside one OpenMP task per result tile) timed on the host over the
    same inputs: norms, scaling, shape estimate and every tile GEMM."""
    import numpy as np
    import torch

    import oracle

    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    oracle.set_num_threads(threads)
    Ah, Bh = A.pool.cpu().numpy(), B.pool.cpu().numpy()
    vols = A.trange.tile_volumes()
    ext = [r.extents for r in A.trange.ranges]
    nt = A.trange.tiles_shape[0]
    c_off_gpu = c.c_off.cpu().numpy()
    gpu_C = c.C.cpu().numpy()
    times = []
    for rep in range(2):
        t0 = time.perf_counter()
        if sparse:
            na = oracle.tile_norms(Ah, A.offsets, vols).reshape(nt, nt)
            nb = oracle.tile_norms(Bh, B.offsets, vols).reshape(nt, nt)
            sa, _ = oracle.scale_tile_norms(na, ext, thr)
            sb, _ = oracle.scale_tile_norms(nb, ext, thr)
            est, _ = oracle.shape_gemm(sa, sb, 0, 0, nt, nt, nt, oracle.k_sizes([ext[0]]), 1.0, thr)
            a_off = np.where(sa.reshape(-1) >= thr, A.offsets, -1)
            b_off = np.where(sb.reshape(-1) >= thr, B.offsets, -1)
            present = est.reshape(-1) >= thr
        else:
            a_off, b_off = A.offsets, B.offsets
            present = np.ones(nt * nt, bool)
        mext = c.mext_h
        vol_c = np.multiply.outer(mext, c.next_h).reshape(-1)
        c_off = np.full(nt * nt, -1, np.int64)
        c_off[present] = np.concatenate([[0], np.cumsum(vol_c[present])])[:-1]
        ref, npairs = oracle.contract(0, 0, mext, c.next_h, c.kext_h, Ah, a_off, Bh, b_off, c_off,
                                      int(vol_c[present].sum()))
        times.append(time.perf_counter() - t0)
    t = min(times)
    pattern_equal = bool(np.array_equal(present, c_off_gpu >= 0))
    # relative Frobenius error of the GPU result over the common pattern
    num = den = 0.0
    for tt in np.nonzero(present & (c_off_gpu >= 0))[0]:
        v = vol_c[tt]
        g = gpu_C[c_off_gpu[tt]:c_off_gpu[tt] + v]
        r = ref[c_off[tt]:c_off[tt] + v]
        num += float(np.sum((g - r) ** 2))
        den += float(np.sum(r * r))
    rel = (num / den) ** 0.5 if den > 0 else 0.0
    cpu = {
        "value": round(gpu_flops / t / 1e12, 4),
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": "full workload (norms, scaling, shape estimate, all tile GEMMs), best of 2",
        "seconds": round(t, 3),
        "blas": oracle.blas_name(),
        "pattern_equal": pattern_equal,
        "tile_pairs": npairs,
    }
    return cpu, {"frobenius": rel}


if __name__ == "__main__":
    main()
