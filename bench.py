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
PEAK_BF16_TFLOPS = 2500.0
PEAK_INT8_TOPS = 5000.0
PEAK_FP32_TFLOPS = 157.3
# HBM bytes per launch of the roofline kernel from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes, each scaled by what the counter reads for a plain copy
# of 1 GiB (524302 FETCH_SIZE units read, 1048576 WRITE_SIZE units written:
# 2 KiB and 1 KiB per unit).  Keyed by workload, precision and macro tile,
# with the profile summary the counters come from; a line whose
# configuration has no entry reports traffic null.
TRAFFIC_BYTES = {
    ("blocksparse", "f64", (64, 64, 16)): (4668000 * 2048 + 1006000 * 1024,
                                           "profiles/grouped_gemm_f64_blocksparse16384.md"),
    ("blocksparse", "i8x5", (64, 64, 16)): (2202000 * 2048 + 905900 * 1024,
                                            "profiles/split_gemm_i8x5_blocksparse16384.md"),
}


def traffic(workload, precision, tile):
    """(bytes, source) of the profiled HBM traffic of this configuration, or
    (None, None)."""
    hit = TRAFFIC_BYTES.get((workload, precision, tuple(tile) if tile else None))
    return hit if hit else (None, None)


def host_cpu():
    """Threads this process may run on, the machine's CPU count and model."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads_usable": usable, "nproc": os.cpu_count(), "model": model}


def cpu_threads(args):
    return args.cpu_threads or host_cpu()["threads_usable"]


def median_time(fn, reps=3, warmup=1):
    """Median wall time (s) of `reps` calls of fn after `warmup` calls; also
    the last call's return value."""
    import statistics
    out = None
    for _ in range(warmup):
        out = fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), out


def limb_products(precision):
    """MFMA products per FP64-equivalent product and the pipe's published
    peak for a split precision."""
    if precision == "f32x2":
        return 3, PEAK_FP32_TFLOPS, "FP32 MFMA"
    s = int(precision.split("x")[-1])
    if precision.startswith("i8"):
        return s * (s + 1) // 2, PEAK_INT8_TOPS, "INT8 MFMA"
    return s * (s + 1) // 2, PEAK_BF16_TFLOPS, "BF16 MFMA"


def split_roofline(precision, flops, gemm_ms):
    """The split GEMM against its own pipe: limb products (FP64-equivalent
    flops times products per FP64 product) per second over the published
    peak, and the FP64-equivalent rate against the FP64 peak."""
    prods, peak, pipe = limb_products(precision)
    limb = prods * flops / (gemm_ms * 1e-3) / 1e12
    eq = flops / (gemm_ms * 1e-3) / 1e12
    return {"bound": "mfma", "pipe": pipe, "products_per_fp64_product": prods,
            "achieved": round(limb, 3), "peak": peak, "unit": "TFLOP/s" if "FP" in pipe or
            "BF16" in pipe else "TOP/s", "frac": round(limb / peak, 4),
            "fp64_equivalent_frac": round(eq / PEAK_FP64_TFLOPS, 4)}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    # the GEMMs settle over their first ten or so launches (3.9 -> 3.25 ms
    # for the block-sparse FP64 GEMM in a per-launch trace), so the default
    # warm-up covers them
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--workload", default="blocksparse",
                   choices=["blocksparse", "dense", "ladder", "tot"])
    p.add_argument("--o", type=int, default=32, help="ladder: occupied extent")
    p.add_argument("--v", type=int, default=256, help="ladder: virtual extent")
    p.add_argument("--nbatch", type=int, default=1024, help="tot: independent products")
    p.add_argument("--dtype", default="f32", help="tot: f32 or f64")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only copy 1 GiB three times (scales FETCH_SIZE/WRITE_SIZE to bytes)")
    p.add_argument("--n", type=int, default=None, help="matrix extent (default per workload)")
    p.add_argument("--tile", default="auto",
                   help="macro tile BMxBN[xBK[xWGS]] of the grouped GEMM, or 'auto' (autotuner)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--precision", default="f64",
                   help="f64, f32x2, bf16xS or i8xS (S BF16 or INT8 limbs)")
    p.add_argument("--no-split", action="store_true", help="skip the split-precision runs")
    p.add_argument("--split", default="bf16x3,bf16x4,bf16x5,i8x4,i8x5,f32x2",
                   help="split precisions measured beside the FP64 run")
    p.add_argument("--ladder-split", default="bf16x3,f32x2,i8x4,i8x5",
                   help="ladder: split precisions measured beside the main precision")
    p.add_argument("--split-tile", default=None,
                   help="macro tile BMxBN of the split GEMM (64x64, 128x64 or 64x128)")
    p.add_argument("--v-permute", default="auto", choices=["auto", "copy", "map"],
                   help="ladder: permute V into a copy each step, or read it through the "
                        "permutation's index map (auto: the map for split precisions, the "
                        "copy for FP64)")
    return p.parse_args()


def fixed_tile(args):
    """The --tile argument as a tuple, or None for 'auto'."""
    if args.tile in (None, "auto"):
        return None
    return tuple(int(x) for x in args.tile.split("x"))


def autotune(args, c, step_fn, reps=3):
    """With --tile auto: one untimed step builds the work list, then the
    autotuner times every compiled tile shape on it and the contraction
    keeps the fastest.  Returns the tile to report."""
    import torch
    if args.tile != "auto":
        return list(fixed_tile(args)) if args.tile else None
    step_fn(None)
    torch.cuda.synchronize()
    tile, _ = c.tune(reps=reps)
    return list(tile[:3])


def time_steps(c, step_fn, steps, warmup):
    """Runs warmup + timed steps; returns (ms per step, mean GEMM launch ms)."""
    import torch
    events = []
    for _ in range(warmup):
        step_fn(None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step_fn(events)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gemm_ms = sum(a.elapsed_time(b) for a, b in events) / max(1, len(events))
    return (t1 - t0) * 1e3 / steps, gemm_ms


def time_kernel(fn, reps=5):
    """Mean duration (ms) of fn's launches on the current stream, by events."""
    import torch
    fn()
    ev = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        ev.append((e0, e1))
    torch.cuda.synchronize()
    return sum(a.elapsed_time(b) for a, b in ev) / reps


def bandwidth(name_bytes_ms):
    """Roofline entries of the bandwidth-bound kernels: {name: (bytes, ms)}."""
    out = {}
    for name, (nbytes, ms) in name_bytes_ms.items():
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"bound": "hbm", "bytes": int(nbytes), "avg_launch_ms": round(ms, 4),
                     "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(gbs / PEAK_HBM_GBS, 4)}
    return out


def stored_bytes(X):
    """Bytes of the stored (non-zero) tiles of a tiled array."""
    import numpy as np
    vols = X.trange.tile_volumes()
    return int(np.sum(np.where(X.offsets >= 0, vols, 0))) * X.pool.element_size()


def calibrate_copy():
    import torch
    n = 1 << 27  # 2^27 doubles = 1 GiB
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copied_bytes_per_launch": n * 8, "launches": 3}), flush=True)


def main():
    args = parse()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ladder":
        return bench_ladder(args)
    if args.workload == "tot":
        return bench_tot(args)
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
    tile = fixed_tile(args)

    sizes = {}
    if sparse:
        E.make_sparse(ctx, A, thr)
        E.make_sparse(ctx, B, thr)
        for X in (A, B):
            sizes[id(X)] = E._sizes_vector(X.trange, X.pool.device)

    # per-tile max |a| from the norm pass, for the limb split's scales
    maxabs = {id(X): torch.zeros(X.trange.ntiles, dtype=torch.float64, device=X.pool.device)
              for X in (A, B)}

    def screen(X, with_max):
        E.tile_norms(ctx, X, out=X.shape_norms, maxabs=maxabs[id(X)] if with_max else None)
        E.scale_norms(ctx, X.shape_norms, X.trange, thr, sizes=sizes[id(X)])

    def make_step(c):
        def step(events):
            if sparse and c.nslices:
                # each operand's limb split starts on a side stream once its
                # norm pass (which gives the tile maxima) is enqueued: A's
                # split runs beside B's norms, both beside the shape
                # estimate and the work list
                for w, X in enumerate((A, B)):
                    screen(X, True)
                    c.split_operand_async(w, maxabs[id(X)])
                c.plan_and_join()
            elif c.nslices:
                c.split_and_plan()
            else:
                if sparse:
                    for X in (A, B):
                        screen(X, False)
                c.shape_step()
                c.worklist_step()
            if events is not None:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                c.gemm_step()
                e1.record()
                events.append((e0, e1))
            else:
                c.gemm_step()
        return step

    c = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=thr, tile=tile,
                      precision=args.precision)
    used_tile = (autotune(args, c, make_step(c))
                 if args.precision in ("f64", "f32") or args.precision.startswith(("bf16", "i8"))
                 else None)
    ms, gemm_ms = time_steps(c, make_step(c), args.steps, args.warmup)
    cnt = c.counters()
    flops = cnt["flops"]
    value = flops / (ms * 1e-3) / 1e12
    achieved = flops / (gemm_ms * 1e-3) / 1e12
    peak = (PEAK_FP64_TFLOPS if args.precision == "f64"
            else PEAK_INT8_TOPS if args.precision.startswith("i8") else PEAK_BF16_TFLOPS)
    tr_bytes, tr_src = traffic(args.workload, args.precision, used_tile or [64, 64, 16])

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
        "dtype": "f64" if args.precision == "f64" else args.precision,
        "data": "synthetic",
        "config": {
            "workload": workload,
            "tiles_per_mode": A.trange.tiles_shape[0],
            "result_tiles": cnt["ntasks"],
            "tile_pairs": cnt["npairs"],
            "flops_per_step": flops,
            "macro_tile": used_tile if used_tile else ([64, 64, 32] if args.precision != "f64"
                                                       else [64, 64, 16]),
            "tile_choice": "autotuner" if args.tile == "auto" else "fixed",
            "seed": info["seed"],
            "generator": info["generator"],
        },
        "roofline": ({
            "bound": "mfma",
            "kernel": f"grouped_gemm_kernel<double,{','.join(map(str, used_tile or [64, 64, 16]))}>",
            "achieved": round(achieved, 3),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": tr_bytes,
            "traffic_source": tr_src,
            "avg_launch_ms": round(gemm_ms, 4),
        } if args.precision == "f64" else dict(
            split_roofline(args.precision, flops, gemm_ms), kernel="split GEMM",
            traffic=tr_bytes, traffic_source=tr_src, avg_launch_ms=round(gemm_ms, 4))),
    }
    if sparse:
        out["config"]["near_threshold_estimates"] = near_threshold(ctx, c)

    ref = None
    if not args.no_cpu:
        out["cpu_baseline"], out["rel_error"], ref = cpu_leg(args, c, A, B, thr, sparse, flops)

    if args.precision == "f64" and not args.no_split:
        # split precision on the same inputs: every limb count's step time and
        # its error against the FP64 oracle (or the FP64 result without the
        # CPU leg)
        base = ref if ref is not None else c.C.cpu().numpy()
        used = used_mask(c)
        r = base[used]
        rn, rmax = float(np.linalg.norm(r)), float(np.abs(r).max())
        split = {}
        for prec in args.split.split(","):
            limbs = int(prec.split("x")[-1])
            wide_ok = (prec.startswith("bf16") and limbs <= 4) or (prec.startswith("i8")
                                                                  and 3 <= limbs <= 4)
            st = (tuple(int(x) for x in args.split_tile.split("x"))
                  if args.split_tile and wide_ok else None)
            cs = E.Contraction(ctx, "m,n", A, "m,k", B, "k,n", threshold=thr, precision=prec,
                               tile=st)
            if st is None and args.tile == "auto" and prec.startswith(("bf16", "i8")):
                make_step(cs)(None)
                torch.cuda.synchronize()
                st = tuple(cs.tune(reps=2)[0][:3])
            ms_s, gemm_s = time_steps(cs, make_step(cs), args.steps, args.warmup)
            g = cs.C.cpu().numpy()[used]
            split[prec] = {
                "value": round(flops / (ms_s * 1e-3) / 1e12, 3),
                "ms_per_step": round(ms_s, 4),
                "gemm_tflops": round(flops / (gemm_s * 1e-3) / 1e12, 3),
                "gemm_ms": round(gemm_s, 4),
                "speedup_vs_f64": round(ms / ms_s, 3),
                "macro_tile": list(st) if st else [64, 64],
                "rel_err_frobenius": float(np.linalg.norm(g - r) / rn),
                "rel_err_max": float(np.abs(g - r).max() / rmax),
                "error_against": "oracle" if ref is not None else "fp64 result",
                "roofline": split_roofline(prec, flops, gemm_s),
            }
            if cs.nslices:
                # the limb split: every stored FP64 element read once, S
                # limbs of the padded limb layout written
                mx = (maxabs[id(A)], maxabs[id(B)]) if sparse else None
                nbytes = split_bytes(cs, A, B)
                split[prec]["bandwidth"] = bandwidth(
                    {"split_limbs": (nbytes, time_kernel(lambda: cs.split_step(mx)))})
            del cs
        out["split_precision"] = split
    if sparse:
        # screening: every stored element read once per norm pass
        nb = stored_bytes(A) + stored_bytes(B)
        ms_n = time_kernel(lambda: [E.tile_norms(ctx, X, out=X.shape_norms) for X in (A, B)])
        out["bandwidth"] = bandwidth({"tile_norms": (nb, ms_n)})
    print(json.dumps(out), flush=True)


def base_line(args, value, ms, dtype, workload, config, roofline):
    return {"metric": METRIC, "value": round(value, 3), "unit": "TFLOP/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic", "config": dict(workload=workload, **config), "roofline": roofline}


def bench_ladder(args):
    """R(a,b,i,j) = T(c,d,i,j) * V(a,c,b,d) with V stored to force a tile
    permutation; one step = tile norms and shapes of T and V, (the limb
    split), the result estimate, the work list and the grouped GEMM with the
    result transpose fused into its store.  V is permuted into a copy in the
    step (as the reference permutes on read) or read through the
    permutation's index map by the limb split or the GEMM (--v-permute).

    The line's precision (--precision, FP64 by default) is the main entry;
    the split precisions of --split run on the same inputs beside it, each
    with its speed-up and its max and Frobenius error against the oracle on
    a sample of whole result tile columns."""
    import numpy as np
    import torch

    from tiledarray_amd import engine as E
    from tiledarray_amd import workloads as W

    ctx = E.Context(0)
    thr = 1e-10
    T, V, info = W.ladder(o=args.o, v=args.v, otiles=2, wlo=48, whi=96)
    E.make_sparse(ctx, T, thr)
    E.make_sparse(ctx, V, thr)
    sizes = {id(X): E._sizes_vector(X.trange, X.pool.device) for X in (T, V)}
    maxabs = {id(X): torch.zeros(X.trange.ntiles, dtype=torch.float64, device=X.pool.device)
              for X in (T, V)}
    perm_max = torch.zeros(V.trange.ntiles, dtype=torch.float64, device=V.pool.device)

    def measure(precision):
        c = E.Contraction(ctx, "a,b,i,j", T, "c,d,i,j", V, "a,c,b,d", threshold=thr,
                          precision=precision, tile=fixed_tile(args),
                          fuse_permute={"auto": None, "copy": False,
                                        "map": True}[args.v_permute])
        # the permuted V's shape follows the source's; refresh it each step
        perm_norms = c.B.shape_norms
        inv = [int(i) for i in np.argsort(list(c.plan.right_perm[:4]))]

        def step(events):
            for X in (T, V):
                E.tile_norms(ctx, X, out=X.shape_norms,
                             maxabs=maxabs[id(X)] if c.nslices else None)
                E.scale_norms(ctx, X.shape_norms, X.trange, thr, sizes=sizes[id(X)])
            c.permute_step()
            perm_norms.copy_(V.shape_norms.view(V.trange.tiles_shape).permute(*inv).reshape(-1))
            if c.nslices:
                perm_max.copy_(maxabs[id(V)].view(V.trange.tiles_shape).permute(*inv)
                               .reshape(-1))
                c.split_and_plan(maxabs=(maxabs[id(T)], perm_max))
            else:
                c.shape_step()
                c.worklist_step()
            if events is not None:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                c.gemm_step()
                e1.record()
                events.append((e0, e1))
            else:
                c.gemm_step()

        tunable = precision in ("f64", "f32") or precision.startswith(("bf16", "i8"))
        used_tile = autotune(args, c, step, reps=1) if tunable else None
        ms, gemm_ms = time_steps(c, step, args.steps, args.warmup)
        # bandwidth-bound kernels: the V permute (each stored element read
        # and written once), the norm passes over T and V, the limb split
        # (a split precision reads V through an index map while splitting
        # it: no permute pass, no permuted copy)
        bw = {}
        if hasattr(c.B, "_perm_plan"):
            bw["permute"] = (2 * stored_bytes(V), time_kernel(c.permute_step, reps=3))
        if c.nslices:
            bw["split_limbs"] = (split_bytes(c, T, V),
                                 time_kernel(lambda: c.split_step(
                                     maxabs=(maxabs[id(T)], perm_max)), reps=3))
        return c, ms, gemm_ms, used_tile, bw

    main_prec = args.precision
    c, ms, gemm_ms, used_tile, bw = measure(main_prec)
    cnt = c.counters()
    flops = cnt["flops"]
    peak = (PEAK_FP64_TFLOPS if main_prec == "f64"
            else PEAK_INT8_TOPS if main_prec.startswith("i8") else PEAK_BF16_TFLOPS)
    achieved = flops / (gemm_ms * 1e-3) / 1e12
    out = base_line(args, flops / (ms * 1e-3) / 1e12, ms,
                    "f64" if main_prec == "f64" else main_prec,
                    f"ladder R(a,b,i,j)=T(c,d,i,j)*V(a,c,b,d), o={args.o}, v={args.v}, "
                    f"v tiles U[48,96], T~N(0,1e-2), V~N(0,1)exp(-|a-c|/64), threshold 1e-10",
                    {"v_tiles": info["v_tiles"], "result_tiles": cnt["ntasks"],
                     "macro_tile": used_tile or ([64, 64, 16] if main_prec == "f64"
                                                 else [64, 64, 32]),
                     "tile_choice": "autotuner" if args.tile == "auto" else "fixed",
                     "tile_pairs": cnt["npairs"], "flops_per_step": flops,
                     "precision": main_prec, "seed": info["seed"],
                     "generator": info["generator"],
                     "v_permute": "copy" if hasattr(c.B, "_perm_plan") else "index map"},
                    ({"bound": "mfma", "achieved": round(achieved, 3),
                      "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                      "traffic": None, "avg_launch_ms": round(gemm_ms, 4)} if main_prec == "f64"
                     else dict(split_roofline(main_prec, flops, gemm_ms), traffic=None,
                               avg_launch_ms=round(gemm_ms, 4))))
    if main_prec == "f64" and not hasattr(c.B, "_perm_plan"):
        # the permute kernel on V outside the step (into a scratch copy)
        Y = E.permute_array(ctx, V, list(c.plan.right_perm[:4]))
        bw["permute_standalone"] = (2 * stored_bytes(V),
                                    time_kernel(lambda: Y._perm_plan.run(V.pool, Y.pool), reps=3))
        del Y
    bw["tile_norms"] = (stored_bytes(T) + stored_bytes(V),
                        time_kernel(lambda: [E.tile_norms(ctx, X, out=X.shape_norms)
                                             for X in (T, V)], reps=3))
    out["bandwidth"] = bandwidth(bw)
    sample = None
    if not args.no_cpu:
        sample = LadderSample(args, c, T, V)
        out["cpu_baseline"] = sample.cpu_baseline()
        out["rel_error"] = sample.errors(c)
    if not args.no_split and args.ladder_split:
        # the split precisions need the device memory the main run holds
        # (at o=48, v=320 a permuted V alone is 84 GB)
        del c
        torch.cuda.empty_cache()
        split = {}
        for prec in args.ladder_split.split(","):
            cs, ms_s, gemm_s, tile_s, bw_s = measure(prec)
            entry = {"value": round(flops / (ms_s * 1e-3) / 1e12, 3),
                     "ms_per_step": round(ms_s, 4),
                     "gemm_tflops": round(flops / (gemm_s * 1e-3) / 1e12, 3),
                     "gemm_ms": round(gemm_s, 4),
                     "speedup_vs_main": round(ms / ms_s, 3),
                     "macro_tile": tile_s,
                     "v_permute": "copy" if hasattr(cs.B, "_perm_plan") else "index map",
                     "roofline": split_roofline(prec, flops, gemm_s),
                     "bandwidth": bandwidth(bw_s)}
            assert cs.counters()["npairs"] == cnt["npairs"]
            if sample is not None:
                e = sample.errors(cs)
                entry["rel_err_frobenius"] = e["frobenius_sample"]
                entry["rel_err_max"] = e["max_sample"]
                entry["error_against"] = "oracle, " + e["sample"]
            split[prec] = entry
            del cs
            torch.cuda.empty_cache()
        out["split_precision"] = split
    print(json.dumps(out), flush=True)


class LadderSample:
    """The oracle over a sample of whole result tile columns R(a,b,.,.) of
    the ladder (the (a,b) tiles of smallest area, as many columns as keep the
    host's threads busy, one task per result tile): per tile permute of the V
    tiles the sample needs (as TA permutes on read), then the tile GEMMs.
    Gives the CPU rate (median of 3 after 1 warm-up) and the GPU results'
    errors on the sample."""

    def __init__(self, args, c, T, V):
        import numpy as np

        import oracle
        self.threads = cpu_threads(args)
        oracle.set_num_threads(self.threads)
        self.Th = T.pool.cpu().numpy()
        self.rp = rp = [int(x) for x in c.plan.right_perm[:4]]
        self.Mt, self.Nt, self.Kt = Mt, Nt, Kt = c.Mt, c.Nt, c.Kt
        # the fused structure only (the contraction itself is not kept, so
        # its device memory can be released)
        self.ops = (c.op_a, c.op_b)
        self.ext = (c.mext_h, c.next_h, c.kext_h)
        ncols = max(1, min(Nt, self.threads // max(1, Mt)))
        self.cols = cols = sorted(int(x) for x in np.argsort(c.next_h, kind="stable")[:ncols])
        c_off_gpu = c.c_off.cpu().numpy()
        vext = V.trange.tile_extents_table()
        b_pool, self.b_off, pos = [], np.full(Kt * Nt, -1, np.int64), 0
        for j in cols:
            for k in range(Kt):
                bt = k * Nt + j  # ordinal in the permuted grid (c,d,a,b)
                idx_p = np.unravel_index(bt, c.B.trange.tiles_shape)
                st = int(np.ravel_multi_index([idx_p[rp[d]] for d in range(4)],
                                              V.trange.tiles_shape))
                off = V.offsets[st]
                if off < 0:
                    continue
                e = [int(x) for x in vext[st]]
                # the stored (unpermuted) tile; the oracle permutes it in the
                # timed region, as TA permutes on read
                b_pool.append((bt, V.pool[off:off + int(np.prod(e))].cpu().numpy().reshape(e)))
        self.src_tiles = b_pool
        self.c_off = np.full(Mt * Nt, -1, np.int64)
        vol = 0
        for j in cols:
            for i in range(Mt):
                if c_off_gpu[i * Nt + j] >= 0:
                    self.c_off[i * Nt + j] = vol
                    vol += int(c.mext_h[i] * c.next_h[j])
        self.vol = max(1, vol)
        self.a_off = np.where(T.shape_norms.cpu().numpy() >= T.shape_threshold, T.offsets, -1)
        self.seconds, self.ref = median_time(self._run, reps=3, warmup=1)
        self.flops = 0
        for j in cols:
            kk = int(sum(c.kext_h[k] for k in range(Kt) if self.b_off[k * Nt + j] >= 0))
            for i in range(Mt):
                if self.c_off[i * Nt + j] >= 0:
                    self.flops += 2 * int(c.mext_h[i]) * int(c.next_h[j]) * kk

    def _run(self):
        import numpy as np

        import oracle
        parts, pos = [], 0
        for bt, tile in self.src_tiles:
            pt = oracle.permute(tile, self.rp).reshape(-1)
            self.b_off[bt] = pos
            parts.append(pt)
            pos += pt.size
        Bh = np.concatenate(parts) if parts else np.zeros(1)
        ref, _ = oracle.contract(*self.ops, *self.ext, self.Th, self.a_off, Bh, self.b_off,
                                 self.c_off, self.vol)
        return ref

    def cpu_baseline(self):
        import oracle
        return {"value": round(self.flops / self.seconds / 1e12, 4), "unit": "TFLOP/s",
                "cores": min(self.threads, len(self.cols) * self.Mt), "kind": "port",
                "blas": oracle.blas_name(), "seconds": round(self.seconds, 3),
                "host": host_cpu(),
                "sample": f"{len(self.cols)} of {self.Nt} result tile columns (per-tile V "
                          f"permute + tile GEMMs, {self.flops / 1e12:.2f} TFLOP), median of 3 "
                          f"after 1 warm-up"}

    def errors(self, c):
        """Relative Frobenius and max errors of contraction c's result over
        the sample."""
        import numpy as np
        off = c.c_off.cpu().numpy()
        num = den = mx = rmax = 0.0
        for j in self.cols:
            for i in range(self.Mt):
                o = self.c_off[i * self.Nt + j]
                if o < 0:
                    continue
                m, n = int(c.mext_h[i]), int(c.next_h[j])
                r = self.ref[o:o + m * n].reshape(m, n)
                og = int(off[i * self.Nt + j])
                g = c.C[og:og + m * n].cpu().numpy().reshape(n, m).T  # stored (a,b,i,j)
                num += float(np.sum((g - r) ** 2))
                den += float(np.sum(r * r))
                mx = max(mx, float(np.abs(g - r).max()))
                rmax = max(rmax, float(np.abs(r).max()))
        return {"frobenius_sample": (num / den) ** 0.5 if den else 0.0,
                "max_sample": mx / rmax if rmax else 0.0,
                "sample": f"{len(self.cols)} result tile columns"}


def bench_tot(args):
    """1024 independent tensor-of-tensor products in one persistent grouped
    launch; one step = cell norms, shapes, batched estimate, work list and the
    grouped GEMM, replayed from a hipGraph."""
    import numpy as np
    import torch

    import oracle
    from tiledarray_amd import engine as E
    from tiledarray_amd import workloads as W

    ctx = E.Context(0)
    dt = torch.float64 if args.dtype == "f64" else torch.float32
    w = W.tot_batch(nbatch=args.nbatch, dtype=dt)
    g = w["grid"]
    c = E.BatchedContraction(ctx, args.nbatch, g, g, g, w["a_dims"], w["b_dims"], w["c_dims"],
                             w["A"], w["a_off"], w["B"], w["b_off"],
                             tile=fixed_tile(args))
    used_tile = list(fixed_tile(args) or (64, 64, 16))[:3]
    if args.tile == "auto":
        c.run()
        torch.cuda.synchronize()
        used_tile = list(c.tune(reps=3)[0][:3])
    graph = c.capture()
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.replay()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    cnt = c.counters()
    flops = cnt["flops"]
    # grouped GEMM alone, eager, for the roofline
    ev = []
    for _ in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        c.gemm_step()
        e1.record()
        ev.append((e0, e1))
    torch.cuda.synchronize()
    gemm_ms = sum(a.elapsed_time(b) for a, b in ev) / len(ev)
    peak = PEAK_FP64_TFLOPS if dt == torch.float64 else PEAK_FP32_TFLOPS
    out = base_line(args, flops / (ms * 1e-3) / 1e12, ms, args.dtype,
                    f"{args.nbatch} tensor-of-tensor products C_n(i,j;a,b)=A_n(i,k;a,c)*B_n(k,j;c,b),"
                    f" 8x8 outer grids 50% filled, inner dims U[16,96], hipGraph replay",
                    {"result_cells": cnt["ntasks"], "cell_pairs": cnt["npairs"],
                     "flops_per_step": flops, "macro_tile": used_tile,
                     "tile_choice": "autotuner" if args.tile == "auto" else "fixed",
                     "seed": w["seed"], "generator": w["generator"]},
                    {"bound": "mfma", "achieved": round(flops / (gemm_ms * 1e-3) / 1e12, 3),
                     "peak": peak, "unit": "TFLOP/s",
                     "frac": round(flops / (gemm_ms * 1e-3) / 1e12 / peak, 4), "traffic": None,
                     "avg_launch_ms": round(gemm_ms, 4)})
    # screening: cell norms (every stored element read once), scaling and
    # the batched estimate
    el = w["A"].element_size()
    nb = (int(np.sum(np.where(w["a_off"] >= 0, c.a_plan.volumes, 0))) +
          int(np.sum(np.where(w["b_off"] >= 0, c.b_plan.volumes, 0)))) * el
    out["bandwidth"] = bandwidth({"screening": (nb, time_kernel(c.shape_step))})
    if not args.no_cpu:
        threads = cpu_threads(args)
        oracle.set_num_threads(threads)
        Ah, Bh = w["A"].cpu().numpy(), w["B"].cpu().numpy()
        c_off = c.c_off.cpu().numpy()
        gotC = c.C.cpu().numpy()

        def run():
            return [one(b) for b in range(args.nbatch)]

        def one(b):
            sl = slice(b * g * g, (b + 1) * g * g)
            co = c_off[sl]
            base = int(co[co >= 0].min()) if (co >= 0).any() else 0
            span = int(sum(w["a_dims"][b][t // g] * w["b_dims"][b][t % g]
                           for t in range(g * g) if co[t] >= 0))
            local = np.where(co >= 0, co - base, -1)
            part, _ = oracle.contract(0, 0, w["a_dims"][b], w["b_dims"][b], w["c_dims"][b], Ah,
                                      w["a_off"][sl], Bh, w["b_off"][sl], local, max(1, span))
            return base, part

        secs, parts = median_time(run, reps=3, warmup=1)
        num = den = 0.0
        for b in range(args.nbatch):
            base, part = parts[b]
            for t in range(g * g):
                o = c_off[b * g * g + t]
                if o >= 0:
                    n = int(w["a_dims"][b][t // g] * w["b_dims"][b][t % g])
                    r = part[o - base:o - base + n].astype(np.float64)
                    num += float(np.sum((gotC[o:o + n] - r) ** 2))
                    den += float(np.sum(r * r))
        out["cpu_baseline"] = {"value": round(flops / secs / 1e12, 4), "unit": "TFLOP/s",
                               "cores": threads, "kind": "port", "blas": oracle.blas_name(),
                               "seconds": round(secs, 3), "host": host_cpu(),
                               "sample": "all instances, tile GEMMs only (shapes not timed), "
                                         "median of 3 after 1 warm-up"}
        out["rel_error"] = {"frobenius": (num / den) ** 0.5 if den else 0.0}
    print(json.dumps(out), flush=True)


def split_bytes(c, A, B):
    """Algorithmic bytes of one limb split of both operands: the stored FP64
    elements read, the padded limb planes written."""
    rd = stored_bytes(A) + stored_bytes(B)
    wr = sum(int(sp.limbs.numel()) * sp.limbs.element_size() for sp in (c.a_split, c.b_split))
    return rd + wr


def near_threshold(ctx, c):
    """Result-shape estimates within 1e-5 relative of the threshold (the
    tiles whose presence would hinge on the estimate's last bits), from one
    unscreened estimate outside the timed steps."""
    import ctypes
    import torch
    from tiledarray_amd import engine as E
    est = torch.empty_like(c.c_norms)
    E.call("tamx_shape_gemm_f32", ctx.ptr, c.A.shape_norms.data_ptr(),
           c.B.shape_norms.data_ptr(), c.op_a, c.op_b, c.Mt, c.Nt, c.Kt, 1,
           c.k_sizes.data_ptr(), ctypes.c_float(abs(c.alpha)), ctypes.c_float(0.0),
           est.data_ptr(), None, ctx.stream())
    e = est.cpu().numpy().astype("float64")
    return int(((e - c.threshold) ** 2 <= (1e-5 * c.threshold) ** 2).sum())


def used_mask(c):
    import numpy as np
    m = np.zeros(c.c_capacity, bool)
    off = c.c_off.cpu().numpy()
    for t, o in enumerate(off):
        if o >= 0:
            m[o:o + c.mext_h[t // c.Nt] * c.next_h[t % c.Nt]] = True
    return m


def cpu_leg(args, c, A, B, thr, sparse, gpu_flops):
    """The oracle (the CPU restatement of TiledArray's one-rank path, MKL
    serial inside one OpenMP task per result tile) timed on the host over the
    same inputs: norms, scaling, shape estimate and every tile GEMM."""
    import numpy as np
    import torch

    import oracle

    threads = cpu_threads(args)
    oracle.set_num_threads(threads)
    Ah, Bh = A.pool.cpu().numpy(), B.pool.cpu().numpy()
    vols = A.trange.tile_volumes()
    ext = [r.extents for r in A.trange.ranges]
    nt = A.trange.tiles_shape[0]
    c_off_gpu = c.c_off.cpu().numpy()
    gpu_C = c.C.cpu().numpy()

    def run():
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
        return present, vol_c, c_off, ref, npairs

    t, (present, vol_c, c_off, ref, npairs) = median_time(run, reps=3, warmup=1)
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
        "sample": "full workload (norms, scaling, shape estimate, all tile GEMMs), "
                  "median of 3 after 1 warm-up",
        "seconds": round(t, 3),
        "blas": oracle.blas_name(),
        "host": host_cpu(),
        "pattern_equal": pattern_equal,
        "tile_pairs": npairs,
    }
    # the oracle result at the GPU result's offsets (identical packing when
    # the patterns agree)
    ref_at_gpu = None
    if pattern_equal:
        ref_at_gpu = np.zeros(c.c_capacity, ref.dtype)
        ref_at_gpu[:ref.size] = ref
    return cpu, {"frobenius": rel}, ref_at_gpu


if __name__ == "__main__":
    main()
