"""Builds libtamx.so (the HIP kernels and the C ABI of include/tamx.h) for
gfx950 with hipcc, in-tree, so the library travels with the source tree.

Every .hip and .cpp file under csrc/ is compiled to an object in parallel and linked
into tiledarray_amd/libtamx.so.  HIP code is compiled with packed FP32 ops
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtamx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-pass-failed"]


# per-file flags: screening arithmetic decides sparsity patterns and must not
# be contracted into FMAs; MFMA kernels keep their accumulators in VGPRs
# (gfx950 MFMAs read and write VGPRs directly), which avoids AGPR<->VGPR
# copies around every MFMA chain
FILE_FLAGS = {"tamx_shape.hip": ["-ffp-contract=off"],
              "tamx_gemm.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
              "tamx_gemm_gather.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
              "tamx_gemm_gather_f32.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
              "tamx_split.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
              "tamx_split_wide.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
              "tamx_split_i8.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src, obj, verbose):
    cmd = [HIPCC] + COMMON + FILE_FLAGS.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n{msgs}")
    if verbose and msgs.strip():
        print(msgs, flush=True)
    return obj


def build(verbose: bool = False, jobs: int = 8) -> str:
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    tmpdir = tempfile.mkdtemp(prefix="tamx_build_")
    try:
        objs = [os.path.join(tmpdir, os.path.basename(s) + ".o") for s in srcs]
        with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
            list(ex.map(lambda so: _compile(so[0], so[1], verbose), zip(srcs, objs)))
        tmp = LIB + ".tmp"
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, LIB)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
