"""Build recipe for the C++ tests under tests/cpp (test infrastructure).

test_contract.bin: the host interface of include/tiledarray_amd.hpp, linked
against tiledarray_amd/libtamx.so.
test_abi_threads.bin: the C ABI of include/tamx.h driven as TiledArray's
callers drive it (per-pair calls, concurrent host threads), checked against
the CPU oracle, so it also links oracle/libta_oracle.so.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = ["-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
LDFLAGS = ["-L", os.path.join(ROOT, "tiledarray_amd"), "-ltamx",
           "-Wl,-rpath,$ORIGIN/../../tiledarray_amd"]

TESTS = {
    "test_contract": [],
    "test_abi_threads": ["-L", os.path.join(ROOT, "oracle"), "-lta_oracle", "-lpthread",
                         "-Wl,-rpath,$ORIGIN/../../oracle"],
}


def build(verbose: bool = False) -> list:
    """Compiles every C++ test; libtamx.so and libta_oracle.so must be built."""
    outs = []
    for name, extra in TESTS.items():
        src = os.path.join(HERE, name + ".cpp")
        out = os.path.join(HERE, name + ".bin")
        cmd = [HIPCC] + CFLAGS + ["-o", out + ".tmp", src] + LDFLAGS + extra
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ test build failed ({name}):\n" + r.stdout + r.stderr)
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
