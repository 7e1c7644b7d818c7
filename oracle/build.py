"""Build recipe for the CPU oracle (test infrastructure only).

Compiles ta_oracle.c into oracle/libta_oracle.so with gcc.  The flags keep
IEEE semantics (-ffp-contract=off, no -ffast-math, no -march) so the float
and double operations happen in the order the C source writes them.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libta_oracle.so")


def build(verbose: bool = False) -> str:
    src = os.path.join(HERE, "ta_oracle.c")
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp",
           "-ffp-contract=off", "-Wall", "-Wno-unknown-pragmas",
           "-o", tmp, src, "-ldl", "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
