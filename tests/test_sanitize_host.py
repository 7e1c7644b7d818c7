"""The host code of the path built with AddressSanitizer and
UndefinedBehaviorSanitizer (g++ -fsanitize=address,undefined, every
sanitizer finding fatal) and run on the CPU: the contraction planner
(tiledarray_amd/csrc/tamx_plan.cpp) and the oracle's C restatement
(oracle/ta_oracle.c), driven by tests/cpp/sanitize_host.cpp.  No GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                    reason="needs gcc and g++")
def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-g", "-O1"]
    ora = tmp_path / "ta_oracle.o"
    subprocess.run(["gcc", "-std=c11", "-fopenmp", "-ffp-contract=off", "-c",
                    os.path.join(ROOT, "oracle", "ta_oracle.c"), "-o", str(ora)] + san,
                   check=True, capture_output=True, text=True)
    exe = tmp_path / "sanitize_host"
    r = subprocess.run(["g++", "-std=c++17", "-w", "-D__HIP_PLATFORM_AMD__",
                        "-I", "/opt/rocm/include", "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "tiledarray_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cpp", "sanitize_host.cpp"),
                        os.path.join(ROOT, "tiledarray_amd", "csrc", "tamx_plan.cpp"),
                        str(ora), "-o", str(exe), "-fopenmp", "-ldl"] + san,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", TA_ORACLE_BLAS="loops")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "SANITIZE OK" in r.stdout
