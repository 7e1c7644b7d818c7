"""Runs the C++ tests under tests/cpp, built by __graft_entry__.build():
test_contract (the host interface of include/tiledarray_amd.hpp) and
test_abi_threads (the C ABI driven per pair and from four host threads,
against the oracle)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp")


@pytest.mark.parametrize("name", ["test_contract", "test_abi_threads"])
def test_cpp(name):
    binary = os.path.join(HERE, name + ".bin")
    if not os.path.exists(binary):
        pytest.fail(f"{binary} is missing: run __graft_entry__.build()")
    r = subprocess.run([binary], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
