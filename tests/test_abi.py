"""CPU-side checks of the C ABI: the library loads, exports every function
include/tamx.h declares, and its host-only planner agrees with the oracle's
restatement of GEMMPermutationOptimizer on hand-derived and random cases.
No compute entry point is called here (no GPU)."""
import ctypes
import itertools
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import permopt
from tiledarray_amd import _lib
from tiledarray_amd._lib import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tamx.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tamx_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_path():
    names = declared_functions()
    for required in ["tamx_context_create", "tamx_grouped_gemm_f64", "tamx_tile_norms_f64",
                     "tamx_shape_gemm_f32", "tamx_contraction_worklist", "tamx_permute_f64",
                     "tamx_plan_contraction"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIBPATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (tamx_[a-z0-9_]+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_every_declared_symbol():
    missing = [n for n in declared_functions() if n not in _lib.SIGNATURES]
    assert not missing, missing
    L = lib()
    for n in declared_functions():
        assert getattr(L, n) is not None


def test_status_strings_and_version():
    L = lib()
    assert L.tamx_status_string(0) == b"ok"
    assert L.tamx_status_string(5) == b"capacity exceeded"
    assert b"gfx950" in L.tamx_version()


def _plan(left, right, target, prefer=True):
    p = _lib.ContractionPlan()
    rc = lib().tamx_plan_contraction(left.encode(), right.encode(), target.encode(),
                                     1 if prefer else 0, ctypes.byref(p))
    return rc, p


def _check_against_oracle(left, right, target, prefer=True):
    L, R, T = left.split(","), right.split(","), target.split(",")
    tl, tr, res, lt, rt = permopt.compute_index_list_contraction(L, R, prefer)
    rc, p = _plan(left, right, target, prefer)
    assert rc == 0, lib().tamx_last_error()
    assert (p.left_permtype, p.right_permtype) == (lt, rt)
    assert (p.op_a, p.op_b) == (permopt.to_cblas_op(lt), permopt.to_cblas_op(rt))
    assert p.result_natural.decode().split(",") == res
    # left/right tiles are permuted to the oracle's target lists when general
    if lt == permopt.GENERAL:
        perm = list(p.left_perm[:len(L)])
        assert [L[perm.index(i)] for i in range(len(L))] == tl
    if rt == permopt.GENERAL:
        perm = list(p.right_perm[:len(R)])
        assert [R[perm.index(i)] for i in range(len(R))] == tr
    rp = list(p.result_perm[:len(res)])
    moved = [None] * len(res)
    for d, x in enumerate(res):
        moved[rp[d]] = x
    assert moved == T
    assert p.result_perm_identity == (res == T)


@pytest.mark.parametrize("left,right,target", [
    ("m,k", "k,n", "m,n"),
    ("m,k", "k,n", "n,m"),
    ("k,m", "k,n", "m,n"),
    ("m,k", "n,k", "m,n"),
    ("c,d,i,j", "a,b,c,d", "a,b,i,j"),
    ("c,d,i,j", "a,c,b,d", "a,b,i,j"),
    ("i,k,j", "k,l", "i,j,l"),
    ("i,k,l", "l,k,j", "i,j"),
])
def test_planner_matches_oracle_hand_cases(left, right, target):
    _check_against_oracle(left, right, target)


def test_planner_ladder_case_details():
    # R(a,b,i,j) = T(c,d,i,j) * V(a,c,b,d): T folds into Trans, V is permuted
    # (a,c,b,d) -> (c,d,a,b), the result (i,j,a,b) -> (a,b,i,j) is the
    # transpose of the fused result tile
    rc, p = _plan("c,d,i,j", "a,c,b,d", "a,b,i,j")
    assert rc == 0
    assert p.op_a == 1 and p.op_b == 0
    assert list(p.right_perm[:4]) == [2, 0, 3, 1]
    assert p.right_target == b"c,d,a,b"
    assert p.result_perm_matrix_transpose == 1
    assert list(p.left_gemm_is_inner[:4]) == [1, 1, 0, 0]
    assert list(p.right_gemm_is_inner[:4]) == [1, 1, 0, 0]


def test_planner_random_cases():
    rng = np.random.default_rng(7)
    letters = list("abcdefgh")
    for _ in range(300):
        n_lo, n_in, n_ro = rng.integers(1, 3), rng.integers(1, 3), rng.integers(1, 3)
        names = list(rng.permutation(letters)[:n_lo + n_in + n_ro])
        lo, inn, ro = names[:n_lo], names[n_lo:n_lo + n_in], names[n_lo + n_in:]
        L = list(rng.permutation(lo + inn))
        R = list(rng.permutation(inn + ro))
        T = list(rng.permutation(lo + ro))
        _check_against_oracle(",".join(L), ",".join(R), ",".join(T), bool(rng.integers(0, 2)))


@pytest.mark.parametrize("left,right,target", [
    ("", "k,n", "m,n"),          # empty operand
    ("m,k", "k,n", "m,q"),       # target does not match the free indices
    ("m,m", "m,n", "n"),         # repeated index
    ("m", "n", "m,n"),           # outer product: not a contraction
])
def test_planner_rejects_bad_annotations(left, right, target):
    rc, _ = _plan(left, right, target)
    assert rc == _lib.TAMX_ERR_BAD_ARG if hasattr(_lib, "TAMX_ERR_BAD_ARG") else rc == 1
    assert lib().tamx_last_error()


def test_plan_constructors_reject_bad_arguments_before_touching_a_device():
    # argument checks come before any device call, so they run without a GPU
    L = _lib.lib()
    p = ctypes.c_void_p()
    off = np.zeros(1, np.int64)
    ext = np.array([4], np.int32)
    st = np.array([1], np.int64)
    # unknown limb type, limb count out of range
    assert L.tamx_split_plan_create_typed(None, 7, 3, 0, 0, 1, off.ctypes.data, ext.ctypes.data,
                                          ext.ctypes.data, ctypes.byref(p)) == _lib.TAMX_ERR_BAD_ARG
    assert L.tamx_split_plan_create_typed(None, _lib.LIMB_I8, 7, 0, 0, 1, off.ctypes.data,
                                          ext.ctypes.data, ext.ctypes.data,
                                          ctypes.byref(p)) == _lib.TAMX_ERR_BAD_ARG
    assert L.tamx_split_plan_create_general_typed(None, _lib.LIMB_BF16, 3, 1, off.ctypes.data, 5,
                                                  1, ext.ctypes.data, st.ctypes.data,
                                                  ext.ctypes.data, st.ctypes.data,
                                                  ctypes.byref(p)) == _lib.TAMX_ERR_BAD_ARG
    # gather plans: 1 to 4 modes each side
    assert L.tamx_gather_plan_create(None, 1, 0, 1, ext.ctypes.data, st.ctypes.data,
                                     ext.ctypes.data, st.ctypes.data,
                                     ctypes.byref(p)) == _lib.TAMX_ERR_BAD_ARG
    assert L.tamx_gather_plan_create(None, -1, 1, 1, ext.ctypes.data, st.ctypes.data,
                                     ext.ctypes.data, st.ctypes.data,
                                     ctypes.byref(p)) == _lib.TAMX_ERR_BAD_ARG
    assert p.value is None
    assert L.tamx_split_plan_limb_type(None) == _lib.LIMB_BF16


@pytest.mark.parametrize("target,left,right", [
    ("i,j", "i,k", "k,j"),          # no fused index: a plain contraction
    ("n,i,j", "n,k,k", "n,k,j"),    # an index repeated within an operand
    ("n,i,i", "n,i,k", "n,k,j"),    # a repeated result index
    ("n,i,j,z", "n,i,k", "n,k,j"),  # a result index in no operand
])
def test_fused_contract_rejects_bad_index_patterns(target, left, right):
    # the index analysis runs before any device work
    from tiledarray_amd import engine as E
    with pytest.raises(ValueError):
        E.fused_contract(None, target, None, left, None, right)
