"""ctypes binding of libtamx.so (the C ABI declared in include/tamx.h).

The library is required: importing the bindings without a built library, or
calling a compute entry point without a GPU, raises instead of falling back
to any CPU path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libtamx.so")

TAMX_OK = 0
TAMX_ERR_BAD_ARG, TAMX_ERR_OOM, TAMX_ERR_UNSUPPORTED, TAMX_ERR_DEVICE, TAMX_ERR_CAPACITY = 1, 2, 3, 4, 5
TAMX_NO_TRANS, TAMX_TRANS = 0, 1
TAMX_TASK_BETA1, TAMX_TASK_STORE_TRANS = 1, 2
WL_NTASKS, WL_NPAIRS, WL_CELEMS, WL_STATUS, WL_FLOPS, WL_NCOUNTERS = 0, 1, 2, 3, 4, 8
PLAN_MAX_RANK, PLAN_NAME_LEN = 8, 64
PREC_F64, PREC_F32, PREC_BF16X3, PREC_BF16XS, PREC_F32X2 = 0, 1, 2, 3, 4
TUNE_FORCE, TUNE_NO_WARMUP = 1, 2
NOALLOC_MAX_TILES, NOALLOC_MAX_CHUNKS = 262144, 1048576
LIMB_BF16, LIMB_I8 = 0, 1


class TamxError(RuntimeError):
    def __init__(self, status, where, message):
        super().__init__(f"{where}: {message} (status {status})")
        self.status = status


class GemmTask(ctypes.Structure):
    _fields_ = [("c_off", ctypes.c_int64), ("m", ctypes.c_int32), ("n", ctypes.c_int32),
                ("ldc", ctypes.c_int32), ("pair_begin", ctypes.c_int32),
                ("pair_count", ctypes.c_int32), ("flags", ctypes.c_int32)]


class GemmPair(ctypes.Structure):
    _fields_ = [("a_off", ctypes.c_int64), ("b_off", ctypes.c_int64), ("k", ctypes.c_int32),
                ("lda", ctypes.c_int32), ("ldb", ctypes.c_int32), ("a_tile", ctypes.c_int32),
                ("b_tile", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ContractionDesc(ctypes.Structure):
    _fields_ = [("op_a", ctypes.c_int32), ("op_b", ctypes.c_int32),
                ("Mt", ctypes.c_int64), ("Nt", ctypes.c_int64), ("Kt", ctypes.c_int64),
                ("nbatch", ctypes.c_int64),
                ("mext", ctypes.c_void_p), ("next", ctypes.c_void_p), ("kext", ctypes.c_void_p),
                ("a_off", ctypes.c_void_p), ("b_off", ctypes.c_void_p),
                ("a_norms", ctypes.c_void_p), ("b_norms", ctypes.c_void_p),
                ("c_norms", ctypes.c_void_p),
                ("a_threshold", ctypes.c_float), ("b_threshold", ctypes.c_float),
                ("c_threshold", ctypes.c_float), ("pair_threshold", ctypes.c_float),
                ("store_trans", ctypes.c_int32),
                ("c_off", ctypes.c_void_p), ("c_capacity", ctypes.c_int64)]


class GemmConfig(ctypes.Structure):
    _fields_ = [("bm", ctypes.c_int32), ("bn", ctypes.c_int32), ("bk", ctypes.c_int32),
                ("wgs_per_cu", ctypes.c_int32)]

    def key(self):
        return (self.bm, self.bn, self.bk, self.wgs_per_cu)


def gemm_configs():
    out = []
    for i in range(lib().tamx_gemm_config_count()):
        c = GemmConfig()
        call("tamx_gemm_config_get", i, ctypes.byref(c))
        out.append(c)
    return out


class ContractionPlan(ctypes.Structure):
    _fields_ = [("left_rank", ctypes.c_int32), ("right_rank", ctypes.c_int32),
                ("result_rank", ctypes.c_int32), ("inner_rank", ctypes.c_int32),
                ("left_permtype", ctypes.c_int32), ("right_permtype", ctypes.c_int32),
                ("op_a", ctypes.c_int32), ("op_b", ctypes.c_int32),
                ("left_perm", ctypes.c_int32 * PLAN_MAX_RANK),
                ("right_perm", ctypes.c_int32 * PLAN_MAX_RANK),
                ("result_perm", ctypes.c_int32 * PLAN_MAX_RANK),
                ("result_perm_identity", ctypes.c_int32),
                ("result_perm_matrix_transpose", ctypes.c_int32),
                ("left_gemm_is_inner", ctypes.c_int32 * PLAN_MAX_RANK),
                ("right_gemm_is_inner", ctypes.c_int32 * PLAN_MAX_RANK),
                ("left_target", ctypes.c_char * PLAN_NAME_LEN),
                ("right_target", ctypes.c_char * PLAN_NAME_LEN),
                ("result_natural", ctypes.c_char * PLAN_NAME_LEN)]


_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_f = ctypes.c_float
_d = ctypes.c_double

# name -> argtypes (restype is int status unless listed in _RESTYPES)
SIGNATURES = {
    "tamx_status_string": [_i],
    "tamx_last_error": [],
    "tamx_version": [],
    "tamx_context_create": [_i, _vp],
    "tamx_context_destroy": [_vp],
    "tamx_context_num_cus": [_vp],
    "tamx_worklist_create": [_vp, _i64, _i64, _i64, _vp],
    "tamx_worklist_destroy": [_vp],
    "tamx_worklist_buffers": [_vp, _vp, _vp, _vp],
    "tamx_worklist_read_counters": [_vp, _vp, _vp],
    "tamx_plan_contraction": [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _i,
                              ctypes.POINTER(ContractionPlan)],
    "tamx_norm_plan_create": [_vp, _vp, _vp, _i64, _vp],
    "tamx_norm_plan_create_cells": [_vp, _i64, _vp, _vp, _vp, _vp],
    "tamx_norm_plan_destroy": [_vp],
    "tamx_tile_norms_f64": [_vp, _vp, _vp, _vp, _vp],
    "tamx_tile_norms_f32": [_vp, _vp, _vp, _vp, _vp],
    "tamx_tile_norms_f64_noalloc": [_vp, _vp, _vp, _i64, _vp, _i64, _vp, _vp],
    "tamx_shape_scale": [_vp, _vp, _i, _vp, _i64, _vp, _i, _f, _vp, _vp],
    "tamx_shape_gemm_f32": [_vp, _vp, _vp, _i, _i, _i64, _i64, _i64, _i64, _vp, _f, _f, _vp,
                            _vp, _vp],
    "tamx_contraction_worklist": [_vp, ctypes.POINTER(ContractionDesc), _vp, _vp],
    "tamx_permute_plan_create": [_vp, _i, _vp, _i64, _vp, _vp, _vp, _vp],
    "tamx_permute_plan_destroy": [_vp],
    "tamx_permute_plan_volume": [_vp],
    "tamx_permute_f64": [_vp, _vp, _vp, _vp, _vp],
    "tamx_permute_f32": [_vp, _vp, _vp, _vp, _vp],
    "tamx_grouped_gemm_f64": [_vp, _vp, _i, _i, _d, _vp, _vp, _vp, _vp],
    "tamx_gemm_config_count": [],
    "tamx_gemm_config_get": [_i, _vp],
    "tamx_grouped_gemm_f64_config": [_vp, _vp, _i, _i, _d, _vp, _vp, _vp, _vp, _vp],
    "tamx_grouped_gemm_f32_config": [_vp, _vp, _i, _i, _f, _vp, _vp, _vp, _vp, _vp],
    "tamx_grouped_gemm_f32": [_vp, _vp, _i, _i, _f, _vp, _vp, _vp, _vp],
    "tamx_split_plan_create": [_vp, _i, _i, _i, _i64, _vp, _vp, _vp, _vp],
    "tamx_split_plan_destroy": [_vp],
    "tamx_split_plan_create_general": [_vp, _i, _i64, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp],
    "tamx_split_plan_limb_elems": [_vp],
    "tamx_split_plan_sigmas": [_vp, _vp],
    "tamx_split_f64_premax": [_vp, _vp, _vp, _vp, _vp, _vp],
    "tamx_tile_norms_max_f64": [_vp, _vp, _vp, _vp, _vp, _vp],
    "tamx_split_f64": [_vp, _vp, _vp, _vp, _vp],
    "tamx_grouped_gemm_split": [_vp, _vp, _d, _vp, _vp, _vp, _vp, _vp, _vp],
    "tamx_grouped_gemm_f32x2": [_vp, _vp, _i, _i, _d, _vp, _vp, _vp, _vp],
    "tamx_grouped_gemm_split_config": [_vp, _vp, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "tamx_autotune_grouped_gemm": [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp],
    "tamx_autotune_clear": [],
    "tamx_autotune_grouped_gemm_split": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp,
                                         _vp],
    "tamx_split_plan_nslices": [_vp],
    "tamx_split_plan_create_typed": [_vp, _i, _i, _i, _i, _i64, _vp, _vp, _vp, _vp],
    "tamx_split_plan_create_general_typed": [_vp, _i, _i, _i64, _vp, _i, _i, _vp, _vp, _vp, _vp,
                                             _vp],
    "tamx_split_plan_limb_type": [_vp],
    "tamx_gather_plan_create": [_vp, _i64, _i, _i, _vp, _vp, _vp, _vp, _vp],
    "tamx_gather_plan_destroy": [_vp],
    "tamx_grouped_gemm_gather_f64": [_vp, _vp, _i, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "tamx_grouped_gemm_gather_f32": [_vp, _vp, _i, _i, _f, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "tamx_grouped_gemm_scatter_f64": [_vp, _vp, _i, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                      _vp],
    "tamx_grouped_gemm_scatter_f32": [_vp, _vp, _i, _i, _f, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                      _vp],
    "tamx_grouped_gemm_split_scatter": [_vp, _vp, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "tamx_grouped_gemm_f32x2_scatter": [_vp, _vp, _i, _i, _d, _vp, _vp, _vp, _vp, _vp],
    "tamx_autotune_grouped_gemm_gather": [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i,
                                          _vp, _vp, _vp],
    "tamx_graph_begin": [_vp, _vp],
    "tamx_graph_end": [_vp, _vp, _vp],
    "tamx_graph_launch": [_vp, _vp],
    "tamx_graph_destroy": [_vp],
}
_RESTYPES = {
    "tamx_gemm_config_count": ctypes.c_int,
    "tamx_status_string": ctypes.c_char_p,
    "tamx_last_error": ctypes.c_char_p,
    "tamx_version": ctypes.c_char_p,
    "tamx_permute_plan_volume": _i64,
    "tamx_split_plan_limb_elems": _i64,
}

_lib = None


def lib():
    """Loads libtamx.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(
                f"{LIBPATH} is missing: build it with `python tiledarray_amd/build.py` "
                "(there is no CPU fallback for the contraction path)")
        L = ctypes.CDLL(LIBPATH)
        for name, argtypes in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = argtypes
            f.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(rc, where):
    if rc != TAMX_OK:
        msg = lib().tamx_last_error()
        raise TamxError(rc, where, msg.decode() if msg else lib().tamx_status_string(rc).decode())
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args), name)
