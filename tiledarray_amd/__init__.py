"""tiledarray_amd -- MI355X-native TiledArray contraction hot path.

The compute path is libtamx.so (HIP kernels for gfx950 behind the C ABI of
include/tamx.h).  This package holds its Python binding and the single-rank
host driver; there is no CPU fallback for any compute step.
"""
from ._lib import TamxError, lib  # noqa: F401
from .tiled import DistArray, TiledRange, TiledRange1  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # the engine imports torch's GPU runtime lazily
    if name in ("Context", "Contraction", "contract", "make_sparse", "tile_norms",
                "plan_contraction", "permute_array", "WorkList", "PermutePlan",
                "NormPlan", "scale_norms", "k_sizes", "FLOAT_EPS"):
        from . import engine
        return getattr(engine, name)
    raise AttributeError(name)
