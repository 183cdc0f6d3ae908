"""MI355X (gfx950) implementation of the STAQ-DIC hot path.

The compute lives in libstaqdic.so (HIP kernels behind the C ABI of
include/staq_dic.h); this package is its Python binding and the host-side
mirror of the reference's MATLAB functions for the path.
"""
from ._lib import Batch, Context, StaqError, load, params  # noqa: F401
from .api import (  # noqa: F401
    IntegerSearch, LocalICGN, LocalICGNQuadtree, Subpb1, Subpb1Quadtree, ba_interp2,
    classify_local, classify_subpb1, default_context, funImgGradient, funIntegerSearch,
    local_icgn_raw, subpb1_raw,
)
from .global_step import Subpb2Quadtree  # noqa: F401
from .initguess import Init, regularizeNd, uniform_initial_guess  # noqa: F401
from .mesh import GenerateQuadtreeMesh, MeshSetUp  # noqa: F401
from .mgsearch import funIntegerSearchMg, gridfit  # noqa: F401

__all__ = [
    "Batch", "Context", "StaqError", "load", "params", "IntegerSearch", "LocalICGN",
    "LocalICGNQuadtree", "Subpb1", "Subpb1Quadtree", "ba_interp2", "classify_local",
    "classify_subpb1", "default_context", "funImgGradient", "funIntegerSearch",
    "local_icgn_raw", "subpb1_raw", "Subpb2Quadtree", "funIntegerSearchMg", "gridfit",
    "GenerateQuadtreeMesh", "MeshSetUp", "Init", "regularizeNd", "uniform_initial_guess",
]
