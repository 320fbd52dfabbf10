"""owl_amd: MI355X implementation of OWL's staggered-grid FD solvers and
FWI gradient.  The compute path is libowl_amd.so (HIP, gfx950) behind the C
ABI in include/owl_amd.h; this package is its Python host interface."""
from ._lib import (OWL_CKPT_BINOMIAL, OWL_CKPT_FULL, OWL_CKPT_UNIFORM, OWL_GRAD_EXACT,
                   OWL_GRAD_OWL, OWL_STORE_BF16, OWL_STORE_F16, OWL_STORE_F32, CflError,
                   OwlError, lib)
from .acoustic2d import Acoustic2DBatch, WaveSolverAcousticIso2D
from .acoustic3d import Acoustic3DBatch
from .elastic2d import Elastic2DBatch
from .elastic3d import Elastic3DBatch
from .fsg2d import FSG2DBatch
from .geometry import Geometry, Receiver, Source, Window, set_adaptive_range

__all__ = [
    "Acoustic2DBatch", "Acoustic3DBatch", "Elastic2DBatch", "Elastic3DBatch", "FSG2DBatch", "WaveSolverAcousticIso2D",
    "Geometry", "Receiver", "Source", "Window",
    "set_adaptive_range", "OwlError", "CflError", "lib", "OWL_GRAD_OWL", "OWL_GRAD_EXACT",
    "OWL_CKPT_FULL", "OWL_CKPT_UNIFORM", "OWL_CKPT_BINOMIAL", "OWL_STORE_F32", "OWL_STORE_F16",
    "OWL_STORE_BF16",
]
