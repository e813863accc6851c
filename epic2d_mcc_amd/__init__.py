"""MI355X-native per-timestep particle loop of ePIC2D_mcc.

libepic2d.so (hand-written gfx950 HIP kernels behind the C-ABI of
include/epic2d.h) does the work; this package binds it (ctypes), reads the
reference's input.ini format and derives its normalisation.
"""
from . import config  # noqa: F401
from ._lib import (BC, FP32, FP64, OP_DIRICHLET_WALL, OP_REF_GHOST, SHAPE, SOLVER_GS, SOLVER_PCG,  # noqa: F401
                   SOLVER_PCG_MIXED, SOLVER_SPECTRAL, Context, DomainDesc, EmitterDesc, EpicError, MccDesc,
                   SpeciesDesc, device_name, header_symbols, lib)

__all__ = ["config", "Context", "DomainDesc", "SpeciesDesc", "EmitterDesc", "MccDesc", "EpicError", "lib",
           "device_name", "header_symbols"]
