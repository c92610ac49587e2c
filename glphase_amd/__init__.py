"""glphase_amd: gfx950 (MI355X) engine for GLPhase's Metropolis-Hastings
surrogate-parent update.

libglphase_amd.so holds the HIP kernels, the extern "C" boundary
(include/glphase_amd.h) and the C++ host classes (HMMLike, GLPack, GLVQ,
samplers).  This package loads it through ctypes for tests and benchmarks.
"""
from ._lib import (GlpError, PREC_F32, PREC_MIXED, PREC_MIXED_FB, PROP_HOST, PROP_KLIST,
                   PROP_UNIFORM, lib)

__all__ = ["GlpError", "lib", "PROP_UNIFORM", "PROP_KLIST", "PROP_HOST", "PREC_F32",
           "PREC_MIXED", "PREC_MIXED_FB"]
