"""ctypes bindings of libglphase_amd.so (include/glphase_amd.h).

The library is the product: every call here goes to the HIP engine or to its
C++ host helpers.  Loading fails loudly when the shared library is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libglphase_amd.so")

GLP_OK = 0
PROP_UNIFORM, PROP_KLIST, PROP_HOST = 0, 1, 2
PREC_F32, PREC_MIXED, PREC_MIXED_FB = 0, 1, 2
GLP_NOALLOC_MAX_WINDOWS = 1024  # include/glphase_amd.h

c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_f32p = ctypes.POINTER(ctypes.c_float)
c_f64p = ctypes.POINTER(ctypes.c_double)
c_u8p = ctypes.POINTER(ctypes.c_uint8)


class GlpWindow(ctypes.Structure):
    _fields_ = [
        ("num_sites", ctypes.c_uint),
        ("words_per_hap", ctypes.c_uint),
        ("panel", ctypes.c_void_p),
        ("packed", ctypes.c_void_p),
        ("tran", ctypes.c_void_p),
        ("parents", ctypes.c_void_p),
        ("like", ctypes.c_void_p),
    ]


class GlpWorkWindow(ctypes.Structure):
    _fields_ = [
        ("num_sites", ctypes.c_uint),
        ("words_per_hap", ctypes.c_uint),
        ("panel", ctypes.c_void_p),
        ("parents", ctypes.c_void_p),
        ("tran", ctypes.c_void_p),
        ("het_aa", ctypes.c_void_p),
        ("emit", ctypes.c_void_p),
        ("hnew", ctypes.c_void_p),
    ]


class GlpWorkDesc(ctypes.Structure):
    _fields_ = [
        ("num_windows", ctypes.c_uint),
        ("windows", ctypes.POINTER(GlpWorkWindow)),
        ("num_samples", ctypes.c_uint),
        ("num_haps", ctypes.c_uint),
        ("pc16", ctypes.c_float * 16),
        ("pen", ctypes.c_float),
        ("seed", ctypes.c_uint64),
        ("gen", ctypes.c_uint),
        ("workspace", ctypes.c_void_p),
    ]


class GlpKnnDesc(ctypes.Structure):
    _fields_ = [
        ("panel", ctypes.c_void_p),
        ("num_haps", ctypes.c_uint),
        ("num_sites", ctypes.c_uint),
        ("num_samples", ctypes.c_uint),
        ("K", ctypes.c_uint),
        ("freq_lb", ctypes.c_double),
        ("freq_ub", ctypes.c_double),
        ("using_maf", ctypes.c_int),
        ("klist", ctypes.c_void_p),
        ("workspace", ctypes.c_void_p),
    ]


class GlpMhDesc(ctypes.Structure):
    _fields_ = [
        ("num_windows", ctypes.c_uint),
        ("windows", ctypes.POINTER(GlpWindow)),
        ("num_samples", ctypes.c_uint),
        ("num_haps", ctypes.c_uint),
        ("bits_per_code", ctypes.c_uint),
        ("emit_lut", ctypes.c_void_p),
        ("num_cycles", ctypes.c_uint),
        ("k0", ctypes.c_uint),
        ("pen", ctypes.c_float),
        ("seed", ctypes.c_uint64),
        ("gen", ctypes.c_uint),
        ("prop_mode", ctypes.c_int),
        ("klist", ctypes.c_void_p),
        ("K", ctypes.c_uint),
        ("extra", ctypes.c_void_p),
        ("precision", ctypes.c_int),
        ("trace_like", ctypes.c_void_p),
        ("trace_acc", ctypes.c_void_p),
        ("kickstart_from_ref", ctypes.c_int),
    ]


# (name, restype, argtypes) of every symbol include/glphase_amd.h declares
SIGNATURES = [
    ("glp_last_error", ctypes.c_char_p, []),
    ("glp_version", ctypes.c_char_p, []),
    ("glp_check_device", ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    ("glp_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("glp_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("glp_set_tran", ctypes.c_int, [ctypes.c_void_p, c_f32p, ctypes.c_size_t]),
    ("glp_set_mutation_mat", ctypes.c_int, [ctypes.c_void_p, c_f32p]),
    ("glp_set_codebook", ctypes.c_int, [ctypes.c_void_p, c_f32p, ctypes.c_uint]),
    ("glp_set_packed_gls", ctypes.c_int, [ctypes.c_void_p, c_u32p, ctypes.c_size_t, ctypes.c_uint]),
    ("glp_set_pen", ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
    ("glp_run_hmm_on_device", ctypes.c_int,
     [ctypes.c_void_p, c_u64p, ctypes.c_size_t, c_u32p, ctypes.c_uint, ctypes.c_uint,
      ctypes.c_uint, c_u32p, ctypes.c_ulong]),
    ("glp_hmm_like", ctypes.c_int,
     [ctypes.c_void_p, c_u64p, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
      c_u32p, c_u32p, c_f32p]),
    ("glp_emit_lut_bytes", ctypes.c_size_t, [ctypes.c_uint, ctypes.c_int]),
    ("glp_build_emit_lut", ctypes.c_int, [c_f32p, ctypes.c_uint, c_f32p, ctypes.c_int, ctypes.c_void_p]),
    ("glp_plain_init", ctypes.c_int, [ctypes.POINTER(GlpMhDesc), ctypes.c_void_p]),
    ("glp_plain_step", ctypes.c_int, [ctypes.POINTER(GlpMhDesc), ctypes.c_uint, ctypes.c_void_p]),
    ("glp_plan_create", ctypes.c_int,
     [ctypes.POINTER(GlpMhDesc), ctypes.c_uint, ctypes.POINTER(ctypes.c_void_p)]),
    ("glp_plan_run", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint, ctypes.c_float, ctypes.c_uint64, ctypes.c_void_p]),
    ("glp_mh_run_noalloc", ctypes.c_int,
     [ctypes.POINTER(GlpMhDesc), ctypes.c_uint, ctypes.c_void_p, c_u64p]),
    ("glp_plan_run_plain", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint, ctypes.c_float, ctypes.c_uint64, ctypes.c_void_p]),
    ("glp_plan_stats", ctypes.c_int, [ctypes.c_void_p, c_u64p, c_u64p]),
    ("glp_plan_lanes", ctypes.c_int, [ctypes.c_void_p, c_u32p, c_u32p]),
    ("glp_plan_set_lanes", ctypes.c_int, [ctypes.c_void_p, c_u32p]),
    ("glp_plan_set_waves", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint]),
    ("glp_plan_waves", ctypes.c_int, [ctypes.c_void_p, c_u32p]),
    ("glp_plan_set_split", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("glp_plan_split", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    ("glp_plan_autotune", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint, ctypes.c_float, ctypes.c_uint64, ctypes.c_void_p, c_u32p]),
    ("glp_plan_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("glp_knn_k_out", ctypes.c_uint, [ctypes.c_uint, ctypes.c_uint]),
    ("glp_knn_workspace_bytes", ctypes.c_size_t, [ctypes.POINTER(GlpKnnDesc)]),
    ("glp_knn_build", ctypes.c_int, [ctypes.POINTER(GlpKnnDesc), ctypes.c_void_p]),
    ("glp_plan_graph_run", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_float,
                                          ctypes.c_uint64, ctypes.c_void_p]),
    ("glp_plan_graph_run_plain", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_float,
                                                ctypes.c_uint64, ctypes.c_void_p]),
    ("glp_work_workspace_bytes", ctypes.c_size_t, [ctypes.POINTER(GlpWorkDesc)]),
    ("glp_hmm_work", ctypes.c_int, [ctypes.POINTER(GlpWorkDesc), ctypes.c_void_p]),
    ("glp_work_emit", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, c_f32p,
                                     ctypes.c_void_p, ctypes.c_void_p]),
    ("glp_selftest_recip", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.POINTER(ctypes.c_uint64),
                                          ctypes.POINTER(ctypes.c_uint32)]),
    ("glp_init_haps", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint64,
                                     ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p]),
    ("glp_replace", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p,
                                   ctypes.c_void_p]),
    ("glp_stbin_read", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t,
                                      ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("glp_stbin_dims", ctypes.c_int, [ctypes.c_void_p, c_u32p, c_u32p,
                                      ctypes.POINTER(ctypes.c_size_t)]),
    ("glp_stbin_name", ctypes.c_char_p, [ctypes.c_void_p, ctypes.c_uint]),
    ("glp_stbin_site", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint,
                                      ctypes.POINTER(ctypes.c_char_p), c_u32p,
                                      ctypes.POINTER(ctypes.c_char_p),
                                      ctypes.POINTER(ctypes.c_char_p)]),
    ("glp_stbin_gls", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_f32p]),
    ("glp_stbin_free", None, [ctypes.c_void_p]),
    ("glp_tran_from_positions", ctypes.c_int, [c_u32p, ctypes.c_uint, ctypes.c_uint,
                                               ctypes.c_double, ctypes.c_double, c_f32p]),
    ("glp_write_vcf", ctypes.c_int,
     [ctypes.c_char_p, ctypes.c_uint, ctypes.POINTER(ctypes.c_char_p), ctypes.c_uint,
      ctypes.POINTER(ctypes.c_char_p), c_u32p, ctypes.POINTER(ctypes.c_char_p),
      ctypes.POINTER(ctypes.c_char_p), c_u32p, ctypes.c_uint, ctypes.c_char_p]),
    ("glp_model_params", ctypes.c_int,
     [ctypes.c_uint, ctypes.c_uint, c_f64p, ctypes.c_double, c_f32p, c_f32p, c_f32p]),
    ("glp_vq_train", ctypes.c_int,
     [c_f32p, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint64, ctypes.c_size_t, c_f32p, c_u32p]),
    ("glp_vq_codes", ctypes.c_int, [c_f32p, ctypes.c_size_t, c_f32p, ctypes.c_uint, c_u8p]),
    ("glp_vq_train_device", ctypes.c_int,
     [c_f32p, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint64, ctypes.c_size_t, c_f32p, c_u32p]),
    ("glp_vq_codes_device", ctypes.c_int,
     [c_f32p, ctypes.c_size_t, c_f32p, ctypes.c_uint, c_u8p]),
    ("glp_vq_assign_device", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, c_f32p, ctypes.c_uint, ctypes.c_void_p,
      ctypes.c_void_p]),
    ("glp_pack_codes", ctypes.c_int,
     [c_u8p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, c_u32p]),
    ("glp_pack_codes_device", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
      ctypes.c_void_p, ctypes.c_void_p]),
    ("glp_pack_nonvq", ctypes.c_int,
     [c_f32p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, c_u32p]),
]

_lib = None


class GlpError(RuntimeError):
    pass


def lib():
    """The loaded library; raises when it has not been built."""
    global _lib
    if _lib is None:
        # torch ships its own HIP runtime; loading it first makes this
        # library bind to the same libamdhip64 instance in one process
        try:
            import torch  # noqa: F401
        except Exception:
            pass
        if not os.path.exists(LIB_PATH):
            raise GlpError(
                "libglphase_amd.so is not built (run __graft_entry__.build() or "
                "python glphase_amd/build.py)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != GLP_OK:
        msg = lib().glp_last_error().decode(errors="replace")
        raise GlpError(f"{what}: status {rc}: {msg}")


def ptr(a, ctype):
    """ctypes pointer to a contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctype)
