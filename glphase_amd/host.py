"""Python wrappers of the library's host-side helpers (section 3 of
include/glphase_amd.h): model parameters, GL codebook training, VQ codes and
code packing.  All work is done by the C++ code in libglphase_amd.so."""
import ctypes

import numpy as np

from ._lib import c_f32p, c_f64p, c_u32p, c_u8p, check, lib, ptr


def model_params(n_haps, r_morgans, inflation=1.0):
    """(tran [M][3] f32, pc [4][4] f32, mu) as Insti::initialize builds them."""
    r = np.ascontiguousarray(r_morgans, dtype=np.float64)
    M = r.shape[0]
    tran = np.zeros((M, 3), np.float32)
    pc = np.zeros((4, 4), np.float32)
    mu = ctypes.c_float()
    check(lib().glp_model_params(n_haps, M, ptr(r, c_f64p), inflation, ptr(tran, c_f32p),
                                 ptr(pc, c_f32p), ctypes.byref(mu)), "glp_model_params")
    return tran, pc, mu.value


def vq_train(gls, n_codes=256, key=0, training_size=0, device=False):
    """(codebook [n_codes][2] f32, Lloyd iterations).  device=True runs the
    nearest-codeword search of each iteration on the current GPU; the result
    is the same."""
    g = np.ascontiguousarray(gls, dtype=np.float32).reshape(-1, 3)
    cb = np.zeros((n_codes, 2), np.float32)
    it = ctypes.c_uint32()
    name = "glp_vq_train_device" if device else "glp_vq_train"
    check(getattr(lib(), name)(ptr(g, c_f32p), g.shape[0], n_codes, key, training_size,
                               ptr(cb, c_f32p), ctypes.byref(it)), name)
    return cb, it.value


def vq_codes(gls, codebook, device=False):
    g = np.ascontiguousarray(gls, dtype=np.float32)
    shape = g.shape[:-1]
    g = g.reshape(-1, 3)
    cb = np.ascontiguousarray(codebook, dtype=np.float32)
    out = np.zeros(g.shape[0], np.uint8)
    name = "glp_vq_codes_device" if device else "glp_vq_codes"
    check(getattr(lib(), name)(ptr(g, c_f32p), g.shape[0], ptr(cb, c_f32p), cb.shape[0],
                               ptr(out, c_u8p)), name)
    return out.reshape(shape)


def pack_codes(codes, bits=8, s0=0, s1=None):
    """codes [N][M] uint8 -> packed words [ceil(M*bits/32)][s1-s0] uint32."""
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    N, M = c.shape
    s1 = N if s1 is None else s1
    per = 32 // bits
    out = np.zeros(((M + per - 1) // per, s1 - s0), np.uint32)
    check(lib().glp_pack_codes(ptr(c, c_u8p), N, M, bits, s0, s1, ptr(out, c_u32p)), "glp_pack_codes")
    return out


def pack_nonvq(gls, s0=0, s1=None):
    g = np.ascontiguousarray(gls, dtype=np.float32)
    N, M, _ = g.shape
    s1 = N if s1 is None else s1
    out = np.zeros(((M + 3) // 4, s1 - s0), np.uint32)
    check(lib().glp_pack_nonvq(ptr(g, c_f32p), N, M, s0, s1, ptr(out, c_u32p)), "glp_pack_nonvq")
    return out
