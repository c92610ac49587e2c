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


def pack_codes(codes, bits=8, s0=0, s1=None, device=False):
    """codes [N][M] uint8 -> packed words [ceil(M*bits/32)][s1-s0] uint32.
    device=True packs on the current GPU (glp_pack_codes_device); the words
    are the same."""
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    N, M = c.shape
    s1 = N if s1 is None else s1
    per = 32 // bits
    if device:
        import torch
        cd = torch.from_numpy(c).to("cuda")
        od = torch.empty(((M + per - 1) // per, s1 - s0), dtype=torch.int32, device="cuda")
        check(lib().glp_pack_codes_device(cd.data_ptr(), N, M, bits, s0, s1, od.data_ptr(),
                                          torch.cuda.current_stream().cuda_stream),
              "glp_pack_codes_device")
        return od.cpu().numpy().view(np.uint32)
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


def read_stbin(path, region=None, size_limit=0, keep_names=None, drop_first=True):
    """A SNPTools .bin file through glp_stbin_read: dict(names, chrom, pos,
    ref, alt, gls, not_read); gls is [M][N][2] (Het, AA) / sum with
    drop_first, else [M][N][3] (RR, Het, AA)."""
    L = lib()
    h = ctypes.c_void_p()
    enc = (lambda s: None if s is None else str(s).encode())
    check(L.glp_stbin_read(enc(path), enc(region), size_limit, enc(keep_names), ctypes.byref(h)),
          "glp_stbin_read")
    try:
        n, m, nr = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_size_t()
        check(L.glp_stbin_dims(h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(nr)),
              "glp_stbin_dims")
        N, M = n.value, m.value
        names = [L.glp_stbin_name(h, i).decode() for i in range(N)]
        chrom, pos, ref, alt = [], np.zeros(M, np.uint32), [], []
        c, r, a, p = ctypes.c_char_p(), ctypes.c_char_p(), ctypes.c_char_p(), ctypes.c_uint32()
        for j in range(M):
            check(L.glp_stbin_site(h, j, ctypes.byref(c), ctypes.byref(p), ctypes.byref(r),
                                   ctypes.byref(a)), "glp_stbin_site")
            chrom.append(c.value.decode())
            pos[j] = p.value
            ref.append(r.value.decode())
            alt.append(a.value.decode())
        k = 2 if drop_first else 3
        gls = np.zeros((M, N, k), np.float32)
        if M and N:
            check(L.glp_stbin_gls(h, int(drop_first), ptr(gls, c_f32p)), "glp_stbin_gls")
        return dict(names=names, chrom=chrom, pos=pos, ref=ref, alt=alt, gls=gls,
                    not_read=nr.value)
    finally:
        L.glp_stbin_free(h)


def tran_from_positions(pos, n_haps, density=1.0, inflation=1.0):
    """tran [M][3] from site positions (no genetic map), as Insti::initialize."""
    p = np.ascontiguousarray(pos, dtype=np.uint32)
    out = np.zeros((p.shape[0], 3), np.float32)
    check(lib().glp_tran_from_positions(ptr(p, c_u32p), p.shape[0], n_haps, density, inflation,
                                        ptr(out, c_f32p)), "glp_tran_from_positions")
    return out


def write_vcf(path, names, chrom, pos, ref, alt, hsum, sampling_gens, command_line=""):
    """Impute::result + Insti::save_vcf from allele tallies hsum [2N][M] (uint32)."""
    h = np.ascontiguousarray(hsum, dtype=np.uint32)
    N, M = len(names), len(pos)
    assert h.shape == (2 * N, M)
    strs = lambda xs: (ctypes.c_char_p * len(xs))(*[str(x).encode() for x in xs])
    p = np.ascontiguousarray(pos, dtype=np.uint32)
    check(lib().glp_write_vcf(str(path).encode(), N, strs(names), M, strs(chrom), ptr(p, c_u32p),
                              strs(ref), strs(alt), ptr(h, c_u32p), sampling_gens,
                              command_line.encode()), "glp_write_vcf")
