"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's cpu_baseline
leg, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liboracle.so")

u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
f32p = ctypes.POINTER(ctypes.c_float)
f64p = ctypes.POINTER(ctypes.c_double)
u8p = ctypes.POINTER(ctypes.c_uint8)


class OrcWindow(ctypes.Structure):
    _fields_ = [("haps", u64p), ("wn", ctypes.c_uint), ("nhaps", ctypes.c_uint),
                ("packed", u32p), ("nsamp", ctypes.c_uint), ("M", ctypes.c_uint),
                ("bits", ctypes.c_uint), ("lut", f32p), ("tran", f32p)]


class OrcMhArgs(ctypes.Structure):
    _fields_ = [("ncycles", ctypes.c_uint), ("pen", ctypes.c_float), ("seed", ctypes.c_uint64),
                ("window_id", ctypes.c_uint32), ("gen", ctypes.c_uint32), ("k0", ctypes.c_uint32),
                ("prop_mode", ctypes.c_int), ("precision", ctypes.c_int), ("klist", u32p),
                ("K", ctypes.c_uint), ("init_parents", u32p), ("extra", u32p),
                ("kickstart", ctypes.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_L = None


def L():
    global _L
    if _L is None:
        if not os.path.exists(SO):
            build()
        lib = ctypes.CDLL(SO)
        lib.orc_norm.restype = ctypes.c_float
        lib.orc_logf_cr.restype = ctypes.c_float
        lib.orc_logf_cr.argtypes = [ctypes.c_float]
        lib.orc_uniform_int.restype = ctypes.c_uint32
        lib.orc_unif_excl.restype = ctypes.c_uint32
        lib.orc_hmm_like_f32.restype = ctypes.c_float
        for f in ("orc_hmm_like_mixed", "orc_hmm_like_mixed_pass", "orc_hmm_like_f64",
                  "orc_hmm_like_brute", "orc_hmm_like_fb", "orc_hmm_like_fb_pass"):
            getattr(lib, f).restype = ctypes.c_double
        lib.orc_vq_find.restype = ctypes.c_uint
        lib.orc_mu.restype = ctypes.c_float
        lib.orc_mu.argtypes = [ctypes.c_uint]
        lib.orc_pc.argtypes = [ctypes.c_float, f32p]
        lib.orc_vq_find.argtypes = [f32p, ctypes.c_uint, ctypes.c_float, ctypes.c_float]
        lib.orc_vq_train.argtypes = [f32p, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint64,
                                     ctypes.c_size_t, f32p, u32p]
        lib.orc_vq_codes.argtypes = [f32p, ctypes.c_size_t, f32p, ctypes.c_uint, u8p]
        lib.orc_draw.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.c_uint32, u32p]
        lib.orc_tran_from_morgans.argtypes = [f64p, ctypes.c_uint, ctypes.c_double, f32p]
        lib.orc_hmm_work_weights.argtypes = [u64p, ctypes.c_uint, u32p, f32p, f32p, f32p, f32p,
                                             ctypes.c_float, ctypes.c_uint, f32p, f32p]
        lib.orc_hmm_work_sample.argtypes = [f32p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint64,
                                            ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                            u64p, u64p]
        lib.orc_prob_emit.argtypes = [f32p, ctypes.c_uint, f32p, f32p, f32p]
        lib.orc_knn_k_out.restype = ctypes.c_uint
        lib.orc_knn_k_out.argtypes = [ctypes.c_uint, ctypes.c_uint]
        lib.orc_knn_mask.argtypes = [u64p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_int, u64p, f64p]
        lib.orc_knn.argtypes = [u64p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                ctypes.c_uint, ctypes.c_double, ctypes.c_double, ctypes.c_int, u32p]
        lib.orc_knn_rows.argtypes = [u64p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                     ctypes.c_uint, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     u32p, ctypes.c_uint, u32p, ctypes.c_int]
        lib.orc_hmm_work.argtypes = [u64p, ctypes.c_uint, u32p, f32p, f32p, f32p, f32p,
                                     ctypes.c_float, ctypes.c_uint, ctypes.c_uint64,
                                     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, u64p, u64p]
        lib.orc_tran_from_positions.argtypes = [u32p, ctypes.c_uint, ctypes.c_uint,
                                                ctypes.c_double, ctypes.c_double, f32p]
        lib.orc_init_haps.argtypes = [f32p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint64,
                                      ctypes.c_uint32, u64p]
        lib.orc_replace.argtypes = [u64p, ctypes.c_uint, ctypes.c_uint, u32p]
        _L = lib
    return _L


def P(a, t):
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return a.ctypes.data_as(t)


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    L().orc_philox(P(c, u32p), P(k, u32p), P(o, u32p))
    return o


def draw(seed, k, sample, window, gen):
    o = np.zeros(4, np.uint32)
    L().orc_draw(seed, k, sample, window, gen, P(o, u32p))
    return o


def emit_lut(codebook, pc):
    cb = np.ascontiguousarray(codebook, np.float32)
    p = np.ascontiguousarray(pc, np.float32).reshape(16)
    out = np.zeros((cb.shape[0], 4), np.float32)
    L().orc_emit_lut(P(cb, f32p), ctypes.c_uint(cb.shape[0]), P(p, f32p), P(out, f32p))
    return out


def unpack_gl(word, codebook, bits, slot):
    cb = np.ascontiguousarray(codebook, np.float32)
    out = np.zeros(3, np.float32)
    L().orc_unpack_gl(ctypes.c_uint32(word), P(cb, f32p), ctypes.c_uint(bits), ctypes.c_uint(slot),
                      P(out, f32p))
    return out


def sample_emit(packed, sample, M, bits, lut):
    pk = np.ascontiguousarray(packed, np.uint32)
    lt = np.ascontiguousarray(lut, np.float32)
    out = np.zeros((M, 4), np.float32)
    L().orc_sample_emit(P(pk, u32p), ctypes.c_uint(pk.shape[1]), ctypes.c_uint(sample),
                        ctypes.c_uint(M), ctypes.c_uint(bits), P(lt, f32p), P(out, f32p))
    return out


def _like(fn, haps, parents, emit, tran):
    h = np.ascontiguousarray(haps, np.uint64)
    pr = np.ascontiguousarray(parents, np.uint32)
    e = np.ascontiguousarray(emit, np.float32)
    t = np.ascontiguousarray(tran, np.float32)
    return fn(P(h, u64p), ctypes.c_uint(h.shape[1]), P(pr, u32p), P(e, f32p), P(t, f32p),
              ctypes.c_uint(e.shape[0]))


def hmm_like_f32(haps, parents, emit, tran):
    return _like(L().orc_hmm_like_f32, haps, parents, emit, tran)


def hmm_like_mixed(haps, parents, emit, tran):
    return _like(L().orc_hmm_like_mixed, haps, parents, emit, tran)


def hmm_like_mixed_pass(haps, parents, emit, tran, every):
    """One mixed-precision pass (scaled every 8th site or every site) and
    whether a sum fell below 2^-99 in it, or its last sum is not finite."""
    h = np.ascontiguousarray(haps, np.uint64)
    pr = np.ascontiguousarray(parents, np.uint32)
    e = np.ascontiguousarray(emit, np.float32)
    t = np.ascontiguousarray(tran, np.float32)
    tiny = ctypes.c_int(0)
    v = L().orc_hmm_like_mixed_pass(P(h, u64p), ctypes.c_uint(h.shape[1]), P(pr, u32p),
                                    P(e, f32p), P(t, f32p), ctypes.c_uint(e.shape[0]),
                                    ctypes.c_uint(every), ctypes.byref(tiny))
    return v, bool(tiny.value)


def hmm_like_fb(haps, parents, emit, tran):
    return _like(L().orc_hmm_like_fb, haps, parents, emit, tran)


def hmm_like_fb_pass(haps, parents, emit, tran, every):
    """One forward-backward pass (both halves scaled every 8th step or every
    step) and whether either half flagged a tiny or non-finite sum."""
    h = np.ascontiguousarray(haps, np.uint64)
    pr = np.ascontiguousarray(parents, np.uint32)
    e = np.ascontiguousarray(emit, np.float32)
    t = np.ascontiguousarray(tran, np.float32)
    tiny = ctypes.c_int(0)
    v = L().orc_hmm_like_fb_pass(P(h, u64p), ctypes.c_uint(h.shape[1]), P(pr, u32p),
                                 P(e, f32p), P(t, f32p), ctypes.c_uint(e.shape[0]),
                                 ctypes.c_uint(every), ctypes.byref(tiny))
    return v, bool(tiny.value)


def hmm_like_f64(haps, parents, emit, tran):
    return _like(L().orc_hmm_like_f64, haps, parents, emit, tran)


def hmm_like_brute(haps, parents, emit, tran):
    return _like(L().orc_hmm_like_brute, haps, parents, emit, tran)


class MH:
    """orc_mh_window over one window; keeps the numpy buffers alive."""

    def __init__(self, haps, packed, bits, lut, tran, nsamp):
        self.h = np.ascontiguousarray(haps, np.uint64)
        self.pk = np.ascontiguousarray(packed, np.uint32)
        self.lut = np.ascontiguousarray(lut, np.float32)
        self.t = np.ascontiguousarray(tran, np.float32)
        self.w = OrcWindow(P(self.h, u64p), self.h.shape[1], self.h.shape[0], P(self.pk, u32p),
                           nsamp, self.t.shape[0], bits, P(self.lut, f32p), P(self.t, f32p))
        self.nsamp = nsamp

    def run(self, ncycles, pen, seed, window_id=0, gen=0, k0=0, prop_mode=0, precision=0,
            klist=None, init_parents=None, extra=None, s0=0, s1=None, trace=False,
            nthreads=0, kickstart=False):
        s1 = self.nsamp if s1 is None else s1
        a = OrcMhArgs()
        a.ncycles, a.pen, a.seed = ncycles, pen, seed
        a.window_id, a.gen, a.k0 = window_id, gen, k0
        a.prop_mode, a.precision = prop_mode, precision
        a.kickstart = int(kickstart)
        keep = []
        if klist is not None:
            kl = np.ascontiguousarray(klist, np.uint32)
            keep.append(kl)
            a.klist, a.K = P(kl, u32p), kl.shape[-1]
        if init_parents is not None:
            ip = np.ascontiguousarray(init_parents, np.uint32)
            keep.append(ip)
            a.init_parents = P(ip, u32p)
        if extra is not None:
            ex = np.ascontiguousarray(extra, np.uint32)
            keep.append(ex)
            a.extra = P(ex, u32p)
        par = np.zeros((self.nsamp, 4), np.uint32)
        cur = np.zeros(self.nsamp, np.float64)
        if nthreads:
            L().orc_mh_window_omp(ctypes.byref(self.w), ctypes.byref(a), s0, s1, P(par, u32p),
                                  P(cur, f64p), nthreads)
            return par, cur
        tl = np.full((self.nsamp, ncycles + 1), np.nan) if trace else None
        ta = np.full((self.nsamp, max(ncycles, 1)), 255, np.uint8) if trace else None
        L().orc_mh_window(ctypes.byref(self.w), ctypes.byref(a), s0, s1, P(par, u32p),
                          P(cur, f64p), P(tl, f64p) if trace else None,
                          P(ta, u8p) if trace else None)
        if trace:
            return par, cur, tl, ta
        return par, cur


def vq_train(gls, n_codes, seed, training_size=0):
    g = np.ascontiguousarray(gls, np.float32).reshape(-1, 3)
    cb = np.zeros((n_codes, 2), np.float32)
    it = np.zeros(1, np.uint32)
    rc = L().orc_vq_train(P(g, f32p), g.shape[0], n_codes, seed, training_size, P(cb, f32p), P(it, u32p))
    return rc, cb, int(it[0])


def vq_find(codebook, RR, Het):
    cb = np.ascontiguousarray(codebook, np.float32)
    return L().orc_vq_find(P(cb, f32p), cb.shape[0], RR, Het)


def vq_codes(gls, codebook):
    g = np.ascontiguousarray(gls, np.float32)
    shape = g.shape[:-1]
    g = g.reshape(-1, 3)
    cb = np.ascontiguousarray(codebook, np.float32)
    out = np.zeros(g.shape[0], np.uint8)
    L().orc_vq_codes(P(g, f32p), g.shape[0], P(cb, f32p), cb.shape[0], P(out, u8p))
    return out.reshape(shape)


def pack_codes(codes, bits, s0=0, s1=None):
    c = np.ascontiguousarray(codes, np.uint8)
    N, M = c.shape
    s1 = N if s1 is None else s1
    per = 32 // bits
    out = np.zeros(((M + per - 1) // per, s1 - s0), np.uint32)
    L().orc_pack_codes(P(c, u8p), ctypes.c_uint(N), ctypes.c_uint(M), ctypes.c_uint(bits),
                       ctypes.c_uint(s0), ctypes.c_uint(s1), P(out, u32p))
    return out


def pack_nonvq(gls, s0=0, s1=None):
    g = np.ascontiguousarray(gls, np.float32)
    N, M, _ = g.shape
    s1 = N if s1 is None else s1
    out = np.zeros(((M + 3) // 4, s1 - s0), np.uint32)
    L().orc_pack_nonvq(P(g, f32p), ctypes.c_uint(N), ctypes.c_uint(M), ctypes.c_uint(s0),
                       ctypes.c_uint(s1), P(out, u32p))
    return out


def mu(hn):
    return L().orc_mu(hn)


def pc(mu_):
    out = np.zeros((4, 4), np.float32)
    L().orc_pc(mu_, P(out, f32p))
    return out


def tran_from_morgans(r, inflation=1.0):
    rr = np.ascontiguousarray(r, np.float64)
    out = np.zeros((rr.shape[0], 3), np.float32)
    L().orc_tran_from_morgans(P(rr, f64p), rr.shape[0], inflation, P(out, f32p))
    return out


def tran_from_positions(pos, hn, density=1.0, inflation=1.0):
    p = np.ascontiguousarray(pos, np.uint32)
    out = np.zeros((p.shape[0], 3), np.float32)
    L().orc_tran_from_positions(P(p, u32p), p.shape[0], hn, density, inflation, P(out, f32p))
    return out


def init_haps(het_aa, seed, window=0):
    """Initial haplotypes [2N][wn] of het_aa [M][N][2]."""
    g = np.ascontiguousarray(het_aa, np.float32)
    M, N, _ = g.shape
    out = np.zeros((2 * N, (M + 63) // 64), np.uint64)
    L().orc_init_haps(P(g, f32p), N, M, seed, window, P(out, u64p))
    return out


def replace(haps, N, M, hsum):
    """Impute::replace for samples [0, N) on hsum [2N][M] (updated in place)."""
    h = np.ascontiguousarray(haps, np.uint64)
    assert hsum.dtype == np.uint32 and hsum.flags["C_CONTIGUOUS"] and hsum.shape == (2 * N, M)
    L().orc_replace(P(h, u64p), N, M, P(hsum, u32p))
    return hsum


def hmm_work(haps, parents, emit, tran, prob, pc_, S, seed, sample, window, gen):
    h = np.ascontiguousarray(haps, np.uint64)
    pr = np.ascontiguousarray(parents, np.uint32)
    e = np.ascontiguousarray(emit, np.float32)
    t = np.ascontiguousarray(tran, np.float32)
    p = np.ascontiguousarray(prob, np.float32)
    c = np.ascontiguousarray(pc_, np.float32).reshape(16)
    wn = h.shape[1]
    ha = np.zeros(wn, np.uint64)
    hb = np.zeros(wn, np.uint64)
    L().orc_hmm_work(P(h, u64p), wn, P(pr, u32p), P(e, f32p), P(t, f32p), P(p, f32p), P(c, f32p),
                     S, e.shape[0], seed, sample, window, gen, P(ha, u64p), P(hb, u64p))
    return ha, hb


def hmm_work_weights(haps, parents, emit, tran, prob, pc_, S):
    """Genotype weights c [M][4] and unnormalised state posteriors [M][4]."""
    h = np.ascontiguousarray(haps, np.uint64)
    pr = np.ascontiguousarray(parents, np.uint32)
    e = np.ascontiguousarray(emit, np.float32)
    t = np.ascontiguousarray(tran, np.float32)
    p = np.ascontiguousarray(prob, np.float32)
    c16 = np.ascontiguousarray(pc_, np.float32).reshape(16)
    M = e.shape[0]
    c = np.zeros((M, 4), np.float32)
    post = np.zeros((M, 4), np.float32)
    L().orc_hmm_work_weights(P(h, u64p), h.shape[1], P(pr, u32p), P(e, f32p), P(t, f32p),
                             P(p, f32p), P(c16, f32p), S, M, P(c, f32p), P(post, f32p))
    return c, post


def hmm_work_sample(c, wn, seed, sample, window, gen):
    cc = np.ascontiguousarray(c, np.float32)
    ha = np.zeros(wn, np.uint64)
    hb = np.zeros(wn, np.uint64)
    L().orc_hmm_work_sample(P(cc, f32p), cc.shape[0], wn, seed, sample, window, gen,
                            P(ha, u64p), P(hb, u64p))
    return ha, hb


def prob_emit(het_aa, pc_):
    """Insti::initialize's (prob [M][3], emit [M][4]) from (Het, AA) pairs [M][2]."""
    q = np.ascontiguousarray(het_aa, np.float32)
    c16 = np.ascontiguousarray(pc_, np.float32).reshape(16)
    M = q.shape[0]
    prob = np.zeros((M, 3), np.float32)
    emit = np.zeros((M, 4), np.float32)
    L().orc_prob_emit(P(q, f32p), M, P(c16, f32p), P(prob, f32p), P(emit, f32p))
    return prob, emit


def knn_mask(haps, M, lb=0.0, ub=1.0, using_maf=False):
    """(site mask words, per-site variant allele frequencies)."""
    h = np.ascontiguousarray(haps, np.uint64)
    mask = np.zeros(h.shape[1], np.uint64)
    af = np.zeros(M, np.float64)
    L().orc_knn_mask(P(h, u64p), h.shape[1], h.shape[0], M, lb, ub, int(using_maf),
                     P(mask, u64p), P(af, f64p))
    return mask, af


def knn(haps, M, nsamp, K, lb=0.0, ub=1.0, using_maf=False):
    """Candidate lists [nsamp][2 ceil(K/2)] (K capped at nhaps - 2)."""
    h = np.ascontiguousarray(haps, np.uint64)
    kout = L().orc_knn_k_out(K, h.shape[0])
    out = np.zeros((nsamp, kout), np.uint32)
    L().orc_knn(P(h, u64p), h.shape[1], h.shape[0], M, nsamp, K, lb, ub, int(using_maf),
                P(out, u32p))
    return out


def knn_rows(haps, M, rows, K, lb=0.0, ub=1.0, using_maf=False, nthreads=8):
    """The lists of the given samples only, over nthreads threads."""
    h = np.ascontiguousarray(haps, np.uint64)
    r = np.ascontiguousarray(rows, np.uint32)
    kout = L().orc_knn_k_out(K, h.shape[0])
    out = np.zeros((r.shape[0], kout), np.uint32)
    L().orc_knn_rows(P(h, u64p), h.shape[1], h.shape[0], M, K, lb, ub, int(using_maf),
                     P(r, u32p), r.shape[0], P(out, u32p), nthreads)
    return out
