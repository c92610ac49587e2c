"""Synthetic founder-mosaic haplotypes and genotype likelihoods.

The workload generator of the benchmark configurations: F founder haplotypes
with per-site alt frequency ~ Beta(0.5, 0.5); every sample haplotype is a
mosaic of founders switching at 1 cM/Mb over sites spaced Exp(mean 200 bp),
with 1e-3 allele flips; read depth ~ Poisson(depth), alt-read probability
(eps, 1/2, 1 - eps) per genotype, and GL triples = normalised binomial
likelihoods in fp32 (the reference's own GL simulation is
samples/hapGen/generateGLs.R:27-40).  Deterministic for a given seed.
"""
import numpy as np


def make_dataset(n_samples, n_sites, n_founders=40, depth=4.0, err=0.01, flip=1e-3,
                 mean_spacing=200.0, cm_per_mb=1.0, n_ref_haps=0, seed=0):
    rng = np.random.default_rng(seed)
    M = n_sites
    spacing = np.maximum(1, np.rint(rng.exponential(mean_spacing, M))).astype(np.int64)
    spacing[0] = 0
    pos = 10_000 + np.cumsum(spacing)
    # Morgans between consecutive sites (1 cM/Mb = 1e-8 M/bp)
    r = np.zeros(M)
    r[1:] = np.diff(pos) * (cm_per_mb * 1e-8)
    freq = rng.beta(0.5, 0.5, M)
    founders = rng.random((n_founders, M)) < freq
    n_haps = 2 * n_samples + n_ref_haps
    switch = rng.random((n_haps, M)) < r[None, :]
    seg = np.cumsum(switch, axis=1)
    seg_founder = rng.integers(0, n_founders, size=(n_haps, int(seg.max()) + 1))
    fidx = np.take_along_axis(seg_founder, seg, axis=1)
    haps = founders[fidx, np.arange(M)[None, :]]
    haps ^= rng.random((n_haps, M)) < flip
    truth = haps[: 2 * n_samples]
    g = truth[0::2].astype(np.int8) + truth[1::2].astype(np.int8)
    d = rng.poisson(depth, (n_samples, M))
    p_alt = np.array([err, 0.5, 1.0 - err])
    alt = rng.binomial(d, p_alt[g])
    lp = np.log(p_alt)[None, None, :]
    lq = np.log1p(-p_alt)[None, None, :]
    ll = alt[..., None] * lp + (d - alt)[..., None] * lq
    ll -= ll.max(axis=2, keepdims=True)
    gl = np.exp(ll)
    gl /= gl.sum(axis=2, keepdims=True)
    return {
        "pos": pos,
        "r_morgans": r,
        "truth": truth,                        # [2N][M] bool
        "ref_haps": haps[2 * n_samples:],      # [R][M] bool
        "gls": gl.astype(np.float32),          # [N][M][3] (RR, Het, AA)
    }


def pack_haps(bits):
    """[H][M] bool -> [H][ceil(M/64)] uint64, site s = bit s & 63 of word s >> 6
    (impute.hpp:53-66)."""
    H, M = bits.shape
    wn = (M + 63) // 64
    padded = np.zeros((H, wn * 64), dtype=np.uint8)
    padded[:, :M] = bits
    by = np.packbits(padded.reshape(H, wn * 8, 8), axis=2, bitorder="little").reshape(H, wn * 8)
    return np.ascontiguousarray(by).view("<u8").reshape(H, wn).copy()


def unpack_haps(words, M):
    H, wn = words.shape
    by = np.ascontiguousarray(words).view(np.uint8).reshape(H, wn * 8)
    return np.unpackbits(by, axis=1, bitorder="little")[:, :M].astype(bool)


def initial_haps(gls, seed=0):
    """Haplotypes drawn from the GLs with a flat prior and random phase, as
    Insti::initialize does (insti.cpp:1353-1372)."""
    rng = np.random.default_rng(seed)
    N, M, _ = gls.shape
    u = rng.random((N, M, 3))
    t0, t1, t2 = gls[..., 0], gls[..., 1], gls[..., 2]
    hom_ref = u[..., 0] < t0
    het = ~hom_ref & (u[..., 1] < t1 / np.maximum(t1 + t2, 1e-30))
    hom_alt = ~hom_ref & ~het
    a = hom_alt | (het & (u[..., 2] < 0.5))
    b = hom_alt | (het & ~(u[..., 2] < 0.5))
    h = np.empty((2 * N, M), dtype=bool)
    h[0::2] = a
    h[1::2] = b
    return h
