"""Shared builders of test problems (host arrays only)."""
import json
import os

import numpy as np

from glphase_amd import host, synth

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def synthetic(N, M, bits=8, seed=0, depth=4.0, training_size=1 << 15, n_ref=0, n_founders=40):
    """Founder-mosaic problem: panel, packed codes, tran, codebook, pc."""
    d = synth.make_dataset(N, M, depth=depth, seed=seed, n_ref_haps=n_ref, n_founders=n_founders)
    hb = synth.initial_haps(d["gls"], seed=seed + 1)
    if n_ref:
        hb = np.concatenate([hb, d["ref_haps"]], axis=0)
    panel = synth.pack_haps(hb)
    tran, pc, mu = host.model_params(2 * N, d["r_morgans"])
    cb, _ = host.vq_train(d["gls"], 1 << bits, key=seed, training_size=training_size)
    codes = host.vq_codes(d["gls"], cb)
    packed = host.pack_codes(codes, bits)
    return dict(panel=panel, packed=packed, tran=tran, codebook=cb, pc=pc, codes=codes,
                gls=d["gls"], truth=d["truth"], bits=bits, N=N, M=M, H=panel.shape[0],
                r_morgans=d["r_morgans"])


def planted(num_samps=2, seed=112):
    """The planted-parent data set of cuda_tests.cpp:224-330 / 500-560: 12
    haplotypes of 1024 sites, some words all ones, and samples whose GLs are
    ALT/ALT over blocks that only a planted pair explains."""
    k = KATS["planted_parents"]
    M, H, word = k["num_sites"], k["num_haps"], 64
    wn = M // word
    panel = np.zeros((H, wn), np.uint64)
    for h, words in k["panel_words_all_ones"].items():
        for w in words:
            panel[int(h), w] = np.uint64(0xFFFFFFFFFFFFFFFF)
    rng = np.random.default_rng(seed)
    gls = rng.random((num_samps, M, 3)).astype(np.float32)
    for s in range(num_samps):
        for w in k["alt_words"][str(s)]:
            gls[s, w * word:(w + 1) * word, 2] = k["high_exp"]
    # mutation matrix and tran of cuda_tests.cpp:237-271
    big = H
    mu = np.float32(0)
    for i in range(1, big):
        mu = np.float32(mu + 1.0 / i)
    mu = np.float32(1) / mu
    rho = np.float32(0.5 * mu * (M - 1) / (M * 1000) / 1)
    mu = np.float32(mu / (np.float32(big) + mu))
    pc = np.zeros((4, 4), np.float32)
    for i in range(4):
        for j in range(4):
            d = i ^ j
            pc[i, j] = (1 - mu) * (1 - mu) if d == 0 else (mu * mu if d == 3 else mu * (1 - mu))
    tran = np.zeros((M, 3), np.float32)
    rt = np.float32(rho * np.float32(1000))
    r = np.float32(rt / (rt + np.float32(4)))
    tran[1:, 0] = (1 - r) * (1 - r)
    tran[1:, 1] = r * (1 - r)
    tran[1:, 2] = r * r
    cb, _ = host.vq_train(gls, 16, key=seed)
    codes = host.vq_codes(gls, cb)
    packed = host.pack_codes(codes, 4)
    return dict(panel=panel, packed=packed, tran=tran, codebook=cb, pc=pc, codes=codes, gls=gls,
                bits=4, N=num_samps, M=M, H=H, expect=k["expect"])


def check_planted(parents, expect, samples):
    for s in samples:
        e = expect[str(s)]
        for pair in (parents[s][0:2], parents[s][2:4]):
            lo, hi = sorted(int(x) for x in pair)
            assert hi == e["other"] and lo in e["pair"], (s, list(parents[s]))


def tiny_emission_problem(N=48, M=301, seed=0, row_scale=(1.0, 1.1, 1.2, 1.3)):
    """A problem whose site sums reach the subnormal range: every emission is
    about 2^-42 (the mutation matrix row idx is row_scale[idx] * 2^-42.9), so
    the sum falls by that factor at each site.  After a rescale (sum near 1)
    two sites of GL code 0 (Het = 1, emission 2c) keep it at or above the
    rescale threshold 2^-84, and a third site of code 2 (AA = 1, emission c)
    takes it below 2^-126."""
    rng = np.random.default_rng(seed)
    bits = 4
    H = 2 * N
    wn = (M + 63) // 64
    panel = rng.integers(0, np.iinfo(np.uint64).max, (H, wn), dtype=np.uint64, endpoint=True)
    if M % 64:
        panel[:, -1] &= np.uint64((1 << (M % 64)) - 1)
    cb = np.zeros((16, 2), np.float32)
    cb[0] = (0.0, 1.0)
    cb[1] = (0.0, 0.95)
    cb[2] = (0.0, 0.0)
    rr = rng.random(13).astype(np.float32)
    cb[3:, 0] = rr
    cb[3:, 1] = (rng.random(13).astype(np.float32) * (np.float32(1) - rr)).astype(np.float32)
    prob = np.full(16, 0.02)
    prob[:3] = (0.5, 0.08, 0.16)
    codes = rng.choice(16, size=(N, M), p=prob / prob.sum()).astype(np.uint8)
    packed = host.pack_codes(codes, bits)
    c = np.float32(2.0 ** -42.9)
    pc = np.array([[np.float32(s) * c] * 4 for s in row_scale], np.float32)
    r = np.full(M, 1e-4)
    r[0] = 0.0
    tran, _, _ = host.model_params(H, r)
    return dict(panel=panel, packed=packed, tran=tran, codebook=cb, pc=pc, codes=codes,
                bits=bits, N=N, M=M, H=H)


def count_subnormal_sums(p, samples, parents, emit_of):
    """Float32 restatement of the site-sum sequence (impute.cpp:339-372
    order), used only to check that a test reaches subnormal sums: the
    number of passes with a rescale whose sum is below 2^-126."""
    f32 = np.float32
    norm = f32(np.finfo(np.float32).tiny) ** f32(2.0 / 3.0)
    hits = 0
    for s, P in zip(samples, parents):
        e = emit_of(int(s))
        rows = [p["panel"][int(h)] for h in P]

        def bit(row, m):
            return int((int(row[m >> 6]) >> (m & 63)) & 1)

        def ix(m):
            f0, f1, m0, m1 = (bit(r_, m) for r_ in rows)
            return (f0 << 1) | m0, (f0 << 1) | m1, (f1 << 1) | m0, (f1 << 1) | m1

        i = ix(0)
        l = [f32(0.25) * e[0][i[k]] for k in range(4)]
        found = False
        for m in range(1, p["M"]):
            t0, t1, t2 = p["tran"][m]
            l00, l01, l10, l11 = l
            b = [l00 * t0 + (l01 + l10) * t1 + l11 * t2, l01 * t0 + (l00 + l11) * t1 + l10 * t2,
                 l10 * t0 + (l00 + l11) * t1 + l01 * t2, l11 * t0 + (l01 + l10) * t1 + l00 * t2]
            i = ix(m)
            l = [f32(b[k] * e[m][i[k]]) for k in range(4)]
            sm = f32(f32(f32(l[0] + l[1]) + l[2]) + l[3])
            if sm < norm:
                if sm < f32(2.0 ** -126):
                    found = True
                sc = f32(1) / sm
                l = [f32(x * sc) for x in l]
        hits += int(found)
    return hits


def hom_alt_run_problem(hn, N=64, M=300, runs=((40, 12), (150, 5), (262, 30)), seed=0):
    """A problem whose emissions reach the smallest value of the reference's
    model: the mutation matrix of a panel of hn haplotypes (mu = 1/H_{hn-1},
    then mu/(hn + mu), insti.cpp:1254-1270) and, over runs of sites (start,
    length), GLs certain of the alternative homozygote (codeword (0, 0):
    AA = 1) while every haplotype of the panel carries the reference allele.
    There every parent state emits pc[0][3] = mu^2 (2^-33.2 at hn = 10,000,
    2^-37.5 at 40,000), so eight such sites take the state sum down by more
    than 2^-199 (from 2^100 below 2^-99).
    Elsewhere codes and haplotypes are random."""
    rng = np.random.default_rng(seed)
    bits = 4
    H = 2 * N
    wn = (M + 63) // 64
    panel = rng.integers(0, np.iinfo(np.uint64).max, (H, wn), dtype=np.uint64, endpoint=True)
    if M % 64:
        panel[:, -1] &= np.uint64((1 << (M % 64)) - 1)
    cb = np.zeros((16, 2), np.float32)
    cb[0] = (0.0, 1.0)
    cb[1] = (0.0, 0.0)  # AA = 1
    cb[2] = (1.0, 0.0)
    rr = rng.random(13).astype(np.float32)
    cb[3:, 0] = rr
    cb[3:, 1] = (rng.random(13).astype(np.float32) * (np.float32(1) - rr)).astype(np.float32)
    codes = rng.integers(0, 16, size=(N, M)).astype(np.uint8)
    bitmask = np.zeros(M, bool)
    for a, n in runs:
        codes[:, a:a + n] = 1
        bitmask[a:a + n] = True
    for m in np.nonzero(bitmask)[0]:
        panel[:, m >> 6] &= ~np.uint64(1 << (int(m) & 63))
    packed = host.pack_codes(codes, bits)
    r = np.full(M, 2e-4)
    r[0] = 0.0
    tran, pc, mu = host.model_params(hn, r)
    return dict(panel=panel, packed=packed, tran=tran, codebook=cb, pc=pc, codes=codes,
                bits=bits, N=N, M=M, H=H, mu=mu)
