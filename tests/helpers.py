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
