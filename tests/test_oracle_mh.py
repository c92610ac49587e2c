"""Pins the oracle's MH loop (Insti::solve, insti.cpp:1478-1551, with
MHSampler::SampleHap, MHSampler.hpp:82-96) and checks the library's host
helpers (parameters, VQ, packing) against the oracle bit for bit."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from glphase_amd import host
from tests import helpers


def u01(x):
    return float(x) / 4294967296.0


@pytest.mark.parametrize("pen", [1.0, 0.04])
def test_trace_replays_accept_rule(orc, pen):
    p = helpers.synthetic(40, 300, seed=3)
    lut = orc.emit_lut(p["codebook"], p["pc"])
    mh = orc.MH(p["panel"], p["packed"], 8, lut, p["tran"], p["N"])
    C, seed = 30, 77
    par, cur, tl, ta = mh.run(C, pen, seed, trace=True)
    for i in range(p["N"]):
        curr = tl[i, 0]
        for k in range(C):
            x = orc.draw(seed, k, i, 0, 0)
            prop = tl[i, 1 + k]
            acc = prop >= curr or u01(x[2]) < math.exp((prop - curr) * float(np.float32(pen)))
            assert bool(ta[i, k]) == acc
            if prop >= curr:
                assert ta[i, k] == 1  # better or equal is always accepted
            if acc:
                curr = prop
        assert cur[i] == curr
        assert curr <= 0


def test_fresh_parents_and_exclusion(orc):
    p = helpers.synthetic(30, 100, seed=4)
    lut = orc.emit_lut(p["codebook"], p["pc"])
    mh = orc.MH(p["panel"], p["packed"], 8, lut, p["tran"], p["N"])
    par, _ = mh.run(20, 1.0, 5)
    for i in range(p["N"]):
        assert all(h // 2 != i and h < p["H"] for h in par[i])
    # with no cycles the parents are the fresh draws of counter kInitTag
    par0, cur0 = mh.run(0, 1.0, 5)
    for i in range(p["N"]):
        x = orc.draw(5, 0xFFFFFFFF, i, 0, 0)
        assert [orc.L().orc_unif_excl(int(v), p["H"], i) for v in x] == list(par0[i])


def test_much_worse_rejected_better_accepted(orc):
    # other_unittests.cpp:256-283 in trace form: a proposal at twice the
    # current log-likelihood (-100 vs -200) is accepted with prob e^-100
    assert math.exp(-100.0) < 1e-40
    p = helpers.synthetic(10, 200, seed=6)
    lut = orc.emit_lut(p["codebook"], p["pc"])
    mh = orc.MH(p["panel"], p["packed"], 8, lut, p["tran"], p["N"])
    _, _, tl, ta = mh.run(50, 1.0, 9, trace=True)
    curr = tl[:, 0].copy()
    for k in range(50):
        prop = tl[:, 1 + k]
        much_worse = prop < curr - 60
        assert not ta[much_worse, k].any()
        curr = np.where(ta[:, k] == 1, prop, curr)


def test_planted_parents_found(orc):
    # cuda_tests.cpp:118-656: after enough MH cycles each sample's father and
    # mother pairs are {planted pair member, the single-block haplotype}.  The
    # reference runs 100 cycles under fixed GSL seeds; under Philox streams
    # 100 cycles reach the planted sets in 84% of (seed, sample) cases and
    # 400 cycles in 160 of 160, so the test runs 400.
    for nsamp, samples in ((2, (0, 1)), (4, (2, 3))):
        p = helpers.planted(num_samps=nsamp)
        lut = orc.emit_lut(p["codebook"], p["pc"])
        mh = orc.MH(p["panel"], p["packed"], 4, lut, p["tran"], nsamp)
        par, _ = mh.run(400, 1.0, 2024)
        helpers.check_planted(par, p["expect"], samples)
        # a bad parent set scores below the planted one (cuda_tests.cpp:393-429)
        emit = orc.sample_emit(p["packed"], 0, p["M"], 4, lut)
        bad = orc.hmm_like_f32(p["panel"], np.array([2, 2, 2, 2], np.uint32), emit, p["tran"])
        good = orc.hmm_like_f32(p["panel"], np.array([5, 10, 6, 10], np.uint32), emit, p["tran"])
        assert good > bad


def test_host_mode_uses_given_candidates(orc):
    p = helpers.synthetic(8, 64, seed=8)
    lut = orc.emit_lut(p["codebook"], p["pc"])
    mh = orc.MH(p["panel"], p["packed"], 8, lut, p["tran"], p["N"])
    C = 12
    rng = np.random.default_rng(0)
    init = rng.integers(0, p["H"], (p["N"], 4)).astype(np.uint32)
    extra = rng.integers(0, p["H"], (C, p["N"])).astype(np.uint32)
    par, _, tl, ta = mh.run(C, 1.0, 3, prop_mode=2, init_parents=init, extra=extra, trace=True)
    for i in range(p["N"]):
        allowed = set(init[i]) | set(extra[:, i])
        assert set(par[i]) <= allowed


# ---- library host helpers vs oracle ----------------------------------------

def test_lib_model_params_match_oracle(orc):
    r = np.random.default_rng(1).uniform(0, 0.01, 500)
    r[0] = 0
    tran, pc, mu = host.model_params(2000, r, inflation=1.5)
    assert np.float32(mu) == np.float32(orc.mu(2000))
    assert (pc == orc.pc(np.float32(mu))).all()
    assert (tran == orc.tran_from_morgans(r, 1.5)).all()


@pytest.mark.parametrize("bits,training", [(4, 0), (8, 0), (8, 3000)])
def test_lib_vq_matches_oracle(orc, bits, training):
    p = helpers.synthetic(20, 400, bits=4, seed=2)
    gls = p["gls"]
    cb, it = host.vq_train(gls, 1 << bits, key=99, training_size=training)
    rc, cbo, ito = orc.vq_train(gls, 1 << bits, 99, training)
    assert rc == 0 and it == ito
    assert (cb == cbo).all()
    codes = host.vq_codes(gls, cb)
    assert (codes == orc.vq_codes(gls, cb)).all()
    assert (host.pack_codes(codes, bits) == orc.pack_codes(codes, bits)).all()
    assert (host.pack_codes(codes, bits, 5, 13) == orc.pack_codes(codes, bits, 5, 13)).all()


def test_lib_pack_nonvq_matches_oracle(orc):
    gls = np.random.default_rng(3).random((7, 13, 3)).astype(np.float32)
    assert (host.pack_nonvq(gls) == orc.pack_nonvq(gls)).all()
    assert (host.pack_nonvq(gls, 2, 6) == orc.pack_nonvq(gls, 2, 6)).all()


def test_library_exports_header_symbols():
    from glphase_amd import _lib
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "glphase_amd.h")).read()
    names = set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(glp_[a-z_0-9]+)\s*\(", hdr, re.M))
    names |= set(re.findall(r"\b(glp_[a-z_0-9]+)\(", hdr))
    assert len(names) >= 25
    L = _lib.lib()
    for n in sorted(names):
        assert hasattr(L, n), n
    assert {s[0] for s in _lib.SIGNATURES} == names


def test_library_errors_without_gpu():
    # argument checks answer before any device work
    from glphase_amd import _lib
    L = _lib.lib()
    out = ctypes.c_void_p()
    assert L.glp_plan_create(None, 0, ctypes.byref(out)) == 1
    assert b"windows" in L.glp_last_error()
    assert L.glp_emit_lut_bytes(8, 0) == 256 * 16 * 16
    assert L.glp_emit_lut_bytes(3, 0) == 0


def test_kickstart_needs_reference_haplotypes():
    # main.cpp:430-433: kickstarting without a reference panel is an error,
    # refused by the descriptor check before any device work
    from glphase_amd import _lib
    L = _lib.lib()
    fake = ctypes.c_void_p(4096)
    w = _lib.GlpWindow(64, 1, fake, fake, fake, fake, fake)
    d = _lib.GlpMhDesc()
    d.num_windows, d.windows = 1, ctypes.pointer(w)
    d.num_samples, d.num_haps, d.bits_per_code = 10, 20, 8
    d.emit_lut, d.num_cycles, d.precision = fake, 4, 1
    d.kickstart_from_ref = 1
    out = ctypes.c_void_p()
    assert L.glp_plan_create(ctypes.byref(d), 0, ctypes.byref(out)) == 1
    assert b"kickstart" in L.glp_last_error()
    assert L.glp_plain_init(ctypes.byref(d), None) == 1


def test_plan_rejects_arrays_past_32_bit_offsets():
    # the kernels read a window's panel and packed codes through buffer
    # descriptors with 32-bit byte offsets: 2^18 sites x 2^14 samples at
    # 8-bit codes are 4 GiB of codes, and 2^19 haplotypes of 2^16 sites are
    # 4 GiB of panel; the descriptor check refuses both before any device
    # work (the pointers are never read)
    from glphase_amd import _lib
    L = _lib.lib()
    fake = ctypes.c_void_p(4096)

    def desc(M, N, H):
        w = _lib.GlpWindow(M, (M + 63) // 64, fake, fake, fake, fake, fake)
        d = _lib.GlpMhDesc()
        d.num_windows, d.windows = 1, ctypes.pointer(w)
        d.num_samples, d.num_haps, d.bits_per_code = N, H, 8
        d.emit_lut, d.num_cycles, d.precision = fake, 4, 1
        return d, w

    out = ctypes.c_void_p()
    d, w = desc(1 << 18, 1 << 14, 1 << 15)  # w keeps the window alive
    assert L.glp_plan_create(ctypes.byref(d), 0, ctypes.byref(out)) == 1
    assert b"codes reach 4 GiB" in L.glp_last_error()
    d, w = desc(1 << 16, 1 << 10, 1 << 19)
    assert L.glp_plan_create(ctypes.byref(d), 0, ctypes.byref(out)) == 1
    assert b"panel reaches 4 GiB" in L.glp_last_error()


def test_kickstart_draws_from_reference_haplotypes(orc):
    # insti.cpp:1510-1513: with kickstarting, proposal 0 of a generation is
    # drawn from the reference haplotypes [2N, H) only.  With pen = 0 every
    # proposal is accepted (u < exp(0) = 1), so after one proposal every
    # chain holds a reference haplotype; without kickstarting only the
    # chains whose draws happen to reach the 10 reference haplotypes do
    from tests import helpers
    N, R = 100, 10
    p = helpers.synthetic(N, 64, seed=2, n_ref=R, training_size=4096)
    lut = orc.emit_lut(p["codebook"], p["pc"])
    mh = orc.MH(p["panel"], p["packed"], 8, lut, p["tran"], N)
    par_k, _ = mh.run(1, 0.0, 77, kickstart=True)
    par_n, _ = mh.run(1, 0.0, 77)
    assert (par_k >= 2 * N).any(axis=1).all()
    assert (par_n >= 2 * N).any(axis=1).mean() < 0.5
    # only proposal 0: with k0 = 1 (the generation's second proposal) the
    # flag changes nothing
    a, _ = mh.run(1, 0.0, 77, k0=1, kickstart=True)
    b, _ = mh.run(1, 0.0, 77, k0=1)
    assert (a == b).all()
    # K-lists: proposal 0 takes a list entry among the reference haplotypes;
    # a list with none keeps the usual draw
    rng = np.random.default_rng(3)
    K = 8
    kl = np.stack([rng.choice([h for h in range(2 * N + R) if h // 2 != i], K, replace=False)
                   for i in range(N)]).astype(np.uint32)
    kl[::2, :] = np.minimum(kl[::2, :], 2 * N - 1)  # even samples: no reference entries
    for i in range(0, N, 2):
        kl[i] = rng.choice([h for h in range(2 * N) if h // 2 != i], K, replace=False)
    pk, _ = mh.run(1, 0.0, 77, prop_mode=1, klist=kl, kickstart=True)
    pn, _ = mh.run(1, 0.0, 77, prop_mode=1, klist=kl)
    odd = np.arange(1, N, 2)
    has_ref = (kl[odd] >= 2 * N).any(axis=1)
    assert (pk[odd][has_ref] >= 2 * N).any(axis=1).all()
    assert (pk[0::2] == pn[0::2]).all()
