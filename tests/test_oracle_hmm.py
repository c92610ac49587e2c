"""Pins the oracle's forward pass (Impute::hmm_like, impute.cpp:327-375) with
closed forms, a 4^M brute-force path sum, the FP64 forward oracle and the
model's symmetries.  A dropped term, a wrong transition weight or a swapped
parent index fails at least one of these."""
import numpy as np
import pytest

from glphase_amd.synth import pack_haps


def problem(M, H=8, seed=0, r=None, emit=None):
    rng = np.random.default_rng(seed)
    haps = pack_haps(rng.random((H, M)) < 0.5)
    if emit is None:
        emit = rng.uniform(0.05, 1.0, (M, 4)).astype(np.float32)
    if r is None:
        r = rng.uniform(0, 0.3, M)
    tran = np.zeros((M, 3), np.float32)
    rr = np.float32(r)
    tran[:, 0] = (1 - rr) * (1 - rr)
    tran[:, 1] = rr * (1 - rr)
    tran[:, 2] = rr * rr
    tran[0] = 0
    return haps, emit, tran


def bits(haps, M):
    from glphase_amd.synth import unpack_haps
    return unpack_haps(haps, M).astype(int)


@pytest.mark.parametrize("M", [1, 2, 3, 5, 8])
def test_brute_force(orc, M):
    for seed in range(4):
        haps, emit, tran = problem(M, seed=seed)
        P = np.array([1, 6, 3, 2], np.uint32)
        b = orc.hmm_like_brute(haps, P, emit, tran)
        assert abs(orc.hmm_like_f64(haps, P, emit, tran) - b) <= 1e-12 * abs(b) + 1e-14
        assert abs(orc.hmm_like_f32(haps, P, emit, tran) - b) <= 1e-5 * abs(b) + 1e-6
        assert abs(orc.hmm_like_mixed(haps, P, emit, tran) - b) <= 1e-5 * abs(b) + 1e-6


def test_half_recombination_closed_form(orc):
    # r = 0.5 -> tran = (1/4, 1/4, 1/4): sites are independent and
    # loglik = sum_s log(1/4 sum_xy e_s[idx_xy(s)])
    M = 300
    haps, emit, tran = problem(M, r=np.full(M, 0.5), seed=3)
    P = np.array([0, 5, 2, 7], np.uint32)
    h = bits(haps, M)
    want = 0.0
    for s in range(M):
        tot = 0.0
        for x in (0, 1):
            for y in (0, 1):
                tot += emit[s, (h[P[x], s] << 1) | h[P[2 + y], s]]
        want += np.log(0.25 * tot)
    got = orc.hmm_like_f32(haps, P, emit, tran)
    assert abs(got - want) <= 1e-5 * abs(want)


def test_no_recombination_closed_form(orc):
    # r = 0 -> the state never changes: loglik = log(1/4 sum_xy prod_s e_s)
    M = 200
    haps, emit, tran = problem(M, r=np.zeros(M), seed=4)
    P = np.array([3, 1, 4, 6], np.uint32)
    h = bits(haps, M)
    terms = []
    for x in (0, 1):
        for y in (0, 1):
            terms.append(sum(np.log(float(emit[s, (h[P[x], s] << 1) | h[P[2 + y], s]])) for s in range(M)))
    m = max(terms)
    want = m + np.log(0.25 * sum(np.exp(t - m) for t in terms))
    got = orc.hmm_like_f32(haps, P, emit, tran)
    assert abs(got - want) <= 1e-5 * abs(want)


def test_constant_emission_closed_form(orc):
    # e_s[idx] = c_s for every idx: the transition matrix is doubly stochastic
    # (t0 + 2 t1 + t2 = 1), so loglik = sum_s log c_s
    M = 2000
    rng = np.random.default_rng(9)
    c = rng.uniform(0.01, 0.9, M).astype(np.float32)
    haps, _, tran = problem(M, seed=9)
    emit = np.repeat(c[:, None], 4, axis=1)
    P = np.array([0, 1, 2, 3], np.uint32)
    want = float(np.sum(np.log(c.astype(np.float64))))
    got = orc.hmm_like_f32(haps, P, emit, tran)
    assert abs(got - want) <= 2e-5 * abs(want)
    assert abs(orc.hmm_like_f64(haps, P, emit, tran) - want) <= 1e-6 * abs(want)


def test_rescaling_path_matches_fp64(orc):
    # emissions ~1e-3 force the rescale branch (sum < FLT_MIN^(2/3)) every
    # ~8 sites; FP32 must stay within 1e-5 of FP64
    M = 4096
    rng = np.random.default_rng(1)
    emit = rng.uniform(1e-4, 2e-3, (M, 4)).astype(np.float32)
    haps, _, tran = problem(M, seed=1, emit=emit)
    P = np.array([2, 4, 6, 0], np.uint32)
    f64 = orc.hmm_like_f64(haps, P, emit, tran)
    assert abs(orc.hmm_like_f32(haps, P, emit, tran) - f64) <= 1e-5 * abs(f64)
    assert abs(orc.hmm_like_mixed(haps, P, emit, tran) - f64) <= 1e-7 * abs(f64)


def test_parent_swap_invariance(orc):
    # f0 <-> f1 and m0 <-> m1 relabel the hidden states only
    M = 1000
    haps, emit, tran = problem(M, seed=2)
    P = np.array([1, 2, 3, 4], np.uint32)
    base = orc.hmm_like_f64(haps, P, emit, tran)
    for Q in ([2, 1, 3, 4], [1, 2, 4, 3], [2, 1, 4, 3]):
        assert abs(orc.hmm_like_f64(haps, np.array(Q, np.uint32), emit, tran) - base) <= 1e-10 * abs(base)
        assert abs(orc.hmm_like_f32(haps, np.array(Q, np.uint32), emit, tran) - base) <= 1e-5 * abs(base)
    # swapping the father and mother pairs is a symmetry only when
    # emit[1] == emit[2] (the reference pc makes them equal up to rounding)
    e2 = emit.copy()
    e2[:, 2] = e2[:, 1]
    b2 = orc.hmm_like_f64(haps, P, e2, tran)
    assert abs(orc.hmm_like_f64(haps, np.array([3, 4, 1, 2], np.uint32), e2, tran) - b2) <= 1e-10 * abs(b2)


def test_loglik_not_positive(orc):
    M = 500
    for seed in range(5):
        haps, emit, tran = problem(M, seed=seed)
        assert orc.hmm_like_f32(haps, np.array([0, 1, 2, 3], np.uint32), emit, tran) <= 0


def test_planted_mistakes_fail(orc):
    # the checks above detect plausible mistakes: a transposed transition
    # weight or a swapped parent changes the brute-force answer
    M = 6
    haps, emit, tran = problem(M, seed=11)
    P = np.array([1, 6, 3, 2], np.uint32)
    b = orc.hmm_like_brute(haps, P, emit, tran)
    bad = tran[:, [0, 2, 1]].copy()
    assert abs(orc.hmm_like_f32(haps, P, emit, bad) - b) > 1e-4 * abs(b)
    assert abs(orc.hmm_like_f32(haps, np.array([1, 3, 6, 2], np.uint32), emit, tran) - b) > 1e-4 * abs(b)


def test_sample_emit_decodes_packed_codes(orc):
    rng = np.random.default_rng(0)
    N, M, bits_ = 3, 70, 8
    codes = rng.integers(0, 256, (N, M)).astype(np.uint8)
    packed = orc.pack_codes(codes, bits_)
    lut = rng.random((256, 4)).astype(np.float32)
    for i in range(N):
        e = orc.sample_emit(packed, i, M, bits_, lut)
        assert (e == lut[codes[i]]).all()


@pytest.mark.parametrize("hn", [10_000, 40_000])
def test_mixed_smallest_emissions_match_fp64(orc, hn):
    # runs of sites where every parent state emits mu^2 (a certain
    # alternative homozygote against all-reference parents) take the sum
    # from 2^100 below 2^-99 within 8 sites at these panel sizes (mu^2 =
    # 2^-33.2 and 2^-37.5); the 8-site pass flags it and the mixed result
    # comes from the pass scaled after every site, within 1e-7 of FP64
    from tests import helpers
    p = helpers.hom_alt_run_problem(hn)
    lut = orc.emit_lut(p["codebook"], p["pc"])
    mu2 = np.float32(p["mu"]) * np.float32(p["mu"])
    assert lut[1, 0] == pytest.approx(mu2, rel=1e-6)  # pc[0][3] * AA
    rng = np.random.default_rng(5)
    worst_unguarded = 0.0
    for _ in range(12):
        i = int(rng.integers(p["N"]))
        P = rng.choice([h for h in range(p["H"]) if h // 2 != i], 4).astype(np.uint32)
        emit = orc.sample_emit(p["packed"], i, p["M"], p["bits"], lut)
        f64 = orc.hmm_like_f64(p["panel"], P, emit, p["tran"])
        v8, tiny8 = orc.hmm_like_mixed_pass(p["panel"], P, emit, p["tran"], 8)
        v1, tiny1 = orc.hmm_like_mixed_pass(p["panel"], P, emit, p["tran"], 1)
        assert tiny8 and not tiny1
        got = orc.hmm_like_mixed(p["panel"], P, emit, p["tran"])
        assert got == v1
        assert abs(got - f64) <= 1e-7 * abs(f64)
        worst_unguarded = max(worst_unguarded, abs(v8 - f64) / abs(f64) if np.isfinite(v8) else np.inf)
    # without the guard the 8-site pass is off by more than the 1e-5 bound
    assert worst_unguarded > 1e-5


def test_mixed_guard_keeps_ordinary_passes(orc):
    # a pass whose sums stay above 2^-99 is the 8-site pass itself
    haps, emit, tran = problem(2000, seed=7)
    P = np.array([1, 6, 3, 2], np.uint32)
    v8, tiny = orc.hmm_like_mixed_pass(haps, P, emit, tran, 8)
    assert not tiny
    assert orc.hmm_like_mixed(haps, P, emit, tran) == v8
    v1, _ = orc.hmm_like_mixed_pass(haps, P, emit, tran, 1)
    f64 = orc.hmm_like_f64(haps, P, emit, tran)
    assert abs(v1 - f64) <= 1e-7 * abs(f64)
