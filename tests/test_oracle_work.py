"""The oracle's hmm_work (impute.cpp:380-485) and Insti::initialize's
prob/emit builder (insti.cpp:1345-1378), pinned without the reference:
state posteriors against FP64 brute force over all 4^M state paths, the
genotype weights against closed forms (identity mutation matrix, S = 0),
and the sampling rule at its extremes and quartiles."""
import itertools

import numpy as np
import pytest

from oracle import bind as orc


def _problem(M, seed, H=8):
    rng = np.random.default_rng(seed)
    wn = (M + 63) // 64
    haps = rng.integers(0, 2**63, (H, wn), dtype=np.uint64)
    P = rng.choice(H, 4, replace=False).astype(np.uint32)
    emit = rng.uniform(0.05, 1.0, (M, 4)).astype(np.float32)
    r = rng.uniform(0.001, 0.3, M)
    tran = np.stack([(1 - r) ** 2, r * (1 - r), r * r], 1).astype(np.float32)
    q = rng.dirichlet([1, 1, 1], M).astype(np.float32)
    return haps, P, emit, tran, q


def _alleles(haps, P, M):
    def bit(h, m):
        return int((int(haps[h, m >> 6]) >> (m & 63)) & 1)
    f0, f1, m0, m1 = (int(x) for x in P)
    # state xy uses father hap f_x and mother hap m_y: s_xy = (f_x << 1) | m_y
    return [[(bit([f0, f1][x], m) << 1) | bit([m0, m1][y], m) for x in (0, 1) for y in (0, 1)]
            for m in range(M)]


def brute_posterior(haps, P, emit, tran, M):
    """P(state_m = xy | data) in FP64 by summing every path; states ordered
    00, 01, 10, 11; the transition weight between two states is
    tran[m][Hamming distance of their 2-bit labels]."""
    s = _alleles(haps, P, M)
    post = np.zeros((M, 4))
    for path in itertools.product(range(4), repeat=M):
        w = 0.25 * float(emit[0, s[0][path[0]]])
        for m in range(1, M):
            d = bin(path[m - 1] ^ path[m]).count("1")
            w *= float(tran[m, d]) * float(emit[m, s[m][path[m]]])
        for m in range(M):
            post[m, path[m]] += w
    return post / post.sum(1, keepdims=True)


@pytest.mark.parametrize("M,seed", [(1, 0), (2, 1), (5, 2), (7, 3)])
def test_posterior_matches_brute_force(M, seed):
    haps, P, emit, tran, q = _problem(M, seed)
    pc = np.eye(4, dtype=np.float32)
    prob = np.concatenate([q[:, :1], q[:, 1:2], q[:, 2:3]], 1)
    _, post = orc.hmm_work_weights(haps, P, emit, tran, prob, pc, 1.0)
    got = post / post.sum(1, keepdims=True)
    assert np.allclose(got, brute_posterior(haps, P, emit, tran, M), rtol=2e-5, atol=1e-7)


def test_planted_posterior_mistake_is_caught():
    # a transposed state order (01 <-> 10) must fail the brute-force check
    haps, P, emit, tran, q = _problem(5, 9)
    _, post = orc.hmm_work_weights(haps, P, emit, tran, q, np.eye(4, dtype=np.float32), 1.0)
    got = (post / post.sum(1, keepdims=True))[:, [0, 2, 1, 3]]
    assert not np.allclose(got, brute_posterior(haps, P, emit, tran, 5), rtol=2e-5, atol=1e-7)


def test_identity_mutation_weights():
    # pc = I: the weight of genotype g is prob_g times the posterior mass of
    # the states whose parental alleles equal g (impute.cpp:457-469)
    M = 40
    haps, P, emit, tran, q = _problem(M, 4)
    c, post = orc.hmm_work_weights(haps, P, emit, tran, q, np.eye(4, dtype=np.float32), 1.0)
    s = _alleles(haps, P, M)
    pf = np.stack([q[:, 0], q[:, 1], q[:, 1], q[:, 2]], 1).astype(np.float64)
    want = np.zeros((M, 4))
    for m in range(M):
        for xy in range(4):
            want[m, s[m][xy]] += post[m, xy]
    assert np.allclose(c, pf * want, rtol=1e-6, atol=1e-30)


def test_zero_exponent_gives_equal_weights():
    haps, P, emit, tran, q = _problem(30, 5)
    pc = np.full((4, 4), 0.25, np.float32)
    c, _ = orc.hmm_work_weights(haps, P, emit, tran, q, pc, 0.0)
    assert (c == 1.0).all()


def _uniforms(M, seed, sample, window, gen):
    u = []
    for m in range(M):
        if m % 4 == 0:
            x = orc.draw(seed, 0x80000000 | (m >> 2), sample, window, gen)
        u.append(int(x[m % 4]) / 2.0**32)
    return np.array(u)


def _bits(h, M):
    return np.array([(int(h[m >> 6]) >> (m & 63)) & 1 for m in range(M)])


def test_sampling_rule_quartiles_and_extremes():
    M, seed, sample, window, gen = 200, 77, 3, 1, 2
    wn = (M + 63) // 64
    u = _uniforms(M, seed, sample, window, gen)
    # equal weights: genotype = quartile of u
    ha, hb = orc.hmm_work_sample(np.ones((M, 4), np.float32), wn, seed, sample, window, gen)
    g = 2 * _bits(ha, M) + _bits(hb, M)
    assert (g == np.minimum((u * 4).astype(int), 3)).all()
    # one-hot weights always pick their genotype
    for k in range(4):
        c = np.zeros((M, 4), np.float32)
        c[:, k] = 1
        ha, hb = orc.hmm_work_sample(c, wn, seed, sample, window, gen)
        assert (2 * _bits(ha, M) + _bits(hb, M) == k).all()
    # bits past M stay clear
    ha, _ = orc.hmm_work_sample(np.tile([0, 0, 0, 1], (M, 1)).astype(np.float32), wn, seed,
                                sample, window, gen)
    assert int(ha[-1]) >> (M & 63) == 0


def test_prob_emit_closed_form():
    het_aa = np.array([[0.2, 0.3], [0.9, 0.3], [0.0, 0.0], [-0.1, 1.2]], np.float32)
    pc = np.arange(16, dtype=np.float32).reshape(4, 4) / 16
    prob, emit = orc.prob_emit(het_aa, pc)
    t = np.stack([np.maximum(1 - het_aa[:, 0] - het_aa[:, 1], 0), np.maximum(het_aa[:, 0], 0),
                  np.maximum(het_aa[:, 1], 0)], 1)
    assert np.allclose(prob, t, atol=1e-7)
    g = np.stack([t[:, 0], t[:, 1], t[:, 1], t[:, 2]], 1)
    assert np.allclose(emit, g @ pc.T, rtol=1e-6)
    # the het counts twice: row sums of an all-ones matrix give RR + 2 Het + AA
    _, e1 = orc.prob_emit(het_aa, np.ones((4, 4), np.float32))
    assert np.allclose(e1[:, 0], t[:, 0] + 2 * t[:, 1] + t[:, 2], rtol=1e-6)


def test_hmm_work_is_weights_then_sample():
    M = 100
    haps, P, emit, tran, q = _problem(M, 6)
    pc = np.full((4, 4), 0.01, np.float32) + np.eye(4, dtype=np.float32) * 0.96
    ha, hb = orc.hmm_work(haps, P, emit, tran, q, pc, 0.7, 11, 5, 0, 1)
    c, _ = orc.hmm_work_weights(haps, P, emit, tran, q, pc, 0.7)
    ha2, hb2 = orc.hmm_work_sample(c, haps.shape[1], 11, 5, 0, 1)
    assert (ha == ha2).all() and (hb == hb2).all()
