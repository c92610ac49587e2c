"""Pins the oracle's forward-backward form of the mixed-precision pass
(orc_hmm_like_fb): a forward half over sites 0 .. n-1 and the
emission-weighted backward variable over sites M-1 down to n (n =
ceil(M/2)), joined in FP64 through tran row n.

Checks that do not restate it: the 4^M brute-force path sum (odd and even
M, strong transitions), the FP64 forward pass, the closed forms at r = 0.5,
r = 0 and constant emissions, reversal of the window (the likelihood of the
same data read backwards is the same number), agreement of its forward half
with the forward-only mixed pass, and the underflow guard on runs of
mu^2-emitting sites.  An FP64 statement of the same join with a planted
mistake (join through the identity instead of row n, the backward half's
transitions one row off, the odd window's identity step dropped) misses the
FP64 forward pass by at least ten times the bound the oracle meets."""
import numpy as np
import pytest

from tests.test_oracle_hmm import bits, problem


@pytest.mark.parametrize("M", [1, 2, 3, 4, 5, 6, 7, 8])
def test_fb_brute_force(orc, M):
    for seed in range(5):
        r = np.random.default_rng(seed + 10 * M).uniform(0.0, 0.45, M)
        haps, emit, tran = problem(M, seed=seed, r=r)
        P = np.array([1, 6, 3, 2], np.uint32)
        b = orc.hmm_like_brute(haps, P, emit, tran)
        assert abs(orc.hmm_like_fb(haps, P, emit, tran) - b) <= 1e-6 * abs(b) + 1e-7


@pytest.mark.parametrize("M", [9, 10, 64, 65, 100, 333, 1024, 4095, 4096])
def test_fb_matches_fp64(orc, M):
    r = np.random.default_rng(M).uniform(0.0, 0.3, M)
    haps, emit, tran = problem(M, seed=M, r=r)
    P = np.array([1, 6, 3, 2], np.uint32)
    f64 = orc.hmm_like_f64(haps, P, emit, tran)
    assert abs(orc.hmm_like_fb(haps, P, emit, tran) - f64) <= 1e-7 * abs(f64)


def test_fb_closed_forms(orc):
    # r = 0.5: independent sites; r = 0: the state never changes; constant
    # emission c_s: loglik = sum_s log c_s (doubly stochastic transitions)
    for M in (300, 301):
        haps, emit, tran = problem(M, r=np.full(M, 0.5), seed=3)
        P = np.array([0, 5, 2, 7], np.uint32)
        h = bits(haps, M)
        want = sum(np.log(0.25 * sum(float(emit[s, (h[P[x], s] << 1) | h[P[2 + y], s]])
                                     for x in (0, 1) for y in (0, 1))) for s in range(M))
        assert abs(orc.hmm_like_fb(haps, P, emit, tran) - want) <= 1e-6 * abs(want)

        haps, emit, tran = problem(M, r=np.zeros(M), seed=4)
        P = np.array([3, 1, 4, 6], np.uint32)
        h = bits(haps, M)
        terms = [sum(np.log(float(emit[s, (h[P[x], s] << 1) | h[P[2 + y], s]])) for s in range(M))
                 for x in (0, 1) for y in (0, 1)]
        m = max(terms)
        want = m + np.log(0.25 * sum(np.exp(t - m) for t in terms))
        assert abs(orc.hmm_like_fb(haps, P, emit, tran) - want) <= 1e-6 * abs(want)

        rng = np.random.default_rng(M)
        c = rng.uniform(0.01, 1.0, M).astype(np.float32)
        haps, _, tran = problem(M, seed=6)
        emit = np.repeat(c[:, None], 4, axis=1)
        want = float(np.sum(np.log(c.astype(np.float64))))
        assert abs(orc.hmm_like_fb(haps, P, emit, tran) - want) <= 1e-6 * abs(want)


def reversed_window(haps, emit, tran, M):
    """The same data read from the last site to the first: site j of the
    result is site M-1-j, and the transition into it is the one between
    sites M-j and M-1-j (row M-j of tran)."""
    from glphase_amd.synth import pack_haps
    h = bits(haps, M)[:, ::-1]
    t = np.zeros_like(tran)
    t[1:] = tran[1:][::-1]
    return pack_haps(h.astype(bool)), np.ascontiguousarray(emit[::-1]), t


@pytest.mark.parametrize("M", [2, 3, 7, 64, 65, 301, 1000])
def test_fb_reversal_invariance(orc, M):
    r = np.random.default_rng(M + 1).uniform(0.0, 0.4, M)
    haps, emit, tran = problem(M, seed=M + 2, r=r)
    P = np.array([2, 7, 0, 5], np.uint32)
    a = orc.hmm_like_fb(haps, P, emit, tran)
    hr, er, tr = reversed_window(haps, emit, tran, M)
    b = orc.hmm_like_fb(hr, P, er, tr)
    f = orc.hmm_like_f64(hr, P, er, tr)
    assert abs(a - b) <= 1e-7 * abs(a)
    assert abs(b - f) <= 1e-7 * abs(f)


def test_fb_forward_half_is_the_mixed_pass(orc):
    # with the backward half reduced to identity steps (emission 1 and tran
    # row (1, 0, 0) at every site past n), the join returns the forward
    # half's own sum: the forward-only mixed pass over the first n sites,
    # whose states are the same floats; only the last sum differs (FP64 in
    # the join, three float additions in the mixed pass: <= 3 * 2^-24 of it)
    for M in (64, 65, 200, 201):
        n = (M + 1) // 2
        haps, emit, tran = problem(M, seed=M)
        emit[n:] = 1.0
        tran[n:] = (1.0, 0.0, 0.0)
        P = np.array([1, 6, 3, 2], np.uint32)
        got = orc.hmm_like_fb(haps, P, emit, tran)
        want = orc.hmm_like_mixed(haps, P, emit[:n], tran[:n]) if n % 8 else None
        f64 = orc.hmm_like_f64(haps, P, emit[:n].copy(), tran[:n].copy())
        assert abs(got - f64) <= 1e-7 * abs(f64)
        if want is not None:
            assert abs(got - want) <= 3 * 2.0 ** -24


def fb_fp64(haps, P, emit, tran, M, mistake=None):
    """FP64 statement of the join (normalised halves), with an optional
    planted mistake: 'identity_join', 'backward_row_off_by_one',
    'no_identity_step'."""
    h = bits(haps, M)
    n = (M + 1) // 2
    p = M & 1

    def e_at(s):
        return np.array([emit[s, (h[P[x], s] << 1) | h[P[2 + y], s]]
                         for x in (0, 1) for y in (0, 1)], np.float64)

    def T(row):
        t0, t1, t2 = (float(v) for v in row)
        return np.array([[t0, t1, t1, t2], [t1, t0, t2, t1], [t1, t2, t0, t1], [t2, t1, t1, t0]])

    l = 0.25 * e_at(0)
    logl = 0.0
    for s in range(1, n):
        l = e_at(s) * (T(tran[s]) @ l)
        z = l.sum()
        logl += np.log(z)
        l /= z
    u = np.ones(4)
    logu = 0.0
    steps = n - (0 if mistake == "no_identity_step" else p)
    for j in range(steps):
        s = M - 1 - j
        row = tran[s + 1] if s + 1 < M else (1.0, 0.0, 0.0)
        if mistake == "backward_row_off_by_one" and 0 < s + 1 < M:
            row = tran[s]
        u = e_at(s) * (T(row) @ u)
        z = u.sum()
        logu += np.log(z)
        u /= z
    g = u if mistake == "identity_join" or n >= M else T(tran[n]) @ u
    return logl + logu + np.log(float(l @ g))


@pytest.mark.parametrize("M", [8, 9, 120, 121])
def test_fb_planted_mistakes_fail(orc, M):
    r = np.random.default_rng(M).uniform(0.2, 0.45, M)
    haps, emit, tran = problem(M, seed=M + 5, r=r)
    P = np.array([1, 6, 3, 2], np.uint32)
    f64 = orc.hmm_like_f64(haps, P, emit, tran)
    assert abs(fb_fp64(haps, P, emit, tran, M) - f64) <= 1e-12 * abs(f64)
    assert abs(orc.hmm_like_fb(haps, P, emit, tran) - f64) <= 1e-7 * abs(f64)
    mistakes = ["identity_join", "backward_row_off_by_one"] + (["no_identity_step"] if M & 1 else [])
    # each mistake misses FP64 by at least 10x the oracle's own 1e-7 bound
    # (one transition term of the window is enough for that)
    for m in mistakes:
        assert abs(fb_fp64(haps, P, emit, tran, M, m) - f64) > 1e-6 * abs(f64), m


@pytest.mark.parametrize("hn", [10_000, 40_000])
def test_fb_smallest_emissions_match_fp64(orc, hn):
    # mu^2 runs in both halves: the 8-step pass flags the sum below 2^-99,
    # the result comes from the pass scaled after every step, within 1e-7
    # of FP64
    from tests import helpers
    p = helpers.hom_alt_run_problem(hn, M=300, runs=((20, 12), (150, 9), (262, 30)))
    lut = orc.emit_lut(p["codebook"], p["pc"])
    rng = np.random.default_rng(6)
    for _ in range(10):
        i = int(rng.integers(p["N"]))
        P = rng.choice([h for h in range(p["H"]) if h // 2 != i], 4).astype(np.uint32)
        emit = orc.sample_emit(p["packed"], i, p["M"], p["bits"], lut)
        f64 = orc.hmm_like_f64(p["panel"], P, emit, p["tran"])
        v8, tiny8 = orc.hmm_like_fb_pass(p["panel"], P, emit, p["tran"], 8)
        v1, tiny1 = orc.hmm_like_fb_pass(p["panel"], P, emit, p["tran"], 1)
        assert tiny8 and not tiny1
        got = orc.hmm_like_fb(p["panel"], P, emit, p["tran"])
        assert got == v1
        assert abs(got - f64) <= 1e-7 * abs(f64)


def test_fb_guard_keeps_ordinary_passes(orc):
    haps, emit, tran = problem(2001, seed=7)
    P = np.array([1, 6, 3, 2], np.uint32)
    v8, tiny = orc.hmm_like_fb_pass(haps, P, emit, tran, 8)
    assert not tiny
    assert orc.hmm_like_fb(haps, P, emit, tran) == v8
