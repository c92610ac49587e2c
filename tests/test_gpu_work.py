"""GPU parity of the haplotype resampling step (Impute::hmm_work,
impute.cpp:380-485) and of whole generations (MH update + resampling +
panel swap, insti.cpp:1668-1705 with cudaSolve's order, 1433-1463) against
the CPU oracle: sampled haplotypes bit-exact."""
import sys

import numpy as np
import pytest

from glphase_amd import PREC_F32, PREC_MIXED
from tests import helpers

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def het_aa_of(p):
    return np.ascontiguousarray(p["gls"][:, :, 1:3], np.float32)


def oracle_work(orc, panel, parents, het_aa, tran, pc, S, seed, window, gen):
    N, M = het_aa.shape[:2]
    wn = panel.shape[1]
    out = np.zeros((2 * N, wn), np.uint64)
    for i in range(N):
        prob, emit = orc.prob_emit(het_aa[i], pc)
        ha, hb = orc.hmm_work(panel, parents[i], emit, tran, prob, pc, S, seed, i, window, gen)
        out[2 * i], out[2 * i + 1] = ha, hb
    return out


def job_with_parents(p, parents, C=4, precision=PREC_F32):
    from glphase_amd.engine import Job
    win = dict(panel=p["panel"], packed=p["packed"], tran=p["tran"])
    job = Job([win], p["N"], p["H"], p["codebook"], p["pc"], bits=p["bits"], num_cycles=C,
              precision=precision)
    job.parents_d[0].copy_(job.torch.from_numpy(parents.view(np.int32)).to(job.device))
    return job


@pytest.mark.parametrize("N,M,S", [(60, 300, 1.0), (60, 300, 0.36), (33, 64, 1.0), (20, 1, 0.5),
                                   (70, 1000, 0.81)])
def test_hmm_work_matches_oracle(orc, N, M, S):
    say(f"hmm_work N={N} M={M} S={S}")
    from glphase_amd.engine import Resampler
    p = helpers.synthetic(N, M, seed=M + N)
    rng = np.random.default_rng(N * M)
    parents = rng.integers(0, p["H"], (N, 4)).astype(np.uint32)
    job = job_with_parents(p, parents)
    rs = Resampler(job, [het_aa_of(p)], p["pc"])
    seed, gen = 4242, 3
    rs.run(pen=S, seed=seed, gen=gen)
    got = rs.hnew()[0]
    want = oracle_work(orc, p["panel"], parents, het_aa_of(p), p["tran"], p["pc"], S, seed, 0, gen)
    assert (got == want).all()
    if M % 64:
        assert (got[:, -1] >> np.uint64(M % 64) == 0).all()


@pytest.mark.parametrize("S", [1.0, 0.49])
def test_hmm_work_subnormal_sums(orc, S):
    # mutation-matrix entries near 2^-127.7 make every emission, and so most
    # site sums of both recursions, subnormal or just above: the short
    # reciprocal gives NaN there, the wavefront redoes the recursion with the
    # IEEE division, and the sampled haplotypes still equal the oracle's
    say(f"hmm_work with subnormal site sums S={S}")
    from glphase_amd.engine import Resampler
    p = helpers.synthetic(48, 300, seed=17)
    c = np.float32(2.0 ** -127.7)
    p["pc"] = np.array([[np.float32(r) * c] * 4 for r in (1.0, 1.1, 1.2, 1.3)], np.float32)
    rng = np.random.default_rng(18)
    parents = rng.integers(0, p["H"], (p["N"], 4)).astype(np.uint32)
    job = job_with_parents(p, parents)
    rs = Resampler(job, [het_aa_of(p)], p["pc"])
    seed, gen = 31, 2
    rs.run(pen=S, seed=seed, gen=gen)
    got = rs.hnew()[0]
    want = oracle_work(orc, p["panel"], parents, het_aa_of(p), p["tran"], p["pc"], S, seed, 0, gen)
    assert (got == want).all()


def test_generations_end_to_end(orc):
    # three generations of batched MH + resampling + panel swap, annealed pen
    say("generations end to end")
    from glphase_amd.engine import Resampler
    N, M, C = 80, 400, 12
    p = helpers.synthetic(N, M, seed=21, depth=2.0)
    job = job_with_parents(p, np.zeros((N, 4), np.uint32), C=C, precision=PREC_MIXED)
    job.make_plan(4)
    rs = Resampler(job, [het_aa_of(p)], p["pc"])
    lut = orc.emit_lut(p["codebook"], p["pc"])
    panel = p["panel"].copy()
    for g in range(3):
        pen = min((g + 1.0) / 2.0, 1.0) ** 2
        job.run_batched(gen=g, pen=pen, seed=17)
        rs.run(pen=pen, seed=99, gen=g)
        got = rs.hnew()[0]
        mh = orc.MH(panel, p["packed"], p["bits"], lut, p["tran"], N)
        po, _ = mh.run(C, pen, 17, gen=g, precision=PREC_MIXED)
        assert (job.parents()[0] == po).all(), f"generation {g}: parents differ"
        want = oracle_work(orc, panel, po, het_aa_of(p), p["tran"], p["pc"], pen, 99, 0, g)
        assert (got == want).all(), f"generation {g}: haplotypes differ"
        panel[: 2 * N] = want
        rs.adopt()


def test_hmm_work_full_size_sampled(orc):
    # 5,000 samples x 4,096 sites: a sample of rows against the oracle
    say("hmm_work 5000 x 4096")
    from glphase_amd.engine import Resampler
    N, M = 5000, 4096
    p = helpers.synthetic(N, M, seed=5, depth=2.0, training_size=1 << 16)
    rng = np.random.default_rng(1)
    parents = rng.integers(0, p["H"], (N, 4)).astype(np.uint32)
    job = job_with_parents(p, parents)
    rs = Resampler(job, [het_aa_of(p)], p["pc"])
    rs.run(pen=0.64, seed=7, gen=1)
    got = rs.hnew()[0]
    q = het_aa_of(p)
    for i in rng.choice(N, 12, replace=False):
        prob, emit = orc.prob_emit(q[i], p["pc"])
        ha, hb = orc.hmm_work(p["panel"], parents[i], emit, p["tran"], prob, p["pc"], 0.64, 7,
                              int(i), 0, 1)
        assert (got[2 * i] == ha).all() and (got[2 * i + 1] == hb).all(), f"sample {i}"
