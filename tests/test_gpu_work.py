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


def groups(M):
    return (M + 3) // 4


def posteriors(rs, w):
    """Window w's posteriors l * beta [N][M][4] from the resampling step's
    workspace: per window a beta table, then a scaled-l table, each
    [ceil(M / 4)][N][4 states][4 sites] floats, windows in order."""
    job = rs.job
    N = job.N
    off = sum(2 * groups(job.M[v]) * N * 16 for v in range(w))
    M = job.M[w]
    G = groups(M)
    ws = rs.ws.cpu().numpy().view(np.float32)
    beta = ws[off:off + G * N * 16].reshape(G, N, 4, 4)
    lf = ws[off + G * N * 16:off + 2 * G * N * 16].reshape(G, N, 4, 4)
    # [g][i][state][site j] -> [i][4 g + j][state]
    return (lf * beta).transpose(1, 0, 3, 2).reshape(N, 4 * G, 4)[:, :M, :]


def check_tables(orc, rs, w, p, parents, pc):
    q = het_aa_of(p)
    got = posteriors(rs, w)
    for i in range(p["N"]):
        prob, emit = orc.prob_emit(q[i], pc)
        _, post = orc.hmm_work_weights(p["panel"], parents[i], emit, p["tran"], prob, pc, 1.0)
        assert (got[i].view(np.uint32) == post.view(np.uint32)).all(), f"window {w} sample {i}"


@pytest.mark.parametrize("N,M", [(17, 1), (17, 2), (70, 31), (70, 32), (70, 33), (130, 64),
                                 (33, 255), (33, 256), (33, 257), (65, 511), (64, 513),
                                 (20, 1025), (9, 2100), (1, 100), (1, 33), (2, 64), (5, 8199)])
def test_hmm_work_tables_match_oracle(orc, N, M):
    # every stored posterior l * beta bit-equal to the oracle's, at site
    # counts on either side of the 32-site blocks and 256-site transition
    # chunks, and sample counts that leave workgroups part-filled
    say(f"hmm_work tables N={N} M={M}")
    from glphase_amd.engine import Resampler
    p = helpers.synthetic(N, M, seed=3 * M + N)
    rng = np.random.default_rng(M + 7 * N)
    parents = rng.integers(0, p["H"], (N, 4)).astype(np.uint32)
    job = job_with_parents(p, parents)
    rs = Resampler(job, [het_aa_of(p)], p["pc"])
    rs.run(pen=1.0, seed=1, gen=0)
    rs.hnew()
    check_tables(orc, rs, 0, p, parents, p["pc"])


def test_hmm_work_tables_multi_window(orc):
    say("hmm_work tables, several windows")
    from glphase_amd.engine import Job, Resampler
    N = 40
    sizes = [33, 700, 1, 256, 300]
    probs = [helpers.synthetic(N, M, seed=50 + w) for w, M in enumerate(sizes)]
    cb, pc = probs[0]["codebook"], probs[0]["pc"]
    for q in probs:
        q["codebook"], q["pc"] = cb, pc
    wins = [dict(panel=q["panel"], packed=q["packed"], tran=q["tran"]) for q in probs]
    job = Job(wins, N, probs[0]["H"], cb, pc, bits=probs[0]["bits"], num_cycles=4,
              precision=PREC_F32)
    rng = np.random.default_rng(5)
    parents = []
    for w, q in enumerate(probs):
        par = rng.integers(0, q["H"], (N, 4)).astype(np.uint32)
        job.parents_d[w].copy_(job.torch.from_numpy(par.view(np.int32)).to(job.device))
        parents.append(par)
    rs = Resampler(job, [het_aa_of(q) for q in probs], pc)
    rs.run(pen=0.7, seed=12, gen=4)
    job.torch.cuda.synchronize(job.device)
    for w, q in enumerate(probs):
        check_tables(orc, rs, w, q, parents[w], pc)
        got = rs.hnew_d[w].cpu().numpy().view(np.uint64)
        want = oracle_work(orc, q["panel"], parents[w], het_aa_of(q), q["tran"], pc, 0.7, 12, w, 4)
        assert (got == want).all(), f"window {w}"


def test_emission_table_matches_oracle(orc):
    # glp_work_emit's table equals Insti::initialize's emit arrays (the
    # oracle's prob_emit) bit for bit, including GL pairs whose genotype
    # probabilities clamp at 0 (negative entries, Het + AA > 1)
    say("emission table")
    from glphase_amd.engine import Resampler
    N, M = 37, 301
    p = helpers.synthetic(N, M, seed=8)
    q = het_aa_of(p).copy()
    rng = np.random.default_rng(2)
    q[rng.random(q.shape) < 0.05] = -0.25
    q[:, ::17, :] = (0.7, 0.6)
    job = job_with_parents(p, np.zeros((N, 4), np.uint32))
    rs = Resampler(job, [q], p["pc"])
    got = rs.emit_d[0].cpu().numpy()
    for i in range(N):
        _, emit = orc.prob_emit(q[i], p["pc"])
        assert (got[:, i, :].view(np.uint32) == emit.view(np.uint32)).all(), f"sample {i}"


def test_hmm_work_without_emission_table(orc):
    # a window given no table (emit NULL): glp_hmm_work builds it in the
    # workspace, with the same tables and haplotypes as the prebuilt one
    say("hmm_work, emission table built per call")
    from glphase_amd.engine import Resampler
    N, M = 45, 700
    p = helpers.synthetic(N, M, seed=31)
    rng = np.random.default_rng(32)
    parents = rng.integers(0, p["H"], (N, 4)).astype(np.uint32)
    job = job_with_parents(p, parents)
    outs = []
    for table in (True, False):
        rs = Resampler(job, [het_aa_of(p)], p["pc"], emit_table=table)
        assert rs.ws.numel() == 2 * groups(M) * N * 64 + (0 if table else M * N * 16)
        rs.run(pen=0.8, seed=4, gen=2)
        outs.append((rs.hnew()[0], posteriors(rs, 0)))
        check_tables(orc, rs, 0, p, parents, p["pc"])
    assert (outs[0][0] == outs[1][0]).all()
    assert (outs[0][1].view(np.uint32) == outs[1][1].view(np.uint32)).all()


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
    check_tables(orc, rs, 0, p, parents, p["pc"])


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
