"""GPU parity at the shapes of the five benchmark workloads, through the C
ABI, against the CPU oracle:

  1. 100 samples x 1,024 sites, the whole loop (initial haplotypes, 10
     burn-in + 10 sampling generations of MH update, hmm_work, panel swap
     and tallies), C = 2N = 200 proposals, the reference's FP32 arithmetic;
     the oracle builds its own transitions and mutation matrix;
  2. the plain path (one launch per proposal) at 1,000 x 1,024, FP32,
     every proposal's log-likelihood and accept decision;
  3. 2,500 samples over 32 windows of lognormal site counts (median 1,024,
     clipped to 256..4,096), 32 proposals, FP32, autotuned depths, eager and
     HIP-graph runs, sampled chains of every window;
  4. (5,000 x 4,096 x 64, mixed precision: test_gpu_parity.py
     test_large_window_sampled_parity, and its forward-backward form:
     test_gpu_fb.py test_fb_large_window_sampled_parity);
  5. the K-list MH update of 10,000 samples + 20,000 panel haplotypes x
     2,048 sites, K = 200 lists built on the GPU, 64 proposals, mixed
     precision (the forward-only and the forward-backward form), read depth
     8 (confident GLs), every chain.

Parents, accept decisions and FP32 log-likelihoods are compared bit for
bit; mixed-precision log-likelihoods within 1e-12 relative."""
import sys

import numpy as np
import pytest

from glphase_amd import PREC_F32, PREC_MIXED, PREC_MIXED_FB, PROP_KLIST, host, synth
from tests import helpers

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_work(orc, panel, parents, het_aa_nm, tran, pc, S, seed, window, gen):
    N = het_aa_nm.shape[0]
    out = np.zeros((2 * N, panel.shape[1]), np.uint64)
    for i in range(N):
        prob, emit = orc.prob_emit(het_aa_nm[i], pc)
        out[2 * i], out[2 * i + 1] = orc.hmm_work(panel, parents[i], emit, tran, prob, pc, S,
                                                  seed, i, window, gen)
    return out


def test_config1_whole_run(orc):
    say("config 1: 100 x 1024, 10 + 10 generations, C = 200, FP32")
    import torch

    from glphase_amd._lib import check, lib
    from glphase_amd.phase import PhaseRun, genotype_trios
    N, M, C, bn, sn, sa, seed = 100, 1024, 200, 10, 10, 10, 77
    d = synth.make_dataset(N, M, depth=4.0, seed=101)
    het_aa = np.ascontiguousarray(d["gls"][:, :, 1:3].transpose(1, 0, 2))  # [M][N][2]
    het_nm = np.ascontiguousarray(het_aa.transpose(1, 0, 2))               # [N][M][2]
    tran_lib, pc_lib, _ = host.model_params(2 * N, d["r_morgans"])
    run = PhaseRun(het_aa, tran=tran_lib, bits=8, cycles=C, burnin=bn, sampling=sn,
                   sa_burnin=sa, seed=seed, precision=PREC_F32, lanes=0)
    try:
        # the oracle's own parameters (insti.cpp:1254-1315)
        tran_o = orc.tran_from_morgans(d["r_morgans"])
        pc_o = orc.pc(orc.mu(2 * N))
        assert np.array_equal(tran_o, run.tran) and np.array_equal(pc_o, run.pc)
        # initial haplotypes (insti.cpp:1316-1365)
        panel_o = orc.init_haps(het_aa, seed, 0)
        assert (run.haplotypes() == panel_o).all()
        # the oracle codes the GLs with the run's codebook (VQ training parity
        # is tested on its own)
        codes_o = orc.vq_codes(genotype_trios(het_aa), run.codebook)
        packed_o = orc.pack_codes(codes_o, 8)
        lut_o = orc.emit_lut(run.codebook, pc_o)
        hsum_o = np.zeros((2 * N, M), np.uint32)
        stream = run.job.stream_handle()
        for g in range(bn + sn):
            pen = float(np.float32(min((g + 1.0) / sa, 1.0)) ** 2)
            run.job.run_batched(gen=g, pen=pen, seed=seed)
            run.rs.run(pen=pen, seed=seed, gen=g)
            mh = orc.MH(panel_o, packed_o, 8, lut_o, tran_o, N)
            po, _ = mh.run(C, pen, seed, gen=g, precision=PREC_F32, nthreads=8)
            assert (run.job.parents()[0] == po).all(), f"generation {g}: parents differ"
            want = oracle_work(orc, panel_o, po, het_nm, tran_o, pc_o, pen, seed, 0, g)
            assert (run.rs.hnew()[0] == want).all(), f"generation {g}: haplotypes differ"
            run.rs.adopt()
            panel_o = want
            if g >= bn:
                check(lib().glp_replace(run.panel_d.data_ptr(), N, M, run.hsum_d.data_ptr(),
                                        stream), "glp_replace")
                orc.replace(panel_o, N, M, hsum_o)
        torch.cuda.synchronize()
        assert (run.hsum_d.cpu().numpy().view(np.uint32) == hsum_o).all()
        assert (run.haplotypes() == panel_o).all()
    finally:
        run.close()


def test_config2_plain_path(orc):
    say("config 2: plain path 1000 x 1024, FP32, 16 proposals")
    from glphase_amd.engine import Job
    p = helpers.synthetic(1000, 1024, seed=202, depth=4.0)
    C, pen, seed = 16, 1.0, 9090
    win = dict(panel=p["panel"], packed=p["packed"], tran=p["tran"])

    def job():
        return Job([win], p["N"], p["H"], p["codebook"], p["pc"], bits=8, num_cycles=C,
                   precision=PREC_F32, trace=True)

    lut = orc.emit_lut(p["codebook"], p["pc"])
    mh = orc.MH(p["panel"], p["packed"], 8, lut, p["tran"], p["N"])
    po, co, tlo, tao = mh.run(C, pen, seed, gen=5, precision=PREC_F32, trace=True)
    for how in ("plan", "desc", "graph"):
        j = job()
        if how == "plan":
            j.run_plain(gen=5, pen=pen, seed=seed)
        elif how == "desc":
            j.run_plain_desc(gen=5, pen=pen, seed=seed)
        else:
            j.run_plain_graph(gen=5, pen=pen, seed=seed)
        tl, ta = j.trace()
        assert (j.parents()[0] == po).all(), how
        assert (j.like()[0] == co).all(), how
        assert np.array_equal(tl[0], tlo), how  # every proposal's log-likelihood
        assert (ta[0] == tao).all(), how        # every accept decision
        j.close()


def config3_windows(N=2500, W=32, seed=21):
    """32 windows of lognormal site counts cut from one founder-mosaic data
    set of 4,096 sites, each with its own initial panel slice and tran."""
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.round(np.exp(rng.normal(np.log(1024.0), 0.6, W))), 256, 4096).astype(int)
    d = synth.make_dataset(N, 4096, depth=4.0, seed=seed)
    hb = synth.initial_haps(d["gls"], seed=seed + 1)
    cb, _ = host.vq_train(d["gls"], 256, key=seed, training_size=1 << 16, device=True)
    codes = host.vq_codes(d["gls"], cb, device=True)
    _, pc, _ = host.model_params(2 * N, np.zeros(4))
    wins = []
    for M in sizes:
        off = int(rng.integers(0, 4096 - M + 1))
        r = d["r_morgans"][off:off + M].copy()
        r[0] = 0.0
        tran, _, _ = host.model_params(2 * N, r)
        wins.append(dict(panel=synth.pack_haps(hb[:, off:off + M]),
                         packed=host.pack_codes(np.ascontiguousarray(codes[:, off:off + M]), 8),
                         tran=tran))
    return wins, cb, pc, sizes


def test_config3_uneven_windows(orc):
    say("config 3: 2500 samples x 32 uneven windows, 32 proposals, FP32, autotuned")
    from glphase_amd.engine import Job
    N, C, pen, seed = 2500, 32, 0.8, 313
    wins, cb, pc, sizes = config3_windows(N)
    assert sizes.min() >= 256 and sizes.max() <= 4096
    job = Job(wins, N, 2 * N, cb, pc, bits=8, num_cycles=C, precision=PREC_F32, trace=True)
    job.make_plan(0)
    lanes = job.autotune(gen=1, pen=pen, seed=seed)
    say(f"  sites {sizes.min()}..{sizes.max()}, lanes per window {lanes.tolist()}")
    lut = orc.emit_lut(cb, pc)
    rng = np.random.default_rng(3)

    def result():
        par, like = job.parents(), job.like()
        tl, ta = job.trace()
        return par, like, tl, ta

    def check_sampled(res, gen):
        par, like, tl, ta = res
        for w, q in enumerate(wins):
            mh = orc.MH(q["panel"], q["packed"], 8, lut, q["tran"], N)
            pick = {0, N - 1} | set(int(x) for x in rng.choice(N, 10, replace=False))
            for i in sorted(pick):
                po, co, tlo, tao = mh.run(C, pen, seed, gen=gen, window_id=w, trace=True,
                                          precision=PREC_F32, s0=i, s1=i + 1)
                assert (par[w, i] == po[i]).all(), (gen, w, i)
                assert like[w, i] == co[i], (gen, w, i)
                assert np.array_equal(tl[w, i], tlo[i]) and (ta[w, i] == tao[i]).all(), (gen, w, i)

    def same(a, b):
        return all(np.array_equal(x, y, equal_nan=True) for x, y in zip(a, b))

    # generation 1 eagerly, checked against the oracle; the graph replays it
    job.run_batched(gen=1, pen=pen, seed=seed)
    a = result()
    check_sampled(a, 1)
    job.run_batched_graph(gen=1, pen=pen, seed=seed)
    assert same(result(), a)
    # generation 2 through the graph (node arguments rewritten), checked
    # against the oracle; the eager run repeats it
    job.run_batched_graph(gen=2, pen=pen, seed=seed)
    b = result()
    check_sampled(b, 2)
    job.run_batched(gen=2, pen=pen, seed=seed)
    assert same(result(), b)
    job.close()


def test_config5_klist_reference_panel(orc):
    say("config 5: K-list MH, 10000 samples + 20000 panel x 2048, K = 200, mixed and FB")
    import torch

    from glphase_amd.engine import Job, knn_lists
    N, R, M, K, C, pen, seed = 10_000, 20_000, 2048, 200, 64, 1.0, 555
    p = helpers.synthetic(N, M, seed=505, depth=8.0, n_ref=R, training_size=1 << 16)
    assert p["H"] == 2 * N + R
    panel_d = torch.from_numpy(p["panel"].view(np.int64)).cuda()
    kl = knn_lists(panel_d, M, N, K).cpu().numpy().view(np.uint32)
    assert kl.shape == (N, K)
    win = dict(panel=p["panel"], packed=p["packed"], tran=p["tran"])
    lut = orc.emit_lut(p["codebook"], p["pc"])
    mh = orc.MH(p["panel"], p["packed"], 8, lut, p["tran"], N)
    # both mixed forms: the forward-only pass and the forward-backward one
    for prec in (PREC_MIXED, PREC_MIXED_FB):
        say(f"  precision {prec}")
        job = Job([win], N, p["H"], p["codebook"], p["pc"], bits=8, num_cycles=C,
                  prop_mode=PROP_KLIST, klist=kl[None], precision=prec)
        job.make_plan(0)
        job.run_batched(gen=20, pen=pen, seed=seed)
        par, like = job.parents()[0], job.like()[0]
        job.close()
        assert np.isfinite(like).all()
        for i in range(N):
            if i % 1000 == 0:
                assert set(par[i]) <= set(kl[i])
        # every chain (OpenMP over samples), on the GPU-built lists
        po, co = mh.run(C, pen, seed, gen=20, prop_mode=PROP_KLIST, klist=kl, precision=prec,
                        nthreads=8)
        assert (par == po).all()
        assert (np.abs(like - co) <= 1e-12 * np.abs(co)).all()
    # and a sample of the lists themselves against the oracle's
    rng = np.random.default_rng(8)
    rows = np.unique(np.concatenate([[0, N - 1], rng.choice(N, 200, replace=False)])).astype(np.uint32)
    assert (kl[rows] == orc.knn_rows(p["panel"], M, rows, K, nthreads=8)).all()
