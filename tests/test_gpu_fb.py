"""GPU parity of the forward-backward form (GLP_PREC_MIXED_FB): every pass
runs as a forward half and a mirrored backward half on one lane quad,
joined in FP64.  Chosen parents and accept decisions must equal the oracle's
(orc_hmm_like_fb fed the same Philox streams) exactly; log-likelihoods agree
within 1e-12 relative (device vs host libm log).  Site counts cover odd
windows (the backward half's identity first step), the 8-site scaling and
32-site chunk boundaries of each half, and the layout's word edges."""
import sys

import numpy as np
import pytest

from glphase_amd import PREC_MIXED, PREC_MIXED_FB, PROP_KLIST
from tests import helpers
from tests.test_gpu_parity import compare_run, job_for, oracle_for

pytestmark = pytest.mark.gpu

FB = PREC_MIXED_FB


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def strong_tran(p, r=0.3):
    r = np.float32(r)
    p["tran"][1:] = (np.float32(1 - r) * np.float32(1 - r), r * np.float32(1 - r), r * r)


@pytest.mark.parametrize("bits", [4, 8])
@pytest.mark.parametrize("M", [1, 2, 3, 4, 5, 15, 16, 17, 18, 31, 33, 63, 64, 65, 66, 127, 128,
                               129, 130, 257, 301, 1000])
def test_fb_site_counts(orc, bits, M):
    # r = 0.3 so that every transition term, the join's row n included, counts
    say(f"fb site counts bits={bits} M={M}")
    p = helpers.synthetic(40, M, bits=bits, seed=5 * M + bits, training_size=4096)
    strong_tran(p)
    C, pen, seed = 10, 0.7, 23 + M
    job = job_for(p, C, precision=FB, trace=True)
    job.make_plan(3)
    job.run_batched(gen=2, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, FB, gen=2)
    job.run_plain(gen=2, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, FB, gen=2)


@pytest.mark.parametrize("lanes", [1, 2, 3, 5, 8, 16])
def test_fb_batched_lanes(orc, lanes):
    say(f"fb batched lanes={lanes}")
    p = helpers.synthetic(200, 501, seed=lanes + 40, depth=2.0)
    C, pen, seed = 40, 0.5, 99
    job = job_for(p, C, precision=FB, trace=True)
    assert job.make_plan(lanes)[0] == lanes
    job.run_batched(gen=1, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, FB, gen=1)
    ev, useful = job.stats()
    assert useful == p["N"] * (C + 1) and ev >= useful


def test_fb_eight_wave_workgroups_and_graphs(orc):
    say("fb 8-wave workgroups, graphs")
    p = helpers.synthetic(400, 600, seed=77, depth=2.0)
    C, pen, seed = 24, 0.6, 808
    job = job_for(p, C, precision=FB, trace=True)
    job.make_plan(3)
    job.set_waves(8)
    job.run_batched(gen=4, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, FB, gen=4)
    par, like = job.parents(), job.like()
    for _ in range(3):  # replays with eager work between them
        job.run_batched_graph(gen=4, pen=pen, seed=seed)
        assert (job.parents() == par).all() and (job.like() == like).all()
        job.run_plain(gen=4, pen=pen, seed=seed)
        assert (job.parents() == par).all()
        job.run_plain_graph(gen=4, pen=pen, seed=seed)
        assert (job.parents() == par).all()


@pytest.mark.parametrize("lanes", [1, 3, 5, 8, 16])
def test_fb_split_chain_slots(orc, lanes):
    # 8-wave workgroups whose wavefronts 4-7 hold one chain slot fewer: the
    # grid grows to hold every chain, and the chains equal the oracle's,
    # eager and through graph replays with eager work between them
    say(f"fb split chain slots lanes={lanes}")
    p = helpers.synthetic(400, 600, seed=lanes + 500, depth=2.0)
    C, pen, seed = 24, 0.6, 808
    job = job_for(p, C, precision=FB, trace=True)
    job.make_plan(lanes)
    job.set_waves(8)
    job.set_split(True)
    assert job.split()
    s = 16 // lanes
    per_wg = 4 * s + 4 * (s - 1 if s > 1 else s)
    assert job.lanes() == (lanes, -(-400 // per_wg))
    job.run_batched(gen=4, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, FB, gen=4)
    par, like = job.parents(), job.like()
    for _ in range(3):
        job.run_batched_graph(gen=4, pen=pen, seed=seed)
        assert (job.parents() == par).all() and (job.like() == like).all()
        job.run_batched(gen=4, pen=pen, seed=seed)
        assert (job.parents() == par).all()
    ev, useful = job.stats()
    assert useful == p["N"] * (C + 1) and ev >= useful
    # switching the split off again rebuilds the graph with the smaller grid
    job.set_split(False)
    assert job.lanes() == (lanes, -(-(-(-400 // s)) // 8))
    job.run_batched_graph(gen=4, pen=pen, seed=seed)
    assert (job.parents() == par).all() and (job.like() == like).all()


def test_fb_auto_plan_splits_when_deeper(orc):
    # lanes = 0: the deepest depth at one wavefront per SIMD (4-wavefront
    # workgroups), or 8-wavefront workgroups with split chain slots where
    # their deepest depth that fits the device is deeper
    import torch
    say("fb automatic plan shape")
    N, C = 5000, 8
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    q4 = max(Q for Q in range(1, min(16, C + 1) + 1) if -(-N // (16 // Q)) <= 4 * cus)
    fits = [Q for Q in range(1, C + 2)
            if -(-N // (4 * (16 // Q) + 4 * max(1, 16 // Q - 1))) <= cus]
    q8 = max(fits) if fits else 0
    p = helpers.synthetic(N, 40, seed=31, training_size=4096)
    job = job_for(p, C, precision=FB, trace=True)
    lanes, grid = job.make_plan(0)
    if q8 > q4:
        assert (lanes, job.waves(), job.split()) == (q8, 8, True)
    else:
        assert (lanes, job.waves(), job.split()) == (q4, 4, False)
    job.run_batched(gen=1, pen=0.8, seed=17)
    compare_run(orc, p, job, C, 0.8, 17, FB, gen=1)


def test_fb_layout_follows_panel_changes(orc):
    # every run rebuilds the windows' layouts from their current data: a
    # changed panel between runs is seen by the next run, eager and graph
    say("fb layout rebuilt per run")
    p = helpers.synthetic(120, 333, seed=9)
    C, pen, seed = 12, 1.0, 5
    job = job_for(p, C, precision=FB, trace=True)
    job.make_plan(4)
    job.run_batched_graph(gen=0, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, FB)
    rng = np.random.default_rng(3)
    flip = rng.integers(0, 2 ** 63, p["panel"].shape, dtype=np.uint64)
    p["panel"] = p["panel"] ^ (flip & np.uint64(0x0101010101010101))
    job.set_panel(0, p["panel"])
    job.run_batched_graph(gen=1, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, FB, gen=1)
    job.run_batched(gen=2, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, FB, gen=2)


def test_fb_multi_window_uneven_autotune(orc):
    say("fb multi-window, autotune")
    from glphase_amd import host
    from glphase_amd.engine import Job
    rng = np.random.default_rng(11)
    N = 64
    sizes = [int(x) for x in rng.integers(40, 400, 6)] + [129, 1]
    probs = [helpers.synthetic(N, M, seed=300 + w) for w, M in enumerate(sizes)]
    cb, pc = probs[0]["codebook"], probs[0]["pc"]
    for q in probs:
        q["codebook"], q["pc"] = cb, pc
        q["packed"] = host.pack_codes(host.vq_codes(q["gls"], cb), 8)
    wins = [dict(panel=q["panel"], packed=q["packed"], tran=q["tran"]) for q in probs]
    C, pen, seed = 20, 0.8, 42
    job = Job(wins, N, probs[0]["H"], cb, pc, bits=8, num_cycles=C, precision=FB, trace=True)
    job.make_plan(0)
    lanes = job.autotune(gen=2, pen=pen, seed=seed)
    assert all(1 <= q <= 16 for q in lanes)
    job.run_batched(gen=2, pen=pen, seed=seed)
    par, like = job.parents(), job.like()
    tl, ta = job.trace()
    for w, q in enumerate(probs):
        mh, _ = oracle_for(orc, q)
        po, co, tlo, tao = mh.run(C, pen, seed, gen=2, window_id=w, precision=FB, trace=True)
        assert (par[w] == po).all()
        assert np.allclose(like[w], co, rtol=1e-12, atol=0)
        assert np.allclose(tl[w], tlo, rtol=1e-12, atol=0) and (ta[w] == tao).all()


def test_fb_klist_and_kickstart(orc):
    say("fb klist + kickstart")
    from glphase_amd.engine import Job
    N, R, M, C, K = 120, 30, 301, 12, 14
    p = helpers.synthetic(N, M, seed=61, n_ref=R, training_size=4096)
    rng = np.random.default_rng(8)
    kl = np.stack([rng.choice([h for h in range(p["H"]) if h // 2 != i], K, replace=False)
                   for i in range(N)]).astype(np.uint32)
    win = dict(panel=p["panel"], packed=p["packed"], tran=p["tran"])
    job = Job([win], N, p["H"], p["codebook"], p["pc"], bits=8, num_cycles=C, precision=FB,
              trace=True, kickstart=True, prop_mode=PROP_KLIST, klist=kl[None])
    job.make_plan(4)
    job.run_batched(gen=3, pen=0.5, seed=66)
    compare_run(orc, p, job, C, 0.5, 66, FB, gen=3, kickstart=True, prop_mode=1, klist=kl)


@pytest.mark.parametrize("hn", [10_000, 40_000])
def test_fb_smallest_emissions(orc, hn):
    # runs of sites where every state emits mu^2 at hn haplotypes' mutation
    # rate (in both halves of the window): the 8-step scaling of a half
    # falls below 2^-99, the wavefront redoes its passes scaled every step,
    # and the result equals the oracle's, finite
    say(f"fb smallest emissions hn={hn}")
    p = helpers.hom_alt_run_problem(hn, seed=hn % 89, M=300, runs=((20, 12), (150, 9), (262, 30)))
    C, pen, seed = 20, 1.0, 4243
    job = job_for(p, C, precision=FB, trace=True)
    job.make_plan(4)
    job.run_batched(gen=1, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, FB, gen=1)
    tl, _ = job.trace()
    assert np.isfinite(tl[0]).all()
    job.run_plain(gen=1, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, FB, gen=1)


def test_fb_large_window_sampled_parity(orc):
    # 5,000 x 4,096 x 64: every chain's parents bit-exact against the FB
    # oracle; sampled returned log-likelihoods within 1e-5 of FP64 and of
    # the forward-only mixed form
    say("fb large window")
    p = helpers.synthetic(5000, 4096, seed=4, depth=2.0, training_size=1 << 16)
    C, pen, seed = 64, 1.0, 31337
    job = job_for(p, C, precision=FB)
    job.make_plan(0)
    job.run_batched(gen=0, pen=pen, seed=seed)
    par, like = job.parents()[0], job.like()[0]
    mh, lut = oracle_for(orc, p)
    po, co = mh.run(C, pen, seed, precision=FB, nthreads=8)
    assert (par == po).all()
    assert (np.abs(like - co) <= 1e-12 * np.abs(co)).all()
    rng = np.random.default_rng(0)
    for i in sorted(rng.choice(p["N"], 24, replace=False)):
        emit = orc.sample_emit(p["packed"], int(i), p["M"], 8, lut)
        f64 = orc.hmm_like_f64(p["panel"], par[i], emit, p["tran"])
        assert abs(like[i] - f64) <= 1e-5 * abs(f64)
        mixed = orc.hmm_like_mixed(p["panel"], par[i], emit, p["tran"])
        assert abs(like[i] - mixed) <= 1e-5 * abs(mixed)


def test_fb_refusals():
    # the allocation-free entry point and the stateless plain launches need
    # no plan, so they cannot hold the layouts: both refuse the form; a plan
    # refuses more than 16 quads per chain
    say("fb refusals")
    from glphase_amd import GlpError
    p = helpers.synthetic(20, 64, seed=1, training_size=4096)
    job = job_for(p, 4, precision=FB)
    with pytest.raises(GlpError, match="forward-backward"):
        job.run_noalloc(gen=0)
    with pytest.raises(GlpError, match="forward-backward"):
        job.run_plain_desc(gen=0)
    with pytest.raises(GlpError, match="1..16"):
        job.make_plan(17)
    job.make_plan(16)
    job.run_batched(gen=0)
    job2 = job_for(p, 4, precision=PREC_MIXED)
    job2.make_plan(32)
