"""GPU parity: the HIP engine against the CPU oracle on the same seeded
inputs.  Chosen parents, accept decisions and FP32 log-likelihoods must be
identical (bit for bit); FP64-accumulated log-likelihoods within 1e-12
relative (device vs host libm log)."""
import ctypes
import sys

import numpy as np
import pytest

from glphase_amd import PREC_F32, PREC_MIXED, PROP_HOST, PROP_KLIST, PROP_UNIFORM
from tests import helpers

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def job_for(p, C, **kw):
    from glphase_amd.engine import Job
    win = dict(panel=p["panel"], packed=p["packed"], tran=p["tran"])
    return Job([win], p["N"], p["H"], p["codebook"], p["pc"], bits=p["bits"], num_cycles=C, **kw)


def oracle_for(orc, p):
    lut = orc.emit_lut(p["codebook"], p["pc"])
    return orc.MH(p["panel"], p["packed"], p["bits"], lut, p["tran"], p["N"]), lut


@pytest.mark.parametrize("bits", [4, 8])
@pytest.mark.parametrize("M", [1, 2, 31, 32, 33, 64, 65, 100, 257, 1024])
def test_hmm_like_bit_exact(orc, bits, M):
    from glphase_amd import _lib
    from glphase_amd._lib import c_f32p, c_u32p, c_u64p, check, lib, ptr
    say(f"hmm_like bits={bits} M={M}")
    p = helpers.synthetic(16, M, bits=bits, seed=M + bits, training_size=4096)
    L = lib()
    ctx = ctypes.c_void_p()
    check(L.glp_create(0, ctypes.byref(ctx)), "create")
    try:
        check(L.glp_set_tran(ctx, ptr(p["tran"], c_f32p), p["tran"].size), "tran")
        check(L.glp_set_mutation_mat(ctx, ptr(np.ascontiguousarray(p["pc"]), c_f32p)), "pc")
        check(L.glp_set_codebook(ctx, ptr(p["codebook"], c_f32p), 1 << bits), "cb")
        pk = np.ascontiguousarray(p["packed"])
        check(L.glp_set_packed_gls(ctx, ptr(pk, c_u32p), pk.size, bits), "packed")
        rng = np.random.default_rng(M)
        n = 300
        samples = rng.integers(0, p["N"], n).astype(np.uint32)
        parents = rng.integers(0, p["H"], (n, 4)).astype(np.uint32)
        out = np.zeros(n, np.float32)
        panel = np.ascontiguousarray(p["panel"])
        check(L.glp_hmm_like(ctx, ptr(panel, c_u64p), panel.size, M, p["N"], n,
                             ptr(samples, c_u32p), ptr(parents, c_u32p), ptr(out, c_f32p)), "hmm_like")
        lut = orc.emit_lut(p["codebook"], p["pc"])
        for e in range(n):
            emit = orc.sample_emit(p["packed"], int(samples[e]), M, bits, lut)
            want = np.float32(orc.hmm_like_f32(p["panel"], parents[e], emit, p["tran"]))
            assert out[e] == want, (e, out[e], want)
    finally:
        L.glp_destroy(ctx)


def compare_run(orc, p, job, C, pen, seed, precision, gen=0, window=0, **orc_kw):
    mh, _ = oracle_for(orc, p)
    par_o, cur_o, tl_o, ta_o = mh.run(C, pen, seed, gen=gen, window_id=window,
                                      precision=precision, trace=True, **orc_kw)
    par = job.parents()[0]
    like = job.like()[0]
    tl, ta = job.trace()
    assert (par == par_o).all()
    if precision == PREC_F32:
        assert (like == cur_o).all()
        assert np.array_equal(tl[0], tl_o)
    else:
        assert np.allclose(like, cur_o, rtol=1e-12, atol=0)
        assert np.allclose(tl[0], tl_o, rtol=1e-12, atol=0)
    if C:
        assert (ta[0] == ta_o).all()


@pytest.mark.parametrize("prec", [PREC_F32, PREC_MIXED])
def test_plain_path_matches_oracle(orc, prec):
    say(f"plain path prec={prec}")
    p = helpers.synthetic(300, 700, seed=21)
    C, pen, seed = 12, 0.3, 1234
    job = job_for(p, C, precision=prec, trace=True)
    job.run_plain(gen=3, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, prec, gen=3)
    # the stateless per-launch entry points give the same chain
    job2 = job_for(p, C, precision=prec, trace=True)
    job2.run_plain_desc(gen=3, pen=pen, seed=seed)
    assert (job2.parents() == job.parents()).all()


@pytest.mark.parametrize("lanes", [1, 2, 3, 4, 5, 6, 7, 8, 11, 16, 32])
def test_batched_matches_oracle(orc, lanes):
    say(f"batched lanes={lanes}")
    p = helpers.synthetic(200, 500, seed=lanes)
    C, pen, seed = 40, 0.5, 99
    job = job_for(p, C, precision=PREC_F32, trace=True)
    assert job.make_plan(lanes)[0] == lanes
    job.run_batched(gen=1, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, PREC_F32, gen=1)
    ev, useful = job.stats()
    assert useful == p["N"] * (C + 1)
    assert ev >= useful


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("prec,lanes", [(PREC_F32, 1), (PREC_F32, 6), (PREC_MIXED, 3),
                                        (PREC_MIXED, 6)])
def test_batched_eight_wave_workgroups(orc, prec, lanes, split):
    # 8 wavefronts per workgroup sharing the LDS tables (two per SIMD), with
    # and without one chain slot fewer in wavefronts 4-7: the same chains as
    # the plain oracle, eager and through the plan's graph
    say(f"batched, 8-wave workgroups prec={prec} lanes={lanes} split={split}")
    p = helpers.synthetic(400, 600, seed=lanes + 10 * prec, depth=2.0)
    C, pen, seed = 24, 0.6, 808
    job = job_for(p, C, precision=prec, trace=True)
    job.make_plan(lanes)
    job.set_waves(8)
    job.set_split(split)
    assert job.waves() == 8 and job.split() == split
    job.run_batched(gen=4, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, prec, gen=4)
    par, like = job.parents(), job.like()
    job.run_batched_graph(gen=4, pen=pen, seed=seed)
    assert (job.parents() == par).all() and (job.like() == like).all()
    ev, useful = job.stats()
    assert useful == p["N"] * (C + 1)


@pytest.mark.parametrize("mode", ["uniform", "klist"])
@pytest.mark.parametrize("form", ["plain", "batched"])
def test_kickstart_from_reference(orc, mode, form):
    # insti.cpp:1510-1513: proposal 0 of the generation from the reference
    # haplotypes only, uniform over them or over a K-list's entries among
    # them (lists without any keep the usual draw)
    say(f"kickstart mode={mode} form={form}")
    from glphase_amd.engine import Job
    N, R, M, C, K = 120, 30, 300, 10, 12
    p = helpers.synthetic(N, M, seed=31, n_ref=R, training_size=4096)
    kw, okw = {}, {}
    if mode == "klist":
        rng = np.random.default_rng(4)
        kl = np.stack([rng.choice([h for h in range(p["H"]) if h // 2 != i], K, replace=False)
                       for i in range(N)]).astype(np.uint32)
        for i in range(0, N, 3):  # every third list without reference entries
            kl[i] = rng.choice([h for h in range(2 * N) if h // 2 != i], K, replace=False)
        kw = dict(prop_mode=PROP_KLIST, klist=kl[None])
        okw = dict(prop_mode=1, klist=kl)
    win = dict(panel=p["panel"], packed=p["packed"], tran=p["tran"])
    job = Job([win], N, p["H"], p["codebook"], p["pc"], bits=8, num_cycles=C,
              precision=PREC_MIXED, trace=True, kickstart=True, **kw)
    if form == "plain":
        job.run_plain(gen=3, pen=0.5, seed=66)
    else:
        job.make_plan(4)
        job.run_batched(gen=3, pen=0.5, seed=66)
    compare_run(orc, p, job, C, 0.5, 66, PREC_MIXED, gen=3, kickstart=True, **okw)


@pytest.mark.parametrize("lanes", [4, 6, 16])
def test_batched_mixed_precision(orc, lanes):
    say(f"batched mixed lanes={lanes}")
    p = helpers.synthetic(150, 900, seed=7, depth=2.0)
    C, pen, seed = 33, 1.0, 5
    job = job_for(p, C, precision=PREC_MIXED, trace=True)
    job.make_plan(lanes)
    job.run_batched(gen=0, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, PREC_MIXED)


@pytest.mark.parametrize("bits", [4, 8])
@pytest.mark.parametrize("M", [1, 3, 4, 5, 7, 8, 9, 31, 32, 33, 39, 40, 41, 63, 64, 65, 97, 301])
def test_mixed_precision_site_counts(orc, bits, M):
    # the mixed-precision recurrence scales after every 8th site and reads a
    # window's last chunk through clamped loads: site counts around the
    # 8-site and 32-site boundaries, r = 0.3 so that every transition term
    # counts, through the batched and the plain path
    say(f"mixed precision bits={bits} M={M}")
    p = helpers.synthetic(40, M, bits=bits, seed=3 * M + bits, training_size=4096)
    r = np.float32(0.3)
    p["tran"][1:] = (np.float32(1 - r) * np.float32(1 - r), r * np.float32(1 - r), r * r)
    C, pen, seed = 10, 0.7, 17 + M
    job = job_for(p, C, precision=PREC_MIXED, trace=True)
    job.make_plan(3)
    job.run_batched(gen=2, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, PREC_MIXED, gen=2)
    job2 = job_for(p, C, precision=PREC_MIXED, trace=True)
    job2.run_plain(gen=2, pen=pen, seed=seed)
    compare_run(orc, p, job2, C, pen, seed, PREC_MIXED, gen=2)


def test_autotune_keeps_parity_and_inputs(orc):
    say("autotune")
    from glphase_amd.engine import Job
    rng = np.random.default_rng(5)
    N = 48
    sizes = [700, 300, 90]
    probs = [helpers.synthetic(N, M, seed=200 + w) for w, M in enumerate(sizes)]
    cb, pc = probs[0]["codebook"], probs[0]["pc"]
    from glphase_amd import host
    for q in probs:
        q["codebook"], q["pc"] = cb, pc
        q["packed"] = host.pack_codes(host.vq_codes(q["gls"], cb), 8)
    wins = [dict(panel=q["panel"], packed=q["packed"], tran=q["tran"]) for q in probs]
    C = 24
    init = np.stack([rng.integers(0, probs[0]["H"], (N, 4)) for _ in probs]).astype(np.uint32)
    extra = np.stack([rng.integers(0, probs[0]["H"], (C, N)) for _ in probs]).astype(np.uint32)
    job = Job(wins, N, probs[0]["H"], cb, pc, bits=8, num_cycles=C, prop_mode=PROP_HOST,
              init_parents=init, extra=extra, trace=True)
    job.make_plan(0)
    lanes = job.autotune(gen=1, pen=0.5, seed=3)
    assert all(1 <= q <= 32 for q in lanes)
    assert (job.parents() == init).all()  # host-mode inputs restored
    job.run_batched(gen=1, pen=0.5, seed=3)
    par = job.parents()
    for w, q in enumerate(probs):
        mh, _ = oracle_for(orc, q)
        po, _ = mh.run(C, 0.5, 3, gen=1, window_id=w, prop_mode=2, init_parents=init[w],
                       extra=extra[w])
        assert (par[w] == po).all()


def test_batched_zero_cycles(orc):
    say("batched C=0")
    p = helpers.synthetic(50, 64, seed=2)
    job = job_for(p, 0, trace=True)
    job.make_plan(8)
    job.run_batched(gen=0, pen=1.0, seed=3)
    mh, _ = oracle_for(orc, p)
    par_o, cur_o = mh.run(0, 1.0, 3)
    assert (job.parents()[0] == par_o).all() and (job.like()[0] == cur_o).all()


def test_multi_window_uneven(orc):
    say("multi-window")
    from glphase_amd.engine import Job
    rng = np.random.default_rng(0)
    N = 64
    sizes = [int(x) for x in rng.integers(40, 400, 6)]
    probs = [helpers.synthetic(N, M, seed=100 + w) for w, M in enumerate(sizes)]
    # one shared codebook/pc across windows, as one GLPack per run
    cb, pc = probs[0]["codebook"], probs[0]["pc"]
    from glphase_amd import host
    for q in probs:
        q["codebook"], q["pc"] = cb, pc
        q["packed"] = host.pack_codes(host.vq_codes(q["gls"], cb), 8)
    wins = [dict(panel=q["panel"], packed=q["packed"], tran=q["tran"]) for q in probs]
    C, pen, seed = 20, 0.8, 42
    for lanes in (1, 8):
        job = Job(wins, N, probs[0]["H"], cb, pc, bits=8, num_cycles=C, trace=True)
        job.make_plan(lanes)
        job.run_batched(gen=2, pen=pen, seed=seed)
        par, like = job.parents(), job.like()
        tl, ta = job.trace()
        for w, q in enumerate(probs):
            mh, _ = oracle_for(orc, q)
            po, co, tlo, tao = mh.run(C, pen, seed, gen=2, window_id=w, trace=True)
            assert (par[w] == po).all() and (like[w] == co).all()
            assert np.array_equal(tl[w], tlo) and (ta[w] == tao).all()


def test_klist_mode(orc):
    say("klist mode")
    p = helpers.synthetic(120, 300, seed=13)
    K = 17
    rng = np.random.default_rng(1)
    kl = np.zeros((p["N"], K), np.uint32)
    for i in range(p["N"]):
        cand = [h for h in range(p["H"]) if h // 2 != i]
        kl[i] = rng.choice(cand, K, replace=False)
    C, pen, seed = 25, 1.0, 8
    job = job_for(p, C, prop_mode=PROP_KLIST, klist=kl[None], trace=True)
    job.make_plan(8)
    job.run_batched(gen=0, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, PREC_F32, prop_mode=1, klist=kl)
    for i in range(p["N"]):
        assert set(job.parents()[0][i]) <= set(kl[i])


def test_run_hmm_on_device_host_proposals(orc):
    # HMMLikeCUDA::RunHMMOnDevice semantics: host-drawn initial parents and
    # candidates, SoA [4][N] in and out (hmmLike.cu:399-572)
    say("run_hmm_on_device")
    from glphase_amd._lib import c_f32p, c_u32p, c_u64p, check, lib, ptr
    p = helpers.synthetic(90, 600, seed=17)
    C, seed = 30, 777
    rng = np.random.default_rng(2)
    init = np.zeros((p["N"], 4), np.uint32)
    extra = np.zeros((C, p["N"]), np.uint32)
    for i in range(p["N"]):
        cand = np.array([h for h in range(p["H"]) if h // 2 != i], np.uint32)
        init[i] = rng.choice(cand, 4)
        extra[:, i] = rng.choice(cand, C)
    L = lib()
    ctx = ctypes.c_void_p()
    check(L.glp_create(0, ctypes.byref(ctx)), "create")
    try:
        check(L.glp_set_tran(ctx, ptr(p["tran"], c_f32p), p["tran"].size), "tran")
        check(L.glp_set_mutation_mat(ctx, ptr(np.ascontiguousarray(p["pc"]), c_f32p)), "pc")
        check(L.glp_set_codebook(ctx, ptr(p["codebook"], c_f32p), 256), "cb")
        pk = np.ascontiguousarray(p["packed"])
        check(L.glp_set_packed_gls(ctx, ptr(pk, c_u32p), pk.size, 8), "packed")
        check(L.glp_set_pen(ctx, 0.7), "pen")
        soa = np.ascontiguousarray(init.T)
        panel = np.ascontiguousarray(p["panel"])
        check(L.glp_run_hmm_on_device(ctx, ptr(panel, c_u64p), panel.size, ptr(extra, c_u32p),
                                      p["M"], p["N"], C, ptr(soa, c_u32p), seed), "run")
    finally:
        L.glp_destroy(ctx)
    mh, _ = oracle_for(orc, p)
    par_o, _ = mh.run(C, 0.7, seed, prop_mode=2, init_parents=init, extra=extra)
    assert (soa.T == par_o).all()


def test_planted_parents_gpu(orc):
    say("planted parents")
    for nsamp, samples in ((2, (0, 1)), (4, (0, 1, 2, 3))):
        p = helpers.planted(num_samps=nsamp)
        from glphase_amd.engine import Job
        win = dict(panel=p["panel"], packed=p["packed"], tran=p["tran"])
        job = Job([win], nsamp, p["H"], p["codebook"], p["pc"], bits=4, num_cycles=400)
        job.make_plan(16)
        job.run_batched(gen=0, pen=1.0, seed=2024)
        par = job.parents()[0]
        mh, _ = oracle_for(orc, p)
        po, _ = mh.run(400, 1.0, 2024)
        assert (par == po).all()
        helpers.check_planted(par, p["expect"], [s for s in samples if str(s) in p["expect"]])


def test_large_window_sampled_parity(orc):
    # N = 5000 x M = 4096, 64 proposals, mixed precision: bit-exact parents
    # and 1e-12 log-likelihoods for every chain, and 1e-5 relative agreement
    # of sampled returned log-likelihoods with FP64
    say("large window")
    p = helpers.synthetic(5000, 4096, seed=4, depth=2.0, training_size=1 << 16)
    C, pen, seed = 64, 1.0, 31337
    job = job_for(p, C, precision=PREC_MIXED)
    job.make_plan(0)
    job.run_batched(gen=0, pen=pen, seed=seed)
    par, like = job.parents()[0], job.like()[0]
    mh, lut = oracle_for(orc, p)
    # every chain against the oracle (OpenMP over samples)
    po, co = mh.run(C, pen, seed, precision=PREC_MIXED, nthreads=8)
    assert (par == po).all()
    assert (np.abs(like - co) <= 1e-12 * np.abs(co)).all()
    rng = np.random.default_rng(0)
    for i in sorted(rng.choice(p["N"], 24, replace=False)):
        emit = orc.sample_emit(p["packed"], int(i), p["M"], 8, lut)
        f64 = orc.hmm_like_f64(p["panel"], par[i], emit, p["tran"])
        assert abs(like[i] - f64) <= 1e-5 * abs(f64)
    assert (like <= 0).all()
    assert all(h // 2 != i for i in range(p["N"]) for h in par[i])


@pytest.mark.gpu
def test_rescale_reciprocal_is_exact():
    # the rescale step's reciprocal against the IEEE division for every float
    # from the smallest subnormal to the rescale threshold powf(FLT_MIN, 2/3)
    import ctypes

    from glphase_amd._lib import check, lib
    say("reciprocal self-test")
    norm = np.float32(np.finfo(np.float32).tiny) ** np.float32(2.0 / 3.0)
    hi = int(np.array([norm], np.float32).view(np.uint32)[0]) + 1
    # and on to 2^126 for the short reciprocal (hmm_work's site sums)
    for lo_bits, hi_bits in ((1, hi), (hi, 0x7E800001)):
        bad, first = ctypes.c_uint64(), ctypes.c_uint32()
        check(lib().glp_selftest_recip(lo_bits, hi_bits, ctypes.byref(bad), ctypes.byref(first)),
              "glp_selftest_recip")
        assert bad.value == 0, f"{bad.value} mismatches, first at 0x{first.value:08x}"


@pytest.mark.gpu
def test_lanes_above_32_rejected():
    from glphase_amd._lib import GlpError
    p = helpers.synthetic(16, 64, seed=3)
    job = job_for(p, C=4)
    with pytest.raises(GlpError):
        job.make_plan(64)


def test_subnormal_site_sums_hmm_like(orc):
    # emissions near 2^-42 drive site sums below 2^-126 in every pass: the
    # short reciprocal gives NaN there and the pass is redone with the exact
    # one, so the values still equal the oracle's bit for bit
    from glphase_amd._lib import c_f32p, c_u32p, c_u64p, check, lib, ptr
    say("hmm_like with subnormal site sums")
    p = helpers.tiny_emission_problem(seed=3)
    lut = orc.emit_lut(p["codebook"], p["pc"])
    rng = np.random.default_rng(4)
    n = 64
    samples = rng.integers(0, p["N"], n).astype(np.uint32)
    parents = rng.integers(0, p["H"], (n, 4)).astype(np.uint32)
    emit_of = lambda s: orc.sample_emit(p["packed"], s, p["M"], p["bits"], lut)  # noqa: E731
    assert helpers.count_subnormal_sums(p, samples[:8], parents[:8], emit_of) > 0
    L = lib()
    ctx = ctypes.c_void_p()
    check(L.glp_create(0, ctypes.byref(ctx)), "create")
    try:
        check(L.glp_set_tran(ctx, ptr(p["tran"], c_f32p), p["tran"].size), "tran")
        check(L.glp_set_mutation_mat(ctx, ptr(np.ascontiguousarray(p["pc"]), c_f32p)), "pc")
        check(L.glp_set_codebook(ctx, ptr(p["codebook"], c_f32p), 16), "cb")
        pk = np.ascontiguousarray(p["packed"])
        check(L.glp_set_packed_gls(ctx, ptr(pk, c_u32p), pk.size, 4), "packed")
        out = np.zeros(n, np.float32)
        panel = np.ascontiguousarray(p["panel"])
        check(L.glp_hmm_like(ctx, ptr(panel, c_u64p), panel.size, p["M"], p["N"], n,
                             ptr(samples, c_u32p), ptr(parents, c_u32p), ptr(out, c_f32p)),
              "hmm_like")
        want = np.array([orc.hmm_like_f32(p["panel"], parents[e], emit_of(int(samples[e])),
                                          p["tran"]) for e in range(n)], np.float32)
        assert np.isfinite(want).all()
        assert np.array_equal(out, want)
    finally:
        L.glp_destroy(ctx)


@pytest.mark.parametrize("prec", [PREC_F32, PREC_MIXED])
@pytest.mark.parametrize("form", ["plain", "batched"])
def test_subnormal_site_sums_mh(orc, prec, form):
    say(f"MH with subnormal site sums prec={prec} form={form}")
    p = helpers.tiny_emission_problem(seed=5)
    C, pen, seed = 24, 1.0, 77
    job = job_for(p, C, precision=prec, trace=True)
    if form == "plain":
        job.run_plain(gen=2, pen=pen, seed=seed)
    else:
        job.make_plan(4)
        job.run_batched(gen=2, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, prec, gen=2)


@pytest.mark.parametrize("hn", [10_000, 40_000])
@pytest.mark.parametrize("form", ["plain", "batched"])
def test_mixed_smallest_emissions(orc, hn, form):
    # runs of sites where every parent state emits mu^2 (a certain
    # alternative homozygote against all-reference parents) at the mutation
    # rate of hn haplotypes: an 8-site run takes the sum from 2^100 below
    # 2^-99, the wavefront redoes its passes with the scaling after every
    # site, and parents and log-likelihoods equal the oracle's; every
    # log-likelihood is finite
    say(f"mixed precision, smallest emissions hn={hn} form={form}")
    p = helpers.hom_alt_run_problem(hn, seed=hn % 97)
    C, pen, seed = 20, 1.0, 4242
    job = job_for(p, C, precision=PREC_MIXED, trace=True)
    if form == "plain":
        job.run_plain(gen=1, pen=pen, seed=seed)
    else:
        job.make_plan(4)
        job.run_batched(gen=1, pen=pen, seed=seed)
    compare_run(orc, p, job, C, pen, seed, PREC_MIXED, gen=1)
    tl, _ = job.trace()
    assert np.isfinite(tl[0]).all()


@pytest.mark.parametrize("M", [1, 3, 31, 33, 95, 97, 301])
def test_hmm_like_odd_site_count_strong_transitions(orc, M):
    # with an odd site count the last site's t2 is the last float of tran,
    # alone in its pair; r = 0.3 makes every transition term count
    from glphase_amd._lib import c_f32p, c_u32p, c_u64p, check, lib, ptr
    say(f"hmm_like odd M={M}, r=0.3")
    p = helpers.synthetic(16, M, bits=8, seed=M, training_size=4096)
    r = np.float32(0.3)
    p["tran"][1:] = (np.float32(1 - r) * np.float32(1 - r), r * np.float32(1 - r), r * r)
    L = lib()
    ctx = ctypes.c_void_p()
    check(L.glp_create(0, ctypes.byref(ctx)), "create")
    try:
        check(L.glp_set_tran(ctx, ptr(p["tran"], c_f32p), p["tran"].size), "tran")
        check(L.glp_set_mutation_mat(ctx, ptr(np.ascontiguousarray(p["pc"]), c_f32p)), "pc")
        check(L.glp_set_codebook(ctx, ptr(p["codebook"], c_f32p), 256), "cb")
        pk = np.ascontiguousarray(p["packed"])
        check(L.glp_set_packed_gls(ctx, ptr(pk, c_u32p), pk.size, 8), "packed")
        rng = np.random.default_rng(M + 1)
        n = 64
        samples = rng.integers(0, p["N"], n).astype(np.uint32)
        parents = rng.integers(0, p["H"], (n, 4)).astype(np.uint32)
        out = np.zeros(n, np.float32)
        panel = np.ascontiguousarray(p["panel"])
        check(L.glp_hmm_like(ctx, ptr(panel, c_u64p), panel.size, M, p["N"], n,
                             ptr(samples, c_u32p), ptr(parents, c_u32p), ptr(out, c_f32p)),
              "hmm_like")
        lut = orc.emit_lut(p["codebook"], p["pc"])
        want = np.array([orc.hmm_like_f32(p["panel"], parents[e],
                                          orc.sample_emit(p["packed"], int(samples[e]), M, 8, lut),
                                          p["tran"]) for e in range(n)], np.float32)
        assert np.array_equal(out, want)
    finally:
        L.glp_destroy(ctx)
