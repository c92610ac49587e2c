"""Benchmark of the GLPhase MH surrogate-parent update on one MI355X.

A step is one MH generation over every sample: 4 fresh parents scored, then
C proposals each scored and accepted or rejected (Insti::estimate's per-
generation work, insti.cpp:1668-1705).  Site-evals (SE) per step =
N * (C + 1) * M: one forward-recurrence step of one chain at one site.

Default workload: 5,000 samples x 4,096 sites, 64 proposals per sample,
synthetic founder-mosaic haplotypes and Poisson(2)-depth GLs, 8-bit VQ codes
(codebook trained on every GL), batched persistent kernel with FP32
recurrence and FP64 log accumulation, each pass run as a forward and a
backward half joined in FP64 (GLP_PREC_MIXED_FB; --precision mixed runs the
forward-only pass).  The forward-only pass and the plain path (one launch per
MH iteration) run on the same inputs; one generation is replayed on the host
against an FP64 forward pass (log-likelihood error, accept decisions); the
other benchmark workloads run beside it, and the CPU oracle (C, -O3) is
timed on a bounded sample of each.

python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_SE = 29      # 2 pair sums + 4 x (3 mul + 2 add) + 4 emission mul + 3 adds
PEAK_FP32_TFLOPS = 157.3
VQ_TRAINING = 1 << 18  # GL trios the codebook is trained on


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(N, M, C, depth, seed):
    from glphase_amd import host, synth
    t0 = time.time()
    d = synth.make_dataset(N, M, depth=depth, seed=seed)
    hb = synth.initial_haps(d["gls"], seed=seed + 1)
    panel = synth.pack_haps(hb)
    tran, pc, mu = host.model_params(2 * N, d["r_morgans"])
    t1 = time.perf_counter()
    # every GL trains the codebook, as GLVQ does (glVQ.cpp:53-73); the
    # nearest-codeword search of each Lloyd iteration runs on the device
    cb, iters = host.vq_train(d["gls"], 256, key=seed, training_size=0, device=True)
    vq_train_s = time.perf_counter() - t1
    codes = host.vq_codes(d["gls"], cb, device=True)
    packed = host.pack_codes(codes, 8)
    log(f"workload N={N} M={M} C={C}: generated in {time.time() - t0:.1f}s (VQ {iters} iterations)")
    return dict(panel=panel, packed=packed, tran=tran, codebook=cb, pc=pc, N=N, M=M,
                H=panel.shape[0], het_aa=np.ascontiguousarray(d["gls"][:, :, 1:3]),
                gls=d["gls"], codes=codes, vq_iters=iters, vq_train_s=vq_train_s, seed=seed,
                truth=d["truth"])


# ---------------------------------------------------------------------------
# accuracy of a timed generation against FP64 (numpy; no oracle involved)
# ---------------------------------------------------------------------------
_PHILOX_M = (np.uint64(0xD2511F53), np.uint64(0xCD9E8D57))
_PHILOX_W = (np.uint32(0x9E3779B9), np.uint32(0xBB67AE85))


def philox_blocks(k, sample, window, gen, seed):
    """Philox4x32-10 of counters (k, sample, window, gen) under key seed
    (Salmon et al., SC'11), the engine's draw layout; arrays broadcast."""
    c = [np.asarray(x, np.uint32) for x in np.broadcast_arrays(k, sample, window, gen)]
    k0, k1 = np.uint32(seed & 0xFFFFFFFF), np.uint32(seed >> 32)
    m32 = np.uint64(0xFFFFFFFF)
    for r in range(10):
        if r:
            k0 = np.uint32((int(k0) + int(_PHILOX_W[0])) & 0xFFFFFFFF)
            k1 = np.uint32((int(k1) + int(_PHILOX_W[1])) & 0xFFFFFFFF)
        p0 = _PHILOX_M[0] * c[0].astype(np.uint64)
        p1 = _PHILOX_M[1] * c[2].astype(np.uint64)
        hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & m32).astype(np.uint32)
        hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & m32).astype(np.uint32)
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    return c


def unif_excl(x, nhaps, sample):
    r = ((x.astype(np.uint64) * np.uint64(nhaps - 2)) >> np.uint64(32)).astype(np.int64)
    return r + 2 * (r >= 2 * np.asarray(sample, np.int64))


def emission_table(codebook, pc):
    """FillEmit per code in float32 with the engine's operation order."""
    g0, g1 = codebook[:, 0].astype(np.float32), codebook[:, 1].astype(np.float32)
    g2 = np.maximum(np.float32(1) - g0 - g1, np.float32(0))
    pc = np.asarray(pc, np.float32)
    return np.stack([((pc[i, 0] * g0 + pc[i, 1] * g1) + pc[i, 2] * g1) + pc[i, 3] * g2
                     for i in range(4)], 1).astype(np.float32)


def fp64_forward(panel, codes, E, tran, samples, parents):
    """Log-likelihoods of (sample, 4 parents) rows in FP64, normalised at
    every site: the same model as the engine's passes."""
    n, M = len(samples), codes.shape[1]
    bits = np.unpackbits(np.ascontiguousarray(panel).view(np.uint8).reshape(panel.shape[0], -1),
                         axis=1, bitorder="little")[:, :M]
    # site-major copies, so that each site reads contiguous rows
    h = [np.ascontiguousarray(bits[parents[:, q]].T).astype(np.int64) for q in range(4)]
    cs = np.ascontiguousarray(codes[samples].T)
    E64 = E.astype(np.float64).reshape(-1)
    l = np.zeros((n, 4))
    score = np.zeros(n)
    t = tran.astype(np.float64)
    for m in range(M):
        f0, f1, m0, m1 = (x[m] for x in h)
        base = 4 * cs[m].astype(np.int64)
        e00, e01 = E64[base + 2 * f0 + m0], E64[base + 2 * f0 + m1]
        e10, e11 = E64[base + 2 * f1 + m0], E64[base + 2 * f1 + m1]
        if m == 0:
            b = np.full((n, 4), 0.25)
        else:
            t0, t1, t2 = t[m]
            b = np.stack([l[:, 0] * t0 + (l[:, 1] + l[:, 2]) * t1 + l[:, 3] * t2,
                          l[:, 1] * t0 + (l[:, 0] + l[:, 3]) * t1 + l[:, 2] * t2,
                          l[:, 2] * t0 + (l[:, 0] + l[:, 3]) * t1 + l[:, 1] * t2,
                          l[:, 3] * t0 + (l[:, 1] + l[:, 2]) * t1 + l[:, 0] * t2], 1)
        l = b * np.stack([e00, e01, e10, e11], 1)
        s = l.sum(1)
        score += np.log(s)
        l /= s[:, None]
    return score


def accuracy_vs_fp64(w, C, pen, seed, lanes, precision, gen=0, n_chains=160, tol=1e-5):
    """Runs generation `gen` of the timed workload once more with traces and
    replays n_chains chains on the host (their Philox draws, the device's
    accept decisions): every proposal's log-likelihood against FP64, and the
    accept decisions FP64 values would take, outside a band of relative
    width tol around the accept threshold."""
    import torch  # noqa: F401
    from glphase_amd.engine import Job
    N, H = w["N"], w["H"]
    job = Job([dict(panel=w["panel"], packed=w["packed"], tran=w["tran"])], N, H, w["codebook"],
              w["pc"], bits=8, num_cycles=C, precision=precision, trace=True)
    job.make_plan(lanes)
    job.run_batched(gen=gen, pen=pen, seed=seed)
    tl, ta = job.trace()
    job.close()
    rng = np.random.default_rng(1)
    S = np.sort(rng.choice(N, n_chains, replace=False))
    init = philox_blocks(0xFFFFFFFF, S, 0, gen, seed)
    P = np.stack([unif_excl(init[q], H, S) for q in range(4)], 1)
    kk = np.arange(C)[None, :]
    d = philox_blocks(kk, S[:, None], 0, gen, seed)
    rp = (d[0] & 3).astype(np.int64)
    cand = unif_excl(d[1], H, S[:, None])
    u = d[2].astype(np.float64) * 2.0 ** -32
    rows_s, rows_p = [S.copy()], [P.copy()]
    cur = P.copy()
    for c in range(C):
        prop = cur.copy()
        prop[np.arange(len(S)), rp[:, c]] = cand[:, c]
        rows_s.append(S.copy())
        rows_p.append(prop)
        acc = ta[0, S, c].astype(bool)
        cur = np.where(acc[:, None], prop, cur)
    samples = np.concatenate(rows_s)
    parents = np.concatenate(rows_p)
    E = emission_table(w["codebook"], w["pc"])
    f64 = fp64_forward(w["panel"], w["codes"], E, w["tran"], samples, parents)
    f64 = f64.reshape(C + 1, len(S)).T  # [chain][init + C]
    dev = tl[0, S]
    rel = np.abs(dev - f64) / np.abs(f64)
    # accept decisions along the device's path, decided with FP64 values
    cur64 = f64[:, 0].copy()
    agree = band = total = 0
    for c in range(C):
        p64 = f64[:, c + 1]
        with np.errstate(over="ignore"):
            d64 = (p64 >= cur64) | (u[:, c] < np.exp((p64 - cur64) * pen))
        # accept <=> p > cur + log(u) / pen: decisions whose FP64 value lies
        # within the two values' tolerance of that threshold are in the band
        thr = cur64 + np.log(np.maximum(u[:, c], 1e-300)) / pen if pen > 0 else cur64
        inside = np.abs(p64 - thr) <= 2 * tol * np.abs(cur64)
        gpu = ta[0, S, c].astype(bool)
        band += int(inside.sum())
        total += int((~inside).sum())
        agree += int(((d64 == gpu) & ~inside).sum())
        cur64 = np.where(gpu, p64, cur64)
    return {"chains": int(len(S)), "proposals": int(len(S) * C), "generation": gen,
            "max_rel_err": float(rel.max()), "p99_rel_err": float(np.percentile(rel, 99)),
            "median_rel_err": float(np.median(rel)), "tolerance": tol,
            "accept_agreement": agree / max(total, 1), "decisions_outside_band": total,
            "decisions_inside_band": band}


def time_steps(torch, fn, steps, warmup, stream_events=True):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return start.elapsed_time(stop) / 1e3, wall


def cpu_baseline(w, C, pen, seed, budget_s=12.0, precision=1):
    """The oracle (C restatement of Insti::solve, -O3) on this host's cores:
    whole generations of the same workload under OpenMP over samples,
    repeated until about budget_s seconds of work, plus a single-threaded run
    over a prefix of the samples."""
    from oracle import bind as orc
    lut = orc.emit_lut(w["codebook"], w["pc"])
    mh = orc.MH(w["panel"], w["packed"], 8, lut, w["tran"], w["N"])
    se_per_chain = (C + 1) * w["M"]
    # single core, about budget_s / 4 seconds
    n1 = 8
    t0 = time.perf_counter()
    mh.run(C, pen, seed, precision=precision, s0=0, s1=n1)
    dt1 = time.perf_counter() - t0
    n1 = int(min(w["N"], max(8, n1 * budget_s / 4 / max(dt1, 1e-6))))
    t0 = time.perf_counter()
    mh.run(C, pen, seed, precision=precision, s0=0, s1=n1)
    dt1 = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    cores = min(cores, os.cpu_count() or 1)
    gens, chains = 0, 0
    t0 = time.perf_counter()
    while gens == 0 or time.perf_counter() - t0 < budget_s:
        mh.run(C, pen, seed, gen=gens, precision=precision, s0=0, s1=w["N"], nthreads=cores)
        gens += 1
        chains += w["N"]
    dta = time.perf_counter() - t0
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": se_per_chain * chains / dta, "unit": "site-evals/s", "cores": cores,
        "kind": "port", "cpu_model": model, "host_cpus": os.cpu_count(),
        "sample": f"{gens} generation(s) of all {w['N']} chains x {C + 1} evaluations x "
                  f"{w['M']} sites ({dta:.1f}s, OpenMP over samples)",
        "single_core": {"value": se_per_chain * n1 / dt1, "cores": 1,
                        "sample": f"{n1} chains of one generation ({dt1:.1f}s)"},
    }


def oracle_rates(windows, cb, pc, N, C, pen, seed, precision, prop_mode=0, klist=None,
                 budget_s=3.0):
    """CPU baseline of one workload: the oracle (C restatement of the
    reference's Impute::hmm_like + Insti::solve, -O3) on this host.  One core
    over chains of the first window in order (the reference's serial loop),
    then every core over whole generations of every window (OpenMP over
    samples), each for about budget_s / 2 seconds."""
    from oracle import bind as orc
    lut = orc.emit_lut(cb, pc)
    kw = dict(precision=precision, prop_mode=prop_mode)
    if klist is not None:
        kw["klist"] = klist
    mhs = [orc.MH(w["panel"], w["packed"], 8, lut, w["tran"], N) for w in windows]
    se_gen = sum(N * (C + 1) * int(w["tran"].shape[0]) for w in windows)
    se0 = (C + 1) * int(windows[0]["tran"].shape[0])
    n1, t1 = 0, 0.0
    while n1 < N and t1 < budget_s / 2:
        b = min(N, n1 + 8)
        t0 = time.perf_counter()
        mhs[0].run(C, pen, seed, s0=n1, s1=b, **kw)
        t1 += time.perf_counter() - t0
        n1 = b
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    cores = min(cores, os.cpu_count() or 1)
    gens, ta = 0, 0.0
    while gens == 0 or ta < budget_s / 2:
        t0 = time.perf_counter()
        for k, mh in enumerate(mhs):
            mh.run(C, pen, seed, gen=gens, window_id=k, nthreads=cores, **kw)
        ta += time.perf_counter() - t0
        gens += 1
    return {"cpu_1core": se0 * n1 / t1, "cpu_allcore": se_gen * gens / ta, "cores": cores,
            "kind": "port",
            "sample": f"1 core: {n1} chains ({t1:.1f}s); {cores} cores: {gens} generation(s) "
                      f"of {N} samples x {len(windows)} window(s) ({ta:.1f}s)"}


def aux_knn(torch, steps, seed=11, cpu=True):
    """kNN candidate lists (KNN::ClusterHaps) of the reference-panel workload:
    10,000 samples plus 20,000 panel haplotypes x 2,048 sites, K = 200, panel
    resident on the device.  Cost unit: one 64-bit word compared between a
    query haplotype and a panel haplotype (20,000 x 40,000 x 32 per build).
    The oracle's kNN is timed on the host on a prefix of the samples."""
    from glphase_amd import synth
    from glphase_amd.engine import knn_lists
    from oracle import bind as orc
    N, M, R, K = 10000, 2048, 20000, 200
    t0 = time.time()
    d = synth.make_dataset(N, M, depth=4.0, seed=seed, n_ref_haps=R)
    hb = np.concatenate([synth.initial_haps(d["gls"], seed=seed + 1), d["ref_haps"]], axis=0)
    panel = synth.pack_haps(hb)
    H, wn = panel.shape
    log(f"kNN workload: {N} samples + {R} panel haplotypes x {M} sites in {time.time() - t0:.1f}s")
    pd = torch.from_numpy(panel.view(np.int64)).to("cuda:0")
    out = knn_lists(pd, M, N, K)
    secs, _ = time_steps(torch, lambda: knn_lists(pd, M, N, K), steps, 1)
    ms = secs / steps * 1e3
    words = 2 * N * H * wn
    ns = 200
    t0 = time.perf_counter()
    want = orc.knn(panel, M, ns, K)
    cpu_s = time.perf_counter() - t0
    same = bool((out[:ns].cpu().numpy().view(np.uint32) == want).all())
    log(f"kNN: {ms:.3f} ms/build, {words / (ms / 1e3):.3e} word-compares/s; "
        f"oracle {ns} samples in {cpu_s:.2f}s; first {ns} rows identical: {same}")
    mh = aux_reference_panel_mh(torch, d, panel, out, steps, seed, cpu=cpu)
    return {"ms_per_build": ms, "unit": "word-compares/s", "value": words / (ms / 1e3),
            "mh_klist": mh,
            "config": {"samples": N, "panel_haplotypes": R, "sites": M, "K": K},
            "identical_to_oracle_rows": ns if same else 0,
            "cpu_baseline": {"value": 2 * ns * H * wn / cpu_s, "unit": "word-compares/s",
                             "cores": 1, "kind": "port",
                             "sample": f"oracle kNN lists of {ns} samples ({cpu_s:.2f}s)"}}


def aux_reference_panel_mh(torch, d, panel, klist, steps, seed, C=64, cpu=True):
    """The MH update of the reference-panel workload: 10,000 samples whose
    candidates are their K = 200 nearest haplotypes of panel + samples
    (lists from the kNN build), 64 proposals each, mixed precision in the
    forward-backward form, autotuned depth.  Cost unit: the headline's
    site-eval."""
    from glphase_amd import PREC_MIXED_FB as PREC, PROP_KLIST, host
    from glphase_amd.engine import Job
    N, M = d["gls"].shape[0], d["gls"].shape[1]
    tran, pc, _ = host.model_params(2 * N, d["r_morgans"])  # hn = 2N (insti.cpp:1205)
    cb, _ = host.vq_train(d["gls"], 256, key=seed, training_size=VQ_TRAINING, device=True)
    packed = host.pack_codes(host.vq_codes(d["gls"], cb, device=True), 8)
    job = Job([dict(panel=panel, packed=packed, tran=tran)], N, panel.shape[0], cb, pc, bits=8,
              num_cycles=C, prop_mode=PROP_KLIST, klist=klist.cpu().numpy().view(np.uint32)[None],
              precision=PREC)
    job.make_plan(0)
    lanes = int(job.autotune(gen=0, pen=1.0, seed=seed)[0])
    gen = [0]

    def step():
        job.run_batched(gen=gen[0], pen=1.0, seed=seed)
        gen[0] += 1

    secs, _ = time_steps(torch, step, steps, 1)
    ms = secs / steps * 1e3
    se = N * (C + 1) * M

    def pstep():
        job.run_plain(gen=gen[0], pen=1.0, seed=seed)
        gen[0] += 1

    psecs, _ = time_steps(torch, pstep, 1, 0)
    pms = psecs * 1e3
    log(f"reference-panel MH (K-lists): {ms:.3f} ms/step, {se / (ms / 1e3):.3e} SE/s, "
        f"lanes/chain {lanes}; plain {pms:.3f} ms/step")
    job.close()
    out = {"ms_per_step": ms, "value": se / (ms / 1e3), "unit": "site-evals/s",
           "lanes_per_chain": lanes,
           "plain": {"ms_per_step": pms, "value": se / (pms / 1e3)},
           "config": {"samples": N, "panel_haplotypes": panel.shape[0] - 2 * N, "sites": M,
                      "proposals": C, "K": int(klist.shape[1]),
                      "precision": "mixed, forward-backward"}}
    if cpu:
        out["cpu"] = oracle_rates([dict(panel=panel, packed=packed, tran=tran)], cb, pc, N, C,
                                  1.0, seed, PREC, prop_mode=PROP_KLIST,
                                  klist=klist.cpu().numpy().view(np.uint32))
    return out


def aux_batching(torch, steps, seed=21, N=2500, W=32, C=32, cpu=True):
    """Uneven windows: 2,500 samples over 32 windows whose site counts are
    lognormal (median 1,024, clipped to 256..4,096), 32 proposals per sample,
    FP32 forward and FP32 log accumulation (the reference's arithmetic); one
    persistent launch per generation with depths autotuned per window-size
    bucket, the same through its HIP graph, and the plain path."""
    from glphase_amd import PREC_F32, host, synth
    from glphase_amd.engine import Job
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.round(np.exp(rng.normal(np.log(1024.0), 0.6, W))), 256, 4096).astype(int)
    t0 = time.time()
    data = [synth.make_dataset(N, int(M), depth=4.0, seed=seed + k) for k, M in enumerate(sizes)]
    allg = np.concatenate([x["gls"].reshape(-1, 3) for x in data])
    cb, _ = host.vq_train(allg, 256, key=seed, training_size=VQ_TRAINING, device=True)
    del allg
    wins = []
    pc = None
    for k, x in enumerate(data):
        panel = synth.pack_haps(synth.initial_haps(x["gls"], seed=seed + 100 + k))
        tran, pc, _ = host.model_params(2 * N, x["r_morgans"])
        packed = host.pack_codes(host.vq_codes(x["gls"], cb, device=True), 8)
        wins.append(dict(panel=panel, packed=packed, tran=tran))
    del data
    log(f"batching workload: {N} samples x {W} windows, sites {sizes.min()}..{sizes.max()} "
        f"(total {int(sizes.sum())}) in {time.time() - t0:.1f}s")
    se = N * (C + 1) * int(sizes.sum())
    job = Job(wins, N, 2 * N, cb, pc, bits=8, num_cycles=C, precision=PREC_F32)
    job.make_plan(0)
    tuned = [int(q) for q in job.autotune(gen=0, pen=1.0, seed=seed)]
    out = {"config": {"samples": N, "windows": W, "sites_total": int(sizes.sum()),
                      "sites_min": int(sizes.min()), "sites_median": int(np.median(sizes)),
                      "sites_max": int(sizes.max()), "proposals": C, "precision": "fp32"},
           "lanes_per_window": tuned, "waves_per_workgroup": job.waves(),
           "split_slots": job.split(),
           "unit": "site-evals/s"}
    for name, fn in (("batched", job.run_batched), ("batched_graph", job.run_batched_graph),
                     ("plain", job.run_plain)):
        gen = [0]

        def step(fn=fn):
            fn(gen=gen[0], pen=1.0, seed=seed)
            gen[0] += 1

        n = steps if name != "plain" else max(1, steps // 2)
        secs, _ = time_steps(torch, step, n, 1)
        ms = secs / n * 1e3
        out[name] = {"ms_per_step": ms, "value": se / (ms / 1e3)}
        log(f"batching {name}: {ms:.3f} ms/step, {se / (ms / 1e3):.3e} SE/s")
    out["batched_over_plain"] = out["batched"]["value"] / out["plain"]["value"]
    job.close()
    if cpu:
        out["cpu"] = oracle_rates(wins, cb, pc, N, C, 1.0, seed, PREC_F32)
    return out


def aux_whole_run(torch, w, C, seed=7, burnin=56, sampling=200):
    """A whole phasing run of the MH workload's samples (glphase_amd.phase:
    Insti::estimate's schedule): initial haplotypes from the GLs, 56
    annealed burn-in + 200 sampling generations of MH update (C proposals),
    hmm_work, panel update and tallies, all on the device; genotype
    discordance of the output calls against the simulated truth."""
    from glphase_amd.phase import PhaseRun, called_genotypes
    het_aa = np.ascontiguousarray(w["gls"][:, :, 1:3].transpose(1, 0, 2))
    run = PhaseRun(het_aa, tran=w["tran"], cycles=C, burnin=burnin, sampling=sampling, seed=seed,
                   training_size=VQ_TRAINING)
    hsum = run.run()
    a, b = called_genotypes(hsum, run.sn)
    truth = w["truth"]
    tdos = (truth[0::2].astype(np.int8) + truth[1::2].astype(np.int8)).T  # [M][N]
    disc = float(((a + b) != tdos).mean())
    g = w["gls"]
    gl_disc = float((np.argmax(g, axis=2).T != tdos).mean())
    gens = burnin + sampling
    log(f"whole run: {gens} generations in {run.gen_s:.2f}s ({run.gen_s / gens * 1e3:.2f} ms each), "
        f"setup {run.setup_s:.2f}s; discordance {disc:.4f} (GL argmax {gl_disc:.4f})")
    out = {"generations": gens, "burnin": burnin, "sampling": sampling, "proposals": C,
           "seconds": run.gen_s, "ms_per_generation": run.gen_s / gens * 1e3,
           "setup_s": run.setup_s, "genotype_discordance": disc,
           "gl_argmax_discordance": gl_disc}
    run.close()
    return out


def aux_small(torch, steps, N, M, C, seed, cpu=True):
    """A one-window workload in the reference's FP32 arithmetic: the plain
    path (one launch per MH iteration, one proposal per sample per launch)
    and the batched kernel (autotuned depth) on the same inputs, and the CPU
    oracle.  Per-launch time of the plain path and site-evals/s of both."""
    from glphase_amd import PREC_F32
    from glphase_amd.engine import Job
    p = make_workload(N, M, C, 4.0, seed)
    win = dict(panel=p["panel"], packed=p["packed"], tran=p["tran"])
    job = Job([win], N, p["H"], p["codebook"], p["pc"], bits=8, num_cycles=C, precision=PREC_F32)
    job.make_plan(0)
    lanes = int(job.autotune(gen=0, pen=1.0, seed=seed)[0])
    se = N * (C + 1) * M
    res = {}
    for name, fn in (("plain", job.run_plain), ("batched", job.run_batched)):
        gen = [0]

        def step(fn=fn):
            fn(gen=gen[0], pen=1.0, seed=seed)
            gen[0] += 1

        secs, _ = time_steps(torch, step, steps, 1)
        res[name] = secs / steps * 1e3
    job.close()
    ms = res["plain"]
    log(f"{N} x {M} x {C}: plain {ms / (C + 1):.4f} ms/launch, {se / (ms / 1e3):.3e} SE/s; "
        f"batched {res['batched']:.3f} ms/step ({lanes} lanes/chain)")
    out = {"ms_per_launch": ms / (C + 1), "launches_per_step": C + 1, "value": se / (ms / 1e3),
           "unit": "site-evals/s",
           "plain_roofline_frac": FLOP_PER_SE * se / (ms / 1e3) / 1e12 / PEAK_FP32_TFLOPS,
           "batched": {"ms_per_step": res["batched"], "value": se / (res["batched"] / 1e3),
                       "lanes_per_chain": lanes},
           "config": {"samples": N, "sites": M, "proposals": C, "proposals_per_launch": 1,
                      "precision": "fp32"}}
    if cpu:
        out["cpu"] = oracle_rates([win], p["codebook"], p["pc"], N, C, 1.0, seed, PREC_F32)
    return out

def aux_vq(torch, w, steps):
    """GL codebook training (GLVQ) and coding of the MH workload's GLs with
    the nearest-codeword search on the device.  Coding cost unit: one
    (GL, codeword) distance; the GLs are resident on the device."""
    from glphase_amd import host
    from glphase_amd._lib import c_f32p, check, lib, ptr
    gls, cb = w["gls"], w["codebook"]
    # host and device-assisted training give the same codebook (checked on a
    # subsample of VQ_TRAINING GLs: the host search over every GL is slow)
    t0 = time.perf_counter()
    cbd, itd = host.vq_train(gls, 256, key=w["seed"], training_size=VQ_TRAINING, device=True)
    dev_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    cbh, ith = host.vq_train(gls, 256, key=w["seed"], training_size=VQ_TRAINING)
    host_s = time.perf_counter() - t0
    same_cb = bool(cbd.tobytes() == cbh.tobytes() and itd == ith)
    flat = np.ascontiguousarray(gls.reshape(-1, 3))
    n = flat.shape[0]
    gd = torch.from_numpy(flat).to("cuda:0")
    out = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream

    def code():
        check(lib().glp_vq_assign_device(gd.data_ptr(), n, 1, ptr(cb, c_f32p), cb.shape[0],
                                         out.data_ptr(), stream), "glp_vq_assign_device")

    secs, _ = time_steps(torch, code, steps, 1)
    ms = secs / steps * 1e3
    t0 = time.perf_counter()
    host_codes = host.vq_codes(gls, cb)
    host_codes_s = time.perf_counter() - t0
    same_codes = bool((out.cpu().numpy() == host_codes.reshape(-1)).all())
    # GLPack::GetPackedGLs of those codes on the device ([M/4][N] words)
    N, M = gls.shape[0], gls.shape[1]
    packed = torch.empty(((M + 3) // 4, N), dtype=torch.int32, device="cuda:0")

    def pack():
        check(lib().glp_pack_codes_device(out.data_ptr(), N, M, 8, 0, N, packed.data_ptr(),
                                          stream), "glp_pack_codes_device")

    psecs, _ = time_steps(torch, pack, steps, 1)
    pack_ms = psecs / steps * 1e3
    t0 = time.perf_counter()
    host_packed = host.pack_codes(host_codes, 8)
    host_pack_s = time.perf_counter() - t0
    same_packed = bool((packed.cpu().numpy().view(np.uint32) == host_packed).all())
    log(f"GL packing: {pack_ms:.3f} ms device / {host_pack_s * 1e3:.1f} ms host "
        f"(identical: {same_packed}), {2 * n / (pack_ms / 1e3) / 1e9:.0f} GB/s")
    log(f"VQ: training on all {n} GLs {w['vq_train_s']:.2f}s ({w['vq_iters']} iterations, "
        f"device search); on {VQ_TRAINING} GLs {host_s:.2f}s host / {dev_s:.2f}s device "
        f"(identical: {same_cb}); codes of {n} GLs {ms:.3f} ms device / "
        f"{host_codes_s * 1e3:.1f} ms host (identical: {same_codes})")
    return {"codes_ms": ms, "unit": "distances/s", "value": n * cb.shape[0] / (ms / 1e3),
            "gls": n, "codewords": int(cb.shape[0]), "codes_identical": same_codes,
            "host_codes_ms": host_codes_s * 1e3,
            "pack": {"ms": pack_ms, "host_ms": host_pack_s * 1e3, "identical": same_packed,
                     "bytes_per_s": 2 * n / (pack_ms / 1e3)},
            "train": {"training_gls": n, "iterations": w["vq_iters"], "device_s": w["vq_train_s"],
                      "subsample_check": {"training_gls": VQ_TRAINING, "iterations": itd,
                                          "device_s": dev_s, "host_s": host_s,
                                          "identical": same_cb}}}


TRAFFIC_FILES = ("k_batched_hbm_traffic.md", "k_batched_fb_hbm_traffic.md")


def recorded_traffic(workload):
    """HBM bytes per dispatch of the batched kernel for this workload (plan
    key "NxMxC[ fb] qQ wW"), from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summaries of the forward-only and forward-backward passes
    (scaled by the copy-kernel calibration), or None."""
    prefix = "traffic_bytes_per_dispatch:"
    for name in TRAFFIC_FILES:
        try:
            with open(os.path.join(ROOT, "profiles", name)) as f:
                for line in f:
                    if line.startswith(prefix) and "=" in line:
                        key, val = line[len(prefix):].rsplit("=", 1)
                        if key.strip() == workload:
                            return float(val)
        except (OSError, ValueError):
            continue
    return None


def copy_calibration(nbytes):
    """A plain device-to-device copy of known size, for scaling the HBM byte
    counters (FETCH_SIZE / WRITE_SIZE) a profiler reports for the MH kernels."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "read_bytes": nbytes,
                      "written_bytes": nbytes, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--samples", type=int, default=5000)
    ap.add_argument("--sites", type=int, default=4096)
    ap.add_argument("--cycles", type=int, default=64)
    ap.add_argument("--depth", type=float, default=2.0)
    ap.add_argument("--lanes", type=int, default=0, help="lanes per chain (0 = auto)")
    ap.add_argument("--waves", type=int, default=0,
                    help="wavefronts per workgroup with --lanes (4 or 8; 0 = the plan's default)")
    ap.add_argument("--split", type=int, choices=(0, 1), default=0,
                    help="with --lanes and --waves 8: wavefronts 4-7 of each workgroup hold one "
                         "chain slot fewer (glp_plan_set_split)")
    ap.add_argument("--precision", choices=("fb", "mixed"), default="fb",
                    help="fb: two half-window passes joined in FP64 (GLP_PREC_MIXED_FB); "
                         "mixed: the forward-only pass (GLP_PREC_MIXED)")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-aux", action="store_true", help="skip the kNN and VQ measurements")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES bytes (counter scaling)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch

    from glphase_amd import PREC_MIXED, PREC_MIXED_FB
    from glphase_amd.engine import Job

    N, M, C = args.samples, args.sites, args.cycles
    pen, seed = 1.0, 20240601
    prec = PREC_MIXED_FB if args.precision == "fb" else PREC_MIXED
    w = make_workload(N, M, C, args.depth, args.seed)
    win = dict(panel=w["panel"], packed=w["packed"], tran=w["tran"])
    job = Job([win], N, w["H"], w["codebook"], w["pc"], bits=8, num_cycles=C, precision=prec)
    lanes, grid = job.make_plan(args.lanes)
    if args.lanes and args.waves:
        job.set_waves(args.waves)
        job.set_split(args.split)
        lanes, grid = job.lanes()
    if not args.lanes:
        t0 = time.time()
        tuned = job.autotune(gen=0, pen=pen, seed=seed)
        lanes, grid = job.lanes()
        log(f"autotune: lanes/chain {list(tuned)}, {job.waves()} wavefronts per workgroup"
            f"{' (split chain slots)' if job.split() else ''} "
            f"in {time.time() - t0:.2f}s")
    se_step = N * (C + 1) * M
    gen = [0]

    def step():
        job.run_batched(gen=gen[0], pen=pen, seed=seed)
        gen[0] += 1

    log(f"batched: lanes/chain={lanes} grid={grid}")
    secs, wall = time_steps(torch, step, args.steps, args.warmup)
    ms = secs / args.steps * 1e3
    evals, useful = job.stats()
    value = se_step * args.steps / secs
    log(f"batched: {ms:.3f} ms/step, {value:.3e} SE/s (wall {wall:.3f}s); "
        f"evaluated {evals} chain passes for {useful} useful")

    # the same generation through the plan's HIP graph
    ggen = [0]

    def gstep():
        job.run_batched_graph(gen=ggen[0], pen=pen, seed=seed)
        ggen[0] += 1

    gsecs, _ = time_steps(torch, gstep, args.steps, 1)
    gms = gsecs / args.steps * 1e3
    log(f"batched (HIP graph): {gms:.3f} ms/step")

    # haplotype resampling (Impute::hmm_work) after the generation, and the
    # whole generation (MH update + resampling)
    from glphase_amd.engine import Resampler
    rs = Resampler(job, [w["het_aa"]], w["pc"])
    wgen = [0]

    def wstep():
        rs.run(pen=pen, seed=seed, gen=wgen[0])
        wgen[0] += 1

    wsecs, _ = time_steps(torch, wstep, args.steps, 1)
    wms = wsecs / args.steps * 1e3
    log(f"hmm_work: {wms:.3f} ms/step")

    # a whole sampling generation as Insti::estimate runs it: MH update,
    # hmm_work, the new haplotypes into the panel, Impute::replace's tallies
    from glphase_amd._lib import check as lib_check, lib as glp_lib
    hsum = torch.zeros((2 * N, M), dtype=torch.int32, device="cuda:0")
    fgen = [0]

    def fstep():
        job.run_batched(gen=fgen[0], pen=pen, seed=seed)
        rs.run(pen=pen, seed=seed, gen=fgen[0])
        rs.adopt()
        lib_check(glp_lib().glp_replace(job.keep[0].data_ptr(), N, M, hsum.data_ptr(),
                                        job.stream_handle()), "glp_replace")
        fgen[0] += 1

    fsecs, _ = time_steps(torch, fstep, args.steps, 1)
    fms = fsecs / args.steps * 1e3
    log(f"full generation (MH + hmm_work + adopt + replace): {fms:.3f} ms/step")

    # plain path on the same inputs: one launch per MH iteration, in the
    # headline's arithmetic and in the forward-only one
    def plain_rate(precision):
        job_p = Job([win], N, w["H"], w["codebook"], w["pc"], bits=8, num_cycles=C,
                    precision=precision)
        job_p.make_plan(1)
        gp = [0]

        def pstep():
            job_p.run_plain(gen=gp[0], pen=pen, seed=seed)
            gp[0] += 1

        psecs, _ = time_steps(torch, pstep, args.plain_steps, 1)
        pg = [0]

        def pgstep():
            job_p.run_plain_graph(gen=pg[0], pen=pen, seed=seed)
            pg[0] += 1

        pgsecs, _ = time_steps(torch, pgstep, args.plain_steps, 1)
        job_p.close()
        return psecs / args.plain_steps * 1e3, pgsecs / args.plain_steps * 1e3

    pms, pgms = plain_rate(prec)
    pvalue = se_step / (pms / 1e3)
    log(f"plain: {pms:.3f} ms/step, {pvalue:.3e} SE/s; HIP graph {pgms:.3f} ms/step")
    forward_only = None
    if prec == PREC_MIXED_FB:
        # the forward-only mixed pass (GLP_PREC_MIXED) on the same inputs
        job_m = Job([win], N, w["H"], w["codebook"], w["pc"], bits=8, num_cycles=C,
                    precision=PREC_MIXED)
        job_m.make_plan(args.lanes)
        if not args.lanes:
            job_m.autotune(gen=0, pen=pen, seed=seed)
        mgen = [0]

        def mstep():
            job_m.run_batched(gen=mgen[0], pen=pen, seed=seed)
            mgen[0] += 1

        msecs, _ = time_steps(torch, mstep, args.steps, 1)
        mms = msecs / args.steps * 1e3
        mlanes = job_m.lanes()[0]
        job_m.close()
        mpms, _ = plain_rate(PREC_MIXED)
        forward_only = {"ms_per_step": mms, "value": se_step / (mms / 1e3),
                        "lanes_per_chain": mlanes, "plain_ms_per_step": mpms,
                        "plain_value": se_step / (mpms / 1e3),
                        "fb_speedup": mms / ms, "batched_fb_over_plain_forward_only": mpms / ms}
        log(f"forward-only mixed pass: batched {mms:.3f} ms/step (lanes {mlanes}), plain "
            f"{mpms:.3f} ms/step; the forward-backward form is {mms / ms:.2f}x faster")

    achieved = FLOP_PER_SE * se_step / (ms / 1e3) / 1e12
    workload_key = (f"{N}x{M}x{C}" + (" fb" if prec == PREC_MIXED_FB else "") +
                    f" q{lanes} w{job.waves()}" + (" split" if job.split() else ""))
    traffic = recorded_traffic(workload_key)
    out = {
        "metric": "HMM site-evals/s (sample x proposal x site), batched MH generation",
        "value": value,
        "unit": "site-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 recurrence, fp64 log accumulation" +
                 (" (forward and backward halves joined in fp64)" if prec == PREC_MIXED_FB else ""),
        "data": "synthetic founder-mosaic haplotypes and GLs",
        "config": {"workload": f"{N} samples x {M} sites, {C} proposals/sample, 8-bit VQ GLs, "
                               "batched persistent kernel (mixed precision" +
                               (", forward-backward form)" if prec == PREC_MIXED_FB else ")"),
                   "precision": args.precision,
                   "samples": N, "sites": M, "proposals": C, "lanes_per_chain": lanes,
                   "waves_per_workgroup": job.waves(), "split_slots": job.split(),
                   "vq_training_gls": N * M, "vq_iterations": w["vq_iters"]},
        "mh_iterations_per_s": 1e3 / ms,
        "proposals_per_s": N * C / (ms / 1e3),
        "plain": {"value": pvalue, "ms_per_step": pms, "launches_per_step": C + 1,
                  "batched_over_plain": value / pvalue, "graph_ms_per_step": pgms,
                  "roofline_frac": FLOP_PER_SE * se_step / (pms / 1e3) / 1e12 / PEAK_FP32_TFLOPS},
        "graph": {"batched_ms_per_step": gms, "plain_ms_per_step": pgms},
        "forward_only": forward_only,
        "hmm_work": {"ms_per_step": wms, "generation_ms": ms + wms, "full_generation_ms": fms,
                     "note": "Impute::hmm_work for all samples after the MH update"},
        "speculation": {"chain_passes": evals, "useful_passes": useful},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "flop_per_site_eval": FLOP_PER_SE,
                     "hbm_gbps": None if traffic is None else traffic / (ms / 1e3) / 1e9},
    }
    out["accuracy"] = accuracy_vs_fp64(w, C, pen, seed, lanes, prec)
    a = out["accuracy"]
    log(f"accuracy vs FP64 ({a['chains']} chains, {a['proposals']} proposals of generation "
        f"{a['generation']}): max {a['max_rel_err']:.2e}, p99 {a['p99_rel_err']:.2e}; accept "
        f"agreement {a['accept_agreement']:.6f} outside the band ({a['decisions_inside_band']} inside)")
    cpu = not args.no_cpu
    if not args.no_aux:
        out["vq"] = aux_vq(torch, w, args.steps)
        out["knn"] = aux_knn(torch, args.steps, cpu=cpu)
        out["batching"] = aux_batching(torch, args.steps, cpu=cpu)
        out["plain_small"] = aux_small(torch, args.steps, 1000, 1024, 16, 31, cpu=cpu)
        out["small_run"] = aux_small(torch, args.steps, 100, 1024, 200, 41, cpu=cpu)
        out["whole_run"] = aux_whole_run(torch, w, C)
    if cpu:
        out["cpu_baseline"] = cpu_baseline(w, C, pen, seed, precision=prec)
    if not args.no_aux:
        out["configs"] = config_table(out)
    print(json.dumps(out), flush=True)


def config_table(out):
    """The five benchmark workloads side by side: CPU oracle on one core and
    on every core, the GPU plain path and the batched form (site-evals/s),
    and the batched form's (config 2: the plain path's) fraction of the FP32
    VALU peak at 29 FLOP per site-eval."""
    def frac(v):
        return None if v is None else FLOP_PER_SE * v / 1e12 / PEAK_FP32_TFLOPS

    def row(cpu, plain, batched, primary):
        return {"cpu_1core": cpu and cpu["cpu_1core"], "cpu_allcore": cpu and cpu["cpu_allcore"],
                "gpu_plain": plain, "gpu_batched": batched,
                "roofline_frac": frac(batched if primary == "batched" else plain)}

    c1, c2, c3 = out["small_run"], out["plain_small"], out["batching"]
    c5 = out["knn"]["mh_klist"]
    cb = out.get("cpu_baseline")
    return {
        "1": dict(row(c1.get("cpu"), c1["value"], c1["batched"]["value"], "batched"),
                  workload="100 x 1024, C = 200, fp32"),
        "2": dict(row(c2.get("cpu"), c2["value"], c2["batched"]["value"], "plain"),
                  workload="1000 x 1024, C = 16, fp32, plain path"),
        "3": dict(row(c3.get("cpu"), c3["plain"]["value"], c3["batched"]["value"], "batched"),
                  workload="2500 x 32 windows (lognormal sites), C = 32, fp32"),
        "4": dict(row(cb and {"cpu_1core": cb["single_core"]["value"], "cpu_allcore": cb["value"]},
                      out["plain"]["value"], out["value"], "batched"),
                  workload="5000 x 4096, C = 64, mixed precision, forward-backward form"),
        "5": dict(row(c5.get("cpu"), c5["plain"]["value"], c5["value"], "batched"),
                  workload="10000 + 20000 panel x 2048, K = 200, C = 64, mixed precision, "
                           "forward-backward form"),
    }


if __name__ == "__main__":
    main()
