"""Benchmark of the GLPhase MH surrogate-parent update on one MI355X.

A step is one MH generation over every sample: 4 fresh parents scored, then
C proposals each scored and accepted or rejected (Insti::estimate's per-
generation work, insti.cpp:1668-1705).  Site-evals (SE) per step =
N * (C + 1) * M: one forward-recurrence step of one chain at one site.

Default workload: 5,000 samples x 4,096 sites, 64 proposals per sample,
synthetic founder-mosaic haplotypes and Poisson(2)-depth GLs, 8-bit VQ codes,
batched persistent kernel with FP32 recurrence and FP64 log accumulation.
The plain path (one launch per MH iteration) runs on the same inputs, and the
CPU oracle (C, -O3) is timed on a bounded sample of the same chains.

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
    cb, iters = host.vq_train(d["gls"], 256, key=seed, training_size=VQ_TRAINING)
    vq_host_s = time.perf_counter() - t1
    codes = host.vq_codes(d["gls"], cb)
    packed = host.pack_codes(codes, 8)
    log(f"workload N={N} M={M} C={C}: generated in {time.time() - t0:.1f}s (VQ {iters} iterations)")
    return dict(panel=panel, packed=packed, tran=tran, codebook=cb, pc=pc, N=N, M=M,
                H=panel.shape[0], het_aa=np.ascontiguousarray(d["gls"][:, :, 1:3]),
                gls=d["gls"], codes=codes, vq_iters=iters, vq_host_s=vq_host_s, seed=seed,
                truth=d["truth"])


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


def aux_knn(torch, steps, seed=11):
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
    mh = aux_reference_panel_mh(torch, d, panel, out, steps, seed)
    return {"ms_per_build": ms, "unit": "word-compares/s", "value": words / (ms / 1e3),
            "mh_klist": mh,
            "config": {"samples": N, "panel_haplotypes": R, "sites": M, "K": K},
            "identical_to_oracle_rows": ns if same else 0,
            "cpu_baseline": {"value": 2 * ns * H * wn / cpu_s, "unit": "word-compares/s",
                             "cores": 1, "kind": "port",
                             "sample": f"oracle kNN lists of {ns} samples ({cpu_s:.2f}s)"}}


def aux_reference_panel_mh(torch, d, panel, klist, steps, seed, C=64):
    """The MH update of the reference-panel workload: 10,000 samples whose
    candidates are their K = 200 nearest haplotypes of panel + samples
    (lists from the kNN build), 64 proposals each, mixed precision,
    autotuned depth.  Cost unit: the headline's site-eval."""
    from glphase_amd import PREC_MIXED, PROP_KLIST, host
    from glphase_amd.engine import Job
    N, M = d["gls"].shape[0], d["gls"].shape[1]
    tran, pc, _ = host.model_params(2 * N, d["r_morgans"])  # hn = 2N (insti.cpp:1205)
    cb, _ = host.vq_train(d["gls"], 256, key=seed, training_size=VQ_TRAINING, device=True)
    packed = host.pack_codes(host.vq_codes(d["gls"], cb, device=True), 8)
    job = Job([dict(panel=panel, packed=packed, tran=tran)], N, panel.shape[0], cb, pc, bits=8,
              num_cycles=C, prop_mode=PROP_KLIST, klist=klist.cpu().numpy().view(np.uint32)[None],
              precision=PREC_MIXED)
    job.make_plan(0)
    lanes = int(job.autotune(gen=0, pen=1.0, seed=seed)[0])
    gen = [0]

    def step():
        job.run_batched(gen=gen[0], pen=1.0, seed=seed)
        gen[0] += 1

    secs, _ = time_steps(torch, step, steps, 1)
    ms = secs / steps * 1e3
    se = N * (C + 1) * M
    log(f"reference-panel MH (K-lists): {ms:.3f} ms/step, {se / (ms / 1e3):.3e} SE/s, "
        f"lanes/chain {lanes}")
    job.close()
    return {"ms_per_step": ms, "value": se / (ms / 1e3), "unit": "site-evals/s",
            "lanes_per_chain": lanes,
            "config": {"samples": N, "panel_haplotypes": panel.shape[0] - 2 * N, "sites": M,
                       "proposals": C, "K": int(klist.shape[1])}}


def aux_batching(torch, steps, seed=21, N=2500, W=32, C=32):
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
           "lanes_per_window": tuned, "unit": "site-evals/s"}
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


def aux_plain_small(torch, steps, seed=31, N=1000, M=1024, C=16):
    """The plain path at 1,000 samples x 1,024 sites: one launch per MH
    iteration, one proposal per sample per launch (C launches after the
    initial one), FP32 arithmetic.  Per-launch time and site-evals/s."""
    from glphase_amd import PREC_F32
    p = make_workload(N, M, C, 4.0, seed)
    from glphase_amd.engine import Job
    job = Job([dict(panel=p["panel"], packed=p["packed"], tran=p["tran"])], N, p["H"],
              p["codebook"], p["pc"], bits=8, num_cycles=C, precision=PREC_F32)
    job.make_plan(1)
    gen = [0]

    def step():
        job.run_plain(gen=gen[0], pen=1.0, seed=seed)
        gen[0] += 1

    secs, _ = time_steps(torch, step, steps, 1)
    ms = secs / steps * 1e3
    se = N * (C + 1) * M
    log(f"plain 1000 x 1024: {ms / (C + 1):.4f} ms/launch, {se / (ms / 1e3):.3e} SE/s")
    job.close()
    return {"ms_per_launch": ms / (C + 1), "launches_per_step": C + 1, "value": se / (ms / 1e3),
            "unit": "site-evals/s",
            "config": {"samples": N, "sites": M, "proposals_per_launch": 1, "precision": "fp32"}}


def aux_vq(torch, w, steps):
    """GL codebook training (GLVQ) and coding of the MH workload's GLs with
    the nearest-codeword search on the device.  Coding cost unit: one
    (GL, codeword) distance; the GLs are resident on the device."""
    from glphase_amd import host
    from glphase_amd._lib import c_f32p, check, lib, ptr
    gls, cb = w["gls"], w["codebook"]
    t0 = time.perf_counter()
    cbd, itd = host.vq_train(gls, 256, key=w["seed"], training_size=VQ_TRAINING, device=True)
    dev_s = time.perf_counter() - t0
    same_cb = bool(cbd.tobytes() == cb.tobytes() and itd == w["vq_iters"])
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
    same_codes = bool((out.cpu().numpy() == w["codes"].reshape(-1)).all())
    t0 = time.perf_counter()
    host.vq_codes(gls, cb)
    host_codes_s = time.perf_counter() - t0
    log(f"VQ: training on {VQ_TRAINING} GLs {w['vq_host_s']:.2f}s host / {dev_s:.2f}s device "
        f"(identical: {same_cb}); codes of {n} GLs {ms:.3f} ms device / "
        f"{host_codes_s * 1e3:.1f} ms host (identical: {same_codes})")
    return {"codes_ms": ms, "unit": "distances/s", "value": n * cb.shape[0] / (ms / 1e3),
            "gls": n, "codewords": int(cb.shape[0]), "codes_identical": same_codes,
            "host_codes_ms": host_codes_s * 1e3,
            "train": {"training_gls": VQ_TRAINING, "iterations": itd,
                      "device_s": dev_s, "host_s": w["vq_host_s"], "identical": same_cb}}


def recorded_traffic(workload):
    """HBM bytes per dispatch of the batched kernel for this workload, from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the
    copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "k_batched_hbm_traffic.md")
    prefix = "traffic_bytes_per_dispatch:"
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(prefix) and "=" in line:
                    key, val = line[len(prefix):].rsplit("=", 1)
                    if key.strip() == workload:
                        return float(val)
    except (OSError, ValueError):
        return None
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

    from glphase_amd import PREC_MIXED
    from glphase_amd.engine import Job

    N, M, C = args.samples, args.sites, args.cycles
    pen, seed = 1.0, 20240601
    w = make_workload(N, M, C, args.depth, args.seed)
    win = dict(panel=w["panel"], packed=w["packed"], tran=w["tran"])
    job = Job([win], N, w["H"], w["codebook"], w["pc"], bits=8, num_cycles=C,
              precision=PREC_MIXED)
    lanes, grid = job.make_plan(args.lanes)
    if not args.lanes:
        t0 = time.time()
        tuned = job.autotune(gen=0, pen=pen, seed=seed)
        lanes, grid = job.lanes()
        log(f"autotune: lanes/chain {list(tuned)} in {time.time() - t0:.2f}s")
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

    # plain path on the same inputs: one launch per MH iteration
    job_p = Job([win], N, w["H"], w["codebook"], w["pc"], bits=8, num_cycles=C, precision=PREC_MIXED)
    job_p.make_plan(1)
    gp = [0]

    def pstep():
        job_p.run_plain(gen=gp[0], pen=pen, seed=seed)
        gp[0] += 1

    psecs, _ = time_steps(torch, pstep, args.plain_steps, 1)
    pms = psecs / args.plain_steps * 1e3
    pvalue = se_step * args.plain_steps / psecs
    log(f"plain: {pms:.3f} ms/step, {pvalue:.3e} SE/s")

    pg = [0]

    def pgstep():
        job_p.run_plain_graph(gen=pg[0], pen=pen, seed=seed)
        pg[0] += 1

    pgsecs, _ = time_steps(torch, pgstep, args.plain_steps, 1)
    pgms = pgsecs / args.plain_steps * 1e3
    log(f"plain (HIP graph): {pgms:.3f} ms/step")

    achieved = FLOP_PER_SE * se_step / (ms / 1e3) / 1e12
    workload_key = f"{N}x{M}x{C}"
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
        "dtype": "fp32 recurrence, fp64 log accumulation",
        "data": "synthetic founder-mosaic haplotypes and GLs",
        "config": {"workload": f"{N} samples x {M} sites, {C} proposals/sample, 8-bit VQ GLs, "
                               "batched persistent kernel (mixed precision)",
                   "samples": N, "sites": M, "proposals": C, "lanes_per_chain": lanes},
        "mh_iterations_per_s": 1e3 / ms,
        "proposals_per_s": N * C / (ms / 1e3),
        "plain": {"value": pvalue, "ms_per_step": pms, "launches_per_step": C + 1,
                  "batched_over_plain": value / pvalue, "graph_ms_per_step": pgms},
        "graph": {"batched_ms_per_step": gms, "plain_ms_per_step": pgms},
        "hmm_work": {"ms_per_step": wms, "generation_ms": ms + wms, "full_generation_ms": fms,
                     "note": "Impute::hmm_work for all samples after the MH update"},
        "speculation": {"chain_passes": evals, "useful_passes": useful},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "flop_per_site_eval": FLOP_PER_SE,
                     "hbm_gbps": None if traffic is None else traffic / (ms / 1e3) / 1e9},
    }
    if not args.no_aux:
        out["vq"] = aux_vq(torch, w, args.steps)
        out["knn"] = aux_knn(torch, args.steps)
        out["batching"] = aux_batching(torch, args.steps)
        out["plain_small"] = aux_plain_small(torch, args.steps)
        out["whole_run"] = aux_whole_run(torch, w, C)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, C, pen, seed)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
