"""A whole phasing run on the device: Insti::initialize + Insti::estimate's
simulated-annealing MH schedule (insti.cpp:1195-1410, 1557-1728) around the
HIP kernels of libglphase_amd.so, and the VCF of Insti::save_vcf.

Per generation g of bn + sn (defaults 56 + 200, main.cpp:45, insti.cpp:64-65):
  pen = min((g + 1) / sa_burnin, 1)^2                      (insti.cpp:1668-1671)
  MH update of every sample, C proposals, pen in the accept test
  Impute::hmm_work for every sample (annealing exponent pen)
  the new haplotypes replace the old ones                  (insti.cpp:1455)
  from generation bn on, Impute::replace tallies them      (insti.cpp:1706-1708)
and at the end the tallies over the sn sampling generations give the output's
genotype and allele posteriors (Impute::result, Insti::save_vcf).

Everything per generation runs on one HIP stream; the host only sequences
launches.  Random draws are Philox streams keyed by `seed` (the reference's
GSL/curand draws cannot be reproduced).
"""
import ctypes
import time

import numpy as np

from . import host
from ._lib import PREC_MIXED_FB, PROP_KLIST, PROP_UNIFORM, check, lib
from .engine import Job, Resampler, _dev, _torch, knn_lists


def genotype_trios(het_aa):
    """Insti::initialize's per-site genotype probabilities (insti.cpp:1345-1349)
    from (Het, AA) pairs [M][N][2]: (max(1 - Het - AA, 0), max(Het, 0),
    max(AA, 0)), returned sample-major [N][M][3] float32 (GLPack's input)."""
    g = np.asarray(het_aa, np.float32)
    het, aa = g[..., 0], g[..., 1]
    r = np.float32(1) - het - aa
    t = np.stack([np.where(r < 0, np.float32(0), r), np.where(het < 0, np.float32(0), het),
                  np.where(aa < 0, np.float32(0), aa)], axis=-1).astype(np.float32)
    return np.ascontiguousarray(t.transpose(1, 0, 2))


class PhaseRun:
    """One window of N samples x M sites (plus optional reference haplotypes).

    het_aa: [M][N][2] GL pairs (host.read_stbin's drop_first layout).
    pos: [M] site positions (transitions without a genetic map) or tran: [M][3].

    kickstart=True draws the first proposal of every generation from ref_haps
    only (Insti::s_bKickStartFromRef, insti.cpp:1510-1513; it needs them).

    knn_k > 0 restricts candidates to K-lists built on the device from the
    current panel (sample haplotypes plus ref_haps) at generation sa_burnin
    and, with recluster_every, every n sampling generations.  This is an
    extension: the reference builds its kNN sampler only from a scaffold at
    the start of Insti::estimate and stops with "kNN sampler with no
    scaffold is not implemented yet" without one (insti.cpp:1578-1586).
    """

    def __init__(self, het_aa, pos=None, tran=None, ref_haps=None, bits=8, cycles=None,
                 burnin=56, sampling=200, sa_burnin=28, seed=1, knn_k=0, recluster_every=0,
                 training_size=0, precision=PREC_MIXED_FB, lanes=0, kickstart=False,
                 device="cuda:0"):
        torch = _torch()
        self.torch = torch
        g = np.ascontiguousarray(het_aa, np.float32)
        M, N, _ = g.shape
        self.N, self.M = N, M
        self.wn = (M + 63) // 64
        R = 0 if ref_haps is None else int(ref_haps.shape[0])
        self.H = 2 * N + R
        self.bn, self.sn, self.sa = burnin, sampling, sa_burnin
        self.C = cycles if cycles is not None else 2 * N  # nn * in (insti.cpp:261-264)
        self.seed = seed
        self.knn_k, self.recluster_every = knn_k, recluster_every
        t0 = time.perf_counter()
        if tran is None:
            tran = host.tran_from_positions(pos, 2 * N)
        # mu from the sample haplotypes only: hn = 2N (insti.cpp:1205,
        # 1256-1270); reference haplotypes only widen the candidate range
        _, pc, _ = host.model_params(2 * N, np.zeros(M))
        trios = genotype_trios(g)
        cb, self.vq_iters = host.vq_train(trios, 1 << bits, key=seed, training_size=training_size,
                                          device=True)
        codes = host.vq_codes(trios, cb, device=True)
        packed = host.pack_codes(codes, bits)
        panel = np.zeros((self.H, self.wn), np.uint64)
        if R:
            panel[2 * N:] = ref_haps
        self.job = Job([dict(panel=panel, packed=packed, tran=tran)], N, self.H, cb, pc, bits=bits,
                       num_cycles=self.C, precision=precision,
                       prop_mode=PROP_KLIST if knn_k else PROP_UNIFORM,
                       klist=np.zeros((1, N, max(2, self._kout())), np.uint32) if knn_k else None,
                       kickstart=kickstart, device=device)
        self.panel_d = self.job.keep[0]
        self.het_aa_d = _dev(torch, g, self.job.device)
        # initial haplotypes (insti.cpp:1316-1365) straight into the device panel
        check(lib().glp_init_haps(self.het_aa_d.data_ptr(), N, M, seed, 0,
                                  self.panel_d.data_ptr(), self.job.stream_handle()),
              "glp_init_haps")
        self.job.make_plan(lanes)
        self.uniform_plan = self.job.plan_with(PROP_UNIFORM, lanes) if knn_k else None
        self.rs = Resampler(self.job, [g.transpose(1, 0, 2)], pc)
        self.hsum_d = torch.zeros((2 * N, M), dtype=torch.int32, device=self.job.device)
        self.codebook, self.tran, self.pc = cb, tran, pc
        self.setup_s = time.perf_counter() - t0
        self.gen_s = None

    def _kout(self):
        return lib().glp_knn_k_out(self.knn_k, self.H)

    def _recluster(self):
        kl = knn_lists(self.panel_d, self.M, self.N, self.knn_k, device=self.job.device)
        self.job.set_klist(kl)

    def run(self, progress=None):
        """Every generation; returns the allele tallies [2N][M] (uint32)."""
        torch = self.torch
        t0 = time.perf_counter()
        stream = self.job.stream_handle()
        for g in range(self.bn + self.sn):
            pen = min((g + 1.0) / self.sa, 1.0)
            pen = float(np.float32(pen) * np.float32(pen))
            plan = None
            if self.knn_k:
                if g == self.sa or (self.recluster_every and g > self.bn
                                    and (g - self.bn) % self.recluster_every == 0):
                    self._recluster()
                # uniform candidates before the clustering generation (insti.cpp:1686-1691)
                plan = None if g >= self.sa else self.uniform_plan
            self.job.run_batched(gen=g, pen=pen, seed=self.seed, plan=plan)
            self.rs.run(pen=pen, seed=self.seed, gen=g)
            self.rs.adopt()
            if g >= self.bn:
                check(lib().glp_replace(self.panel_d.data_ptr(), self.N, self.M,
                                        self.hsum_d.data_ptr(), stream), "glp_replace")
            if progress and (g + 1) % progress == 0:
                torch.cuda.synchronize(self.job.device)
                print(f"generation {g + 1}/{self.bn + self.sn}", flush=True)
        torch.cuda.synchronize(self.job.device)
        self.gen_s = time.perf_counter() - t0
        return self.hsum_d.cpu().numpy().view(np.uint32)

    def haplotypes(self):
        """Current sample haplotypes [2N][wn] (uint64)."""
        self.torch.cuda.synchronize(self.job.device)
        return self.panel_d[: 2 * self.N].cpu().numpy().view(np.uint64)

    def close(self):
        self.job.close()


def phase_stbin(path, out_vcf=None, region=None, command_line="", **kw):
    """Reads a SNPTools .bin file, phases it on the device and (optionally)
    writes the VCF; returns (reader dict, tallies, PhaseRun)."""
    d = host.read_stbin(path, region=region)
    run = PhaseRun(d["gls"], pos=d["pos"], **kw)
    hsum = run.run()
    if out_vcf:
        host.write_vcf(out_vcf, d["names"], d["chrom"], d["pos"], d["ref"], d["alt"], hsum, run.sn,
                       command_line)
    return d, hsum, run


def called_genotypes(hsum, sn):
    """GT of Insti::save_vcf from tallies [2N][M]: ([M][N] allele of hap a,
    [M][N] allele of hap b)."""
    norm = np.float32(1.0 / sn)
    p0 = hsum[0::2].astype(np.float32) * norm
    p1 = hsum[1::2].astype(np.float32) * norm
    one = np.float32(1)
    prr, pra, paa = (one - p0) * (one - p1), (one - p0) * p1 + p0 * (one - p1), p0 * p1
    rr = (prr >= pra) & (prr >= paa)
    het = ~rr & (pra >= prr) & (pra >= paa)
    a = np.where(rr, 0, np.where(het, (p0 > p1).astype(int), 1))
    b = np.where(rr, 0, np.where(het, (p0 <= p1).astype(int), 1))
    return a.T, b.T
