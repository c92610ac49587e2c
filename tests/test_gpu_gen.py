"""Generation bookkeeping on the GPU through the C ABI -- initial haplotypes
(glp_init_haps) and allele tallies (glp_replace) bit-exact against the
oracle -- and a whole phasing run (glphase_amd.phase) on the reference's
example data set (samples/hapGen/ex.bin: 60 samples x 1,000 sites) scored
against its simulated truth (ex.vcf.gz)."""
import gzip
import os
import sys

import numpy as np
import pytest

from oracle import bind as orc

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "samples")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gl_pairs(M, N, seed):
    rng = np.random.default_rng(seed)
    het = rng.random((M, N)).astype(np.float32) * np.float32(0.7)
    aa = (rng.random((M, N)).astype(np.float32) * (np.float32(1) - het)).astype(np.float32)
    g = np.stack([het, aa], 2).astype(np.float32)
    g[::7, ::3] = (0, 0)
    g[::11, ::5] = (1, 0)
    return g


def gpu_init_haps(g, seed, window):
    import torch
    from glphase_amd._lib import check, lib
    M, N, _ = g.shape
    wn = (M + 63) // 64
    gd = torch.from_numpy(np.ascontiguousarray(g)).cuda()
    out = torch.full((2 * N + 3, wn), -1, dtype=torch.int64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    check(lib().glp_init_haps(gd.data_ptr(), N, M, seed, window, out.data_ptr(), s),
          "glp_init_haps")
    torch.cuda.synchronize()
    o = out.cpu().numpy().view(np.uint64)
    assert (o[2 * N:] == np.uint64(0xFFFFFFFFFFFFFFFF)).all()  # rows past 2N untouched
    return o[: 2 * N]


@pytest.mark.parametrize("N,M,seed,window", [(1, 1, 0, 0), (37, 130, 5, 3), (300, 1000, 9, 0),
                                             (5000, 4096, 4, 1)])
def test_init_haps_matches_oracle(N, M, seed, window):
    say(f"init haps N={N} M={M}")
    g = _gl_pairs(M, N, N + M)
    got = gpu_init_haps(g, seed, window)
    want = orc.init_haps(g, seed, window)
    assert (got == want).all()


def gpu_replace_rounds(haps_list, N, M, offset=0):
    # offset: the tallies start `offset` int32 into their buffer (offset 1
    # leaves them 4 bytes off 16-byte alignment, the per-site form)
    import torch
    from glphase_amd._lib import check, lib
    buf = torch.zeros(offset + 2 * N * M, dtype=torch.int32, device="cuda")
    hs = buf[offset:]
    s = torch.cuda.current_stream().cuda_stream
    for h in haps_list:
        hd = torch.from_numpy(h.view(np.int64)).cuda()
        check(lib().glp_replace(hd.data_ptr(), N, M, hs.data_ptr(), s), "glp_replace")
    torch.cuda.synchronize()
    return hs.cpu().numpy().view(np.uint32).reshape(2 * N, M)


# site counts that are multiples of 4 take the 16-byte form (4 sites per
# lane), with chunks of 256 sites whole, partial (1000, 2000, 5000, 260) or
# shorter than one (4, 100); the others the per-site form
@pytest.mark.parametrize("N,M,rounds", [(1, 1, 3), (13, 130, 6), (200, 1000, 8), (9, 1025, 4),
                                         (9, 2000, 3), (5000, 4096, 3), (7, 5000, 4), (3, 4, 5),
                                         (11, 100, 4), (5, 260, 4)])
def test_replace_matches_oracle(N, M, rounds, offset=0):
    say(f"replace N={N} M={M} offset={offset}")
    rng = np.random.default_rng(N + M)
    wn = (M + 63) // 64
    hl = []
    for r in range(rounds):
        h = rng.integers(0, 2**63, (2 * N, wn), dtype=np.uint64)
        if M % 64:
            h[:, -1] &= np.uint64((1 << (M % 64)) - 1)
        if r % 2:
            h = h.reshape(N, 2, wn)[:, ::-1].reshape(2 * N, wn).copy()
        hl.append(h)
    got = gpu_replace_rounds(hl, N, M, offset)
    want = np.zeros((2 * N, M), np.uint32)
    for h in hl:
        orc.replace(h, N, M, want)
    assert (got == want).all()


@pytest.mark.parametrize("N,M", [(13, 130), (11, 100), (5, 260), (9, 2000)])
def test_replace_unaligned_tallies(N, M):
    # tallies 4 bytes off 16-byte alignment: the per-site form at every size
    test_replace_matches_oracle(N, M, 4, offset=1)


def _truth(path):
    pos, gt = [], []
    with gzip.open(path, "rt") as f:
        for line in f:
            if line.startswith("#"):
                continue
            t = line.rstrip("\n").split("\t")
            pos.append(int(t[1]))
            gt.append([(int(x[0]), int(x[2])) for x in t[9:]])
    return np.array(pos), np.array(gt)  # [M], [M][N][2]


def test_phase_reference_example(tmp_path):
    say("phase ex.bin (60 samples x 1000 sites, 56 + 200 generations)")
    from glphase_amd import host
    from glphase_amd.phase import called_genotypes, phase_stbin
    out = tmp_path / "ex.phased.vcf.gz"
    d, hsum, run = phase_stbin(os.path.join(GOLD, "ex.bin"), out_vcf=str(out), seed=3,
                               command_line="test")
    say(f"setup {run.setup_s:.2f}s, {run.bn + run.sn} generations in {run.gen_s:.2f}s")
    N, M = len(d["names"]), len(d["pos"])
    assert hsum.shape == (2 * N, M)
    assert (hsum.reshape(N, 2, M).sum(1) <= 2 * run.sn).all()
    tpos, truth = _truth(os.path.join(GOLD, "ex.vcf.gz"))
    keep = np.isin(tpos, d["pos"])
    assert keep.sum() == M
    truth = truth[keep]
    a, b = called_genotypes(hsum, run.sn)
    dosage = a + b
    tdos = truth.sum(2)
    disc = float((dosage != tdos).mean())
    # the GLs alone: most likely genotype of (1 - Het - AA, Het, AA)
    g = d["gls"]
    argmax = np.argmax(np.stack([1 - g[..., 0] - g[..., 1], g[..., 0], g[..., 1]], 2), 2)
    disc_gl = float((argmax != tdos).mean())
    # switch errors between consecutive heterozygous sites called correctly
    sw, n = 0, 0
    for i in range(N):
        het = np.where((tdos[:, i] == 1) & (dosage[:, i] == 1))[0]
        phase = a[het, i] == truth[het, i, 0]
        sw += int((phase[1:] != phase[:-1]).sum())
        n += max(len(het) - 1, 0)
    say(f"genotype discordance {disc:.4f} (GL argmax {disc_gl:.4f}); switch error "
        f"{sw / max(n, 1):.4f} over {n} het pairs")
    assert disc < disc_gl
    assert sw / max(n, 1) < 0.2
    # the written VCF carries the same calls
    lines = [l for l in gzip.open(out, "rt").read().split("\n") if l and not l.startswith("#")]
    assert len(lines) == M
    for m in (0, M // 2, M - 1):
        calls = [c.split(":")[0] for c in lines[m].split("\t")[9:]]
        assert calls == [f"{x}|{y}" for x, y in zip(a[m], b[m])]
    run.close()


def test_phase_is_reproducible():
    say("phase reproducibility (ex.bin, short schedule)")
    from glphase_amd.phase import phase_stbin
    kw = dict(seed=5, burnin=6, sampling=4, sa_burnin=3)
    _, h1, r1 = phase_stbin(os.path.join(GOLD, "ex.bin"), **kw)
    _, h2, r2 = phase_stbin(os.path.join(GOLD, "ex.bin"), **kw)
    assert (h1 == h2).all()
    assert (r1.haplotypes() == r2.haplotypes()).all()
    r1.close()
    r2.close()


def test_phase_with_knn_lists():
    say("phase ex.bin with K = 20 candidate lists")
    from glphase_amd.phase import called_genotypes, phase_stbin
    d, hsum, run = phase_stbin(os.path.join(GOLD, "ex.bin"), seed=7, knn_k=20, recluster_every=20,
                               burnin=40, sampling=60, sa_burnin=20)
    tpos, truth = _truth(os.path.join(GOLD, "ex.vcf.gz"))
    truth = truth[np.isin(tpos, d["pos"])]
    a, b = called_genotypes(hsum, run.sn)
    g = d["gls"]
    argmax = np.argmax(np.stack([1 - g[..., 0] - g[..., 1], g[..., 0], g[..., 1]], 2), 2)
    disc = float(((a + b) != truth.sum(2)).mean())
    say(f"K-list run: discordance {disc:.4f} (GL argmax {float((argmax != truth.sum(2)).mean()):.4f})")
    assert disc < float((argmax != truth.sum(2)).mean())
    run.close()
