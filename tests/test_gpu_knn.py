"""GPU kNN candidate lists (glp_knn_build) against the oracle: the
reference's known answer, random tie-heavy panels, a reference-panel run at
10,000 samples + 20,000 panel haplotypes x 2,048 sites with K = 200, and the
MH update in K-list mode on lists built on the GPU."""
import sys

import numpy as np
import pytest

from glphase_amd import PREC_MIXED, PROP_KLIST
from tests import helpers
from tests.test_oracle_knn import (NUM_CLUSTERS, NUM_HAPS, NUM_SITES, _words, kat_haplotypes,
                                   reference_shuffle)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_knn(panel, M, nsamp, K, lb=0.0, ub=1.0, maf=False):
    from glphase_amd.engine import knn_lists
    return knn_lists(panel, M, nsamp, K, lb, ub, maf).cpu().numpy().view(np.uint32)


def test_knn_reference_known_answer():
    say("kNN known answer")
    perm = reference_shuffle()
    for second, lb, want in ((False, 0.0, [2, 15, 3, 14]), (True, 0.4375, [5, 6, 8, 2])):
        haps = kat_haplotypes(second)
        words = _words([haps[perm[j]] for j in range(NUM_HAPS)])
        k = gpu_knn(words, NUM_SITES, NUM_HAPS // 2, NUM_CLUSTERS, lb=lb)
        assert [int(perm[x]) for x in k[0]] == want


@pytest.mark.parametrize("H,M,nsamp,K,lb,ub,maf", [(40, 100, 20, 10, 0.0, 1.0, False),
                                                   (300, 700, 150, 40, 0.05, 0.5, True),
                                                   (64, 64, 10, 61, 0.0, 1.0, False),
                                                   (5000, 1000, 1200, 200, 0.01, 1.0, False),
                                                   # M >= 2,560: histogram bins of 2 distances,
                                                   # then the refining histogram in the bin
                                                   (600, 3000, 300, 30, 0.0, 1.0, False),
                                                   (402, 9000, 201, 24, 0.0, 1.0, True)])
def test_knn_matches_oracle(orc, H, M, nsamp, K, lb, ub, maf):
    say(f"kNN H={H} M={M} K={K}")
    rng = np.random.default_rng(H * M)
    wn = (M + 63) // 64
    # full 64-bit words: every bit position of a word, bit 63 included
    base = rng.integers(0, 2**64, (6, wn), dtype=np.uint64)
    noise = (rng.integers(0, 2**64, (H, wn), dtype=np.uint64)
             & rng.integers(0, 2**64, (H, wn), dtype=np.uint64)
             & rng.integers(0, 2**64, (H, wn), dtype=np.uint64))
    haps = base[rng.integers(0, 6, H)] ^ noise
    if M % 64:
        haps[:, -1] &= np.uint64((1 << (M % 64)) - 1)
    got = gpu_knn(haps, M, nsamp, K, lb, ub, maf)
    want = orc.knn(haps, M, nsamp, K, lb, ub, maf)
    assert (got == want).all()


def test_knn_reference_panel_size(orc):
    # config 5 shape: 10,000 samples plus 20,000 panel haplotypes, 2,048 sites
    say("kNN 10000 samples + 20000 panel, K=200")
    p = helpers.synthetic(10000, 2048, seed=3, n_ref=20000, training_size=1 << 14)
    got = gpu_knn(p["panel"], 2048, 10000, 200)
    assert got.shape == (10000, 200)
    want = orc.knn(p["panel"], 2048, 40, 200)
    assert (got[:40] == want).all()
    # every row, through the row-parallel restatement of the same lists
    rows = np.arange(10000, dtype=np.uint32)
    assert (got == orc.knn_rows(p["panel"], 2048, rows, 200, nthreads=16)).all()


def test_mh_on_gpu_built_lists(orc):
    say("MH on GPU-built K-lists")
    from glphase_amd.engine import Job
    N, M, C, K = 300, 600, 24, 40
    p = helpers.synthetic(N, M, seed=9, n_ref=200)
    klist = gpu_knn(p["panel"], M, N, K)
    assert (klist == orc.knn(p["panel"], M, N, K)).all()
    win = dict(panel=p["panel"], packed=p["packed"], tran=p["tran"])
    job = Job([win], N, p["H"], p["codebook"], p["pc"], bits=8, num_cycles=C,
              prop_mode=PROP_KLIST, klist=klist[None], precision=PREC_MIXED)
    job.make_plan(4)
    job.run_batched(gen=2, pen=1.0, seed=8)
    lut = orc.emit_lut(p["codebook"], p["pc"])
    mh = orc.MH(p["panel"], p["packed"], 8, lut, p["tran"], N)
    po, _ = mh.run(C, 1.0, 8, gen=2, prop_mode=1, klist=klist, precision=PREC_MIXED)
    assert (job.parents()[0] == po).all()
