"""The oracle's kNN candidate lists (kNN.cpp:8-225) pinned by the
reference's own known answers (other_unittests.cpp:103-255, KNN.clustersOK):
neighbour order of sample 0, allele frequencies and the site threshold, under
the reference test's own shuffle of haplotypes 2..15 and under random
relabellings; plus an independent numpy ranking on random panels."""
import numpy as np
import pytest

from oracle import bind as orc

NUM_CLUSTERS, NUM_HAPS = 4, 16
NUM_SITES = NUM_HAPS + NUM_CLUSTERS * 2


def _words(haps):
    out = np.zeros((len(haps), 1), np.uint64)
    for h, sites in enumerate(haps):
        w = 0
        for s in sites:
            w |= 1 << s
        out[h, 0] = np.uint64(w)
    return out


def kat_haplotypes(second):
    haps = [set(range(i, i + NUM_CLUSTERS + 2)) for i in range(NUM_HAPS)]
    # hap 1 moved to the far end (other_unittests.cpp:127-131)
    haps[1] = set(range(NUM_SITES - 6, NUM_SITES))
    if second:
        haps[1] = set(range(1, 7))  # undone (164-168)
        for h in (1, 5, 6, 8, 9, 10):
            haps[h].add(0)
        for h in (5, 6, 8, 11, 12):
            haps[h].add(2)
        for h in (0, 1, 5, 7, 15, 14, 13):
            haps[h].add(NUM_SITES - 1)
        for h in (0, 1, 2, 3, 4, 5, 8):
            haps[h].add(NUM_SITES - 2)
    return haps


def reference_shuffle():
    """shuffledIndexes of other_unittests.cpp:114-137: std::srand(1), then
    libstdc++'s random_shuffle over positions 2..15 (for i = 1..n-1, swap
    element i with element rand() % (i + 1)) -- glibc's rand() is called
    through ctypes."""
    import ctypes
    import ctypes.util
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.srand(1)
    perm = list(range(NUM_HAPS))
    sub = perm[2:]
    for i in range(1, len(sub)):
        j = libc.rand() % (i + 1)
        sub[i], sub[j] = sub[j], sub[i]
    return np.array(perm[:2] + sub)


def _neighbours(perm, second, lb):
    haps = kat_haplotypes(second)
    words = _words([haps[perm[j]] for j in range(NUM_HAPS)])
    k = orc.knn(words, NUM_SITES, 1, NUM_CLUSTERS, lb=lb, ub=1.0)
    return [int(perm[x]) for x in k[0]], words


def test_reference_clusters_ok_exact():
    # with the reference's own shuffle: std::sort over 14 candidates is
    # libstdc++'s insertion sort, stable, so equal distances keep the
    # passHaps order -- the tie rule of this restatement
    perm = reference_shuffle()
    assert _neighbours(perm, False, 0.0)[0] == [2, 15, 3, 14]
    got, words = _neighbours(perm, True, 0.4375)
    assert got == [5, 6, 8, 2]
    mask, af = orc.knn_mask(words, NUM_SITES, lb=0.4375, ub=1.0)
    assert af[0] == 0.4375 and af[1] == 0.125 and af[2] == 0.5
    assert bin(int(mask[0])).count("1") == 4


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_reference_clusters_ok_relabelled(seed):
    # any labelling of haplotypes 2..15: the first case has no ties among the
    # chosen neighbours; in the second, haplotypes 2 and 6 tie at distance 2
    # behind 5 and 8, so only their order follows the labelling
    rng = np.random.default_rng(seed)
    perm = np.arange(NUM_HAPS)
    perm[2:] = 2 + rng.permutation(NUM_HAPS - 2)
    assert _neighbours(perm, False, 0.0)[0] == [2, 15, 3, 14]
    got, _ = _neighbours(perm, True, 0.4375)
    assert got[0] == 5 and got[2] == 8 and {got[1], got[3]} == {2, 6}
    first = 6 if list(perm).index(6) < list(perm).index(2) else 2
    assert got[1] == first


def numpy_knn(haps, M, nsamp, K, lb, ub, maf):
    H, wn = haps.shape
    bits = np.unpackbits(haps.view(np.uint8).reshape(H, -1), axis=1, bitorder="little")[:, :M]
    af = bits.sum(0) / H
    if maf:
        af = np.abs(np.abs(af - 0.5) - 0.5)
    keep = (af >= lb) & (af <= ub)
    K = min(K, H - 2)
    kh = (K + 1) // 2
    out = []
    for u in range(nsamp):
        others = np.array([h for h in range(H) if h // 2 != u])
        lists = []
        for q in (2 * u, 2 * u + 1):
            d = (bits[others] != bits[q])[:, keep].sum(1)
            lists.append(others[np.lexsort((others, d))])
        row = list(lists[0][:kh])
        b = [x for x in lists[1] if x not in set(row)][:kh]
        out.append([v for pair in zip(row, b) for v in pair])
    return np.array(out, np.uint32)


@pytest.mark.parametrize("H,M,K,lb,ub,maf", [(40, 100, 10, 0.0, 1.0, False),
                                             (60, 130, 7, 0.05, 0.5, True),
                                             (24, 64, 30, 0.1, 0.9, False)])
def test_knn_matches_numpy_ranking(H, M, K, lb, ub, maf):
    rng = np.random.default_rng(H + M)
    wn = (M + 63) // 64
    # low-entropy haplotypes so that many distances tie
    base = rng.integers(0, 2**63, (4, wn), dtype=np.uint64)
    haps = base[rng.integers(0, 4, H)] ^ (rng.integers(0, 2**63, (H, wn), dtype=np.uint64)
                                         & rng.integers(0, 2**63, (H, wn), dtype=np.uint64)
                                         & rng.integers(0, 2**63, (H, wn), dtype=np.uint64))
    if M % 64:
        haps[:, -1] &= np.uint64((1 << (M % 64)) - 1)
    nsamp = H // 2
    got = orc.knn(haps, M, nsamp, K, lb, ub, maf)
    assert (got == numpy_knn(haps, M, nsamp, K, lb, ub, maf)).all()
    assert got.shape[1] == 2 * ((min(K, H - 2) + 1) // 2)
    for u in range(nsamp):
        assert len(set(got[u].tolist())) == got.shape[1]
        assert all(int(x) // 2 != u for x in got[u])
