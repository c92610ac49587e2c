"""The oracle's generation bookkeeping: Impute::replace (impute.cpp:564-597)
against an independent numpy statement of the same tally rule, and
Insti::initialize's initial haplotypes (insti.cpp:1316-1365) against the
cases the GLs decide on their own (certain genotypes), the genotype
frequencies the GLs imply, and a per-site replay of the documented draws."""
import numpy as np
import pytest

from oracle import bind as orc


def _bits(haps, M):
    H = haps.shape[0]
    return np.unpackbits(haps.view(np.uint8).reshape(H, -1), axis=1, bitorder="little")[:, :M]


def numpy_replace(haps, N, M, hsum):
    b = _bits(haps, M).astype(np.int64)
    out = hsum.astype(np.int64).copy()
    for i in range(N):
        a, c = b[2 * i], b[2 * i + 1]
        ha, hb = out[2 * i], out[2 * i + 1]
        sis = int((a * ha).sum() + (c * hb).sum())
        tra = int((a * hb).sum() + (c * ha).sum())
        if sis > tra:
            out[2 * i] += a
            out[2 * i + 1] += c
        else:
            out[2 * i] += c
            out[2 * i + 1] += a
    return out.astype(np.uint32)


@pytest.mark.parametrize("N,M", [(1, 1), (5, 64), (13, 130), (40, 1000)])
def test_replace_matches_numpy_over_generations(N, M):
    rng = np.random.default_rng(N * M)
    wn = (M + 63) // 64
    hsum = np.zeros((2 * N, M), np.uint32)
    want = hsum.copy()
    total = 0
    for g in range(6):
        haps = rng.integers(0, 2**63, (2 * N, wn), dtype=np.uint64)
        if M % 64:
            haps[:, -1] &= np.uint64((1 << (M % 64)) - 1)
        if g % 2:  # swap each pair's orientation: the tally must follow it back
            haps = haps.reshape(N, 2, wn)[:, ::-1].reshape(2 * N, wn).copy()
        want = numpy_replace(haps, N, M, want)
        orc.replace(haps, N, M, hsum)
        total += int(_bits(haps, M).sum())
        assert (hsum == want).all()
    assert int(hsum.sum()) == total  # every allele counted once per generation


def test_replace_follows_orientation():
    # a sample whose haplotypes swap every generation keeps one orientation;
    # with empty tallies sis == tra, so the first pair lands swapped (hb)
    N, M = 1, 64
    a = np.array([[0xFFFF], [0x0]], np.uint64)
    hsum = np.zeros((2, M), np.uint32)
    for g in range(5):
        h = a if g % 2 == 0 else a[::-1].copy()
        orc.replace(h, N, M, hsum)
    assert (hsum[1, :16] == 5).all() and hsum[1, 16:].sum() == 0 and (hsum[0] == 0).all()


def test_init_haps_certain_genotypes():
    M, N = 200, 3
    g = np.zeros((M, N, 2), np.float32)
    g[:, 0] = (0, 0)  # RR certain
    g[:, 1] = (0, 1)  # AA certain
    g[:, 2] = (1, 0)  # het certain
    b = _bits(orc.init_haps(g, seed=7), M)
    assert b[0].sum() == 0 and b[1].sum() == 0
    assert b[2].all() and b[3].all()
    assert (b[4] + b[5] == 1).all()
    assert 60 < b[4].sum() < 140  # random phase


def test_init_haps_frequencies_follow_gls():
    rng = np.random.default_rng(1)
    M, N = 2000, 150
    het = rng.random((M, N)).astype(np.float32) * np.float32(0.6)
    aa = (rng.random((M, N)).astype(np.float32) * (np.float32(1) - het)).astype(np.float32)
    g = np.stack([het, aa], 2).astype(np.float32)
    b = _bits(orc.init_haps(g, seed=3), M)
    a, c = b[0::2].T, b[1::2].T  # [M][N]
    t0 = np.maximum(1 - het - aa, 0)
    n = M * N
    assert abs((a + c == 0).mean() - t0.mean()) < 4 * np.sqrt(0.25 / n)
    assert abs((a + c == 1).mean() - het.mean()) < 4 * np.sqrt(0.25 / n)
    assert abs((a + c == 2).mean() - aa.mean()) < 4 * np.sqrt(0.25 / n)
    hets = a + c == 1
    assert abs(a[hets].mean() - 0.5) < 4 * np.sqrt(0.25 / hets.sum())


def test_init_haps_replays_documented_draws():
    rng = np.random.default_rng(5)
    M, N, seed, window = 70, 4, 11, 2
    het = rng.random((M, N)).astype(np.float32) * np.float32(0.5)
    aa = rng.random((M, N)).astype(np.float32) * np.float32(0.5)
    g = np.stack([het, aa], 2).astype(np.float32)
    b = _bits(orc.init_haps(g, seed, window), M)
    for i in range(N):
        for m in range(M):
            x = orc.draw(seed, 0xFFFFFFFE, i, m, window).astype(np.float64) / 2.0**32
            t0 = max(np.float32(1) - het[m, i] - aa[m, i], np.float32(0))
            t1, t2 = het[m, i], aa[m, i]
            if x[0] < t0:
                want = (0, 0)
            elif x[1] < np.float32(t1 / (t1 + t2)):
                want = (1, 0) if x[2] < 0.5 else (0, 1)
            else:
                want = (1, 1)
            assert (b[2 * i, m], b[2 * i + 1, m]) == want, (i, m)
