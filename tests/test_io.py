"""Data formats around the phasing run (host code of libglphase_amd.so): the
SNPTools .bin reader against the reference's own known answers
(glReader_tests.cpp:16-140,144-185 on the reference's sample file), its error
behaviour (glReader.cpp:215-285, bio.cpp:6-75), the position-based
transitions (insti.cpp:1253-1300) and the VCF writer (insti.cpp:2098-2167)
on tallies whose genotype posteriors and Phred values are worked out by hand."""
import gzip
import math
import os

import numpy as np
import pytest

from glphase_amd import host
from glphase_amd._lib import GlpError
from oracle import bind as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden", "samples")
ONLY_THREE = os.path.join(GOLD, "20_011976121_012173018.bin.onlyThree.bin")
EX_BIN = os.path.join(GOLD, "ex.bin")


def _site7_and_first(d):
    g = d["gls"]
    return g[7].ravel().tolist(), g[0].ravel()


def test_stbin_reference_known_answers():
    # glReader_tests.cpp:18-70 (sampleBinTests) on loadsSTBin's file
    d = host.read_stbin(ONLY_THREE, drop_first=False)
    assert d["names"] == ["samp1", "samp2", "samp3"]
    assert (d["chrom"][7], int(d["pos"][7]), d["ref"][7], d["alt"][7]) == ("20", 11977595, "G", "A")
    site7, first = _site7_and_first(d)
    assert site7 == [0, 0, 1, 0, 1, 0, 1, 0, 0]
    assert first[0] == 0 and first[1] == 0 and first[2] == 1
    assert abs(first[3] - 0.8) < 1e-3 and abs(first[4] - 0.2) < 1e-3 and first[5] == 0
    assert first[6] == 1 and first[7] == 0 and first[8] == 0
    # ST_DROP_FIRST (glReader_tests.cpp:59-69)
    g = host.read_stbin(ONLY_THREE, drop_first=True)["gls"].ravel()
    assert g[0] == 0 and g[1] == 1 and abs(g[2] - 0.2) < 1e-3 and g[3] == 0
    assert g[4] == 0 and g[5] == 0


def test_stbin_region_known_answers():
    # glReader_tests.cpp:174-185 (loadsSTBinRegion -> sampleBinTestsFirst3)
    d = host.read_stbin(ONLY_THREE, region="20:11976121-11976813", drop_first=False)
    assert d["names"] == ["samp1", "samp2", "samp3"]
    assert len(d["pos"]) == 3
    assert 11977595 not in d["pos"].tolist()
    assert (d["chrom"][1], int(d["pos"][1]), d["ref"][1], d["alt"][1]) == ("20", 11976755, "C", "T")
    assert d["gls"][1].ravel().tolist() == [0, 0, 1, 0, 1, 0, 1, 0, 0]
    first = d["gls"][0].ravel()
    assert abs(first[3] - 0.8) < 1e-3 and abs(first[4] - 0.2) < 1e-3


def test_stbin_limits_subsets_and_regions(tmp_path):
    full = host.read_stbin(ONLY_THREE)
    M = len(full["pos"])
    d = host.read_stbin(ONLY_THREE, size_limit=10)
    assert len(d["pos"]) == 10 and d["not_read"] == M - 10
    assert (d["gls"] == full["gls"][:10]).all()
    keep = tmp_path / "keep.txt"
    keep.write_text("samp3\nsamp1\n")
    d = host.read_stbin(ONLY_THREE, keep_names=str(keep))
    assert d["names"] == ["samp1", "samp3"]
    assert (d["gls"] == full["gls"][:, [0, 2]]).all()
    # open-ended regions, other chromosome
    assert len(host.read_stbin(ONLY_THREE, region="20:11977595-")["pos"]) == \
        int((full["pos"] >= 11977595).sum())
    assert len(host.read_stbin(ONLY_THREE, region="21:1-")["pos"]) == 0
    # the drop-first pairs are the standard trio normalised by its sum
    std = host.read_stbin(ONLY_THREE, drop_first=False)["gls"]
    s = std.sum(2, dtype=np.float32)
    assert np.array_equal(full["gls"][..., 0], std[..., 1] / s)
    assert np.array_equal(full["gls"][..., 1], std[..., 2] / s)


def _write(tmp_path, name, lines):
    p = tmp_path / name
    with gzip.open(p, "wt") as f:
        f.write("\n".join(lines) + "\n")
    return str(p)


@pytest.mark.parametrize("body,message", [
    (["20\t10\tAC\t0 1"], "correct number of columns"),
    (["20\t10\tAC\t0 1\t0 0", "20\t9\tAC\t0 1\t0 0"], "less than last position"),
    (["20\t10\tAC\t0 1\t0 0", "21\t11\tAC\t0 1\t0 0"], "wrong chromosome"),
    (["20\t10\tAC\t0 1\t0 0", "20\t10\tAC\t0 1\t0 0"], "already exists"),
    (["20\t10\tACG\t0 1\t0 0"], "Could not parse alleles"),
])
def test_stbin_errors(tmp_path, body, message):
    p = _write(tmp_path, "bad.bin", ["chr\tpos\tallele\ts1\ts2"] + body)
    with pytest.raises(GlpError, match=message):
        host.read_stbin(p)


def test_stbin_multibase_alleles_and_plain_text(tmp_path):
    p = tmp_path / "plain.bin"
    p.write_text("chr\tpos\tallele\ts1\n20\t5\tAT G\t0.5 0.25\n20\t7\tCA\t0 0\n")
    d = host.read_stbin(str(p), drop_first=False)
    assert (d["ref"], d["alt"]) == (["AT", "C"], ["G", "A"])
    assert d["gls"][0, 0].tolist() == [0.25, 0.5, 0.25]


@pytest.mark.parametrize("region", ["20", "20:0-10", ":1-5", "20:a-5"])
def test_bad_regions(region):
    with pytest.raises(GlpError):
        host.read_stbin(ONLY_THREE, region=region)


def test_tran_from_positions_matches_oracle_and_closed_form():
    d = host.read_stbin(EX_BIN)
    hn = 2 * len(d["names"])
    t = host.tran_from_positions(d["pos"], hn)
    assert t.tobytes() == orc.tran_from_positions(d["pos"], hn).tobytes()
    # closed form in FP64: rho = 0.5 mu (M - 1) / span, r = x / (x + hn)
    pos = d["pos"].astype(np.float64)
    mu = 1 / sum(1.0 / i for i in range(1, hn))
    rho = 0.5 * mu * (len(pos) - 1) / (pos[-1] - pos[0])
    x = np.diff(pos) * rho
    r = x / (x + hn)
    assert np.allclose(t[1:, 0], (1 - r) ** 2, rtol=1e-5)
    assert np.allclose(t[1:, 1], r * (1 - r), rtol=1e-4)
    assert np.allclose(t[1:, 2], r * r, rtol=1e-4)
    # equal positions everywhere -> no recombination
    flat = host.tran_from_positions(np.full(10, 7, np.uint32), 20)
    assert (flat[1:] == [1, 0, 0]).all()


def _read_vcf(path):
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt") as f:
        return f.read().split("\n")


def test_vcf_writer_hand_worked_values(tmp_path):
    # 4 sampling generations; per sample the (ha, hb) tallies at one site give
    # p0 = ha / 4, p1 = hb / 4, GP = Phred of (prr, pra, paa), APP = Phred of
    # (p0, p1); 0 -> 380 (FLT_MAX_10_EXP * 10), 1 -> 0
    tallies = [(0, 0), (4, 0), (2, 2), (3, 1), (4, 4), (0, 3)]
    N = len(tallies)
    hsum = np.zeros((2 * N, 1), np.uint32)
    for i, (a, b) in enumerate(tallies):
        hsum[2 * i, 0], hsum[2 * i + 1, 0] = a, b
    names = [f"s{i}" for i in range(N)]
    out = tmp_path / "o.vcf.gz"
    host.write_vcf(out, names, ["20"], [1234], ["C"], ["T"], hsum, 4, "cmd line")
    lines = _read_vcf(out)
    assert lines[0] == "##fileformat=VCFv4.2"
    assert lines[1].startswith("##source=")
    assert lines[2] == "##phaseAndImputeCall=cmd line"
    assert lines[3] == "##contig=<ID=20>"
    assert lines[7].split("\t") == ["#CHROM", "POS", "ID", "REF", "ALT", "QUAL", "FILTER", "INFO",
                                    "FORMAT"] + names
    assert lines[-1] == ""
    f = lines[8].split("\t")
    assert f[:9] == ["20", "1234", ".", "C", "T", "100", "PASS", ".", "GT:GP:APP"]
    want = ["0|0:0,380,380:380,380",
            "1|0:380,0,380:0,380",
            "0|1:6.02,3.01,6.02:3.01,3.01",       # 0.25, 0.5, 0.25 and 0.5, 0.5
            "1|0:7.27,2.04,7.27:1.25,6.02",       # 0.1875, 0.625, 0.1875 and 0.75, 0.25
            "1|1:380,380,0:0,0",
            "0|1:6.02,1.25,380:380,1.25"]         # p = (0, 0.75): 0.25, 0.75, 0
    assert f[9:] == want
    # plain-text output is the same text
    plain = tmp_path / "o.vcf"
    host.write_vcf(plain, names, ["20"], [1234], ["C"], ["T"], hsum, 4, "cmd line")
    assert _read_vcf(plain) == lines


def test_vcf_writer_random_tallies_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    N, M, sn = 7, 40, 200
    hsum = rng.integers(0, sn + 1, (2 * N, M)).astype(np.uint32)
    pos = np.arange(100, 100 + 10 * M, 10)
    out = tmp_path / "r.vcf"
    host.write_vcf(out, [f"x{i}" for i in range(N)], ["3"] * M, pos, ["A"] * M, ["G"] * M, hsum,
                   sn)
    body = [l for l in _read_vcf(out) if l and not l.startswith("#")]
    assert len(body) == M
    norm = np.float32(1.0 / sn)
    for m, line in enumerate(body):
        f = line.split("\t")
        assert int(f[1]) == pos[m]
        for i in range(N):
            p0 = np.float32(hsum[2 * i, m]) * norm
            p1 = np.float32(hsum[2 * i + 1, m]) * norm
            one = np.float32(1)
            prr, pra, paa = (one - p0) * (one - p1), (one - p0) * p1 + p0 * (one - p1), p0 * p1
            gt, gp, app = f[9 + i].split(":")
            if prr >= pra and prr >= paa:
                assert gt == "0|0"
            elif pra >= prr and pra >= paa:
                assert gt == ("1|0" if p0 > p1 else "0|1")
            else:
                assert gt == "1|1"
            for txt, p in zip(gp.split(",") + app.split(","), (prr, pra, paa, p0, p1)):
                want = 380.0 if p == 0 else (0.0 if p == 1 else -10 * math.log10(float(p)))
                assert abs(float(txt) - want) <= 5e-3 * max(1.0, want), (txt, want)
