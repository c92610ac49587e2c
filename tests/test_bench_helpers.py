"""bench.py's host-side replay of a timed generation (Philox draws, the
emission table and the FP64 forward pass used for its accuracy report)
against the oracle."""
import numpy as np

import bench
from tests import helpers


def test_philox_blocks_match_oracle(orc):
    for seed, k, i, w, g in ((0, 0, 0, 0, 0), (20240601, 5, 17, 0, 3), (2**40 + 7, 0xFFFFFFFF, 4999, 2, 255)):
        got = np.array([x for x in bench.philox_blocks(k, i, w, g, seed)], np.uint32)
        assert (got == orc.draw(seed, k, i, w, g)).all()


def test_unif_excl_matches_oracle(orc):
    rng = np.random.default_rng(0)
    x = rng.integers(0, 2**32, 200, dtype=np.uint64).astype(np.uint32)
    s = rng.integers(0, 50, 200)
    got = bench.unif_excl(x, 100, s)
    want = [orc.L().orc_unif_excl(int(a), 100, int(b)) for a, b in zip(x, s)]
    assert (got == np.array(want)).all()


def test_emission_table_and_fp64_forward_match_oracle(orc):
    p = helpers.synthetic(20, 300, seed=3, training_size=4096)
    E = bench.emission_table(p["codebook"], p["pc"])
    assert np.array_equal(E, orc.emit_lut(p["codebook"], p["pc"]))
    rng = np.random.default_rng(1)
    samples = rng.integers(0, p["N"], 12)
    parents = rng.integers(0, p["H"], (12, 4))
    got = bench.fp64_forward(p["panel"], p["codes"], E, p["tran"], samples, parents)
    for e in range(12):
        emit = orc.sample_emit(p["packed"], int(samples[e]), p["M"], 8, E)
        want = orc.hmm_like_f64(p["panel"], parents[e].astype(np.uint32), emit, p["tran"])
        assert abs(got[e] - want) <= 1e-12 * abs(want)


def test_recorded_traffic_reads_both_pass_forms():
    # the headline plan (forward-backward form) and the forward-only plan
    # each have a committed FETCH_SIZE + WRITE_SIZE record
    fb = bench.recorded_traffic("5000x4096x64 fb q5 w8")
    assert bench.recorded_traffic("5000x4096x64 fb q5 w8 split") is not None
    fwd = bench.recorded_traffic("5000x4096x64 q6 w4")
    assert fb is not None and fwd is not None
    # both lie between the algorithmic bytes (0.686 GB) and 4x that
    for v in (fb, fwd):
        assert 0.686e9 < v < 4 * 0.686e9
    assert bench.recorded_traffic("5000x4096x64 fb q99 w8") is None
