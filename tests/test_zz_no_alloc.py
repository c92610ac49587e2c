"""glp_mh_run_noalloc -- one batched generation with its window table,
order, lanes table and counters in the library's device scratch array --
against the plan path (glp_plan_create + glp_plan_run) on the same inputs:
parents, log-likelihoods, every proposal's value and accept decision and
the pass counters identical, over two calls in a row, the second on a
smaller input (fewer windows, samples and sites)."""
import ctypes
import sys

import numpy as np
import pytest

from glphase_amd import PREC_F32, PREC_MIXED, host
from tests import helpers

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def windows(N, sizes, seed):
    probs = [helpers.synthetic(N, M, seed=seed + w, training_size=4096) for w, M in enumerate(sizes)]
    cb, pc = probs[0]["codebook"], probs[0]["pc"]
    wins = []
    for q in probs:
        wins.append(dict(panel=q["panel"], packed=host.pack_codes(host.vq_codes(q["gls"], cb), 8),
                         tran=q["tran"]))
    return wins, cb, pc, probs[0]["H"]


def run_both(N, sizes, seed, C, prec, gen, pen, mseed, lanes):
    from glphase_amd.engine import Job
    wins, cb, pc, H = windows(N, sizes, seed)

    def job():
        return Job(wins, N, H, cb, pc, bits=8, num_cycles=C, precision=prec, trace=True)

    a = job()
    a.make_plan(lanes)
    a.run_batched(gen=gen, pen=pen, seed=mseed)
    want = (a.parents(), a.like()) + a.trace()
    want_stats = a.stats()
    a.close()
    b = job()
    b.run_noalloc(gen=gen, pen=pen, seed=mseed, lanes=lanes)
    got = (b.parents(), b.like()) + b.trace()
    got_stats = b.run_noalloc(gen=gen, pen=pen, seed=mseed, lanes=lanes, stats=True)
    again = (b.parents(), b.like()) + b.trace()
    b.close()
    return want, want_stats, got, got_stats, again


def same(x, y):
    return all(np.array_equal(p, q, equal_nan=True) for p, q in zip(x, y))


def test_noalloc_matches_plan_two_calls_second_smaller():
    # call 1: 3 windows of 120 samples (300, 77 and 513 sites), FP32,
    # automatic depth; call 2: 1 window of 50 samples x 200 sites, mixed
    # precision, depth 3 -- the scratch array keeps the first call's table
    # and counters, which the second must not see
    for k, (N, sizes, C, prec, lanes) in enumerate(((120, (300, 77, 513), 20, PREC_F32, 0),
                                                     (50, (200,), 12, PREC_MIXED, 3))):
        say(f"noalloc call {k + 1}: N={N} sites={sizes} C={C} prec={prec} lanes={lanes}")
        want, want_stats, got, got_stats, again = run_both(N, sizes, 40 + k, C, prec, gen=k + 1,
                                                           pen=0.7, mseed=99 + k, lanes=lanes)
        assert same(got, want)
        assert same(again, want)  # repeated with the counters read back
        assert got_stats[:2] == want_stats and got_stats[2] == 0
        assert want_stats[1] == N * len(sizes) * (C + 1)


def test_noalloc_refuses_too_many_windows():
    from glphase_amd._lib import GLP_NOALLOC_MAX_WINDOWS, GlpMhDesc, GlpWindow, lib
    n = GLP_NOALLOC_MAX_WINDOWS + 1
    wins = (GlpWindow * n)()
    d = GlpMhDesc()
    d.num_windows = n
    d.windows = ctypes.cast(wins, ctypes.POINTER(GlpWindow))
    rc = lib().glp_mh_run_noalloc(ctypes.byref(d), 0, None, None)
    assert rc != 0
    assert b"GLP_NOALLOC_MAX_WINDOWS" in lib().glp_last_error()
