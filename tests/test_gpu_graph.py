"""HIP-graph generations (glp_plan_graph_run / glp_plan_graph_run_plain):
every replay, with eager launches and synchronisations between replays,
equals the eager result and the oracle."""
import sys

import numpy as np
import pytest

from glphase_amd import PREC_F32, PREC_MIXED
from tests import helpers

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_job(p, C, precision):
    from glphase_amd.engine import Job
    win = dict(panel=p["panel"], packed=p["packed"], tran=p["tran"])
    return Job([win], p["N"], p["H"], p["codebook"], p["pc"], bits=p["bits"], num_cycles=C,
               precision=precision)


@pytest.mark.parametrize("kind,precision", [("batched", PREC_MIXED), ("batched", PREC_F32),
                                            ("plain", PREC_MIXED)])
def test_graph_replays_match_eager_and_oracle(orc, kind, precision):
    say(f"graph {kind} precision={precision}")
    p = helpers.synthetic(150, 700, seed=31, depth=2.0)
    C = 20
    g_job, e_job = make_job(p, C, precision), make_job(p, C, precision)
    g_job.make_plan(4)
    e_job.make_plan(4)
    lut = orc.emit_lut(p["codebook"], p["pc"])
    mh = orc.MH(p["panel"], p["packed"], p["bits"], lut, p["tran"], p["N"])
    graph = g_job.run_batched_graph if kind == "batched" else g_job.run_plain_graph
    eager = e_job.run_batched if kind == "batched" else e_job.run_plain
    for g in range(4):
        pen = [0.25, 0.5, 1.0, 1.0][g]
        graph(gen=g, pen=pen, seed=5)
        eager(gen=g, pen=pen, seed=5)
        gp, gl = g_job.parents(), g_job.like()
        assert (gp == e_job.parents()).all(), f"replay {g}: parents differ from eager"
        assert (gl == e_job.like()).all(), f"replay {g}: likelihoods differ from eager"
        po, _ = mh.run(C, pen, 5, gen=g, precision=precision)
        assert (gp[0] == po).all(), f"replay {g}: parents differ from the oracle"
        if kind == "batched":
            assert g_job.stats() == e_job.stats()
        # eager work on the graph's own buffers between replays
        (g_job.run_batched if kind == "batched" else g_job.run_plain)(gen=g + 10, pen=1.0, seed=6)
        g_job.torch.cuda.synchronize()


def test_graph_rebuilt_after_depth_change(orc):
    say("graph rebuild")
    p = helpers.synthetic(90, 300, seed=8)
    C = 16
    job = make_job(p, C, PREC_F32)
    job.make_plan(2)
    job.run_batched_graph(gen=1, pen=1.0, seed=3)
    first = job.parents()
    job.set_lanes(8)
    job.run_batched_graph(gen=1, pen=1.0, seed=3)
    assert (job.parents() == first).all()
    assert job.lanes()[0] == 8
