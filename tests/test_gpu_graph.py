"""Replay form of the batched run (GPU): the launch chain captured into a HIP
graph gives the eager run's results bit for bit, on every replay, with eager
work and synchronizations between the replays, across model updates and for
each (istotable, isadj) pair."""
import numpy as np
import pytest

import satt
import satt_workloads as W
from test_gpu_batch import shaped_batch
from test_gpu_multi import KEYS, parts

pytestmark = pytest.mark.gpu


def same(a, b, dim):
    for key in KEYS + (("adj_density",) if dim == 3 else ()):
        assert np.array_equal(a[key], b[key]), key
    assert a["chi"] == b["chi"]


@pytest.mark.parametrize("prec", [satt.PREC_FP64, satt.PREC_MIXED])
@pytest.mark.parametrize("dim", [2, 3])
def test_graph_replays_equal_eager(ctx, prec, dim):
    b = shaped_batch(150, 97, 4, dim, seed=21)
    other = W.with_model(b, b["svel"] * 1.03)
    with ctx.plan(b, prec) as p, ctx.plan(other, prec) as q:
        p.run(True, True)
        ref = p.results()
        q.run(True, True)
        ref_other = q.results()
        p.set_graph(True)
        for rep in range(3):
            p.run(True, True)               # rep 0 captures, later ones replay
            same(p.results(), ref, dim)
            q.run(True, True)               # eager work on the same stream between replays
            same(q.results(), ref_other, dim)
        # a model update refills the buffers the graph holds
        p.set_model(other["svel"], other["m11"], other["m12"], other["m22"])
        p.run(True, True)
        same(p.results(), ref_other, dim)
        p.set_model(b["svel"], b["m11"], b["m12"], b["m22"])
        p.run(True, True)
        same(p.results(), ref, dim)
        # another flag pair is another graph; the first one still replays
        p.run(True, False)
        fwd_only = p.results()
        assert np.array_equal(fwd_only["tt_fwd"], ref["tt_fwd"])
        p.run(True, True)
        same(p.results(), ref, dim)
        # eager again
        p.set_graph(False)
        p.run(True, True)
        same(p.results(), ref, dim)


def test_graph_multi_grid_plan(ctx):
    bs = parts(2)
    with ctx.plan_multi(bs, satt.PREC_MIXED) as p:
        p.run(True, True)
        ref = [p.results(k) for k in range(len(bs))]
        p.set_graph(True)
        for rep in range(3):
            p.run(True, True)
            p.wait()
            for k in range(len(bs)):
                same(p.results(k), ref[k], 2)
        assert p.stats()["ms_total"] > 0
