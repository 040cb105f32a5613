"""Teams of waves per item (GPU): the sweep kernels with W > 1 waves per
item -- strips of a sweep shared by the waves, edge columns handed over in
LDS -- give the one-wave results bit for bit, in both precisions, on grids
of one, two, three and seven strips (full and partial), with the FP64
results also within the FP64 bar of the oracle.  So does the deep (four-step)
prefetch, alone and with teams, and the autotuner keeps each stage's fastest
(team size, prefetch depth)."""
import numpy as np
import pytest

import oracle as O
import satt
from conftest import rel_err, rel_l2
from test_gpu_batch import shaped_batch
from test_gpu_multi import KEYS, parts

pytestmark = pytest.mark.gpu


def same(a, b, dim, tag):
    for key in KEYS + (("adj_density",) if dim == 3 else ()):
        assert np.array_equal(a[key], b[key]), (tag, key)
    assert a["chi"] == b["chi"], tag


@pytest.mark.parametrize("prec", [satt.PREC_FP64, satt.PREC_MIXED])
@pytest.mark.parametrize("nx,ny,dim", [(150, 97, 2), (100, 61, 3), (30, 200, 2), (449, 50, 2)])
def test_team_equals_one_wave(ctx, prec, nx, ny, dim):
    b = shaped_batch(nx, ny, 4, dim, seed=31)
    with ctx.plan(b, prec) as p:
        p.run(True, True)
        ref = p.results()
        for W in (2, 3, 4, 5, 8):
            p.set_waves(W)
            p.run(True, True)
            same(p.results(), ref, dim, W)
    assert np.all(ref["item_status"] == 0)
    if prec == satt.PREC_FP64:
        o = O.forward_simulate(b)
        assert rel_err(ref["tt_fwd"], o["tt_fwd"]) < 1e-9
        assert rel_l2(ref["adj_s"], o["adj_s"]) < 1e-8


def test_team_multi_grid_and_graph(ctx):
    bs = parts(2)
    with ctx.plan_multi(bs, satt.PREC_MIXED) as p:
        p.run(True, True)
        ref = [p.results(k) for k in range(len(bs))]
        p.set_waves(4)
        p.set_graph(True)
        for rep in range(2):
            p.run(True, True)
            for k in range(len(bs)):
                same(p.results(k), ref[k], 2, (rep, k))


def check_choice(r):
    """Each sweep stage keeps its fastest (team size, depth): the first
    fastest W of the two-step sweep, replaced by a deep-prefetch trial (at that
    W, then at W = 1) only where one is faster."""
    tried = [w for w in range(8) if r["ms"][w] > 0]
    for k in range(4):
        two = [r["stage_ms"][w][k] for w in tried]
        assert all(t > 0 for t in two), (k, two)
        best, choice = min(two), (tried[int(np.argmin(two))] + 1, 2)
        d0, d1 = r["stage_ms_deep"][0][k], r["stage_ms_deep"][1][k]
        assert d0 > 0 and d1 > 0, (k, d0, d1)
        if d0 < best:
            best, choice = d0, (choice[0], 4)
        if d1 < best:
            best, choice = d1, (1, 4)
        assert (r["stage_waves"][k], r["stage_depth"][k]) == choice, (k, r)


def test_tuner_and_tuned_plan(ctx):
    r = ctx.tune_waves(260, 60, 2, satt.PREC_MIXED)
    assert 1 <= r["waves"] <= 8 and r["ms"][0] > 0
    tried = [m for m in r["ms"] if m > 0]
    assert r["ms"][r["waves"] - 1] == min(tried)
    check_choice(r)
    b = shaped_batch(260, 60, 3, 2, seed=33)
    with ctx.plan(b, satt.PREC_MIXED) as p:
        p.run(True, True)
        ref = p.results()
        p.set_waves(0)                     # tuned per launch group
        p.run(True, True)
        same(p.results(), ref, 2, "tuned")


def test_plan_tuned_on_its_own_inputs(ctx):
    b = shaped_batch(200, 70, 4, 2, seed=36)
    with ctx.plan(b, satt.PREC_MIXED) as p:
        p.run(True, True)
        ref = p.results()
        r = p.tune_waves()
        assert all(1 <= w <= 8 for w in r["stage_waves"])
        check_choice(r)
        same(p.results(), ref, 2, "after tuning")
        p.set_graph(True)
        p.run(True, True)
        same(p.results(), ref, 2, "tuned, replayed")


def test_stage_team_sizes_pinned_equal_one_wave(ctx):
    b = shaped_batch(200, 70, 3, 2, seed=35)
    with ctx.plan(b, satt.PREC_MIXED) as p:
        p.run(True, True)
        ref = p.results()
        p.set_stage_waves([1, 2, 3, 4])
        p.run(True, True)
        same(p.results(), ref, 2, "stage waves")
        with pytest.raises(satt.SattError):
            p.set_stage_waves([1, 0, 1, 1])


@pytest.mark.parametrize("prec", [satt.PREC_FP64, satt.PREC_MIXED])
@pytest.mark.parametrize("nx,ny,dim", [(41, 41, 3), (150, 97, 2), (449, 50, 2)])
def test_deep_prefetch_equals_two_step(ctx, prec, nx, ny, dim):
    b = shaped_batch(nx, ny, 4, dim, seed=37)
    with ctx.plan(b, prec) as p:
        p.run(True, True)
        ref = p.results()
        for waves in ([1, 1, 1, 1], [3, 2, 4, 3]):
            p.set_stage_waves(waves)
            p.set_stage_depth([4, 4, 4, 4])
            p.run(True, True)
            same(p.results(), ref, dim, ("deep", waves))
            p.set_stage_depth([2, 4, 2, 4])
            p.run(True, True)
            same(p.results(), ref, dim, ("mixed depths", waves))
        with pytest.raises(satt.SattError):
            p.set_stage_depth([2, 3, 2, 2])


def test_server_team_results_equal_one_wave(ctx):
    import satt_workloads as W
    reqs = W.serving_requests(6, seed=9, sizes=(67, 150, 200), probs=(0.3, 0.3, 0.4), d=0.01,
                              nrec=20)
    outs = {}
    for w in (1, 4):
        with satt.Server(ctx, satt.PREC_MIXED, waves=w) as srv:
            outs[w] = [srv.wait(srv.submit(q, want_T=True)) for q in reqs]
    for a, b in zip(outs[1], outs[4]):
        assert a["status"] == 0 and b["status"] == 0
        for key in ("tt_rec", "kernel", "T"):
            assert np.array_equal(a[key], b[key]), key
        assert a["iters"] == b["iters"] and a["iters_fp32"] == b["iters_fp32"]
