"""The serving loop (GPU): a request's results are the same bit for bit
whether it runs alone or in a mix of concurrent requests of other sizes, and
equal those of the batched plan on the same one-item batch (which the FP64
oracle checks)."""
import numpy as np
import pytest

import oracle as O
import satt
import satt_workloads as W
from conftest import rel_err, rel_l2

pytestmark = pytest.mark.gpu


def small_requests(n, seed):
    return W.serving_requests(n, seed=seed, sizes=(41, 67, 96), probs=(0.4, 0.3, 0.3), d=0.01,
                              nrec=30)


@pytest.mark.parametrize("prec", [satt.PREC_FP64, satt.PREC_MIXED])
def test_serving_result_same_alone_and_in_the_mix(ctx, prec):
    reqs = small_requests(12, seed=5)
    with satt.Server(ctx, prec, max_batch=4, streams=3) as srv:
        alone = [srv.wait(srv.submit(r, want_T=True)) for r in reqs[:6]]
        tickets = [srv.submit(r, want_T=True) for r in reqs]
        mix = [srv.wait(t) for t in tickets]
        st = srv.stats()
    for k, (a, m) in enumerate(zip(alone, mix)):
        assert a["status"] == 0 and m["status"] == 0, k
        for key in ("tt_rec", "kernel", "T"):
            assert np.array_equal(a[key], m[key]), (k, key)
        assert a["chi"] == m["chi"] and a["iters"] == m["iters"], k
        assert a["iters_fp32"] == m["iters_fp32"], k
    assert st["completed"] == 18 and st["launches"] >= 7


@pytest.mark.parametrize("prec", [satt.PREC_FP64, satt.PREC_MIXED])
def test_serving_equals_batched_plan(ctx, prec):
    reqs = small_requests(3, seed=6)
    with satt.Server(ctx, prec) as srv:
        outs = [srv.wait(srv.submit(r, want_T=True)) for r in reqs]
    for r, o in zip(reqs, outs):
        ref = ctx.forward_simulate(W.request_batch(r), precision=prec)
        assert np.array_equal(o["tt_rec"], ref["tt_fwd"])
        assert np.array_equal(o["kernel"], ref["adj_s"][0])
        assert o["chi"] == ref["chi"]
        assert o["iters"] == tuple(ref["iters"][0, :2])


def test_serving_fp64_matches_oracle(ctx):
    reqs = small_requests(2, seed=7)
    with satt.Server(ctx, satt.PREC_FP64) as srv:
        outs = [srv.wait(srv.submit(r)) for r in reqs]
    for r, o in zip(reqs, outs):
        ref = O.forward_simulate(W.request_batch(r))
        assert rel_err(o["tt_rec"], ref["tt_fwd"]) < 1e-9
        assert rel_l2(o["kernel"], ref["adj_s"][0]) < 1e-8


def test_serving_rejects_malformed_request(ctx):
    r = dict(small_requests(1, seed=8)[0])
    r["nx"] = 2
    with satt.Server(ctx, satt.PREC_FP64) as srv:
        with pytest.raises(satt.SattError) as e:
            srv.submit(r)
        assert e.value.status == satt.E_ARG
