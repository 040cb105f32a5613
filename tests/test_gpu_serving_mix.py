"""The serving loop on the serving workload's own grid sizes (GPU): squared
grids of 101, 201, 401 and 801 nodes at 0.005 deg, 100 receivers each, smooth
random velocity fields (seed 4).

* Every request's results are the same bit for bit alone and in a mix of
  concurrent requests of all four sizes, in both precisions.
* One request per size matches the oracle (src/eikon_solver_2d.f90:764-1126,
  1276-1446 with the loop body of src/tomo2d.f90:168-195): FP64 to the FP64
  bar, mixed precision to the mixed bar.
"""
import numpy as np
import pytest

import oracle as O
import satt
import satt_workloads as W
from conftest import rel_err, rel_l2

pytestmark = pytest.mark.gpu

SIZES = (101, 201, 401, 801)


@pytest.fixture(scope="module")
def requests():
    reqs = W.serving_requests(24, seed=4)
    picked = []
    for m in SIZES:
        picked += [r for r in reqs if r["nx"] == m][:2]
    assert sorted({r["nx"] for r in picked}) == list(SIZES) and len(picked) == 8
    return picked


@pytest.fixture(scope="module")
def oracle_refs(requests):
    first = {}
    for r in requests:
        first.setdefault(r["nx"], r)
    return {m: (r, O.forward_simulate(W.request_batch(r), nthreads=1)) for m, r in first.items()}


@pytest.mark.parametrize("prec", [satt.PREC_FP64, satt.PREC_MIXED])
def test_serving_sizes_alone_and_in_the_mix(ctx, requests, oracle_refs, prec):
    with satt.Server(ctx, prec) as srv:
        alone = [srv.wait(srv.submit(r, want_T=True)) for r in requests]
        # the mix: large requests first, so that the small ones arrive while
        # the 801-node solves are on the device
        order = sorted(range(len(requests)), key=lambda k: -requests[k]["nx"])
        tickets = {k: srv.submit(requests[k], want_T=True) for k in order}
        mix = {k: srv.wait(t) for k, t in tickets.items()}
        st = srv.stats()
    assert st["completed"] == 2 * len(requests)
    assert st["max_in_flight"] >= 2
    for k, a in enumerate(alone):
        m = mix[k]
        assert a["status"] == 0 and m["status"] == 0, k
        for key in ("tt_rec", "kernel", "T"):
            assert np.array_equal(a[key], m[key]), (k, key)
        assert a["chi"] == m["chi"] and a["iters"] == m["iters"], k
        assert a["iters_fp32"] == m["iters_fp32"], k
    for mnx, (r, ref) in oracle_refs.items():
        o = alone[next(k for k, q in enumerate(requests) if q is r)]
        if prec == satt.PREC_FP64:
            assert o["iters"] == tuple(ref["iters"][0, :2]), mnx
            assert rel_err(o["tt_rec"], ref["tt_fwd"]) < 1e-9, mnx
            assert rel_l2(o["kernel"], ref["adj_s"][0]) < 1e-8, mnx
            assert abs(o["chi"] - ref["chi"]) <= 1e-9 * abs(ref["chi"]), mnx
        else:
            et = float(np.max(np.abs(o["tt_rec"] - ref["tt_fwd"])))
            ek = rel_l2(o["kernel"], ref["adj_s"][0])
            print(f"serving {mnx}^2 mixed vs oracle: tt {et * 1e3:.3g} ms, kernel {ek:.3g}")
            assert et <= 1e-3 and ek <= 1e-3, mnx
