"""The headline workload on the device (GPU): the curved-surface grid of
401 x 401 at 0.005 deg with the 8-hill topographic metric, a +-8%
checkerboard, 200 sources x 20 periods = 4000 fwd+adj solves (199 receivers
each) in one plan, as the throughput measurement runs it.

* Eight items spread over sources and periods are planned alone and checked
  against the oracle's loop (src/tomo2d.f90:158-201 over
  src/eikon_solver_2d.f90:764-1126,1276-1446): FP64 to the FP64 bar (receiver
  times 1e-9 relative, kernel 1e-8 relative L2, iteration counts equal),
  mixed precision to the mixed bar (1 ms, 1e-3).
* The same items inside the full 4000-item plan give bit for bit the results
  of the subset plan (travel times, misfit, iterations, T and the masked
  adjoint field), so the full-size run is the checked one.
* The full plan's stacked kernel equals the sum of its 4000 per-item adjoint
  fields divided by svel**2, per period in item order (src/tomo2d.f90:155),
  and its misfit is the sum of the per-item misfits.
"""
import numpy as np
import pytest

import oracle as O
import satt
import satt_workloads as W
from conftest import rel_err, rel_l2

pytestmark = pytest.mark.gpu

NPER = 20
# (source station, period index) pairs: corners of the station list and of
# the period range, and some in between
PICKS = ((0, 0), (0, 19), (37, 7), (71, 13), (103, 3), (142, 16), (181, 10), (199, 19))


def _subset(b, items):
    s = dict(b)
    for k in ("item_pidx", "item_srcx", "item_srcy"):
        s[k] = np.ascontiguousarray(b[k][items])
    sel = np.concatenate([np.arange(b["rec_off"][i], b["rec_off"][i + 1]) for i in items])
    s["rec_off"] = np.concatenate(([0], np.cumsum([b["rec_off"][i + 1] - b["rec_off"][i]
                                                   for i in items]))).astype(np.int32)
    for k in ("recx", "recy", "tt", "weight", "idx_path"):
        s[k] = np.ascontiguousarray(b[k][sel])
    return s


@pytest.fixture(scope="module")
def topo(ctx):
    bt, bi = W.topo_checkerboard(n=401, d=0.005, nsrc=200, nper=NPER, seed=3)
    assert len(bi["item_pidx"]) == 4000 and bi["nx"] == bi["ny"] == 401
    # observed times: FP64 forward on the true model, on the device
    obs = ctx.forward_simulate(bt, istotable=True, isadj=False)["tt_fwd"]
    b = W.set_observed(bi, obs)
    items = np.array([s * NPER + p for s, p in PICKS])
    assert np.array_equal(b["item_pidx"][items], [p for _, p in PICKS])
    sub = _subset(b, items)
    ref = O.forward_simulate(sub, nthreads=8)
    return b, items, sub, ref


def _item_fields(plan, items):
    return [(plan.item_field(int(i), satt.FIELD_T), plan.item_field(int(i), satt.FIELD_TA))
            for i in items]


@pytest.mark.parametrize("prec", [satt.PREC_FP64, satt.PREC_MIXED])
def test_topo401_full_plan(ctx, topo, prec):
    b, items, sub, ref = topo
    paths = sub["idx_path"]
    # the sampled items alone against the oracle
    with ctx.plan(sub, prec) as p:
        p.run(True, True)
        os_ = p.results()
        fs = _item_fields(p, range(len(items)))
    if prec == satt.PREC_FP64:
        assert np.array_equal(os_["iters"], ref["iters"])
        assert rel_err(os_["tt_fwd"][paths], ref["tt_fwd"][paths]) < 1e-9
        assert rel_l2(os_["adj_s"], ref["adj_s"]) < 1e-8
        assert abs(os_["chi"] - ref["chi"]) <= 1e-9 * abs(ref["chi"])
        assert np.array_equal(os_["weight_out"], ref["weight_out"])
    else:
        et = float(np.max(np.abs(os_["tt_fwd"][paths] - ref["tt_fwd"][paths])))
        ek = rel_l2(os_["adj_s"], ref["adj_s"])
        print(f"topo401 mixed vs oracle: tt {et * 1e3:.3g} ms, kernel rel L2 {ek:.3g}")
        assert et <= 1e-3 and ek <= 1e-3
    assert np.all(os_["item_status"] == 0)

    # the full 4000-item plan: the same items give the same bits
    with ctx.plan(b, prec) as p:
        p.run(True, True)
        of = p.results()
        st = p.stats()
        assert st["launches"] == 1
        ff = _item_fields(p, items)
        # stacking at full size: per period, the items' Ta/svel**2 in item order
        acc = np.zeros((NPER, b["nx"], b["ny"]), order="F")
        for i in range(len(b["item_pidx"])):
            q = int(b["item_pidx"][i])
            v = b["svel"][q]
            acc[q] = acc[q] + p.item_field(i, satt.FIELD_TA) / (v * v)
    assert np.all(of["item_status"] == 0)
    assert np.array_equal(of["tt_fwd"][paths], os_["tt_fwd"][paths])
    assert np.array_equal(of["iters"][items], os_["iters"])
    assert np.array_equal(of["iters_fp32"][items], os_["iters_fp32"])
    assert np.array_equal(of["item_chi"][items], os_["item_chi"])
    for k, ((t1, a1), (t2, a2)) in enumerate(zip(ff, fs)):
        assert np.array_equal(t1, t2), k
        assert np.array_equal(a1, a2), k
    assert rel_l2(of["adj_s"], acc) <= 1e-15, rel_l2(of["adj_s"], acc)
    assert abs(of["chi"] - float(np.sum(of["item_chi"]))) <= 1e-12 * of["chi"]
    # every path has a finite positive time; the iteration counts stay far
    # from MaxIter (2000 forward, 1000 adjoint)
    assert np.all(np.isfinite(of["tt_fwd"])) and np.all(of["tt_fwd"] > 0)
    assert of["iters"][:, 0].max() < 2000 and of["iters"][:, 1].max() < 1000
