"""satt_forward_simulate, the drop-in call of forward_simulate (GPU): its
plan stays resident in the context, and a call whose batch differs only in
the model refreshes the model in place (the reference's inversion and line
search call forward_simulate with the same gathers, src/tomo2d.f90:104,272).
Every call's results equal those of a fresh plan on the same batch bit for
bit, whether the plan was reused, replaced by one with other gathers, or
released."""
import numpy as np
import pytest

import satt
from test_loop_oracle import small_batch

pytestmark = pytest.mark.gpu

KEYS = ("tt_fwd", "adj_s", "item_chi", "iters", "weight_out")


def same(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    assert a["chi"] == b["chi"]


@pytest.mark.parametrize("prec", [satt.PREC_FP64, satt.PREC_MIXED])
def test_dropin_reuses_the_plan_across_models(ctx, prec):
    b = small_batch(2)
    b2 = dict(b, svel=b["svel"] * 1.02)
    o1 = ctx.forward_simulate_dropin(b, precision=prec)
    o2 = ctx.forward_simulate_dropin(b2, precision=prec)        # model refreshed in place
    o3 = ctx.forward_simulate_dropin(b2, True, False, prec)     # a forward-only call
    same(o1, ctx.forward_simulate(b, precision=prec))
    same(o2, ctx.forward_simulate(b2, precision=prec))
    r3 = ctx.forward_simulate(b2, True, False, prec)
    assert np.array_equal(o3["tt_fwd"], r3["tt_fwd"]) and o3["chi"] == r3["chi"]
    assert not np.array_equal(o1["tt_fwd"], o2["tt_fwd"])
    ctx.forward_simulate_release()


def test_dropin_replaces_the_plan_for_other_gathers(ctx):
    b, c = small_batch(2), small_batch(3, nsta=5)
    same(ctx.forward_simulate_dropin(b), ctx.forward_simulate(b))
    oc = ctx.forward_simulate_dropin(c)
    rc = ctx.forward_simulate(c)
    same(oc, rc)
    assert np.array_equal(oc["adj_density"], rc["adj_density"])
    # other observed times are other gathers too
    d = dict(b, tt=b["tt"] * 1.01)
    same(ctx.forward_simulate_dropin(d), ctx.forward_simulate(d))
    ctx.forward_simulate_release()
    same(ctx.forward_simulate_dropin(b), ctx.forward_simulate(b))
    ctx.forward_simulate_release()
