"""satt_plan_linesearch (att_tomo_2d%linesearch, src/tomo2d.f90:253-288) on
the device against the oracle's line search on the same case: the same
sub-iteration count and step length, the kept model bit for bit (the same
arithmetic on the same operands), the misfit within the FP64 bar; the plan
keeps the kept model for its next run."""
import numpy as np
import pytest

import oracle as O
import satt
from conftest import rel_err
from test_oracle_linesearch import STEP, ls_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("start,shrink,nmax,neg", [
    (STEP, 0.5, 4, False),        # accepted at once
    (4 * STEP, 0.25, 5, False),   # one shrink, then accepted
    (0.02, 0.5, 3, True),         # uphill: every trial rejected
])
def test_linesearch_matches_oracle(ctx, start, shrink, nmax, neg):
    b, ker, chi0 = ls_case()
    ker = -ker if neg else ker
    want = O.linesearch(b, ker, chi0, start, shrink, nmax)
    with ctx.plan(b, satt.PREC_FP64) as p:
        got = p.linesearch(b["svel"], ker, chi0, start, shrink, nmax)
        assert got["nsub"] == want["nsub"] and got["updatemax"] == want["updatemax"]
        assert np.array_equal(got["svel"], want["svel"])
        assert abs(got["chi"] - want["chi"]) <= 1e-9 * max(abs(want["chi"]), 1e-3 * chi0)
        # the plan now holds the kept model
        p.run(True, True)
        r = p.results()
        b2 = dict(b)
        b2["svel"] = want["svel"]
        o = O.forward_simulate(b2)
        assert rel_err(r["tt_fwd"], o["tt_fwd"]) < 1e-9


def test_linesearch_mixed_and_edge_cases(ctx):
    b, ker, chi0 = ls_case(40, 33)
    want = O.linesearch(b, ker, chi0, 4 * STEP, 0.25, 5)
    with ctx.plan(b, satt.PREC_MIXED) as p:
        p.set_graph(True)
        got = p.linesearch(b["svel"], ker, chi0, 4 * STEP, 0.25, 5)
        assert got["nsub"] == want["nsub"] == 2
        assert np.array_equal(got["svel"], want["svel"])
        none = p.linesearch(b["svel"], ker, chi0, STEP, 0.5, 0)
        assert none["nsub"] == 0 and none["chi"] == chi0
        assert np.array_equal(none["svel"], b["svel"])
        with pytest.raises(satt.SattError) as e:
            p.linesearch(b["svel"], np.zeros_like(ker), chi0, STEP, 0.5, 2)
        assert e.value.status == satt.E_ARG


def test_linesearch_needs_one_grid(ctx):
    b, ker, chi0 = ls_case(25, 21)
    with ctx.plan_multi([b, b], satt.PREC_FP64) as p:
        with pytest.raises(satt.SattError):
            p.linesearch(b["svel"], ker, chi0, STEP, 0.5, 2)
