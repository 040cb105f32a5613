"""The measurement plan's other configurations on the device (GPU):

* the 3-D inversion's 2-D stage (dim 3, density adjoint): 41x41 at 0.05 deg,
  16 periods, 100 stations -> 1600 fwd+adj(+density) solves in one launch;
* uneven batching: grids of 101x101, 201x161, 321x321 and 81x241 in one plan.

FP64 runs match the oracle to the FP64 bar; mixed runs stay within the mixed
bar (1 ms, 1e-3 relative L2) of the oracle."""
import numpy as np
import pytest

import oracle as O
import satt
import satt_workloads as W
from conftest import rel_err, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def acqui():
    bt, bi = W.acqui3d_checkerboard()
    obs = O.forward_simulate(bt, istotable=True, isadj=False, nthreads=8)["tt_fwd"]
    b = W.set_observed(bi, obs)
    return b, O.forward_simulate(b, nthreads=8)


def test_acqui3d_fp64_matches_oracle(ctx, acqui):
    b, ref = acqui
    out = ctx.forward_simulate(b)
    assert out["adj_s"].shape == (16, 41, 41) and len(b["item_pidx"]) == 1600
    assert np.array_equal(out["iters"], ref["iters"])
    assert rel_err(out["tt_fwd"], ref["tt_fwd"]) < 1e-9
    assert rel_l2(out["adj_s"], ref["adj_s"]) < 1e-8
    assert rel_l2(out["adj_density"], ref["adj_density"]) < 1e-8
    assert abs(out["chi"] - ref["chi"]) <= 1e-9 * abs(ref["chi"])


def test_acqui3d_mixed_within_tolerance(ctx, acqui):
    b, ref = acqui
    out = ctx.forward_simulate(b, precision=satt.PREC_MIXED)
    err_t = float(np.max(np.abs(out["tt_fwd"] - ref["tt_fwd"])))
    ek, ed = rel_l2(out["adj_s"], ref["adj_s"]), rel_l2(out["adj_density"], ref["adj_density"])
    print(f"3-D stage mixed: tt {err_t * 1e3:.3g} ms, kernel {ek:.3g}, density {ed:.3g}")
    assert err_t <= 1e-3 and ek <= 1e-3 and ed <= 1e-3
    assert np.all(out["item_status"] == 0)


@pytest.fixture(scope="module")
def uneven():
    mix = W.uneven_mix(nsolve=40, seed=2)
    parts = []
    for bt, bi in mix:
        obs = O.forward_simulate(bt, istotable=True, isadj=False, nthreads=8)["tt_fwd"]
        parts.append(W.set_observed(bi, obs))
    return parts, [O.forward_simulate(b, nthreads=8) for b in parts]


@pytest.mark.parametrize("prec", [satt.PREC_FP64, satt.PREC_MIXED])
def test_uneven_mix_one_plan(ctx, uneven, prec):
    parts, refs = uneven
    with ctx.plan_multi(parts, prec) as p:
        p.run(True, True)
        outs = [p.results(k) for k in range(len(parts))]
        st = p.stats()
    assert st["launches"] == 1
    for k, (out, ref) in enumerate(zip(outs, refs)):
        if prec == satt.PREC_FP64:
            assert np.array_equal(out["iters"], ref["iters"]), k
            assert rel_err(out["tt_fwd"], ref["tt_fwd"]) < 1e-9, k
            assert rel_l2(out["adj_s"], ref["adj_s"]) < 1e-8, k
        else:
            err_t = float(np.max(np.abs(out["tt_fwd"] - ref["tt_fwd"])))
            ek = rel_l2(out["adj_s"], ref["adj_s"])
            print(f"uneven part {k} ({parts[k]['nx']}x{parts[k]['ny']}): tt {err_t * 1e3:.3g} ms, "
                  f"kernel {ek:.3g}")
            assert err_t <= 1e-3 and ek <= 1e-3, k
