"""Several grids in one plan (uneven batching, GPU): every part's results are
those of the part planned alone, bit for bit, in both precisions, and the FP64
parts match the oracle to the FP64 bar."""
import numpy as np
import pytest

import oracle as O
import satt
from conftest import rel_err, rel_l2
from test_gpu_batch import shaped_batch

pytestmark = pytest.mark.gpu

KEYS = ("tt_fwd", "adj_s", "iters", "iters_fp32", "item_chi", "weight_out", "item_status")


def parts(dim):
    return [shaped_batch(150, 97, 4, dim, seed=11), shaped_batch(64, 40, 3, dim, seed=12),
            shaped_batch(30, 200, 3, dim, seed=13), shaped_batch(129, 70, 3, dim, seed=14)]


@pytest.mark.parametrize("prec", [satt.PREC_FP64, satt.PREC_MIXED])
@pytest.mark.parametrize("dim", [2, 3])
def test_multi_grid_plan_equals_parts_planned_alone(ctx, prec, dim):
    bs = parts(dim)
    with ctx.plan_multi(bs, prec) as p:
        p.run(True, True)
        multi = [p.results(k) for k in range(len(bs))]
    for k, b in enumerate(bs):
        alone = ctx.forward_simulate(b, precision=prec)
        for key in KEYS + (("adj_density",) if dim == 3 else ()):
            assert np.array_equal(multi[k][key], alone[key]), (k, key)
        assert multi[k]["chi"] == alone["chi"]


def test_multi_grid_fp64_matches_oracle(ctx):
    bs = parts(2)
    with ctx.plan_multi(bs, satt.PREC_FP64) as p:
        p.run(True, True)
        for k, b in enumerate(bs):
            out, ref = p.results(k), O.forward_simulate(b)
            assert np.array_equal(out["iters"], ref["iters"]), k
            assert rel_err(out["tt_fwd"], ref["tt_fwd"]) < 1e-9, k
            assert rel_l2(out["adj_s"], ref["adj_s"]) < 1e-8, k


def test_multi_grid_model_update_of_one_part(ctx):
    bs = parts(2)[:2]
    with ctx.plan_multi(bs, satt.PREC_FP64) as p:
        p.run(True, False)
        before = [p.results(k)["tt_fwd"] for k in range(2)]
        b1 = bs[1]
        p.set_model(b1["svel"] * 1.02, b1["m11"], b1["m12"], b1["m22"], part=1)
        p.run(True, False)
        after = [p.results(k)["tt_fwd"] for k in range(2)]
    assert np.array_equal(before[0], after[0])
    assert not np.allclose(before[1], after[1])
