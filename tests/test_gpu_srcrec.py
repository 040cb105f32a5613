"""forward_simulate on the reference's example data: the paths of the
examples/04_hawaii_topo sample table (tests/golden), read with satt_srcrec,
on a 61 x 61 grid over the stations with a checkerboard model, against the
oracle to the FP64 bar; the forward times written back with to_csv."""
import os

import numpy as np
import pytest

import oracle as O
import satt
import satt_srcrec as S
import satt_workloads as W
from conftest import rel_err, rel_l2

pytestmark = pytest.mark.gpu

EX04 = os.path.join(os.path.dirname(__file__), "golden", "src_rec_example04_sample.csv")


@pytest.mark.parametrize("prec", [satt.PREC_FP64, satt.PREC_MIXED])
def test_example_table_forward_simulate(ctx, tmp_path, prec):
    t = S.read_src_rec(EX04)
    per = S.periods(t)
    n = 61
    xg = np.linspace(t["stlo"].min() - 0.15, t["stlo"].max() + 0.15, n)
    yg = np.linspace(t["stla"].min() - 0.15, t["stla"].max() + 0.15, n)
    cb = W.checkerboard_2d(n, n, 3, 3, 0.05)
    svel = np.stack([(2.2 + 0.05 * p) * (1 + cb) for p in per])
    one, zero = np.ones((len(per), n, n)), np.zeros((len(per), n, n))
    b = S.batch_from_src_rec(t, 2, xg, yg, svel, one, zero, one)
    got = ctx.forward_simulate(b, precision=prec)
    want = O.forward_simulate(b)
    if prec == satt.PREC_FP64:
        assert rel_err(got["tt_fwd"], want["tt_fwd"]) < 1e-9
        assert rel_l2(got["adj_s"], want["adj_s"]) < 1e-8
    else:
        assert np.max(np.abs(got["tt_fwd"] - want["tt_fwd"])) < 1e-3
        assert rel_l2(got["adj_s"], want["adj_s"]) < 1e-3
    out = tmp_path / "fwd.csv"
    S.to_csv(out, t, got["tt_fwd"])
    back = S.read_src_rec(out)
    assert np.allclose(back["tt"], got["tt_fwd"], rtol=0, atol=5e-7)
