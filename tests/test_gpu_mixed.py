"""Mixed-precision mode of the batched path (GPU).

Two checks:
  - accuracy against the FP64 oracle on the same inputs, to the stated
    mixed-precision bar: receiver travel times within 1 ms, kernels within
    1e-3 relative L2 (and, tighter, to what the mode's design delivers);
  - bit-for-bit agreement with the oracle's restatement of the mixed mode
    when both start from the device's setup fields, which pins every FP32
    operation and the polish.
"""
import numpy as np
import pytest

import oracle as O
import satt
import satt_workloads as W
from conftest import rel_l2
from test_gpu_batch import shaped_batch

pytestmark = pytest.mark.gpu

TOL_T_S = 1e-3      # travel-time error bar of the mixed mode: 1 ms
TOL_KERNEL = 1e-3   # kernel relative L2 bar of the mixed mode
# what the FP32 phase (tau offsets, residual-sized quadratic) plus the FP64
# polish delivers on these grids; a regression guard well inside the bar
GUARD_T_S = 1e-4
GUARD_KERNEL = 1e-4


def check_mixed(out, ref, dim):
    err_t = float(np.max(np.abs(out["tt_fwd"] - ref["tt_fwd"])))
    err_k = rel_l2(out["adj_s"], ref["adj_s"])
    print(f"mixed: max travel-time error {err_t * 1e3:.3g} ms, kernel rel L2 {err_k:.3g}")
    assert err_t <= TOL_T_S and err_k <= TOL_KERNEL
    assert err_t <= GUARD_T_S and err_k <= GUARD_KERNEL
    if dim == 3:
        assert rel_l2(out["adj_density"], ref["adj_density"]) <= GUARD_KERNEL
    assert abs(out["chi"] - ref["chi"]) <= 1e-4 * abs(ref["chi"])
    assert np.all(out["item_status"] == 0)
    assert np.all(out["iters_fp32"][:, 0] >= 1) and np.all(out["iters"][:, 0] >= 1)
    assert np.all(out["iters"][:, 1] == 1)          # one FP64 adjoint polish iteration
    if dim == 3:
        assert np.all(out["iters"][:, 2] == 1)


@pytest.mark.parametrize("nx,ny,dim", [(150, 97, 2), (65, 33, 3), (30, 200, 2)])
def test_mixed_forward_simulate_within_tolerance(ctx, nx, ny, dim):
    b = shaped_batch(nx, ny, 5, dim)
    out = ctx.forward_simulate(b, precision=satt.PREC_MIXED)
    check_mixed(out, O.forward_simulate(b), dim)


def test_mixed_tomo2d_checkerboard(ctx):
    bt, bi = W.tomo2d_checkerboard()
    obs = O.forward_simulate(bt, istotable=True, isadj=False)["tt_fwd"]
    b = W.set_observed(bi, obs)
    check_mixed(ctx.forward_simulate(b, precision=satt.PREC_MIXED), O.forward_simulate(b), 2)


def test_mixed_bitwise_given_device_setup(ctx):
    b = shaped_batch(129, 70, 3, 2)
    xg, yg = b["xgrids"], b["ygrids"]
    with ctx.plan(b, satt.PREC_MIXED) as p:
        p.run(True, True)
        out = p.results()
        for item in range(len(b["item_pidx"])):
            pidx = b["item_pidx"][item]
            fun = 1.0 / b["svel"][pidx]
            spha, sphb, sphc = b["m11"][pidx], b["m22"][pidx], -b["m12"][pidx]
            x0, y0 = b["item_srcx"][item], b["item_srcy"][item]
            T = p.item_field(item, satt.FIELD_T)
            Ta = p.item_field(item, satt.FIELD_TA)
            setup = ctx.fwd_setup_fields(xg, yg, spha, sphb, sphc, fun, x0, y0)
            To, its, _ = O.fwd(xg, yg, spha, sphb, sphc, fun, x0, y0, pre=setup, mixed=True)
            assert its == (out["iters_fp32"][item, 0], out["iters"][item, 0]), item
            assert np.array_equal(T, To), item
            # the adjoint source as the device forms it: w*(t - tobs)
            r0, r1 = b["rec_off"][item], b["rec_off"][item + 1]
            tek = out["tt_fwd"][b["idx_path"][r0:r1]]
            rhs = out["weight_out"][r0:r1] * (tek - b["tt"][r0:r1])
            rx, ry = b["recx"][r0:r1], b["recy"][r0:r1]
            a2, b2, c2, delta = ctx.adj_setup_fields(xg, yg, spha, sphb, sphc, rx, ry, rhs)
            Tao, ita, _ = O.adj_pre(xg, yg, a2, b2, c2, T, delta, mixed=True)
            assert ita == (out["iters_fp32"][item, 1], out["iters"][item, 1]), item
            assert np.array_equal(Ta, O.mask(xg, yg, Tao, x0, y0)), item


def test_mixed_rejects_per_solve_form(ctx):
    b = shaped_batch(40, 30, 3, 2)
    with ctx.plan(b, satt.PREC_MIXED) as p:
        with pytest.raises(satt.SattError) as e:
            p.run(True, True, satt.LAUNCH_PER_SOLVE)
        assert e.value.status == satt.E_ARG
