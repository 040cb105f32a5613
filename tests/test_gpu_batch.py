"""The batched forward_simulate on the device against the oracle's loop (GPU)."""
import numpy as np
import pytest

import oracle as O
import satt
import satt_workloads as W
from conftest import rel_err, rel_l2
from test_loop_oracle import small_batch

pytestmark = pytest.mark.gpu


def check(out, ref, dim):
    assert np.array_equal(out["iters"], ref["iters"])
    assert rel_err(out["tt_fwd"], ref["tt_fwd"]) < 1e-9
    assert abs(out["chi"] - ref["chi"]) <= 1e-9 * abs(ref["chi"])
    assert rel_l2(out["adj_s"], ref["adj_s"]) < 1e-8
    if dim == 3:
        assert rel_l2(out["adj_density"], ref["adj_density"]) < 1e-8
    assert np.array_equal(out["weight_out"], ref["weight_out"])


@pytest.mark.parametrize("dim", [2, 3])
def test_forward_simulate_matches_oracle(ctx, dim):
    b = small_batch(dim)
    out = ctx.forward_simulate(b)
    check(out, O.forward_simulate(b), dim)
    assert np.all(out["item_status"] == 0)


def test_tomo2d_checkerboard_step(ctx):
    """The plain 2-D tomography step: 81x81 at 0.025 deg, 25 stations (seed 0),
    observed times from the true model, outputs tt_fwd, chi and the kernel."""
    bt, bi = W.tomo2d_checkerboard()
    obs = ctx.forward_simulate(bt, istotable=True, isadj=False)["tt_fwd"]
    obs_o = O.forward_simulate(bt, istotable=True, isadj=False)["tt_fwd"]
    assert rel_err(obs, obs_o) < 1e-9
    b = W.set_observed(bi, obs_o)
    check(ctx.forward_simulate(b), O.forward_simulate(b), 2)


def test_per_solve_launch_form_equals_batched(ctx):
    b = small_batch(3, nsta=5)
    with ctx.plan(b) as p:
        p.run(True, True, satt.LAUNCH_BATCHED)
        o1 = p.results()
        p.run(True, True, satt.LAUNCH_PER_SOLVE)
        o2 = p.results()
    for k in ("tt_fwd", "adj_s", "adj_density", "iters", "item_chi"):
        assert np.array_equal(o1[k], o2[k]), k


def test_plan_reuse_and_model_update(ctx):
    b = small_batch(2)
    with ctx.plan(b) as p:
        p.run(True, True)
        o1 = p.results()
        v2 = b["svel"] * 1.01
        p.set_model(v2, b["m11"], b["m12"], b["m22"])
        p.run(True, False)           # line-search form: forward only
        o2 = p.results()
    ref2 = O.forward_simulate(W.with_model(b, v2), istotable=True, isadj=False)
    assert rel_err(o2["tt_fwd"], ref2["tt_fwd"]) < 1e-9
    assert not np.allclose(o1["tt_fwd"], o2["tt_fwd"])


def test_item_errors_are_reported(ctx):
    b = small_batch(2, nsta=4, nper=1)
    b = dict(b); b["item_srcx"] = b["item_srcx"].copy()
    b["item_srcx"][1] = b["xgrids"][0]          # source on the boundary
    with pytest.raises(satt.SattError) as e:
        ctx.forward_simulate(b)
    assert e.value.status == satt.E_SRC_ON_BOUNDARY


def shaped_batch(nx, ny, nsta, dim=2, seed=7, d=0.02):
    xg = 100.0 + d * np.arange(nx); yg = 20.0 + d * np.arange(ny)
    rng = np.random.default_rng(seed)
    sx, sy = W.stations_uniform(rng, nsta, xg, yg)
    periods = np.array([6.0, 15.0])
    paths = W.all_pairs_paths(sx, sy, periods)
    vt = np.stack([(2.9 + 0.03 * p) * (1 + W.checkerboard_2d(nx, ny, 3, 2, 0.07)) for p in periods])
    topo = W.gaussian_hills(xg, yg, 3, rng, h=(0.5, 2.0), w=(5.0, 15.0))
    m11, m12, m22 = W.topo_metric(topo, xg, yg)
    rep = lambda f: np.repeat(f[None], 2, 0)
    bt = W.make_batch(dim, xg, yg, periods, vt, rep(m11), rep(m12), rep(m22), paths)
    obs = O.forward_simulate(bt, istotable=True, isadj=False)["tt_fwd"]
    return W.with_model(W.set_observed(bt, obs), np.stack([(2.9 + 0.03 * p) * np.ones((nx, ny)) for p in periods]))


@pytest.mark.parametrize("nx,ny,dim", [(64, 40, 2), (65, 33, 3), (150, 97, 2), (129, 130, 2), (30, 200, 3)])
def test_strip_pipeline_equals_workgroup_wavefront_bitwise(ctx, nx, ny, dim):
    """The batched strip-pipeline kernels (one wave per item, 64-column
    strips) and the one-solve-per-launch workgroup wavefront kernels run the
    same Gauss-Seidel order: identical bits on grids with full, partial and
    single strips."""
    b = shaped_batch(nx, ny, 4, dim)
    with ctx.plan(b) as p:
        p.run(True, True, satt.LAUNCH_BATCHED)
        o1 = p.results()
        p.run(True, True, satt.LAUNCH_PER_SOLVE)
        o2 = p.results()
    for k in ("iters", "tt_fwd", "adj_s", "item_chi", "weight_out") + (("adj_density",) if dim == 3 else ()):
        assert np.array_equal(o1[k], o2[k]), k
    ref = O.forward_simulate(b)
    check(o1, ref, dim)
