"""The per-(source, period) loop of the oracle and the host-side gather
logic.  CPU only."""
import numpy as np
import pytest

import oracle as O
import satt_workloads as W


def small_batch(dim=2, n=31, nsta=6, nper=2, seed=4):
    d = 0.04
    xg = 100.0 + d * np.arange(n); yg = 20.0 + d * np.arange(n)
    rng = np.random.default_rng(seed)
    sx, sy = W.stations_uniform(rng, nsta, xg, yg)
    periods = np.array([8.0, 20.0][:nper])
    paths = W.all_pairs_paths(sx, sy, periods)
    amp = W.checkerboard_2d(n, n, 2, 2, 0.06)
    vt = np.stack([(2.9 + 0.03 * p) * (1 + amp) for p in periods])
    fx = 0.2 * np.sin(np.arange(n) / 5.0)[:, None] * np.ones((1, n))
    fy = 0.1 * np.cos(np.arange(n) / 3.0)[None, :] * np.ones((n, 1))
    den = 1 + fx ** 2 + fy ** 2
    rep = lambda f: np.repeat(f[None], nper, 0)
    bt = W.make_batch(dim, xg, yg, periods, vt, rep((1 + fy ** 2) / den), rep(-fx * fy / den),
                      rep((1 + fx ** 2) / den), paths)
    obs = O.forward_simulate(bt, istotable=True, isadj=False)["tt_fwd"]
    bi = W.with_model(W.set_observed(bt, obs), np.stack([(2.9 + 0.03 * p) * np.ones((n, n)) for p in periods]))
    return bi


def compose(b):
    """forward_simulate restated from the per-solve oracle entry points
    (src/tomo2d.f90:170-194, src/acqui.f90:122-151, src/measadj.f90:62-104)."""
    nx, ny, npd = b["nx"], b["ny"], b["nperiod"]
    tt_fwd = np.zeros(b["npath"]); chi = 0.0
    adj_s = np.zeros((npd, nx, ny)); adj_d = np.zeros((npd, nx, ny))
    for it in range(len(b["item_pidx"])):
        p = b["item_pidx"][it]
        r0, r1 = b["rec_off"][it], b["rec_off"][it + 1]
        vel = b["svel"][p]; m11, m12, m22 = b["m11"][p], b["m12"][p], b["m22"][p]
        T, _, _ = O.fwd(b["xgrids"], b["ygrids"], m11, m22, -m12, 1.0 / vel, b["item_srcx"][it], b["item_srcy"][it])
        tek = O.interp2(b["xgrids"], b["ygrids"], T, b["recx"][r0:r1], b["recy"][r0:r1])
        tt, w = b["tt"][r0:r1], b["weight"][r0:r1].copy()
        s = 0.0
        for r in range(r1 - r0):
            s = s + w[r] * ((tek[r] - tt[r]) * (tek[r] - tt[r]))
        chi = chi + 0.5 * s
        tt_fwd[b["idx_path"][r0:r1]] = tek
        w[np.abs(tek - tt) / tt > 0.5] = 0.0
        adj, _, _ = O.adj(b["xgrids"], b["ygrids"], m11, m22, -m12, T, b["recx"][r0:r1], b["recy"][r0:r1], w * (tek - tt))
        adj = O.mask(b["xgrids"], b["ygrids"], adj, b["item_srcx"][it], b["item_srcy"][it])
        if b["dim"] == 3:
            kd, _, _ = O.adj(b["xgrids"], b["ygrids"], m11, m22, -m12, T, b["recx"][r0:r1], b["recy"][r0:r1], -1.0 * w)
            kd = O.mask(b["xgrids"], b["ygrids"], kd, b["item_srcx"][it], b["item_srcy"][it])
            adj_s[p] = adj_s[p] + adj / (vel * vel * vel)
            adj_d[p] = adj_d[p] + kd / (vel * vel * vel)
        else:
            adj_s[p] = adj_s[p] + adj / (vel * vel)
    return tt_fwd, chi, adj_s, adj_d


@pytest.mark.parametrize("dim", [2, 3])
def test_loop_matches_composition_bitwise(dim):
    b = small_batch(dim)
    out = O.forward_simulate(b)
    tt_fwd, chi, adj_s, adj_d = compose(b)
    assert np.array_equal(out["tt_fwd"], tt_fwd)
    assert out["chi"] == chi
    assert np.array_equal(out["adj_s"], adj_s)
    if dim == 3:
        assert np.array_equal(out["adj_density"], adj_d)
    assert np.all(out["iters"][:, 0] >= 2) and np.all(out["iters"][:, 1] >= 1)


def test_outliers_count_in_misfit_but_not_in_kernel():
    """chi uses the weights before rejection; the adjoint source zeroes
    |t-tobs|/tobs > 0.5 (src/measadj.f90:72,82-88)."""
    b = small_batch(2)
    b2 = dict(b); b2["tt"] = b["tt"].copy()
    b2["tt"][3] *= 3.0                      # one outlier in the first item
    o1, o2 = O.forward_simulate(b), O.forward_simulate(b2)
    assert o2["weight_out"][3] == 0.0 and np.all(o2["weight_out"][np.arange(len(b["tt"])) != 3] == 1.0)
    tek = o2["tt_fwd"][b["idx_path"][3]]
    extra = 0.5 * (tek - b2["tt"][3]) ** 2 - 0.5 * (tek - b["tt"][3]) ** 2
    assert np.isclose(o2["chi"] - o1["chi"], extra, rtol=1e-9)
    # the outlier's kernel contribution is gone: item 0 differs, others don't
    assert not np.array_equal(o1["adj_s"], o2["adj_s"])


def test_static_block_partition_matches_single_rank():
    b = small_batch(2, nsta=7)
    o1 = O.forward_simulate(b, nthreads=1)
    o3 = O.forward_simulate(b, nthreads=3)
    assert np.array_equal(o1["tt_fwd"], o3["tt_fwd"])
    assert np.isclose(o1["chi"], o3["chi"], rtol=1e-14)
    assert np.allclose(o1["adj_s"], o3["adj_s"], rtol=1e-13, atol=1e-13 * np.abs(o1["adj_s"]).max())


def test_make_batch_orders_follow_the_reference():
    """isrcs: stations in get_sta order, periods in first-appearance order
    per source; gathers in path order (src/src_rec.f90:162-179,252-283,338-369)."""
    cols = {
        "evtname": ["B", "B", "A", "B", "A", "A"],
        "staname": ["A", "C", "B", "C", "C", "B"],
        "evla": [1.0, 1.0, 2.0, 1.0, 2.0, 2.0], "evlo": [10.0, 10.0, 20.0, 10.0, 20.0, 20.0],
        "stla": [2.0, 3.0, 1.0, 3.0, 3.0, 1.0], "stlo": [20.0, 30.0, 10.0, 30.0, 30.0, 10.0],
        "period": [20.0, 20.0, 8.0, 8.0, 20.0, 20.0],
        "tt": np.arange(6.0) + 1, "weight": np.ones(6),
    }
    z = np.zeros((2, 4, 4))
    b = W.make_batch(2, np.arange(4.0), np.arange(4.0), None, z, z, z, z, cols)
    assert b["station_names"] == ["A", "C", "B"]        # receivers first, then events
    # station A (event A): periods 8 then 20; station C: not an event; station B: 20 then 8
    assert list(b["item_sta"]) == [0, 0, 2, 2]
    assert list(b["item_pidx"]) == [0, 1, 1, 0]
    assert list(b["rec_off"]) == [0, 1, 3, 5, 6]
    assert list(b["idx_path"]) == [2, 4, 5, 0, 1, 3]
    assert list(b["recx"]) == [10.0, 30.0, 10.0, 20.0, 30.0, 30.0]
    assert b["item_srcx"][0] == 20.0 and b["item_srcx"][2] == 10.0


def test_scatter_block_partition():
    """scatter_all_i (src/shared/parallel.f90:1877-1903) via the oracle's
    threads: every item computed once whatever the block count."""
    b = small_batch(2, nsta=5, nper=1)
    for nt in (1, 2, 4, 7):
        out = O.forward_simulate(b, nthreads=nt, isadj=False)
        assert np.all(out["iters"][:, 0] > 0)
