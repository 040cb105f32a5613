"""The oracle's line search (att_tomo_2d%linesearch, src/tomo2d.f90:253-288,
with prepare_fwd_linesearch, src/acqui_2d.f90:173-183), pinned by what its
definition fixes.  CPU only.

The case: a checkerboard cb, ker_s = (cb, cb/2) over the two periods,
observed times from the model vt = v0*(1 + ker_s) and the current model v0
homogeneous, so a step of updatemax = max|ker_s| (STEP) lands on the true
model.  The checks are chosen so that a plausible
mistake fails one of them: '>=' instead of '>' in the misfit comparison, the
model kept after a failed search formed with the shrunk step instead of the
last one tried, the base of each trial being the previous trial instead of
svel, the maximum over one period instead of the whole kernel."""
import numpy as np
import pytest

import oracle as O
import satt_workloads as W

PERT = 0.07
STEP = float(np.max(np.abs(W.checkerboard_2d(31, 27, 2, 2, PERT))))


def ls_case(nx=31, ny=27, dim=2, nsta=5, seed=3, d=0.03):
    """(batch at v0 with observed times of vt, ker_s = cb, chi at v0)."""
    xg = 100.0 + d * np.arange(nx)
    yg = 20.0 + d * np.arange(ny)
    rng = np.random.default_rng(seed)
    sx, sy = W.stations_uniform(rng, nsta, xg, yg)
    periods = np.array([6.0, 15.0])
    paths = W.all_pairs_paths(sx, sy, periods)
    cb = W.checkerboard_2d(nx, ny, 2, 2, PERT)
    ker = np.asfortranarray(np.stack([cb, 0.5 * cb]))
    v0 = np.stack([(2.9 + 0.03 * p) * np.ones((nx, ny)) for p in periods])
    vt = v0 * (1 + ker)
    one, zero = np.ones((2, nx, ny)), np.zeros((2, nx, ny))
    bt = W.make_batch(dim, xg, yg, periods, vt, one, zero, one, paths)
    obs = O.forward_simulate(bt, istotable=True, isadj=False)["tt_fwd"]
    b = W.with_model(W.set_observed(bt, obs), v0)
    chi0 = O.forward_simulate(b, istotable=False, isadj=False)["chi"]
    return b, ker, chi0


@pytest.fixture(scope="module")
def case():
    return ls_case()


def test_step_onto_the_true_model_is_taken_at_once(case):
    b, ker, chi0 = case
    r = O.linesearch(b, ker, chi0, STEP, 0.5, 4)
    assert r["nsub"] == 1 and r["updatemax"] == STEP
    assert r["chi"] < 1e-6 * chi0            # the true model up to rounding
    vt = b["svel"] * (1 + ker)
    assert np.max(np.abs(r["svel"] / vt - 1)) < 1e-15


def test_overshoot_shrinks_until_the_misfit_drops(case):
    b, ker, chi0 = case
    r = O.linesearch(b, ker, chi0, 4 * STEP, 0.25, 5)   # 4x too long, then right
    assert r["nsub"] == 2 and r["updatemax"] == STEP
    # the second trial starts from svel again, so it is the true model
    assert r["chi"] < 1e-6 * chi0


def test_equal_misfit_is_accepted():
    """updatemax 0 gives svel itself, whose misfit equals chi0: the search
    stops (the reference shrinks only when the misfit is larger)."""
    b, ker, chi0 = ls_case(25, 21)
    r = O.linesearch(b, ker, chi0, 0.0, 0.5, 3)
    assert r["nsub"] == 1 and r["chi"] == chi0 and r["updatemax"] == 0.0
    assert np.array_equal(r["svel"], b["svel"])


def test_failed_search_keeps_the_last_model_tried(case):
    b, ker, chi0 = case
    n, up, shrink = 3, 0.02, 0.5
    r = O.linesearch(b, -ker, chi0, up, shrink, n)      # uphill: every trial is worse
    assert r["nsub"] == n and r["chi"] > chi0
    assert r["updatemax"] == up * shrink ** n
    # the kept model has the step of the last trial, up*shrink**(n-1), at the
    # kernel's global maximum (in period 0); period 1 moves half as far
    rel = np.abs(r["svel"] / b["svel"] - 1)
    last = up * shrink ** (n - 1)
    assert abs(rel[0].max() - last) < 1e-14
    assert abs(rel[1].max() - 0.5 * last) < 1e-14


def test_no_sub_iteration_keeps_svel(case):
    b, ker, chi0 = case
    r = O.linesearch(b, ker, chi0, STEP, 0.5, 0)
    assert r["nsub"] == 0 and r["chi"] == chi0 and np.array_equal(r["svel"], b["svel"])
