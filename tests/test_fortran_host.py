"""The Fortran host layer (libsatt_fhost.so, built with amdflang): link-level
drop-ins with the reference's subroutine names and argument lists, and
forward_simulate over the reference's src_rec arrays."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
import satt
import satt_workloads as W
from cases import solver_case
from conftest import rel_err

FHOST = os.path.join(os.path.dirname(satt.LIB_PATH), "libsatt_fhost.so")
_d = O._d


def test_fhost_exports_reference_names():
    assert os.path.exists(FHOST), "libsatt_fhost.so not built"
    lib = ctypes.CDLL(FHOST)
    for name in ("fsm_uw_ps_lonlat_2d_", "fsm_o1_jse_lonlat_2d_", "mask_",
                 "satt_fhost_forward_simulate", "satt_fhost_smooth", "satt_fhost_topo_metric",
                 "satt_fhost_steepest_descent", "satt_fhost_linesearch", "inv_grid_iso_",
                 "inv2fwd_iso_"):
        assert hasattr(lib, name), name


def _byref_i(v):
    return ctypes.byref(ctypes.c_int(v))


def _byref_d(v):
    return ctypes.byref(ctypes.c_double(v))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["chk41", "topo33x29", "cb81"])
def test_dropin_solvers_called_like_the_reference(ctx, golden, name):
    """Call fsm_uw_ps_lonlat_2d_ / fsm_o1_jse_lonlat_2d_ / mask_ exactly as
    src/measadj.f90:68-70,85-89 does (Fortran by-reference arguments)."""
    lib = ctypes.CDLL(FHOST)
    c = solver_case(name)
    nx, ny = len(c["xx"]), len(c["yy"])
    f = lambda a: np.asfortranarray(a, dtype=np.float64)
    spha, sphb, sphc, fun = f(c["spha"]), f(c["sphb"]), f(c["sphc"]), f(c["fun"])
    T = np.zeros((nx, ny), order="F"); u = np.zeros((nx, ny), order="F")
    lib.fsm_uw_ps_lonlat_2d_(_d(c["xx"]), _d(c["yy"]), _byref_i(nx), _byref_i(ny), _d(spha),
                             _d(sphb), _d(sphc), _d(T), _d(fun), _byref_d(c["x0"]),
                             _byref_d(c["y0"]), _d(u))
    Tc, _, _ = ctx.fwd(c["xx"], c["yy"], spha, sphb, sphc, fun, c["x0"], c["y0"])
    assert np.array_equal(T, Tc)
    assert rel_err(T, golden[f"{name}/T"]) < 1e-9
    Ta = np.zeros((nx, ny), order="F")
    xr, yr, sa = (np.ascontiguousarray(c[k]) for k in ("xr", "yr", "sa"))
    lib.fsm_o1_jse_lonlat_2d_(_d(c["xx"]), _d(c["yy"]), _byref_i(nx), _byref_i(ny), _d(spha),
                              _d(sphb), _d(sphc), _d(T), _d(Ta), _d(xr), _d(yr), _d(sa),
                              _byref_i(len(xr)))
    Tac, _, _ = ctx.adj(c["xx"], c["yy"], spha, sphb, sphc, T, xr, yr, sa)
    assert np.array_equal(Ta, Tac)
    lib.mask_(_d(c["xx"]), _d(c["yy"]), _byref_i(nx), _byref_i(ny), _d(Ta), _byref_d(c["x0"]),
              _byref_d(c["y0"]))
    assert np.array_equal(Ta, O.mask(c["xx"], c["yy"], Tac, c["x0"], c["y0"]))


def fhost_forward_simulate(b, paths, dim, precision=0):
    lib = ctypes.CDLL(FHOST)
    ids = lambda names: np.array([int(s[1:]) for s in names], dtype=np.int32)
    npd, nx, ny = b["nperiod"], b["nx"], b["ny"]
    npath = len(paths["evtname"])
    tt_fwd = np.zeros(npath); chi = ctypes.c_double(0); st = ctypes.c_int(0)
    adj_s = np.zeros((npd, nx, ny), order="F"); adj_d = np.zeros((npd, nx, ny), order="F")
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    evt, sta = ids(paths["evtname"]), ids(paths["staname"])
    args = [c(paths[k]) for k in ("evla", "evlo", "stla", "stlo", "period", "tt", "weight")]
    lib.satt_fhost_forward_simulate(
        dim, nx, ny, npd, _d(b["xgrids"]), _d(b["ygrids"]), _d(c(b["periods"])), _d(b["svel"]),
        _d(b["m11"]), _d(b["m12"]), _d(b["m22"]), npath, O._i(evt), O._i(sta),
        *(_d(a) for a in args), 1, 1, _d(tt_fwd), ctypes.byref(chi), _d(adj_s), _d(adj_d),
        ctypes.byref(st), int(precision))
    return st.value, tt_fwd, chi.value, adj_s, adj_d


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_fortran_forward_simulate_matches_c_abi(ctx, dim):
    """The Fortran host builds isrcs and the gathers from the path columns
    itself; its result equals the C-ABI batch built by the Python host bit for
    bit, and the oracle within the FP64 bar."""
    n, d = 37, 0.03
    xg = 100.0 + d * np.arange(n); yg = 20.0 + d * np.arange(n)
    rng = np.random.default_rng(21)
    sx, sy = W.stations_uniform(rng, 5, xg, yg)
    periods = np.array([7.0, 12.0])
    paths = W.all_pairs_paths(sx, sy, periods)
    vt = np.stack([(2.9 + 0.03 * p) * (1 + W.checkerboard_2d(n, n, 2, 2, 0.06)) for p in periods])
    one = np.ones((2, n, n)); zero = np.zeros((2, n, n))
    bt = W.make_batch(dim, xg, yg, periods, vt, one, zero, one, paths)
    obs = O.forward_simulate(bt, isadj=False)["tt_fwd"]
    paths["tt"] = obs
    b = W.make_batch(dim, xg, yg, periods, np.stack([(2.9 + 0.03 * p) * np.ones((n, n)) for p in periods]),
                     one, zero, one, paths)
    st, tt_fwd, chi, adj_s, adj_d = fhost_forward_simulate(b, paths, dim)
    assert st >= 0
    out = ctx.forward_simulate(b)
    assert np.array_equal(tt_fwd, out["tt_fwd"]) and chi == out["chi"]
    assert np.array_equal(adj_s, out["adj_s"])
    if dim == 3:
        assert np.array_equal(adj_d, out["adj_density"])
    ref = O.forward_simulate(b)
    assert rel_err(tt_fwd, ref["tt_fwd"]) < 1e-9
    # the mixed-precision form through the same host routine
    st, tt_m, chi_m, adj_m, _ = fhost_forward_simulate(b, paths, dim, satt.PREC_MIXED)
    out_m = ctx.forward_simulate(b, precision=satt.PREC_MIXED)
    assert st >= 0 and np.array_equal(tt_m, out_m["tt_fwd"]) and np.array_equal(adj_m, out_m["adj_s"])


@pytest.mark.gpu
def test_fortran_model_side_routines_match_c_abi(ctx):
    """gaussian_smooth_geo_2, the get_topo metric loop and optimize +
    steepest_descent called through the Fortran host equal the C-ABI calls
    bit for bit; the step length shrinks when the misfit rose
    (src/tomo2d.f90:238-246)."""
    lib = ctypes.CDLL(FHOST)
    nx, ny = 30, 24
    xg = 100.0 + 0.03 * np.arange(nx); yg = 25.0 + 0.03 * np.arange(ny)
    rng = np.random.default_rng(8)
    z = np.asfortranarray(rng.uniform(0, 2, (nx, ny)))
    sm = np.zeros((nx, ny), order="F")
    lib.satt_fhost_smooth(nx, ny, _d(z), _d(xg), _d(yg), ctypes.c_double(0.08), _d(sm))
    assert np.array_equal(sm, ctx.smooth_geo2(xg, yg, z, 0.08))
    periods = np.array([5.0, 10.0, 20.0]); vel = np.array([3.0, 3.3, 3.6])
    outs = [np.zeros((3, nx, ny), order="F") for _ in range(4)]
    lib.satt_fhost_topo_metric(nx, ny, 3, _d(xg), _d(yg), _d(z), _d(vel), _d(periods),
                               ctypes.c_double(2.5), *(_d(o) for o in outs))
    sig = [satt.topo_sigma(v, p, 2.5) for v, p in zip(vel, periods)]
    want = ctx.topo_metric(xg, yg, z, sig)
    assert all(np.array_equal(a, b) for a, b in zip(outs, want))
    adj = np.asfortranarray(rng.uniform(-1, 1, (3, nx, ny)))
    svel = np.asfortranarray(rng.uniform(2.8, 3.6, (3, nx, ny)))
    for now, prev, step in ((2.0, 1.0, 0.02 * 0.8), (1.0, 2.0, 0.02)):
        sv = svel.copy(order="F"); ker = np.zeros((3, nx, ny), order="F")
        upd = ctypes.c_double(0.02)
        lib.satt_fhost_steepest_descent(nx, ny, 3, _d(xg), _d(yg), _d(adj), ctypes.c_double(0.07),
                                        ctypes.c_double(now), ctypes.c_double(prev),
                                        ctypes.c_double(0.8), ctypes.byref(upd), _d(sv), _d(ker))
        assert upd.value == step
        wsv, wker = ctx.steepest_descent(xg, yg, adj, 0.07, step, svel)
        assert np.array_equal(sv, wsv) and np.array_equal(ker, wker)


@pytest.mark.gpu
def test_fortran_linesearch_matches_c_abi(ctx):
    """att_tomo_2d%linesearch through the Fortran host (it builds the batch
    from the path columns and runs satt_plan_linesearch on a plan of its
    own) equals the Python plan's line search bit for bit."""
    lib = ctypes.CDLL(FHOST)
    n, d = 33, 0.03
    xg = 100.0 + d * np.arange(n); yg = 20.0 + d * np.arange(n)
    rng = np.random.default_rng(23)
    sx, sy = W.stations_uniform(rng, 4, xg, yg)
    periods = np.array([8.0, 16.0])
    paths = W.all_pairs_paths(sx, sy, periods)
    cb = W.checkerboard_2d(n, n, 2, 2, 0.06)
    ker = np.asfortranarray(np.stack([cb, 0.5 * cb]))
    v0 = np.stack([(2.9 + 0.03 * p) * np.ones((n, n)) for p in periods])
    one = np.ones((2, n, n)); zero = np.zeros((2, n, n))
    bt = W.make_batch(2, xg, yg, periods, v0 * (1 + ker), one, zero, one, paths)
    paths["tt"] = O.forward_simulate(bt, isadj=False)["tt_fwd"]
    b = W.make_batch(2, xg, yg, periods, v0, one, zero, one, paths)
    chi0 = O.forward_simulate(b, istotable=False, isadj=False)["chi"]
    step = 4 * float(np.max(np.abs(cb)))
    with ctx.plan(b) as p:
        want = p.linesearch(b["svel"], ker, chi0, step, 0.25, 4)
    ids = lambda names: np.array([int(s[1:]) for s in names], dtype=np.int32)
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    sv = np.asfortranarray(b["svel"]).copy(order="F")
    upd, chi = ctypes.c_double(step), ctypes.c_double(0)
    nsub, st = ctypes.c_int(0), ctypes.c_int(0)
    args = [c(paths[k]) for k in ("evla", "evlo", "stla", "stlo", "period", "tt", "weight")]
    lib.satt_fhost_linesearch(
        2, n, n, 2, _d(xg), _d(yg), _d(c(periods)), _d(sv), _d(b["m11"]), _d(b["m12"]),
        _d(b["m22"]), len(paths["tt"]), O._i(ids(paths["evtname"])), O._i(ids(paths["staname"])),
        *(_d(a) for a in args), _d(ker), ctypes.c_double(chi0), ctypes.c_double(0.25), 4,
        ctypes.byref(upd), ctypes.byref(chi), ctypes.byref(nsub), ctypes.byref(st))
    assert st.value >= 0
    assert nsub.value == want["nsub"] == 2 and upd.value == want["updatemax"]
    assert chi.value == want["chi"] and np.array_equal(sv, want["svel"])


@pytest.mark.gpu
def test_projection_dropins_called_like_the_reference(ctx):
    """inv_grid_iso_ / inv2fwd_iso_ of libsatt_fhost.so called with the
    reference's by-reference argument lists (src/acqui.f90:274-279) equal
    the C-ABI calls bit for bit."""
    fh = O.RefSolver(FHOST)            # the same Fortran ABI as the reference build
    vx, vy, vz = np.linspace(100, 101, 21), np.linspace(30, 31, 19), np.linspace(0, 40, 11)
    xi, yi, zi = W.inv_grids(vx, vy, vz, (7, 6, 5), 3)
    adj = np.asfortranarray(np.random.default_rng(2).standard_normal((21, 19, 11)))
    gk = fh.inv_grid_iso(xi, yi, zi, adj, vx, vy, vz)
    assert np.array_equal(gk, ctx.inv_grid_iso(xi, yi, zi, adj, vx, vy, vz))
    up = fh.inv2fwd_iso(xi, yi, zi, gk, vx, vy, vz)
    assert np.array_equal(up, ctx.inv2fwd_iso(xi, yi, zi, gk, vx, vy, vz))
