"""The oracle (CPU restatement) pinned against the reference's own solver
file and against what the mathematics fixes.  CPU only."""
import numpy as np
import pytest

import oracle as O
from cases import CASES, solver_case

PI_F = float(np.float32(np.pi))


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_golden_bitwise(golden, name):
    """FSM_UW_PS_lonlat_2d, FSM_O1_JSE_lonlat_2d and mask of the reference
    (src/eikon_solver_2d.f90, compiled as-is) reproduced bit for bit."""
    c = solver_case(name)
    for k in ("spha", "sphb", "sphc", "fun", "xr", "yr", "sa"):
        assert np.array_equal(golden[f"{name}/in_{k}"], c[k]), f"case generator drifted: {k}"
    T, it, st = O.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["x0"], c["y0"])
    assert st == O.OK and it >= 2
    assert np.array_equal(T, golden[f"{name}/T"])
    Ta, ita, st = O.adj(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], T, c["xr"], c["yr"], c["sa"])
    assert st == O.OK
    assert np.array_equal(Ta, golden[f"{name}/Ta"])
    assert np.array_equal(O.mask(c["xx"], c["yy"], Ta, c["x0"], c["y0"]), golden[f"{name}/Ta_masked"])


@pytest.mark.skipif(not O.RefSolver.available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("seed", [11, 12, 13])
def test_oracle_matches_live_reference(seed):
    """Random heterogeneous media and metrics against the reference build."""
    rng = np.random.default_rng(seed)
    nx, ny = rng.integers(12, 48, 2)
    d = float(rng.uniform(0.01, 0.1))
    xx = 90.0 + d * np.arange(nx); yy = -30.0 + 60 * rng.random() + d * np.arange(ny)
    v = rng.uniform(2.5, 4.0, (nx, ny))
    fx = rng.uniform(-0.4, 0.4, (nx, ny)); fy = rng.uniform(-0.4, 0.4, (nx, ny))
    den = 1 + fx ** 2 + fy ** 2
    a, b, c = (1 + fy ** 2) / den, (1 + fx ** 2) / den, fx * fy / den
    x0 = xx[0] + d * rng.uniform(2, nx - 3); y0 = yy[0] + d * rng.uniform(2, ny - 3)
    ref = O.RefSolver()
    T, _, _ = O.fwd(xx, yy, a, b, c, 1 / v, x0, y0)
    assert np.array_equal(T, ref.fwd(xx, yy, a, b, c, 1 / v, x0, y0))
    nr = 15
    xr = xx[0] + d * rng.uniform(1, nx - 2, nr); yr = yy[0] + d * rng.uniform(1, ny - 2, nr)
    sa = rng.standard_normal(nr)
    Ta, _, _ = O.adj(xx, yy, a, b, c, T, xr, yr, sa)
    assert np.array_equal(Ta, ref.adj(xx, yy, a, b, c, T, xr, yr, sa))


def _geodesic(xx, yy, x0, y0, s):
    xr, yr = xx / 180.0 * PI_F, yy / 180.0 * PI_F
    x0r, y0r = x0 / 180.0 * PI_F, y0 / 180.0 * PI_F
    t = np.sin(yr)[None, :] * np.sin(y0r) + np.cos(yr)[None, :] * np.cos(y0r) * np.cos(xr[:, None] - x0r)
    return np.arccos(np.clip(t, -1, 1)) * 6371.0 * s


@pytest.mark.parametrize("n,d,on_node", [(41, 0.05, True), (81, 0.025, False), (61, 0.1, False)])
def test_forward_homogeneous_is_geodesic(n, d, on_node):
    """Homogeneous slowness, flat metric: the factored scheme is exact, so
    T is the great-circle distance times slowness (P1's pi), up to the
    conditioning of acos near 1 and the REAL(4) source weights (~1e-8)."""
    xx = 100 + d * np.arange(n); yy = 20 + d * np.arange(n)
    one = np.ones((n, n)); s = 1 / 3.0
    x0 = xx[17] + (0 if on_node else 0.37 * d); y0 = yy[11] + (0 if on_node else 0.61 * d)
    T, it, st = O.fwd(xx, yy, one, one, 0 * one, s * one, x0, y0)
    Tc = _geodesic(xx, yy, x0, y0, s)
    m = Tc > 0
    assert st == O.OK
    assert np.max(np.abs(T[m] / Tc[m] - 1)) < 3e-8


def test_forward_frozen_cells_equal_t0_and_tau_bounds():
    c = solver_case("topo33x29")
    T, _, _ = O.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["x0"], c["y0"])
    T0, _, _ = O.t0_fields(c["xx"], c["yy"], c["fun"], c["x0"], c["y0"])
    xr, yr = c["xx"] / 180 * PI_F, c["yy"] / 180 * PI_F
    x0, y0 = c["x0"] / 180 * PI_F, c["y0"] / 180 * PI_F
    fr = (np.abs((xr[:, None] - x0) / (xr[1] - xr[0])) <= 1) & (np.abs((yr[None, :] - y0) / (yr[1] - yr[0])) <= 1)
    assert fr.sum() >= 1
    assert np.array_equal(T[fr], T0[fr])
    tau = T[~fr] / T0[~fr]
    assert np.all(tau <= 10) and np.all(tau > 0.5)


def test_forward_source_on_boundary_is_an_error():
    c = solver_case("chk41")
    with pytest.raises(O.OracleError) as e:
        O.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["xx"][0] + 0.3 * 0.05, c["y0"])
    assert e.value.status == O.E_SRC_ON_BOUNDARY


def test_adjoint_linear_zero_and_boundary():
    c = solver_case("topo33x29")
    T, _, _ = O.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["x0"], c["y0"])
    args = (c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], T, c["xr"], c["yr"])
    Ta1, _, _ = O.adj(*args, c["sa"])
    Ta2, _, _ = O.adj(*args, 2.5 * c["sa"])
    sb = np.random.default_rng(7).standard_normal(len(c["sa"]))
    Tb, _, _ = O.adj(*args, sb)
    Tab, _, _ = O.adj(*args, c["sa"] + sb)
    Tz, _, _ = O.adj(*args, 0 * c["sa"])
    scale = np.abs(Ta1).max()
    # converged to the absolute 1e-9 stop rule from the Ta=100 start
    assert np.abs(Ta2 - 2.5 * Ta1).max() < 1e-6 * scale
    assert np.abs(Tab - (Ta1 + Tb)).max() < 1e-6 * scale
    assert np.abs(Tz).max() < 1e-8
    for t in (Ta1, Tb):
        assert np.all(t[0, :] == 0) and np.all(t[-1, :] == 0) and np.all(t[:, 0] == 0) and np.all(t[:, -1] == 0)


def test_adjoint_single_signed_source_gives_single_signed_field():
    c = solver_case("chk41")
    T, _, _ = O.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["x0"], c["y0"])
    Ta, _, _ = O.adj(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], T, c["xr"], c["yr"], np.abs(c["sa"]))
    assert Ta.min() >= -1e-8 and Ta.max() > 0


def test_mask_pattern():
    xx = 100 + 0.1 * np.arange(20); yy = 30 + 0.2 * np.arange(15)
    t = np.random.default_rng(1).random((20, 15))
    x0, y0 = xx[7] + 0.04, yy[5] - 0.05
    m = O.mask(xx, yy, t, x0, y0)
    want = t.copy()
    sel = (np.abs(xx[:, None] - x0) < 0.15) & (np.abs(yy[None, :] - y0) < 0.3)
    want[sel] = 0
    assert np.array_equal(m, want) and sel.sum() == 9


def test_interp2_exact_for_bilinear_fields_and_nodes():
    xx = 100 + 0.05 * np.arange(30); yy = 20 + 0.04 * np.arange(25)
    X, Y = np.meshgrid(xx, yy, indexing="ij")
    v = 1.5 + 0.3 * X - 0.7 * Y + 0.02 * X * Y
    rng = np.random.default_rng(3)
    xq = rng.uniform(xx[0], xx[-2], 50); yq = rng.uniform(yy[0], yy[-2], 50)
    out = O.interp2(xx, yy, v, xq, yq)
    exact = 1.5 + 0.3 * xq - 0.7 * yq + 0.02 * xq * yq
    assert np.max(np.abs(out - exact)) < 1e-11
    nodes = O.interp2(xx, yy, v, xx[[3, 10]], yy[[4, 9]])
    assert np.allclose(nodes, v[[3, 10], [4, 9]], rtol=0, atol=1e-12)
    with pytest.raises(O.OracleError):
        O.interp2(xx, yy, v, [xx[-1]], [yy[3]])   # last node: the reference reads out of bounds
