"""Model-side steps on the GPU (satt_smooth_geo2, satt_gradient_geo2,
satt_topo_metric, satt_steepest_descent) against the oracle.

Tolerances: the smoothing evaluates a weight per (row pair, column offset),
so it differs from the reference's per-node-pair weight by the rounding of
lon(i')-lon(i) and by the device's exp/atan2/sin/cos; it is held to
1e-12 x max|data|.  The gradient is the reference's arithmetic on the same
operands (cosd from the host) and is compared bit for bit.  The metric is
held to 1e-11 absolute (its entries are O(1)), the dip angle to 1e-9 deg."""
import numpy as np
import pytest

import oracle as O
import satt

pytestmark = pytest.mark.gpu


def axes(nx, ny, d, lon0=100.0, lat0=30.0):
    return lon0 + d * np.arange(nx), lat0 + d * np.arange(ny)


def field(nx, ny, seed, nf=None):
    rng = np.random.default_rng(seed)
    sh = (nx, ny) if nf is None else (nf, nx, ny)
    return rng.uniform(-1, 1, sh)


def smooth_close(got, want, data):
    err = np.max(np.abs(got - want))
    assert err <= 1e-12 * np.max(np.abs(data)), err


@pytest.mark.parametrize("nx,ny,d,sigmas", [
    (37, 29, 0.05, [0.04, 0.3]),        # windows inside the grid, and wider than it
    (70, 45, 0.02, [0.05, 0.011]),
    (130, 9, 0.01, [0.02, 0.5]),
    (1500, 3, 0.001, [0.01]),           # three cells per thread
    (2100, 2, 0.001, [0.004]),          # five cells per thread
])
def test_smoothing_matches_oracle(ctx, nx, ny, d, sigmas):
    lons, lats = axes(nx, ny, d)
    data = field(nx, ny, nx + ny, len(sigmas))
    out = ctx.smooth_geo2(lons, lats, data, sigmas)
    assert out.shape == (len(sigmas), nx, ny)
    for f, s in enumerate(sigmas):
        smooth_close(out[f], O.smooth_geo2(lons, lats, data[f], s), data[f])
    # one field, every sigma (the topography case)
    one = ctx.smooth_geo2(lons, lats, data[0], sigmas)
    for f, s in enumerate(sigmas):
        smooth_close(one[f] if len(sigmas) > 1 else one, O.smooth_geo2(lons, lats, data[0], s),
                     data[0])


def test_smoothing_large_grid_rows_match_oracle(ctx):
    """401 x 401 at 0.005 deg, sigma 0.1 (a 121-node x window): rows at both
    edges and in the middle against the oracle."""
    lons, lats = axes(401, 401, 0.005, 100.0, 20.0)
    z = field(401, 401, 7)
    out = ctx.smooth_geo2(lons, lats, z, 0.1)
    for j in (0, 200, 400):
        want = O.smooth_geo2(lons, lats, z, 0.1, rows=(j, j + 1))
        smooth_close(out[:, j], want[:, j], z)


def test_smoothing_is_deterministic_and_keeps_constants(ctx):
    lons, lats = axes(50, 40, 0.03)
    z = field(50, 40, 9)
    a = ctx.smooth_geo2(lons, lats, z, 0.07)
    b = ctx.smooth_geo2(lons, lats, z, 0.07)
    assert np.array_equal(a, b)
    assert np.all(ctx.smooth_geo2(lons, lats, np.ones((50, 40)), 0.07) == 1.0)


def test_gradient_bitwise(ctx):
    lons, lats = axes(33, 27, 0.04, 95.0, -10.0)
    f = field(33, 27, 3, 3)
    tx, ty = ctx.gradient_geo2(lons, lats, f)
    for p in range(3):
        wx, wy = O.gradient_geo2(lons, lats, f[p])
        assert np.array_equal(tx[p], wx) and np.array_equal(ty[p], wy)
    t1, u1 = ctx.gradient_geo2(lons, lats, f[1])
    assert np.array_equal(t1, tx[1]) and np.array_equal(u1, ty[1])


def hills(lons, lats, seed):
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(lons, lats, indexing="ij")
    z = np.zeros_like(X)
    for _ in range(5):
        cx, cy = rng.uniform(lons[0], lons[-1]), rng.uniform(lats[0], lats[-1])
        z += rng.uniform(0.5, 4) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / rng.uniform(0.005, 0.05))
    return z


def test_topo_metric_matches_oracle(ctx):
    lons, lats = axes(40, 33, 0.02, 100.0, 20.0)
    z = hills(lons, lats, 4)
    sig = [satt.topo_sigma(3.1, T, 2.5) for T in (3.0, 8.0, 20.0)]
    m11, m12, m22, ang = ctx.topo_metric(lons, lats, z, sig)
    w11, w12, w22, wag = O.topo_metric(lons, lats, z, sig)
    for got, want in ((m11, w11), (m12, w12), (m22, w22)):
        assert np.max(np.abs(got - want)) <= 1e-11
    assert np.max(np.abs(ang - wag)) <= 1e-9
    assert np.max(np.abs(m12)) > 1e-3        # the topography is not flat
    m = ctx.topo_metric(lons, lats, z, sig, angle=False)
    assert len(m) == 3 and all(np.array_equal(a, b) for a, b in zip(m, (m11, m12, m22)))


def test_topo_sigma_matches_oracle():
    for v, T, wf in ((3.2, 10.0, 2.5), (3.9, 40.0, 5.0)):
        assert satt.topo_sigma(v, T, wf) == O.topo_sigma(v, T, wf)


def test_steepest_descent_matches_oracle(ctx):
    lons, lats = axes(45, 38, 0.02)
    adj = field(45, 38, 11, 3) * np.array([1.0, 5.0, 0.2])[:, None, None]
    svel = 3.0 + 0.5 * field(45, 38, 12, 3)
    sv, ker = ctx.steepest_descent(lons, lats, adj, 0.06, 0.02, svel)
    wsv, wker = O.steepest_descent(lons, lats, adj, 0.06, 0.02, svel)
    assert np.max(np.abs(ker - wker)) <= 1e-12 * np.max(np.abs(adj))
    assert np.max(np.abs(sv - wsv) / wsv) <= 1e-12
    assert not np.array_equal(sv, svel)


def test_model_side_errors(ctx):
    lons, lats = axes(20, 15, 0.05)
    z = field(20, 15, 1)
    for bad in (0.0, -0.1, np.nan):
        with pytest.raises(satt.SattError) as e:
            ctx.smooth_geo2(lons, lats, z, bad)
        assert e.value.status == satt.E_ARG
    jitter = lons.copy()
    jitter[7] += 0.01
    with pytest.raises(satt.SattError):
        ctx.smooth_geo2(jitter, lats, z, 0.1)
    with pytest.raises(satt.SattError):
        ctx.smooth_geo2(lons, lats[::-1].copy(), z, 0.1)
    wide = 100.0 + 1e-4 * np.arange(4097)     # one node above SATT_GEO_MAX_NX
    with pytest.raises(satt.SattError):
        ctx.smooth_geo2(wide, lats, np.zeros((len(wide), 15)), 0.01)
    # the context is still usable
    assert np.all(np.isfinite(ctx.smooth_geo2(lons, lats, z, 0.1)))
