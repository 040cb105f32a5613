"""The oracle's model-side steps -- gaussian_smooth_geo_2, gradient_2_geo,
get_topo's metric and the steepest-descent update -- pinned by what the
mathematics fixes, since the reference's src/shared/utils.f90 does not build
here (its module needs a CMake-generated header).  CPU only.

Each check is an independent statement of the reference's behaviour (a
closed form, a 1-D reduction or an invariant), chosen so that a plausible
mistake fails one of them: the single-precision pi of the distance, sigma
against 2 sigma, the window half-widths including the row's cos(lat)
scaling of the y half-width, the one-sided edge differences, a swapped fx/fy
in the metric, a per-period rather than global max in the update."""
import numpy as np
import pytest

import oracle as O

PI_F = float(np.float32(np.pi))   # src/shared/shared_par.f90:11
R = 6371.0


def chord_km(lat0, lon0, lat1, lon1, pi=PI_F):
    """Great-circle distance from the chord between unit vectors."""
    d = pi / 180.0

    def vec(la, lo):
        return np.array([np.cos(la * d) * np.cos(lo * d), np.cos(la * d) * np.sin(lo * d),
                         np.sin(la * d)])
    c = np.linalg.norm(vec(lat0, lon0) - vec(lat1, lon1))
    return 2 * R * np.arcsin(c / 2)


def test_gps2dist_is_the_great_circle_with_single_precision_pi():
    rng = np.random.default_rng(0)
    for _ in range(200):
        lat0, lat1 = rng.uniform(-60, 60, 2)
        lon0, lon1 = rng.uniform(80, 120, 2)
        if abs(lat0 - lat1) + abs(lon0 - lon1) < 0.2:
            continue
        d = O.gps2dist(lat0, lon0, lat1, lon1)
        assert d == pytest.approx(chord_km(lat0, lon0, lat1, lon1), rel=1e-12)
    # along the equator the distance is R*dlon*pi/180, 2.8e-8 away from the
    # true pi's: the check above tells the two apart
    d = O.gps2dist(0.0, 100.0, 0.0, 101.5)
    assert d == pytest.approx(R * 1.5 * PI_F / 180, rel=1e-13)
    assert abs(d / (R * 1.5 * np.pi / 180) - 1) > 2e-8


def gauss_1d(x, f, i, h, sigma):
    """Normalised Gaussian of x-distance over |k| <= h around node i."""
    k = np.arange(max(0, i - h), min(len(x) - 1, i + h) + 1)
    w = np.exp(-((x[k] - x[i]) ** 2) / (2 * sigma ** 2))
    return np.sum(w * f[k]) / np.sum(w)


def test_smoothing_along_the_equator_is_a_1d_gaussian_in_degrees():
    """One row at latitude 0 (the other row out of the window): the great-
    circle distance is the longitude difference, the window is
    int(3 sigma/dx)."""
    nx, dx, sigma = 40, 0.01, 0.047
    lons = 100.0 + dx * np.arange(nx)
    lats = np.array([0.0, 10.0])
    f = np.random.default_rng(1).uniform(-1, 1, (nx, 2))
    hx, hy = O.smooth_halfwidths(lons, lats, sigma)
    assert hx == 14 and hy[0] == 0
    out = O.smooth_geo2(lons, lats, f, sigma)
    for i in range(nx):
        assert out[i, 0] == pytest.approx(gauss_1d(lons, f[:, 0], i, 14, sigma), rel=1e-12, abs=1e-13)
    # sigma and 2 sigma, and a window one node wider, are far apart
    assert abs(out[20, 0] - gauss_1d(lons, f[:, 0], 20, 14, 2 * sigma)) > 1e-3
    assert abs(out[20, 0] - gauss_1d(lons, f[:, 0], 20, 15, sigma)) > 1e-9


def test_smoothing_along_a_meridian_widens_the_window_by_one_over_cos_lat():
    """One column (the other column out of the window) at latitude 60: the
    distance is the latitude difference, and the y half-width is
    int(3 sigma/cosd(lat)/dy) -- twice the plain int(3 sigma/dy) here."""
    ny, dy, sigma = 60, 0.01, 0.0087
    lons = np.array([100.0, 110.0])
    lats = 60.0 + dy * np.arange(ny)
    f = np.random.default_rng(2).uniform(-1, 1, (2, ny))
    hx, hy = O.smooth_halfwidths(lons, lats, sigma)
    assert hx == 0
    want_h = [int(3 * sigma / np.cos(np.deg2rad(la)) / dy) for la in lats]
    assert hy == want_h and min(want_h) == 5 and int(3 * sigma / dy) == 2
    out = O.smooth_geo2(lons, lats, f, sigma)
    for j in range(ny):
        assert out[0, j] == pytest.approx(gauss_1d(lats, f[0], j, want_h[j], sigma), rel=1e-12,
                                          abs=1e-13)
    assert abs(out[0, 30] - gauss_1d(lats, f[0], 30, 2, sigma)) > 1e-4


@pytest.mark.parametrize("c", [1.0, 0.5, -2.0])
def test_smoothing_keeps_a_power_of_two_constant_exactly(c):
    lons = 95.0 + 0.03 * np.arange(23)
    lats = 35.0 + 0.04 * np.arange(19)
    out = O.smooth_geo2(lons, lats, np.full((23, 19), c), 0.09)
    assert np.all(out == c)


def test_smoothing_keeps_a_field_linear_in_longitude_inside():
    """East-west symmetric weights: a field linear in longitude is kept
    where the x window is not cut by the edges."""
    nx, ny, dx, sigma = 41, 21, 0.02, 0.05
    lons = 110.0 + dx * np.arange(nx)
    lats = 25.0 + dx * np.arange(ny)
    f = np.repeat((lons - 110.0)[:, None], ny, axis=1)
    hx, _ = O.smooth_halfwidths(lons, lats, sigma)
    out = O.smooth_geo2(lons, lats, f, sigma)
    inner = slice(hx, nx - hx)
    assert np.max(np.abs(out[inner] - f[inner])) < 1e-12


def test_smoothing_spike_reaches_exactly_the_windows_that_hold_it():
    nx, ny, sigma = 31, 27, 0.07
    lons = 100.0 + 0.03 * np.arange(nx)
    lats = 40.0 + 0.025 * np.arange(ny)
    i0, j0 = 12, 14
    f = np.zeros((nx, ny))
    f[i0, j0] = 1.0
    # the spacings are lons(2)-lons(1), lats(2)-lats(1) (src/shared/utils.f90:
    # 1209-1210): 100.03-100 is a little above 0.03, so 0.21/dx truncates to 6
    dx, dy = lons[1] - lons[0], lats[1] - lats[0]
    hx = int(3 * sigma / dx)
    hy = [int(3 * sigma / np.cos(np.deg2rad(la)) / dy) for la in lats]
    assert hx == 6
    assert O.smooth_halfwidths(lons, lats, sigma) == (hx, hy)
    out = O.smooth_geo2(lons, lats, f, sigma)
    for j in range(ny):
        for i in range(nx):
            inside = abs(i - i0) <= hx and abs(j - j0) <= hy[j]
            assert (out[i, j] > 0) == inside, (i, j)


def km_per_deg(lat=None):
    s = R * PI_F / 180.0
    return s if lat is None else s * np.cos(np.deg2rad(lat))


def test_gradient_of_a_linear_field_everywhere():
    nx, ny = 17, 13
    lons = 100.0 + 0.05 * np.arange(nx)
    lats = -20.0 + 0.04 * np.arange(ny)
    al, be = 0.37, -1.3
    f = al * (lons[:, None] - 100.0) + be * (lats[None, :] + 20.0)
    tx, ty = O.gradient_geo2(lons, lats, f)
    np.testing.assert_allclose(tx, al / km_per_deg(lats)[None, :] * np.ones((nx, 1)), rtol=1e-11)
    np.testing.assert_allclose(ty, be / km_per_deg() * np.ones((nx, ny)), rtol=1e-11)


def test_gradient_of_a_quadratic_field_central_inside_one_sided_at_edges():
    nx, ny = 15, 11
    lons = 100.0 + 0.05 * np.arange(nx)
    lats = 10.0 + 0.04 * np.arange(ny)
    u = lons - 100.0
    v = lats - 10.0
    tx, _ = O.gradient_geo2(lons, lats, np.repeat((u ** 2)[:, None], ny, axis=1))
    k = km_per_deg(lats)[None, :]
    want = np.empty((nx, ny))
    want[1:-1] = (2 * u[1:-1])[:, None] / k
    want[0] = (u[0] + u[1]) / k[0]            # forward difference at the first column
    want[-1] = (u[-1] + u[-2]) / k[0]         # backward difference at the last
    np.testing.assert_allclose(tx, want, rtol=1e-10, atol=1e-14)
    _, ty = O.gradient_geo2(lons, lats, np.repeat((v ** 2)[None, :], nx, axis=0))
    wy = np.empty(ny)
    wy[1:-1] = 2 * v[1:-1]
    wy[0], wy[-1] = v[0] + v[1], v[-1] + v[-2]
    np.testing.assert_allclose(ty, np.repeat((wy / km_per_deg())[None, :], nx, axis=0),
                               rtol=1e-10, atol=1e-14)


def hills(lons, lats, seed):
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(lons, lats, indexing="ij")
    z = np.zeros_like(X)
    for _ in range(4):
        cx, cy = rng.uniform(lons[0], lons[-1]), rng.uniform(lats[0], lats[-1])
        z += rng.uniform(0.5, 3) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / rng.uniform(0.01, 0.1))
    return z


def test_topo_metric_inverts_to_the_smoothed_gradient():
    """m11 m22 - m12^2 = 1/(1+fx^2+fy^2) gives back fx^2, fy^2, fx fy from the
    metric alone; they must be those of the smoothed topography."""
    lons = 100.0 + 0.04 * np.arange(26)
    lats = 30.0 + 0.04 * np.arange(22)
    z = hills(lons, lats, 3)
    sig = [0.05, 0.12]
    m11, m12, m22, ang = O.topo_metric(lons, lats, z, sig)
    for p, s in enumerate(sig):
        fx, fy = O.gradient_geo2(lons, lats, O.smooth_geo2(lons, lats, z, s))
        D = 1.0 / (m11[p] * m22[p] - m12[p] ** 2)
        np.testing.assert_allclose(m22[p] * D - 1, fx ** 2, atol=1e-12)
        np.testing.assert_allclose(m11[p] * D - 1, fy ** 2, atol=1e-12)
        np.testing.assert_allclose(-m12[p] * D, fx * fy, atol=1e-12)
        np.testing.assert_allclose(np.tan(np.deg2rad(ang[p])), np.hypot(fx, fy), atol=1e-12)
        assert np.max(np.abs(fx)) > 0.01 and np.max(np.abs(fy)) > 0.01
        assert np.max(np.abs(fx ** 2 - fy ** 2)) > 1e-4


def test_topo_metric_of_flat_topography_is_the_identity():
    lons = 100.0 + 0.05 * np.arange(12)
    lats = 30.0 + 0.05 * np.arange(9)
    m11, m12, m22, ang = O.topo_metric(lons, lats, np.full((12, 9), 2.0), [0.1])
    assert np.all(m11 == 1) and np.all(m22 == 1) and np.all(m12 == 0) and np.all(ang == 0)


def test_topo_sigma_is_the_wavelength_fraction_in_degrees():
    for v, T, wf in ((3.2, 10.0, 2.5), (3.9, 40.0, 5.0)):
        lam_km = v * T
        want = lam_km * wf / (2 * PI_F) / (R * PI_F / 180.0)
        assert O.topo_sigma(v, T, wf) == pytest.approx(want, rel=1e-15)


def test_steepest_descent_scales_by_the_global_max():
    """sigma below a tenth of the spacing leaves the window one node wide, so
    ker_s is adj_s itself and the update is svel*(1 + u*adj/max|adj|) with
    one max over every period."""
    nx, ny = 9, 7
    lons = 100.0 + 0.1 * np.arange(nx)
    lats = 20.0 + 0.1 * np.arange(ny)
    rng = np.random.default_rng(5)
    adj = rng.uniform(-1, 1, (3, nx, ny)) * np.array([1.0, 10.0, 0.1])[:, None, None]
    svel = rng.uniform(2.5, 4, (3, nx, ny))
    sv, ker = O.steepest_descent(lons, lats, adj, 0.01, 0.02, svel)
    assert np.array_equal(ker, adj)
    want = svel * (1 + 0.02 * (adj / np.max(np.abs(adj))))
    assert np.array_equal(sv, want)
    per_period = svel * (1 + 0.02 * (adj / np.max(np.abs(adj), axis=(1, 2), keepdims=True)))
    assert np.max(np.abs(sv - per_period)) > 1e-3

