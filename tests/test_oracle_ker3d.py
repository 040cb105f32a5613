"""The oracle's 3-D kernel steps (combine_kernels, inv_grid_iso, inv2fwd_iso,
post_processing_for_kernel).  CPU only.

inv_grid_iso and inv2fwd_iso live in the reference's solver file, which
oracle/_ref builds as-is: the restatement is compared with it bit for bit,
on the inversion grids the reference builds (get_inv_grids, several shifted
sets) and on grids whose nodes coincide with model nodes, where the
reference's walking search decides which cube a point on a face belongs to.
The rest is pinned by what the mathematics fixes: trilinear interpolation
reproduces trilinear fields, inv_grid_iso is the transpose of inv2fwd_iso on
the interior times the cell volume, the derivatives of Brocher's relations
are the derivatives of the polynomials, and combine_kernels is a signed sum
over periods."""
import numpy as np
import pytest

import oracle as O
import satt_workloads as W

REF = O.RefSolver() if O.RefSolver.available() else None
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built")


def model_axes(nx=13, ny=11, nz=9):
    return (np.linspace(100.0, 101.2, nx), np.linspace(30.0, 31.0, ny),
            np.linspace(0.0, 40.0, nz))


def tie_grids(nset=2):
    """Fine axes 0..20 (step 1) and inversion nodes every 2 from 0 (x), -2
    (y) and 0 (z), shifted by whole fine steps per set: model nodes lie on
    inversion faces everywhere, including the first and last points."""
    vx, vy, vz = np.arange(21.0), np.arange(17.0), np.arange(13.0)
    xi = np.stack([np.arange(-2.0 * s, 23.0, 2.0)[:12] for s in range(nset)], 1)
    yi = np.stack([np.arange(-2.0, 19.0, 2.0)[:11] + s for s in range(nset)], 1)
    zi = np.stack([np.arange(0.0, 13.0, 2.0) for _ in range(nset)], 1)
    return vx, vy, vz, np.asfortranarray(xi), np.asfortranarray(yi), np.asfortranarray(zi)


def field(shape, seed):
    return np.asfortranarray(np.random.default_rng(seed).standard_normal(shape))


@needs_ref
@pytest.mark.parametrize("nset,n_inv", [(1, (6, 5, 5)), (5, (8, 7, 6)), (3, (5, 5, 4))])
def test_projection_bitwise_against_reference_inv_grids(nset, n_inv):
    vx, vy, vz = model_axes()
    xi, yi, zi = W.inv_grids(vx, vy, vz, n_inv, nset)
    adj = field((len(vx), len(vy), len(vz)), nset)
    gk = O.inv_grid_iso(xi, yi, zi, adj, vx, vy, vz)
    assert np.array_equal(gk, REF.inv_grid_iso(xi, yi, zi, adj, vx, vy, vz))
    up = O.inv2fwd_iso(xi, yi, zi, gk, vx, vy, vz)
    assert np.array_equal(up, REF.inv2fwd_iso(xi, yi, zi, gk, vx, vy, vz))


@needs_ref
def test_projection_bitwise_against_reference_on_faces():
    vx, vy, vz, xi, yi, zi = tie_grids(2)
    adj = field((len(vx), len(vy), len(vz)), 4)
    gk = O.inv_grid_iso(xi, yi, zi, adj, vx, vy, vz)
    assert np.array_equal(gk, REF.inv_grid_iso(xi, yi, zi, adj, vx, vy, vz))
    g = field(gk.shape, 5)
    assert np.array_equal(O.inv2fwd_iso(xi, yi, zi, g, vx, vy, vz),
                          REF.inv2fwd_iso(xi, yi, zi, g, vx, vy, vz))


def test_inv2fwd_reproduces_trilinear_fields():
    vx, vy, vz = model_axes()
    xi, yi, zi = W.inv_grids(vx, vy, vz, (6, 5, 5), 3)
    f = lambda x, y, z: 1.5 + 0.3 * x - 2.0 * y + 0.05 * z + 0.7 * x * y - 0.01 * y * z
    gk = np.concatenate([f(*np.meshgrid(xi[:, s], yi[:, s], zi[:, s], indexing="ij")).ravel("F")
                         for s in range(3)])
    up = O.inv2fwd_iso(xi, yi, zi, gk, vx, vy, vz)
    want = 3 * f(*np.meshgrid(vx, vy, vz, indexing="ij"))        # summed over the sets
    assert np.max(np.abs(up - want)) < 1e-10 * np.max(np.abs(want))


def test_inv_grid_is_the_transpose_of_inv2fwd():
    vx, vy, vz = model_axes()
    xi, yi, zi = W.inv_grids(vx, vy, vz, (7, 6, 5), 2)
    a = field((len(vx), len(vy), len(vz)), 6)
    g = field((2 * 7 * 6 * 5,), 7).ravel()
    vol = (vx[1] - vx[0]) * (vy[1] - vy[0]) * (vz[1] - vz[0])
    inner = a[1:-1, 1:-1, 1:-1]
    lhs = np.sum(O.inv2fwd_iso(xi, yi, zi, g, vx, vy, vz)[1:-1, 1:-1, 1:-1] * inner) * vol
    rhs = np.dot(O.inv_grid_iso(xi, yi, zi, a, vx, vy, vz), g)
    assert abs(lhs - rhs) < 1e-12 * (abs(lhs) + 1)
    # the weights of a point sum to one: the total is the interior sum times the volume
    assert abs(O.inv_grid_iso(xi, yi, zi, a, vx, vy, vz).sum() / 2 - inner.sum() * vol) < 1e-12


def test_point_outside_the_inversion_grid_is_an_error():
    vx, vy, vz = model_axes()
    xi, yi, zi = W.inv_grids(vx, vy, vz, (6, 5, 5), 1)
    xi = np.asfortranarray(xi + 0.5)                    # the first model columns fall outside
    with pytest.raises(O.OracleError):
        O.inv2fwd_iso(xi, yi, zi, np.zeros(6 * 5 * 5), vx, vy, vz)


def test_brocher_derivatives():
    """Derivatives of Brocher's Vp(Vs) and rho(Vp) polynomials (the
    reference's empirical_vp / empirical_rho), to the REAL(4) rounding of
    the reference's literals."""
    vp_of_vs = np.polynomial.Polynomial([0.9409, 2.0947, -0.8206, 0.2683, -0.0251])
    rho_of_vp = np.polynomial.Polynomial([0, 1.6612, -0.4721, 0.0671, -0.0043, 0.000106])
    for poly, fn, xs in ((vp_of_vs, O.dalpha_dbeta, (1.0, 2.5, 3.7, 4.6)),
                         (rho_of_vp, O.drho_dalpha, (1.5, 4.0, 6.1, 8.0))):
        d = poly.deriv()
        for x in xs:
            scale = sum(abs(c) * x ** k for k, c in enumerate(d.coef))   # float rounding of each term
            assert abs(fn(x) - d(x)) < 1e-7 * scale


def sens(npd, nx, ny, nz, seed):
    r = np.random.default_rng(seed)
    mk = lambda *s: np.asfortranarray(r.standard_normal(s))
    vs = np.asfortranarray(r.uniform(2.5, 4.5, (nx, ny, nz)))
    return (mk(npd, nx, ny), mk(npd, nx, ny), mk(npd, nx, ny, nz), mk(npd, nx, ny, nz),
            mk(npd, nx, ny, nz), vs, np.asfortranarray(1.73 * vs))


def test_combine_alpha_beta_rho():
    a, ad, svs, svp, srho, vs, vp = sens(4, 5, 6, 3, 1)
    neg = lambda x, s: -np.einsum("pij,pijk->ijk", x, s)
    for rs in (True, False):
        ker, kd = O.combine_kernels(a, ad, svs, svp, srho, vs, vp, True, rs)
        assert ker.shape == (3, 5, 6, 3)
        for c, s in ((0, svs), (1, svp)):
            assert np.allclose(ker[c], neg(a, s), rtol=0, atol=1e-13)
            assert np.allclose(kd[c], neg(ad, s), rtol=0, atol=1e-13)
        if rs:
            assert np.array_equal(ker[2], 0.33 * ker[0]) and np.array_equal(kd[2], kd[0])
        else:
            assert np.allclose(ker[2], neg(a, srho), rtol=0, atol=1e-13)
    # accumulates into what is there
    k2, _ = O.combine_kernels(a, ad, svs, svp, srho, vs, vp, True, False, ker=ker, ker_den=kd)
    assert np.allclose(k2, 2 * ker, rtol=0, atol=1e-12)


def test_combine_through_the_vs_chain_rule():
    a, ad, svs, svp, srho, vs, vp = sens(3, 4, 5, 6, 2)
    ker, kd = O.combine_kernels(a, ad, svs, svp, srho, vs, vp, False, True)
    dab = np.vectorize(O.dalpha_dbeta)(vs)
    dra = np.vectorize(O.drho_dalpha)(vp)
    tot = svs + svp * dab[None] + srho * (dra * dab)[None]
    assert np.allclose(ker[0], -np.einsum("pij,pijk->ijk", a, tot), rtol=0, atol=1e-12)
    assert np.allclose(kd[0], -np.einsum("pij,pijk->ijk", ad, tot), rtol=0, atol=1e-12)


def test_postproc_composition_and_preconditioning():
    vx, vy, vz = model_axes()
    xi, yi, zi = W.inv_grids(vx, vy, vz, (6, 5, 5), 2)
    sh = (len(vx), len(vy), len(vz))
    ker = np.asfortranarray(field(sh, 8)[None])
    kd = np.asfortranarray(np.abs(field(sh, 9))[None])
    g0 = np.asfortranarray(field((1,) + sh, 10))
    k1, g1 = O.kernel_postproc(vx, vy, vz, xi, yi, zi, ker, kd, 0.0, 2.5, 0.7, g0)
    assert np.array_equal(k1, ker)                       # no preconditioning at coe 0
    up = O.inv2fwd_iso(xi, yi, zi, O.inv_grid_iso(xi, yi, zi, ker[0], vx, vy, vz), vx, vy, vz)
    assert np.array_equal(g1[0], g0[0] + up / 2.0 / 2.5 * 0.7)
    kd[0, 3, 4, 5] = 1e3                                # the maximum; others fall below 1e-4
    kd[0, 1, 1, 1] = 0.5
    k2, _ = O.kernel_postproc(vx, vy, vz, xi, yi, zi, ker, kd, 0.5, 2.5, 0.7, g0)
    assert k2[0, 3, 4, 5] == ker[0, 3, 4, 5]             # precond 1 at the maximum
    assert k2[0, 1, 1, 1] == ker[0, 1, 1, 1] * (1 / (0.5 / 1e3) ** 0.5)
    small = np.abs(kd[0]) / 1e3 < 1e-4
    assert np.all(k2[0][small] == 0) and small.sum() > 10
    assert np.all(k2[0][~small] != 0)


def test_interp1_definition():
    """interp1_0: linear between the bracketing nodes, the node value at a
    node, the last value past the last node."""
    x = np.array([0.0, 2.0, 3.0, 7.0])
    v = np.array([1.0, -3.0, 4.0, 8.0])
    assert O.interp1(x, v, 2.5) == (-3.0 * 0.5 + 4.0 * 0.5) / 1.0
    assert O.interp1(x, v, 1.0) == (1.0 * 1.0 + -3.0 * 1.0) / 2.0
    assert O.interp1(x, v, 3.0) == 4.0 and O.interp1(x, v, 0.0) == 1.0
    assert O.interp1(x, v, 9.5) == 8.0 and O.interp1(x, v, 7.0) == 8.0
    lin = 0.3 * x - 2.0
    for q in (0.1, 2.9, 5.5):
        assert abs(O.interp1(x, lin, q) - (0.3 * q - 2.0)) < 1e-15


def test_correct_depth_stretches_profiles():
    """dip 0 keeps every profile; dip 60 degrees (cos 1/2) moves the value at
    depth 2z up to z, a linear profile stays linear until 2z leaves the grid,
    then holds the last value."""
    depz = np.linspace(0.0, 40.0, 21)
    npd, nx, ny = 2, 3, 2
    r = np.random.default_rng(3)
    k = np.asfortranarray(r.standard_normal((npd, nx, ny, 21)))
    lin = np.asfortranarray(np.broadcast_to(1.0 + 0.05 * depz, (npd, nx, ny, 21)))
    dip = np.zeros((npd, nx, ny), order="F")
    out = O.correct_depth(k, lin, k, dip, depz)
    assert np.allclose(out[0], k, rtol=1e-15, atol=1e-15) and np.allclose(out[1], lin, rtol=1e-15)
    dip[:] = 60.0
    out = O.correct_depth(k, lin, k, dip, depz)
    z2 = depz / np.cos(np.deg2rad(60.0))
    want = np.where(z2 <= 40.0, 1.0 + 0.05 * z2, 1.0 + 0.05 * 40.0)
    assert np.allclose(out[1], np.broadcast_to(want, out[1].shape), rtol=0, atol=1e-13)
    # node k goes to node 2k where 2k is a node (cos 60 = 1/2 to rounding)
    assert np.allclose(out[0][..., 3], k[..., 6], rtol=0, atol=1e-12)
