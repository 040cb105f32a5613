"""The 3-D kernel steps on the device (satt_combine_kernels,
satt_inv_grid_iso, satt_inv2fwd_iso, satt_kernel_postproc) against the
oracle, itself equal to the reference's inv_grid_iso / inv2fwd_iso bit for
bit.  The combination and the projections are the reference's arithmetic in
its order and are compared bit for bit, on the reference's shifted
inversion grids and on grids with model points on inversion faces (where
the reference's walking search picks the cube).  The density
preconditioning takes pow(), whose last bit may differ between the device
and the host: 1e-13 relative."""
import numpy as np
import pytest

import oracle as O
import satt
import satt_workloads as W
from test_oracle_ker3d import model_axes, sens, tie_grids

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("abr,rs", [(True, True), (True, False), (False, True)])
def test_combine_bitwise(ctx, abr, rs):
    a, ad, svs, svp, srho, vs, vp = sens(5, 9, 7, 6, 3)
    k0 = np.asfortranarray(np.random.default_rng(4).standard_normal((3 if abr else 1, 9, 7, 6)))
    got = ctx.combine_kernels(a, ad, svs, svp, srho, vs, vp, abr, rs, ker=k0, ker_density=-k0)
    want = O.combine_kernels(a, ad, svs, svp, srho, vs, vp, abr, rs, ker=k0, ker_den=-k0)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert ctx.last_device_ms() > 0


def grids_cases():
    vx, vy, vz = model_axes()
    out = [("ref1", (vx, vy, vz) + W.inv_grids(vx, vy, vz, (6, 5, 5), 1)),
           ("ref5", (vx, vy, vz) + W.inv_grids(vx, vy, vz, (8, 7, 6), 5)),
           ("faces", tie_grids(2))]
    bx, by, bz = np.linspace(100, 102, 41), np.linspace(30, 32, 41), np.linspace(0, 60, 31)
    out.append(("acqui41", (bx, by, bz) + W.inv_grids(bx, by, bz, (10, 10, 8), 5)))
    return out


@pytest.mark.parametrize("name,case", grids_cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_projection_bitwise(ctx, name, case):
    vx, vy, vz, xi, yi, zi = case
    adj = np.asfortranarray(np.random.default_rng(11).standard_normal((len(vx), len(vy), len(vz))))
    gk = ctx.inv_grid_iso(xi, yi, zi, adj, vx, vy, vz)
    assert np.array_equal(gk, O.inv_grid_iso(xi, yi, zi, adj, vx, vy, vz))
    g = np.random.default_rng(12).standard_normal(gk.shape)
    up = ctx.inv2fwd_iso(xi, yi, zi, g, vx, vy, vz)
    assert np.array_equal(up, O.inv2fwd_iso(xi, yi, zi, g, vx, vy, vz))


@pytest.mark.parametrize("nker,coe", [(1, 0.0), (3, 0.0), (1, 0.5), (3, 0.8)])
def test_postproc_matches_oracle(ctx, nker, coe):
    bx, by, bz = np.linspace(100, 102, 41), np.linspace(30, 32, 41), np.linspace(0, 60, 31)
    xi, yi, zi = W.inv_grids(bx, by, bz, (10, 10, 8), 4)
    r = np.random.default_rng(nker + int(10 * coe))
    sh = (nker, 41, 41, 31)
    ker = np.asfortranarray(r.standard_normal(sh))
    kd = np.asfortranarray(r.standard_normal(sh) * np.exp(3 * r.standard_normal(sh)))
    g0 = np.asfortranarray(r.standard_normal(sh))
    k1, g1 = ctx.kernel_postproc(bx, by, bz, xi, yi, zi, ker, kd, coe, 3.7, 0.6, g0)
    wk, wg = O.kernel_postproc(bx, by, bz, xi, yi, zi, ker, kd, coe, 3.7, 0.6, g0)
    if coe == 0:
        assert np.array_equal(k1, ker) and np.array_equal(g1, wg)
    else:
        assert np.max(np.abs(k1 - wk)) <= 1e-13 * np.max(np.abs(wk))
        assert np.max(np.abs(g1 - wg)) <= 1e-13 * np.max(np.abs(wg - g0))
        assert np.any(k1 == 0) and not np.array_equal(k1, ker)


def test_point_outside_is_an_error(ctx):
    vx, vy, vz = model_axes()
    xi, yi, zi = W.inv_grids(vx, vy, vz, (6, 5, 5), 2)
    bad = np.asfortranarray(xi + 0.5)
    with pytest.raises(satt.SattError) as e:
        ctx.inv2fwd_iso(bad, yi, zi, np.zeros(2 * 6 * 5 * 5), vx, vy, vz)
    assert e.value.status == satt.E_ARG
    with pytest.raises(satt.SattError):
        ctx.inv_grid_iso(xi, yi, zi[::-1].copy(order="F"), np.zeros((13, 11, 9), order="F"),
                         vx, vy, vz)
    # still usable
    assert ctx.inv2fwd_iso(xi, yi, zi, np.ones(2 * 6 * 5 * 5), vx, vy, vz).shape == (13, 11, 9)


def test_correct_depth_matches_oracle(ctx):
    depz = np.linspace(0.0, 60.0, 31)
    r = np.random.default_rng(21)
    sh = (5, 7, 6, 31)
    ks = [np.asfortranarray(r.standard_normal(sh)) for _ in range(3)]
    dip = np.asfortranarray(r.uniform(0.0, 40.0, sh[:3]))
    dip[0, 0, 0] = 0.0
    got = ctx.correct_depth(*ks, dip, depz)
    want = O.correct_depth(*ks, dip, depz)
    for g, w in zip(got, want):
        assert np.max(np.abs(g - w)) <= 1e-13 * np.max(np.abs(w))
    assert np.array_equal(got[0][0, 0, 0], ks[0][0, 0, 0]) or \
        np.allclose(got[0][0, 0, 0], ks[0][0, 0, 0], rtol=1e-15)
    with pytest.raises(satt.SattError):
        ctx.correct_depth(*ks, dip + 90.0, depz)


def test_projection_edge_grids(ctx):
    """Inversion nodes on every model node (a tie in every axis at every
    point), a two-node inversion grid, and a model with no interior point
    (inv_grid_iso then gives zeros)."""
    vx, vy, vz = np.arange(9.0), np.arange(7.0), np.arange(5.0)
    cases = [(vx[:, None].copy(order="F"), vy[:, None].copy(order="F"), vz[:, None].copy(order="F")),
             (np.asfortranarray([[-1.0], [9.0]]), np.asfortranarray([[0.0], [6.0]]),
              np.asfortranarray([[-0.5], [4.5]]))]
    rng = np.random.default_rng(31)
    adj = np.asfortranarray(rng.standard_normal((9, 7, 5)))
    for xi, yi, zi in cases:
        gk = ctx.inv_grid_iso(xi, yi, zi, adj, vx, vy, vz)
        assert np.array_equal(gk, O.inv_grid_iso(xi, yi, zi, adj, vx, vy, vz))
        g = rng.standard_normal(gk.shape)
        assert np.array_equal(ctx.inv2fwd_iso(xi, yi, zi, g, vx, vy, vz),
                              O.inv2fwd_iso(xi, yi, zi, g, vx, vy, vz))
    flat = np.asfortranarray(rng.standard_normal((9, 7, 2)))
    xi, yi, zi = cases[1]
    zi2 = np.asfortranarray([[-0.5], [1.5]])
    gk = ctx.inv_grid_iso(xi, yi, zi2, flat, vx, vy, vz[:2])
    assert np.all(gk == 0) and np.array_equal(gk, O.inv_grid_iso(xi, yi, zi2, flat, vx, vy, vz[:2]))
