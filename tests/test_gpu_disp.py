"""Surface-wave dispersion on the GPU (satt_fwdsurf3d) against the reference's
own surfdisp96 (src/surfker/surfdisp96.f, built into oracle/_ref) driven
column by column through the oracle's fwdsurf1d.

Tolerances: the device follows the reference's REAL(4)/REAL(8) typing
variable by variable, so the roots differ only through the device's libm
(sin, cos, exp, log, sqrt, pow) in the last bits; the refinement stops at a
relative bracket of 1e-6.  Phase velocities (rounded to REAL(4) as the
reference returns them) are held to 2e-6 relative, group velocities, a
difference quotient over +-0.5% in period that amplifies root errors about
100-fold, to 1e-4."""
import numpy as np
import pytest

import oracle as O
import satt

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not O.RefSolver.available(), reason="oracle/_ref not built")]


def column_models(nx, ny, nz=31, zmax=60.0, seed=0, pert=0.06):
    """Vs rising from 3.0 to 4.5 km/s over 0..zmax km with a +-pert 3-D
    checkerboard and small random jitter (velocity still rising with depth)."""
    rng = np.random.default_rng(seed)
    depz = np.linspace(0.0, zmax, nz)
    vs1 = 3.0 + 1.5 * depz / zmax
    xp = np.sin(2 * np.pi * np.arange(1, nx + 1) / nx)
    yp = np.sin(3 * np.pi * np.arange(1, ny + 1) / ny)
    zp = np.sin(2 * np.pi * np.arange(1, nz + 1) / nz)
    vs = vs1[None, None, :] * (1 + pert * xp[:, None, None] * yp[None, :, None] * zp[None, None, :])
    vs = vs * (1 + 0.002 * rng.standard_normal((nx, ny, 1)))
    return np.asfortranarray(vs), depz


@pytest.mark.parametrize("iwave", [1, 2])
@pytest.mark.parametrize("igr", [0, 1])
def test_dispersion_matches_reference(ctx, iwave, igr):
    vs, depz = column_models(7, 5, seed=iwave + 2 * igr)
    periods = np.geomspace(4.0, 40.0, 16)
    got = ctx.fwdsurf3d(vs, depz, periods, iwave, igr)
    ref = O.RefSolver()
    tol = 2e-6 if igr == 0 else 1e-4
    worst = 0.0
    for i in range(vs.shape[0]):
        for j in range(vs.shape[1]):
            want = O.fwdsurf1d(ref, vs[i, j], iwave, igr, periods, depz)
            rel = np.max(np.abs(got[:, i, j] - want) / want)
            worst = max(worst, rel)
    assert worst <= tol, worst
    assert ctx.last_device_ms() > 0


def test_dispersion_other_depth_grids_and_period_orders(ctx):
    """Uneven layer thicknesses, more layers, periods short to long and long
    to short (each search starts from the previous period's root)."""
    rng = np.random.default_rng(5)
    depz = np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 4.0, 44))])
    vs1 = 2.8 + 1.8 * depz / depz[-1]
    vs = np.asfortranarray(np.broadcast_to(vs1, (3, 2, len(depz))) *
                           (1 + 0.01 * rng.standard_normal((3, 2, 1))))
    ref = O.RefSolver()
    for periods in (np.geomspace(3.0, 50.0, 12), np.geomspace(50.0, 3.0, 12)):
        got = ctx.fwdsurf3d(vs, depz, periods, 2, 0)
        for i in range(3):
            for j in range(2):
                want = O.fwdsurf1d(ref, vs[i, j], 2, 0, periods, depz)
                assert np.max(np.abs(got[:, i, j] - want) / want) <= 2e-6


def test_dispersion_argument_errors(ctx):
    vs, depz = column_models(2, 2)
    periods = np.array([5.0, 10.0])
    for kw in (dict(iwave=3), dict(igr=-1)):
        with pytest.raises(satt.SattError) as e:
            ctx.fwdsurf3d(vs, depz, periods, **kw)
        assert e.value.status == satt.E_ARG
    with pytest.raises(satt.SattError):
        ctx.fwdsurf3d(vs, depz[::-1].copy(), periods)
    with pytest.raises(satt.SattError):
        ctx.fwdsurf3d(vs, depz, np.array([5.0, -1.0]))
