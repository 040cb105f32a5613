"""The Rayleigh-wave depth kernels on the device (GPU): satt_depthkernel, the
phase-velocity branch of depthkernel1d (src/surfker/surfker.f90:70-103) over
CPS sregn96, for every column of the 3-D stage's model (41 x 41 columns,
31 depths 0-60 km, 16 periods 4-40 s, the checkerboard of src/model.f90:276-300).

* The phase velocities at which the kernels are evaluated are the reference's
  own surfdisp96 (oracle/_ref) on refinegrid's layering, to its REAL(4)
  rounding.
* The kernels of sampled columns equal the oracle's sregn96 at those phase
  velocities to 1e-10 of each column's largest value (the oracle is pinned
  by finite differences of the reference's surfdisp96, tests/test_oracle_depth.py).
* Love waves and group velocities are refused."""
import os

import numpy as np
import pytest

import oracle as O
import satt
import satt_workloads as W

pytestmark = pytest.mark.gpu

HAVE_REF = os.path.exists(O.REF_PATH)


@pytest.fixture(scope="module")
def model():
    return W.acqui3d_model()


@pytest.fixture(scope="module")
def device(ctx, model):
    vs, depz, per = model
    return ctx.depthkernel(vs, depz, per)


def test_depthkernel_shapes_and_signs(device, model):
    sv, sp, sr, cp = device
    vs, depz, per = model
    assert sv.shape == sp.shape == sr.shape == (16, 41, 41, 31) and cp.shape == (16, 41, 41)
    assert np.all(np.isfinite(sv)) and np.all(np.isfinite(sp)) and np.all(np.isfinite(sr))
    # Vs sensitivity is positive near the surface and peaks deeper at longer periods
    assert np.all(sv[:, :, :, 0] > 0)
    depth_of_peak = depz[np.argmax(sv[:, 20, 20, :], axis=1)]
    assert depth_of_peak[-1] > depth_of_peak[0]


@pytest.mark.parametrize("col", [(0, 0), (7, 33), (20, 20), (40, 13), (29, 40)])
def test_depthkernel_columns_match_oracle(device, model, col):
    sv, sp, sr, cp = device
    vs, depz, per = model
    i, j = col
    c = cp[:, i, j]
    if HAVE_REF:
        ref = O.RefSolver()
        thk, vp, vsr, rho = O.refinegrid(vs[i, j], depz)
        want = ref.surfdisp96(thk, vp, vsr, rho, 2, 0, per)
        assert np.all(np.abs(c - want) <= 2.5e-7 * want), (c, want)
    osv, osp, osr = O.depthkernel_rayleigh(vs[i, j], depz, per, c)
    for dev, orc in ((sv, osv), (sp, osp), (sr, osr)):
        d = dev[:, i, j, :]
        scale = np.abs(orc).max(axis=1, keepdims=True)
        assert np.all(np.abs(d - orc) <= 1e-10 * scale), float(np.max(np.abs(d - orc) / scale))


def test_depthkernel_refuses_love_and_group(ctx, model):
    vs, depz, per = model
    for iwave, igr in ((1, 0), (2, 1)):
        with pytest.raises(satt.SattError) as e:
            ctx.depthkernel(vs[:3, :3], depz, per[:2], iwave=iwave, igr=igr)
        assert e.value.status == satt.E_ARG
