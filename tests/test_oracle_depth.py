"""The oracle's Rayleigh-wave depth kernels (CPS sregn96, src/surfker/sregn96.f90,
through depthkernel1d, src/surfker/surfker.f90:70-103), pinned without running
sregn96 itself (whose build needs the reference's CMake-generated version.h):

* finite differences of the reference's own surfdisp96 (oracle/_ref): the
  phase velocity of the layered model with one layer's Vs, Vp or density
  moved by +-1% gives dc/dVs, dc/dVp and dc/drho of that layer;
* the energy-integral group velocity against surfdisp96's group velocity;
* on a homogeneous flat half-space cut into layers (no sphericity), the
  Rayleigh velocity is xi*beta with xi from the Rayleigh equation, so
  sum(alpha dc/dalpha + beta dc/dbeta) = c and sum(rho dc/drho) = 0 exactly
  (Euler's theorem: c is homogeneous of degree 1 in the velocities and of
  degree 0 in density);
* refinegrid's layering: the depths' spacings, the last one repeated, and a
  half-space copy of the last layer.

A dropped term, a swapped index or a wrong sign fails one of these."""
import os

import numpy as np
import pytest

import oracle as O

HAVE_REF = os.path.exists(O.REF_PATH)


def column(nz=31, zmax=60.0):
    depz = np.linspace(0.0, zmax, nz)
    return depz, 3.0 + 1.5 * depz / zmax + 0.2 * np.sin(depz / 7.0)


def test_refinegrid_layers():
    depz, vs = column()
    thk, vp, vsr, rho = O.refinegrid(vs, depz)
    nz = len(depz)
    assert len(thk) == nz + 1
    d32 = depz.astype(np.float32)
    assert np.array_equal(thk[:nz - 1], d32[1:] - d32[:-1])
    assert thk[nz - 1] == thk[nz - 2] and thk[nz] == 0
    assert np.array_equal(vsr[:nz], vs.astype(np.float32))
    for a in (vp, vsr, rho):
        assert a[nz] == a[nz - 1]
    bvp, brho = O.brocher_vp_rho(vs.astype(np.float32))
    assert np.array_equal(vp[:nz], bvp) and np.array_equal(rho[:nz], brho)


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("period", [5.0, 12.0, 25.0, 40.0])
def test_kernels_match_finite_differences_of_surfdisp96(period):
    ref = O.RefSolver()
    depz, vs = column()
    thk, vp, vsr, rho = O.refinegrid(vs, depz)
    t = np.array([period])
    c0 = ref.surfdisp96(thk, vp, vsr, rho, 2, 0, t)[0]
    k = O.sregn96(thk, vp, vsr, rho, period, c0)
    h = 1e-2
    for name, key in (("vs", "dcdb"), ("vp", "dcda"), ("rho", "dcdr")):
        an = k[key]
        scale = np.abs(an[:-1]).max()
        for layer in (0, 2, 5, 9, 14, 20, 26, 30):
            fields = {"vp": vp, "vs": vsr, "rho": rho}
            up, dn = fields[name].copy(), fields[name].copy()
            up[layer] = np.float32(up[layer] * (1 + h))
            dn[layer] = np.float32(dn[layer] * (1 - h))
            pick = lambda a: tuple(a if n == name else fields[n] for n in ("vp", "vs", "rho"))
            cu = ref.surfdisp96(thk, *pick(up), 2, 0, t)[0]
            cd = ref.surfdisp96(thk, *pick(dn), 2, 0, t)[0]
            fd = (cu - cd) / (float(up[layer]) - float(dn[layer]))
            # surfdisp96's root is good to about 1e-6 relative and rounded to
            # REAL(4): a few 1e-5 in the quotient; the second-order truncation
            # stays below 1e-3 of the largest sensitivity
            assert abs(an[layer] - fd) <= 3e-3 * scale + 3e-5, (name, layer, an[layer], fd)
    # the group velocity of the energy integrals against surfdisp96's
    cg = ref.surfdisp96(thk, vp, vsr, rho, 2, 1, t)[0]
    assert abs(k["cg"] - cg) <= 2e-4 * cg


def rayleigh_xi(alpha, beta):
    """The Rayleigh velocity ratio c/beta of a homogeneous half-space: the
    root in (0, 1) of (2 - x^2)^2 = 4 sqrt(1 - x^2 b^2/a^2) sqrt(1 - x^2)."""
    g = (beta / alpha) ** 2
    f = lambda x: (2 - x * x) ** 2 - 4 * np.sqrt(1 - x * x * g) * np.sqrt(1 - x * x)
    lo, hi = 0.5, 0.999999
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(lo) * f(mid) <= 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


@pytest.mark.parametrize("period", [4.0, 20.0])
def test_half_space_euler_identities(period):
    n = 12
    thk = np.full(n, 3.0, dtype=np.float32)
    thk[-1] = 0
    vs = np.full(n, 3.5, dtype=np.float32)
    vp = np.full(n, 6.1, dtype=np.float32)
    rho = np.full(n, 2.7, dtype=np.float32)
    c = rayleigh_xi(float(vp[0]), float(vs[0])) * float(vs[0])
    k = O.sregn96(thk, vp, vs, rho, period, c, iflsph=0)
    assert k["cp"] == c and abs(k["cg"] - c) < 1e-7 * c   # no dispersion
    euler_v = float(np.sum(vp * k["dcda"] + vs * k["dcdb"]))
    euler_r = float(np.sum(rho * k["dcdr"]))
    assert abs(euler_v - c) < 1e-7 * c, euler_v
    assert abs(euler_r) < 1e-7 * c, euler_r
    # positive Vs sensitivity everywhere; at 4 s (a 13 km wavelength) the
    # layers below 30 km sense almost nothing
    assert np.all(k["dcdb"][:-1] > 0)
    if period < 10:
        assert k["dcdb"][10] < 1e-3 * k["dcdb"].max()


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference tree absent)")
def test_depthkernel_column_layout():
    """depthkernel1d's outputs for one column: sen_vs(kmax, nz) = dcdb of
    the first nz layers of each period's sregn96 run."""
    ref = O.RefSolver()
    depz, vs = column()
    per = np.array([6.0, 18.0, 33.0])
    thk, vp, vsr, rho = O.refinegrid(vs, depz)
    cp = ref.surfdisp96(thk, vp, vsr, rho, 2, 0, per)
    sv, sp, sr = O.depthkernel_rayleigh(vs, depz, per, cp)
    assert sv.shape == (3, 31)
    for i, t in enumerate(per):
        k = O.sregn96(thk, vp, vsr, rho, t, cp[i])
        assert np.array_equal(sv[i], k["dcdb"][:31])
        assert np.array_equal(sp[i], k["dcda"][:31])
        assert np.array_equal(sr[i], k["dcdr"][:31])
