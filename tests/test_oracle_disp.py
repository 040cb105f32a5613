"""The oracle's restatement around the reference's dispersion code: get_vprho
(src/surfker/surfker.f90:306-320) pinned against Brocher's (2005)
polynomials.  surfdisp96 itself is the reference's own file, built into
oracle/_ref.  CPU only."""
import numpy as np

import oracle as O


def test_brocher_restatement_matches_the_polynomials():
    """get_vprho's REAL(4) restatement against Brocher's (2005) Vp(Vs) and
    rho(Vp) polynomials in double precision."""
    vs = np.linspace(0.5, 4.8, 200)
    vp, rho = O.brocher_vp_rho(vs)
    vp64 = 0.9409 + 2.0947 * vs - 0.8206 * vs ** 2 + 0.2683 * vs ** 3 - 0.0251 * vs ** 4
    rho64 = (1.6612 * vp64 - 0.4721 * vp64 ** 2 + 0.0671 * vp64 ** 3 - 0.0043 * vp64 ** 4
             + 0.000106 * vp64 ** 5)
    np.testing.assert_allclose(vp, vp64, rtol=2e-6)
    np.testing.assert_allclose(rho, rho64, rtol=2e-6)
    assert vp.dtype == np.float32 and rho.dtype == np.float32
    # Vp/Vs about 1.7 in the crust
    assert 1.65 < float(vp[150] / np.float32(vs[150])) < 1.85
