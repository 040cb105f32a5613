"""The oracle's restatement of the device's mixed-precision mode against the
reference (CPU): the FP32 phase plus FP64 polish must land within the mixed
mode's accuracy bar of the reference's FP64 results on the golden cases, and
on the homogeneous case within float resolution of the closed form."""
import numpy as np
import pytest

import oracle as O
from cases import CASES, solver_case
from conftest import rel_l2


@pytest.mark.parametrize("name", CASES)
def test_mixed_forward_close_to_reference(golden, name):
    c = solver_case(name)
    T, (it32, it64), st = O.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"],
                               c["x0"], c["y0"], mixed=True)
    Tref = golden[f"{name}/T"]
    assert st == 0
    assert it32 >= 1 and 1 <= it64 <= 2
    # seconds; the bar of the mixed mode is 1 ms, the design gives ~1e-6 s
    assert np.max(np.abs(T - Tref)) < 1e-5


@pytest.mark.parametrize("name", CASES)
def test_mixed_adjoint_close_to_reference(golden, name):
    c = solver_case(name)
    T = golden[f"{name}/T"]
    Ta, (it32, it64), st = O.adj(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], T, c["xr"],
                                 c["yr"], c["sa"], mixed=True)
    assert st == 0 and it32 >= 1 and it64 == 1
    assert rel_l2(Ta, golden[f"{name}/Ta"]) < 1e-6


def test_mixed_homogeneous_is_geodesic():
    n, d = 61, 0.03
    xx = 100.0 + d * np.arange(n); yy = 20.0 + d * np.arange(n)
    one = np.ones((n, n), order="F"); zero = np.zeros((n, n), order="F")
    fun = one / 3.2
    x0, y0 = xx[17] + 0.3 * d, yy[40] + 0.6 * d
    T, _, _ = O.fwd(xx, yy, one, one, zero, fun, x0, y0, mixed=True)
    T0, _, _ = O.t0_fields(xx, yy, fun, x0, y0)
    assert np.max(np.abs(T - T0)) <= 3e-8 * np.max(T0)


def test_mixed_source_on_boundary_is_an_error():
    c = solver_case("chk41")
    with pytest.raises(O.OracleError) as e:
        O.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["xx"][0] + 0.01,
              c["y0"], mixed=True)
    assert e.value.status == O.E_SRC_ON_BOUNDARY
