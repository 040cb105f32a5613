"""Per-solve HIP entry points against the oracle and the reference's golden
vectors (GPU)."""
import numpy as np
import pytest

import oracle as O
from cases import CASES, solver_case
from conftest import rel_err

pytestmark = pytest.mark.gpu

# north-star FP64 bar: receiver times within 1e-9 relative, kernels 1e-8 rel L2
TOL_T = 1e-9
TOL_K = 1e-8


@pytest.mark.parametrize("name", CASES)
def test_forward_matches_reference_golden(ctx, golden, name):
    c = solver_case(name)
    T, it, st = ctx.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["x0"], c["y0"])
    _, it_o, _ = O.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["x0"], c["y0"])
    assert st == 0
    assert it == it_o
    assert rel_err(T, golden[f"{name}/T"]) < TOL_T


@pytest.mark.parametrize("name", CASES)
def test_forward_sweeps_bitwise_given_device_setup(ctx, name):
    """With the device's own metric and T0 (its libm differs from glibc in the
    last bits), the wavefront sweeps reproduce the sequential sweeps exactly."""
    c = solver_case(name)
    a, b, cc, T0, T0x, T0y = ctx.fwd_setup_fields(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"],
                                                  c["fun"], c["x0"], c["y0"])
    T, it, _ = ctx.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["x0"], c["y0"])
    To, ito, _ = O.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["x0"], c["y0"],
                       pre=(a, b, cc, T0, T0x, T0y))
    assert it == ito
    assert np.array_equal(T, To)


@pytest.mark.parametrize("name", CASES)
def test_adjoint_matches_reference_golden(ctx, golden, name):
    c = solver_case(name)
    T = golden[f"{name}/T"]
    Ta, it, st = ctx.adj(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], T, c["xr"], c["yr"], c["sa"])
    _, it_o, _ = O.adj(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], T, c["xr"], c["yr"], c["sa"])
    assert st == 0 and it == it_o
    assert rel_err(Ta, golden[f"{name}/Ta"]) < TOL_K
    Tm = ctx.mask(c["xx"], c["yy"], Ta, c["x0"], c["y0"])
    assert np.array_equal(Tm == 0, golden[f"{name}/Ta_masked"] == 0)


@pytest.mark.parametrize("name", CASES)
def test_adjoint_sweeps_bitwise_given_device_setup(ctx, golden, name):
    c = solver_case(name)
    T = golden[f"{name}/T"]
    a, b, cc, delta = ctx.adj_setup_fields(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"],
                                           c["xr"], c["yr"], c["sa"])
    Ta, it, _ = ctx.adj(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], T, c["xr"], c["yr"], c["sa"])
    To, ito, _ = O.adj_pre(c["xx"], c["yy"], a, b, cc, T, delta)
    assert it == ito
    assert np.array_equal(Ta, To)


def test_mask_and_interp2_exact(ctx):
    c = solver_case("topo33x29")
    rng = np.random.default_rng(9)
    t = rng.random((len(c["xx"]), len(c["yy"])))
    assert np.array_equal(ctx.mask(c["xx"], c["yy"], t, c["x0"], c["y0"]),
                          O.mask(c["xx"], c["yy"], t, c["x0"], c["y0"]))
    xq = rng.uniform(c["xx"][0], c["xx"][-2], 200); yq = rng.uniform(c["yy"][0], c["yy"][-2], 200)
    xq[:3] = c["xx"][[0, 5, 9]]; yq[:3] = c["yy"][[2, 7, 1]]     # on nodes
    assert np.array_equal(ctx.interp2(c["xx"], c["yy"], t, xq, yq), O.interp2(c["xx"], c["yy"], t, xq, yq))


def test_error_paths(ctx):
    import satt
    c = solver_case("chk41")
    with pytest.raises(satt.SattError) as e:
        ctx.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["xx"][0] + 0.01, c["y0"])
    assert e.value.status == satt.E_SRC_ON_BOUNDARY
    T, _, _ = ctx.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["x0"], c["y0"])
    with pytest.raises(satt.SattError) as e:
        ctx.adj(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], T, [c["xx"][-1]], [c["y0"]], [1.0])
    assert e.value.status == satt.E_REC_OUT_OF_GRID
    with pytest.raises(satt.SattError) as e:
        ctx.interp2(c["xx"], c["yy"], T, [c["xx"][-1]], [c["y0"]])
    assert e.value.status == satt.E_REC_OUT_OF_GRID
