"""The allocation-free forward solve (GPU): satt_fwd_lonlat_noalloc against
satt_fwd_lonlat, the entry point whose kernels it shares, on two calls in a
row -- a 401 x 401 grid with a topographic metric, then a smaller 81 x 81
grid -- and back.  The same kernels run on the same operands, so T and the
iteration count must be equal bit for bit; a leftover of the larger call in
the shared scratch would show in the smaller one.  Grids beyond the scratch
are refused with SATT_E_ARG before any device work."""
import numpy as np
import pytest

import satt
import satt_workloads as W
from cases import solver_case

pytestmark = pytest.mark.gpu


def topo_case(n=401, d=0.005, seed=3):
    xg, yg, topo, periods, vref = W.topo_hills(n=n, d=d, nper=20, seed=seed)
    m11, m12, m22 = W.topo_metric(topo, xg, yg)
    v = vref[5] * (1 + W.checkerboard_2d(n, n, 4, 4, 0.08))
    rng = np.random.default_rng(seed)
    sx, sy = W.stations_uniform(rng, 1, xg, yg)
    f = lambda a: np.asfortranarray(a, dtype=np.float64)
    return dict(xx=xg, yy=yg, spha=f(m11), sphb=f(m22), sphc=f(-m12), fun=f(1.0 / v),
                x0=float(sx[0]), y0=float(sy[0]))


def _args(c):
    return (c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["x0"], c["y0"])


def test_noalloc_equals_allocating_entry_point(ctx):
    big, small = topo_case(), solver_case("cb81")
    assert big["spha"].size <= satt.NOALLOC_MAX_CELLS
    outs = []
    for c in (big, small, big):
        Tn, itn, stn = ctx.fwd_noalloc(*_args(c))
        T, it, st = ctx.fwd(*_args(c))
        assert stn == st == satt.OK
        assert itn == it
        assert np.array_equal(Tn, T)
        outs.append(Tn)
    assert np.array_equal(outs[0], outs[2])
    print(f"noalloc: 401^2 {itn} iterations, bitwise equal to satt_fwd_lonlat")


def test_noalloc_rejects_grids_beyond_the_scratch(ctx):
    for nx, ny in ((513, 513), (satt.NOALLOC_MAX_AXIS + 1, 3)):
        xx = 100.0 + 0.01 * np.arange(nx)
        yy = 20.0 + 0.01 * np.arange(ny)
        one = np.ones((nx, ny), order="F")
        with pytest.raises(satt.SattError) as e:
            ctx.fwd_noalloc(xx, yy, one, one, 0 * one, one, xx[nx // 2], yy[ny // 2])
        assert e.value.status == satt.E_ARG
