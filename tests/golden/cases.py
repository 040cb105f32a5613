"""Deterministic solver cases shared by the golden-vector script and the
tests (small grids the CPU restatement finishes in milliseconds)."""
import numpy as np


def solver_case(name):
    """Inputs of one forward + adjoint solve: grid axes (deg), metric
    (spha, sphb, sphc) as passed by the call site (m11, m22, -m12), slowness,
    source, receivers and an adjoint source."""
    spec = {
        "chk41": (41, 41, 0.05, False, 1, 24),
        "topo21x17": (21, 17, 0.1, True, 2, 12),
        "cb81": (81, 81, 0.025, False, 3, 24),
        "topo33x29": (33, 29, 0.04, True, 4, 30),
        "topo61": (61, 61, 0.02, True, 5, 40),
    }[name]
    nx, ny, d, topo, seed, nr = spec
    rng = np.random.default_rng(seed)
    xx = 100.0 + d * np.arange(nx)
    yy = 20.0 + d * np.arange(ny)
    I, J = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), indexing="ij")
    v = 3.0 * (1 + 0.05 * np.sin(4 * np.pi * I / nx) * np.sin(4 * np.pi * J / ny))
    v = v * (1 + 0.02 * rng.standard_normal((nx, ny)))
    if topo:
        fx = 0.3 * np.sin(I / 7.0) * np.cos(J / 5.0)
        fy = 0.2 * np.cos(I / 4.0 + J / 9.0)
        den = 1 + fx ** 2 + fy ** 2
        spha, sphb, sphc = (1 + fy ** 2) / den, (1 + fx ** 2) / den, fx * fy / den
    else:
        spha, sphb, sphc = np.ones((nx, ny)), np.ones((nx, ny)), np.zeros((nx, ny))
    x0 = xx[0] + d * (2 + rng.random() * (nx - 5))
    y0 = yy[0] + d * (2 + rng.random() * (ny - 5))
    xr = xx[0] + d * (1.5 + rng.random(nr) * (nx - 4))
    yr = yy[0] + d * (1.5 + rng.random(nr) * (ny - 4))
    sa = rng.standard_normal(nr)
    f = lambda a: np.asfortranarray(a, dtype=np.float64)
    return dict(xx=xx, yy=yy, spha=f(spha), sphb=f(sphb), sphc=f(sphc), fun=f(1.0 / v),
                x0=float(x0), y0=float(y0), xr=xr, yr=yr, sa=sa)


CASES = ["chk41", "topo21x17", "cb81", "topo33x29", "topo61"]
