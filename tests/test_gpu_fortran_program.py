"""A Fortran program that uses the satt_host module with the reference's
CHARACTER name columns (GPU): tests/fortran/satt_host_example04.f90 reads the
examples/04_hawaii_topo sample table (station and event names such as OTLD,
NPOC), and satt_forward_simulate_host matches the names into the station
list, the work list and the gathers itself (src/src_rec.f90:162-217,
252-283,338-369).  Its results equal those of the C-ABI on the batch that
satt_srcrec builds from the same table bit for bit, and the oracle's to the
FP64 bar."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import satt
import satt_srcrec as S
import satt_workloads as W
from conftest import rel_err, rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EX04 = os.path.join(HERE, "golden", "src_rec_example04_sample.csv")
PROG = os.path.join(os.path.dirname(satt.LIB_PATH), "_build", "satt_host_example04")


def model(t, n=61):
    per = S.periods(t)
    xg = np.linspace(t["stlo"].min() - 0.15, t["stlo"].max() + 0.15, n)
    yg = np.linspace(t["stla"].min() - 0.15, t["stla"].max() + 0.15, n)
    cb = W.checkerboard_2d(n, n, 3, 3, 0.05)
    svel = np.stack([(2.2 + 0.05 * p) * (1 + cb) for p in per])
    fx = 0.1 * np.sin(np.arange(n) / 6.0)[:, None] * np.ones((1, n))
    fy = 0.05 * np.cos(np.arange(n) / 4.0)[None, :] * np.ones((n, 1))
    den = 1 + fx ** 2 + fy ** 2
    rep = lambda f: np.repeat(f[None], len(per), 0)
    return xg, yg, svel, rep((1 + fy ** 2) / den), rep(-fx * fy / den), rep((1 + fx ** 2) / den)


@pytest.mark.parametrize("prec", [satt.PREC_FP64, satt.PREC_MIXED])
def test_fortran_program_with_character_names(ctx, tmp_path, prec):
    assert os.path.exists(PROG), "satt_host_example04 not built"
    t = S.read_src_rec(EX04)
    xg, yg, svel, m11, m12, m22 = model(t)
    n, npd = len(xg), svel.shape[0]
    f = lambda a: np.asfortranarray(a, dtype=np.float64).ravel(order="F")
    mod = tmp_path / "model.bin"
    np.concatenate([[n, n, npd], xg, yg, f(svel), f(m11), f(m12), f(m22)]).astype("<f8").tofile(mod)
    out = tmp_path / "out.bin"
    r = subprocess.run([PROG, EX04, str(mod), str(out), str(prec)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    v = np.fromfile(out, dtype="<f8")
    npath = len(t["tt"])
    st, chi, tt_fwd = int(v[0]), v[1], v[2:2 + npath]
    adj_s = v[2 + npath:].reshape((npd, n, n), order="F")
    assert st >= 0
    b = S.batch_from_src_rec(t, 2, xg, yg, svel, m11, m12, m22)
    ref = ctx.forward_simulate(b, precision=prec)
    assert np.array_equal(tt_fwd, ref["tt_fwd"])
    assert np.array_equal(adj_s, ref["adj_s"])
    assert chi == ref["chi"]
    if prec == satt.PREC_FP64:
        o = O.forward_simulate(b)
        assert rel_err(tt_fwd, o["tt_fwd"]) < 1e-9
        assert rel_l2(adj_s, o["adj_s"]) < 1e-8
