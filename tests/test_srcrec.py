"""The source-receiver CSV files (satt_srcrec) on the reference's own example
data: the first 60 paths of examples/03_checkerboard_2d/src_rec_file_ph.csv
and every 100th path of examples/04_hawaii_topo/src_rec_file_raw.csv (a
different column order, no stel/evel), kept as fixtures under tests/golden.
CPU only.

The example-03 file carries tt, coordinates, period, weight and dist with
the decimals SrcRec%to_csv writes (F0.6, F9.4, F0.4, F6.4, F12.4), so
writing it back with its own times reproduces those columns' text, less the
stel/evel columns the reference drops."""
import os

import numpy as np
import pytest

import oracle as O
import satt_srcrec as S

GOLD = os.path.join(os.path.dirname(__file__), "golden")
EX03 = os.path.join(GOLD, "src_rec_example03_head.csv")
EX04 = os.path.join(GOLD, "src_rec_example04_sample.csv")


def raw_rows(path):
    with open(path) as f:
        lines = [ln.rstrip("\n").split(",") for ln in f if ln.strip()]
    return lines[0], lines[1:]


def test_read_example03():
    t = S.read_src_rec(EX03)
    hdr, rows = raw_rows(EX03)
    assert t["npath"] == len(rows) == 60
    assert t["header"] == ["tt", "staname", "stla", "stlo", "evtname", "evla", "evlo", "period",
                           "weight", "dist", "vel"]
    c = {h: k for k, h in enumerate(hdr)}
    for key, name in (("tt", "tt"), ("stla", "stla"), ("evlo", "evlo"), ("period", "period"),
                      ("weight", "weight"), ("dist", "dist"), ("vel", "vel")):
        assert np.array_equal(t[key], np.array([float(r[c[name]]) for r in rows]))
    assert t["staname"][:2] == ["MX01", "MX02"] and t["evtname"][0] == "MX00"


def test_read_example04_other_column_order():
    t = S.read_src_rec(EX04)
    hdr, rows = raw_rows(EX04)
    assert hdr[:3] == ["tt", "stla", "stlo"]          # not the reference's order
    assert t["header"] == ["tt", "staname", "stla", "stlo", "evtname", "evla", "evlo", "period",
                           "weight", "dist", "vel"]
    c = {h: k for k, h in enumerate(hdr)}
    assert t["evtname"] == [r[c["evtname"]] for r in rows]
    assert np.array_equal(t["evla"], [float(r[c["evla"]]) for r in rows])
    assert np.array_equal(t["dist"], [float(r[c["dist"]]) for r in rows])


def test_example03_distances_are_gps2dist():
    """The example-03 dist column is the great-circle distance gps2dist
    computes, to its 4 printed decimals (and the REAL(4) pi's 1e-7)."""
    t = S.read_src_rec(EX03)
    d = np.array([S.gps2dist(a, b, e, f) for a, b, e, f in
                  zip(t["stla"], t["stlo"], t["evla"], t["evlo"])])
    assert np.all(np.abs(d - t["dist"]) <= 5e-5 + 2e-7 * t["dist"])


def test_to_csv_reproduces_the_example_text(tmp_path):
    t = S.read_src_rec(EX03)
    out = tmp_path / "fwd.csv"
    S.to_csv(out, t, t["tt"])
    hdr, rows = raw_rows(EX03)
    keep = [k for k, h in enumerate(hdr) if h not in ("stel", "evel")]
    got = [ln.split(",") for ln in out.read_text().splitlines()]
    assert got[0] == [hdr[k] for k in keep]
    # every column but vel as printed in the file; vel = dist/tt is recomputed
    # from the printed dist, the file's from the unrounded one
    for g, r in zip(got[1:], rows):
        assert g[:-1] == [r[k] for k in keep][:-1]
        assert abs(float(g[-1]) / float(r[hdr.index("vel")]) - 1) < 2e-6


def test_write_raw_round_trip_and_missing_columns(tmp_path):
    t = S.read_src_rec(EX04)
    p = tmp_path / "raw.csv"
    S.write_raw(p, t)
    u = S.read_src_rec(p)
    for key in ("tt", "stla", "stlo", "evla", "evlo", "period", "weight", "dist", "vel"):
        assert np.allclose(u[key], t[key], rtol=0, atol=5e-7), key
    assert u["staname"] == t["staname"] and u["evtname"] == t["evtname"]
    # a file without weight/dist/vel: weight 1, dist from gps2dist, vel = dist/tt
    q = tmp_path / "min.csv"
    q.write_text("tt,staname,stla,stlo,stel,evtname,evla,evlo,evel,period\n"
                 "10.0,A,20.0,100.0,0.0,B,20.5,100.3,0.0,5.0\n")
    m = S.read_src_rec(q)
    assert m["header"] == list(S.REQUIRED) and m["weight"][0] == 1.0
    assert m["dist"][0] == S.gps2dist(20.0, 100.0, 20.5, 100.3)
    assert m["vel"][0] == m["dist"][0] / 10.0
    S.to_csv(tmp_path / "min_fwd.csv", m, np.array([12.5]))
    line = (tmp_path / "min_fwd.csv").read_text().splitlines()
    assert line[0].endswith(",period,weight,dist,vel") and ",1.0000," in line[1]


def test_gps2dist_is_the_oracles():
    for a in ((20.0, 100.0, 20.5, 100.3), (-33.1, 151.2, 35.7, 139.7), (0.0, 0.0, 0.0, 1.0)):
        assert S.gps2dist(*a) == O.gps2dist(*a)


def test_fortran_edit_descriptors():
    assert S.fortran_f(1.0, 6, 4) == "1.0000"
    assert S.fortran_f(0.5, 6, 4) == "0.5000"
    assert S.fortran_f(12.5, 6, 4) == "******"        # does not fit F6.4
    assert S.fortran_f(0.123456789, 0, 6) == "0.123457"
    assert S.fortran_f(-155.109605, 9, 4) == "-155.1096"
    assert S.fortran_f(0.5, 5, 4) == ".5000"          # the optional zero dropped


def test_missing_column_is_an_error(tmp_path):
    p = tmp_path / "bad.csv"
    p.write_text("tt,staname,stla,stlo,evtname,evla,evlo\n1,A,1,1,B,1,1\n")
    with pytest.raises(S.SrcRecError):
        S.read_src_rec(p)


def test_batch_from_the_example_table():
    t = S.read_src_rec(EX04)
    per = S.periods(t)
    n = 12
    xg = np.linspace(t["stlo"].min() - 0.2, t["stlo"].max() + 0.2, n)
    yg = np.linspace(t["stla"].min() - 0.2, t["stla"].max() + 0.2, n)
    one = np.ones((len(per), n, n))
    b = S.batch_from_src_rec(t, 2, xg, yg, 3.0 * one, one, 0 * one, one)
    pairs = {(e, p) for e, p in zip(t["evtname"], t["period"])}
    assert len(b["item_pidx"]) == len(pairs) and b["npath"] == t["npath"]
    assert sorted(b["idx_path"].tolist()) == list(range(t["npath"]))
