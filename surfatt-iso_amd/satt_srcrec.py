"""The reference's source-receiver CSV files, read and written on the host.

    read_src_rec     SrcRecRaw%read_raw_src_rec_file  src/shared/src_rec_parser.f90:20-61
    write_raw        SrcRecRaw%write_raw_src_rec_file src/shared/src_rec_parser.f90:63-96
    to_csv           SrcRec%to_csv (forward times)    src/src_rec.f90:285-323
    batch_from_src_rec  the batch of one forward_simulate over a table

The files are comma-separated with a header row (the reference reads them
with the vendored fortran-csv module, src/fortran-csv/csv_module.F90):
tt, staname, stla, stlo, evtname, evla, evlo, period are required; weight
(default 1), dist (default the great-circle distance, gps2dist) and vel
(default dist/tt) are optional; stel and evel are read past and dropped from
the header.  Numbers are written with the reference's Fortran edit
descriptors (Fw.d: d decimals, at most w characters, '*'s when the value does
not fit; F0.d: as many characters as needed), left-adjusted and trimmed as the
csv module writes them.
"""
import math

import numpy as np

import satt_workloads as W

REQUIRED = ("tt", "staname", "stla", "stlo", "evtname", "evla", "evlo", "period")
ORDER = REQUIRED + ("weight", "dist", "vel")


class SrcRecError(ValueError):
    pass


_DEG2RAD = float(np.float32(3.14159265359)) / 180.0   # src/shared/shared_par.f90:11-15


def gps2dist(lat0, lon0, lat1, lon1):
    """gps2dist_scalar (src/shared/utils.f90:1250-1262), km: the haversine
    distance on a 6371 km sphere with the reference's REAL(4) pi, in its
    evaluation order."""
    dlat = (lat0 - lat1) * _DEG2RAD
    dlon = (lon1 - lon0) * _DEG2RAD
    s1, s2 = math.sin(dlat * 0.5), math.sin(dlon * 0.5)
    a = s1 * s1 + s2 * s2 * math.cos(lat0 * _DEG2RAD) * math.cos(lat1 * _DEG2RAD)
    return 6371.0 * (2.0 * math.atan2(math.sqrt(a), math.sqrt(1.0 - a)))


def _tokens(line):
    out = []
    for t in line.rstrip("\r\n").split(","):
        t = t.strip()
        if len(t) >= 2 and t[0] == '"' and t[-1] == '"':
            t = t[1:-1]
        out.append(t)
    return out


def read_src_rec(path):
    """A table dict: header (the reference's order, stel/evel dropped),
    npath, the name columns as lists of str and the numeric columns as
    float64 arrays."""
    with open(path) as f:
        lines = [ln for ln in f if ln.strip()]
    if not lines:
        raise SrcRecError(f"{path}: empty file")
    header = _tokens(lines[0])
    col = {h: k for k, h in enumerate(header)}
    for name in REQUIRED:
        if name not in col:
            raise SrcRecError(f"{path}: column {name} not found in the header")
    rows = [_tokens(ln) for ln in lines[1:]]
    for r, row in enumerate(rows):
        if len(row) != len(header):
            raise SrcRecError(f"{path}: row {r + 2} has {len(row)} fields, the header {len(header)}")
    num = lambda name: np.array([float(row[col[name]]) for row in rows], dtype=np.float64)
    t = {"npath": len(rows), "tt": num("tt"), "stla": num("stla"), "stlo": num("stlo"),
         "evla": num("evla"), "evlo": num("evlo"), "period": num("period"),
         "staname": [row[col["staname"]] for row in rows],
         "evtname": [row[col["evtname"]] for row in rows]}
    t["weight"] = num("weight") if "weight" in col else np.ones(len(rows))
    if "dist" in col:
        t["dist"] = num("dist")
    else:
        t["dist"] = np.array([gps2dist(a, b, c, d) for a, b, c, d in
                              zip(t["stla"], t["stlo"], t["evla"], t["evlo"])])
    t["vel"] = num("vel") if "vel" in col else t["dist"] / t["tt"]
    # remove_el_from_header, order_header (src_rec_parser.f90:98-152)
    kept = [h for h in header if h not in ("stel", "evel")]
    t["header"] = [h for h in ORDER if h in kept]
    return t


def fortran_f(x, w, d):
    """Fortran Fw.d output of x, left-adjusted and trimmed (F0.d: w = 0)."""
    s = f"{x:.{d}f}"
    if w and len(s) > w:
        if s.startswith("0.") and len(s) - 1 <= w:
            s = s[1:]
        elif s.startswith("-0.") and len(s) - 1 <= w:
            s = "-" + s[2:]
        else:
            s = "*" * w
    return s


def write_raw(path, t):
    """write_raw_src_rec_file: the table's header and columns."""
    hdr = t["header"]
    with open(path, "w") as f:
        f.write(",".join(hdr) + "\n")
        for i in range(t["npath"]):
            cells = [fortran_f(t["tt"][i], 11, 6), t["staname"][i], fortran_f(t["stla"][i], 11, 6),
                     fortran_f(t["stlo"][i], 11, 6), t["evtname"][i],
                     fortran_f(t["evla"][i], 11, 6), fortran_f(t["evlo"][i], 11, 6),
                     fortran_f(t["period"][i], 0, 4)]
            if "weight" in hdr:
                cells.append(fortran_f(t["weight"][i], 6, 4))
            if "dist" in hdr:
                cells.append(fortran_f(t["dist"][i], 11, 6))
            if "vel" in hdr:
                cells.append(fortran_f(t["vel"][i], 11, 6))
            f.write(",".join(cells) + "\n")


def to_csv(path, t, tt_fwd):
    """SrcRec%to_csv: the table with the forward times tt_fwd in place of
    tt, the missing weight/dist/vel columns added to the header, and vel =
    dist/tt_fwd."""
    hdr = list(t["header"]) + [h for h in ("weight", "dist", "vel") if h not in t["header"]]
    with open(path, "w") as f:
        f.write(",".join(hdr) + "\n")
        for i in range(t["npath"]):
            cells = [fortran_f(tt_fwd[i], 0, 6), t["staname"][i], fortran_f(t["stla"][i], 9, 4),
                     fortran_f(t["stlo"][i], 9, 4), t["evtname"][i], fortran_f(t["evla"][i], 9, 4),
                     fortran_f(t["evlo"][i], 9, 4), fortran_f(t["period"][i], 0, 4),
                     fortran_f(t["weight"][i], 6, 4), fortran_f(t["dist"][i], 12, 4),
                     fortran_f(t["dist"][i] / tt_fwd[i], 10, 6)]
            f.write(",".join(cells) + "\n")


def batch_from_src_rec(t, dim, xgrids, ygrids, svel, m11, m12, m22):
    """The batch of forward_simulate over the table: stations, work list and
    gathers in the reference's orders (get_sta, get_periods, isrcs), the
    model (nperiod, nx, ny) over the table's sorted unique periods."""
    paths = {k: t[k] for k in ("evtname", "staname", "evla", "evlo", "stla", "stlo", "period",
                               "tt", "weight")}
    return W.make_batch(dim, xgrids, ygrids, None, svel, m11, m12, m22, paths)


def periods(t):
    """get_periods (src/src_rec.f90:130-160): the sorted unique periods."""
    return np.array(sorted(set(float(p) for p in t["period"])), dtype=np.float64)
