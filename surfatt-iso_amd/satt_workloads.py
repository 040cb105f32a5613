"""Host-side bookkeeping of the reference's source-receiver table, and the
synthetic workloads of the measurement plan.

``make_batch`` turns per-path arrays (the columns of the reference's src_rec
CSV) into the batch the C-ABI takes, with the reference's orders:

* stations: unique receiver names in order of first appearance, then event
  names not yet listed (``get_sta``, src/src_rec.f90:180-217);
* periods: sorted unique periods (``get_periods``, src/src_rec.f90:130-160);
* items: for each station in station order that is an event, its periods in
  order of first appearance (``scatter_src_gather`` / ``get_periods_by_src``,
  src/src_rec.f90:338-369, :162-179); ``item_pidx`` is ``isrcs(:,1)-1``;
* gather: the item's paths in path order, receiver positions from the station
  list (``get_evt_gather``, src/src_rec.f90:252-283); source position from
  ``get_sta_pos`` (src/src_rec.f90:327-336).

The generators follow the reference's own model builders where one exists:
``arange`` (src/shared/utils.f90:645-653), the 2-D checkerboard of
``att_acqui_2d%add_pert`` (src/acqui_2d.f90:101-147) and the topographic
metric of ``att_grid%get_topo`` (src/grid.f90:119-122 with
``gradient_2_geo``, src/shared/utils.f90:1118-1160).
"""
import numpy as np

# default-real literals of the reference, rounded to REAL(4) then widened
PI_F = float(np.float32(3.14159265359))  # constants module pi (src/shared/shared_par.f90:11)
RADIUS = 6371.0
DEG2RAD = PI_F / 180.0


def arange(first, last, step):
    """arange_r8 (src/shared/utils.f90:645-653)."""
    n = int((last - first) / step) + 1
    return np.array([first + i * step for i in range(n)], dtype=np.float64)


def checkerboard_2d(nx, ny, ncb_x, ncb_y, pert):
    """Amplitude field of add_pert with hmarg=0 (src/acqui_2d.f90:129-141):
    pert * sin(ncb_x*pi*i/nx) * sin(ncb_y*pi*j/ny), i, j = 1..n."""
    xp = np.sin(ncb_x * PI_F * np.arange(1, nx + 1) / nx)
    yp = np.sin(ncb_y * PI_F * np.arange(1, ny + 1) / ny)
    return pert * np.outer(xp, yp)


def gradient_2_geo(f, lon, lat):
    """gradient_2_geo (src/shared/utils.f90:1118-1160): d f / d(km) along
    lon and lat, central inside, one-sided at the edges."""
    nx, ny = f.shape
    dx1 = np.zeros(nx); dx1[1:] = np.diff(lon)
    dy1 = np.zeros(ny); dy1[1:] = np.diff(lat)
    dx = RADIUS * np.repeat(dx1[:, None], ny, 1) * DEG2RAD
    dy = RADIUS * np.repeat(dy1[None, :], nx, 0) * DEG2RAD
    dx = dx * np.cos(np.deg2rad(lat))[None, :]
    gfx = np.zeros_like(f); gbx = np.zeros_like(f); gfy = np.zeros_like(f); gby = np.zeros_like(f)
    gfx[1:, :] = f[1:, :] - f[:-1, :]
    gbx[:-1, :] = f[1:, :] - f[:-1, :]
    gfy[:, 1:] = f[:, 1:] - f[:, :-1]
    gby[:, :-1] = f[:, 1:] - f[:, :-1]
    tx = np.zeros_like(f); ty = np.zeros_like(f)
    tx[1:-1, :] = (gfx[1:-1, :] + gbx[1:-1, :]) / (2 * dx[1:-1, :])
    tx[0, :] = gbx[0, :] / dx[1, :]
    tx[-1, :] = gfx[-1, :] / dx[-1, :]
    ty[:, 1:-1] = (gfy[:, 1:-1] + gby[:, 1:-1]) / (2 * dy[:, 1:-1])
    ty[:, 0] = gby[:, 0] / dy[:, 1]
    ty[:, -1] = gfy[:, -1] / dy[:, -1]
    return tx, ty


def topo_metric(topo, lon, lat):
    """m11, m12, m22 of get_topo (src/grid.f90:119-122,140-144):
    a=(1+fy^2)/(1+fx^2+fy^2), b=(1+fx^2)/(..), c=fx*fy/(..); m11=a, m22=b,
    m12=-c."""
    fx, fy = gradient_2_geo(topo, lon, lat)
    den = 1 + fx ** 2 + fy ** 2
    a = (1 + fy ** 2) / den
    b = (1 + fx ** 2) / den
    c = fx * fy / den
    return a, -c, b


def gaussian_hills(lon, lat, nhill, rng, h=(0.5, 4.0), w=(5.0, 30.0)):
    """Sum of Gaussian hills (heights in km, half-widths in km)."""
    X, Y = np.meshgrid(lon, lat, indexing="ij")
    lat0 = np.deg2rad(lat.mean())
    topo = np.zeros(X.shape)
    for _ in range(nhill):
        cx = rng.uniform(lon[0], lon[-1]); cy = rng.uniform(lat[0], lat[-1])
        hh = rng.uniform(*h); ww = rng.uniform(*w)
        ddx = (X - cx) * DEG2RAD * RADIUS * np.cos(lat0)
        ddy = (Y - cy) * DEG2RAD * RADIUS
        topo += hh * np.exp(-(ddx ** 2 + ddy ** 2) / (2 * ww ** 2))
    return topo


def smooth_random_field(nx, ny, corr_cells, rng, lo, hi):
    """Gaussian random field with a Gaussian correlation of corr_cells cells,
    scaled to [lo, hi]."""
    z = rng.standard_normal((nx, ny))
    kx = np.fft.fftfreq(nx)[:, None]; ky = np.fft.fftfreq(ny)[None, :]
    filt = np.exp(-2 * (np.pi * corr_cells) ** 2 * (kx ** 2 + ky ** 2))
    f = np.real(np.fft.ifft2(np.fft.fft2(z) * filt))
    f = (f - f.min()) / (f.max() - f.min())
    return lo + (hi - lo) * f


# ------------------------------------------------------------------ gather

def make_batch(dim, xgrids, ygrids, periods_sorted, svel, m11, m12, m22, paths):
    """Build the C-ABI batch from per-path columns.

    ``paths``: dict of equal-length arrays ``evtname, staname`` (str),
    ``evla, evlo, stla, stlo, period, tt, weight`` (float).  Model arrays are
    (nperiod, nx, ny) over ``periods_sorted``; when ``periods_sorted`` is None
    the sorted unique periods of the table are used (get_periods).
    """
    evt = list(paths["evtname"]); sta = list(paths["staname"])
    per = np.asarray(paths["period"], dtype=np.float64)
    npath = len(evt)
    # get_sta (src/src_rec.f90:180-217)
    names, slat, slon, seen = [], [], [], {}
    for i in range(npath):
        if sta[i] not in seen:
            seen[sta[i]] = len(names); names.append(sta[i])
            slat.append(paths["stla"][i]); slon.append(paths["stlo"][i])
    for i in range(npath):
        if evt[i] not in seen:
            seen[evt[i]] = len(names); names.append(evt[i])
            slat.append(paths["evla"][i]); slon.append(paths["evlo"][i])
    slat = np.asarray(slat, dtype=np.float64); slon = np.asarray(slon, dtype=np.float64)
    # get_periods (src/src_rec.f90:130-160)
    if periods_sorted is None:
        periods_sorted = np.array(sorted(set(per.tolist())), dtype=np.float64)
    pindex = {float(p): k for k, p in enumerate(periods_sorted)}
    # paths per event in path order
    evt_arr = np.array(evt, dtype=object)
    by_evt = {}
    for i, e in enumerate(evt):
        by_evt.setdefault(e, []).append(i)
    items, gathers = [], []
    for j, nm in enumerate(names):           # scatter_src_gather (:345-357)
        if nm not in by_evt:
            continue
        idx = np.asarray(by_evt[nm], dtype=np.int64)
        pe = per[idx]
        _, first = np.unique(pe, return_index=True)   # get_periods_by_src (:171-178)
        for p in pe[np.sort(first)]:
            items.append((pindex[float(p)], j))
            gathers.append(idx[pe == p])                # get_evt_gather
    del evt_arr
    nit = len(items)
    rec_off = np.zeros(nit + 1, dtype=np.int32)
    for k, g in enumerate(gathers):
        rec_off[k + 1] = rec_off[k] + len(g)
    flat = np.concatenate(gathers) if gathers else np.zeros(0, dtype=np.int64)
    sta_idx = np.array([seen[s_] for s_ in sta], dtype=np.int64)
    rx = slon[sta_idx[flat]]
    ry = slat[sta_idx[flat]]
    return {
        "dim": int(dim), "nx": len(xgrids), "ny": len(ygrids),
        "nperiod": len(periods_sorted), "periods": np.asarray(periods_sorted, dtype=np.float64),
        "xgrids": np.ascontiguousarray(xgrids, dtype=np.float64),
        "ygrids": np.ascontiguousarray(ygrids, dtype=np.float64),
        "svel": np.asfortranarray(svel, dtype=np.float64),
        "m11": np.asfortranarray(m11, dtype=np.float64),
        "m12": np.asfortranarray(m12, dtype=np.float64),
        "m22": np.asfortranarray(m22, dtype=np.float64),
        "item_pidx": np.array([p for p, _ in items], dtype=np.int32),
        "item_sta": np.array([j for _, j in items], dtype=np.int32),
        "item_srcx": np.array([slon[j] for _, j in items], dtype=np.float64),
        "item_srcy": np.array([slat[j] for _, j in items], dtype=np.float64),
        "rec_off": rec_off,
        "recx": rx, "recy": ry,
        "tt": np.ascontiguousarray(np.asarray(paths["tt"], dtype=np.float64)[flat]),
        "weight": np.ascontiguousarray(np.asarray(paths["weight"], dtype=np.float64)[flat]),
        "idx_path": flat.astype(np.int32),
        "npath": npath,
        "station_names": names,
    }


def stations_uniform(rng, n, xg, yg, frac=0.9, min_cells=2):
    """n station positions uniform in the interior `frac` of the grid and at
    least `min_cells` cells from every edge."""
    dx, dy = xg[1] - xg[0], yg[1] - yg[0]
    x_lo = max(xg[0] + (1 - frac) / 2 * (xg[-1] - xg[0]), xg[0] + min_cells * dx)
    x_hi = min(xg[-1] - (1 - frac) / 2 * (xg[-1] - xg[0]), xg[-1] - min_cells * dx)
    y_lo = max(yg[0] + (1 - frac) / 2 * (yg[-1] - yg[0]), yg[0] + min_cells * dy)
    y_hi = min(yg[-1] - (1 - frac) / 2 * (yg[-1] - yg[0]), yg[-1] - min_cells * dy)
    return rng.uniform(x_lo, x_hi, n), rng.uniform(y_lo, y_hi, n)


def all_pairs_paths(sx, sy, periods, tt=None):
    """Every station is a source and every other station a receiver, for
    every period (path order: source, period, receiver)."""
    n = len(sx)
    cols = {k: [] for k in ("evtname", "staname", "evla", "evlo", "stla", "stlo",
                            "period", "tt", "weight")}
    for s in range(n):
        for p in periods:
            for r in range(n):
                if r == s:
                    continue
                cols["evtname"].append(f"S{s:04d}"); cols["staname"].append(f"S{r:04d}")
                cols["evla"].append(sy[s]); cols["evlo"].append(sx[s])
                cols["stla"].append(sy[r]); cols["stlo"].append(sx[r])
                cols["period"].append(float(p)); cols["tt"].append(1.0); cols["weight"].append(1.0)
    out = {k: (np.array(v) if k not in ("evtname", "staname") else v) for k, v in cols.items()}
    if tt is not None:
        out["tt"] = np.asarray(tt, dtype=np.float64)
    return out


def set_observed(batch, tt_fwd):
    """Replace the batch's observed times by a forward result (per path)."""
    b = dict(batch)
    b["tt"] = np.ascontiguousarray(np.asarray(tt_fwd)[batch["idx_path"]], dtype=np.float64)
    return b


def with_model(batch, svel, m11=None, m12=None, m22=None):
    b = dict(batch)
    b["svel"] = np.asfortranarray(svel, dtype=np.float64)
    if m11 is not None:
        b["m11"] = np.asfortranarray(m11, dtype=np.float64)
        b["m12"] = np.asfortranarray(m12, dtype=np.float64)
        b["m22"] = np.asfortranarray(m22, dtype=np.float64)
    return b


# -------------------------------------------------------------- workloads

def tomo2d_checkerboard(n=81, d=0.025, nsta=25, seed=0, pert=0.05, ncb=2, v0=3.0,
                        period=10.0):
    """The plain 2-D tomography step: n x n grid at d degrees from
    (100E, 20N), one period; true model v0*(1 + checkerboard), initial model
    v0.  Returns (batch_true, batch_init): run a forward on batch_true to get
    the observed times, then set them on batch_init."""
    xg = 100.0 + d * np.arange(n); yg = 20.0 + d * np.arange(n)
    rng = np.random.default_rng(seed)
    sx, sy = stations_uniform(rng, nsta, xg, yg)
    paths = all_pairs_paths(sx, sy, [period])
    amp = checkerboard_2d(n, n, ncb, ncb, pert)
    vt = (v0 * (1 + amp))[None, :, :]
    ones = np.ones((1, n, n)); zeros = np.zeros((1, n, n))
    bt = make_batch(2, xg, yg, np.array([period]), vt, ones, zeros, ones, paths)
    bi = with_model(bt, v0 * ones)
    return bt, bi


def topo_hills(n=401, d=0.005, nper=20, seed=3, nhill=8):
    """The topography, periods and 1-D dispersion curve of the curved-surface
    workload below (the same draws): (xg, yg, topo km, periods, vref)."""
    xg = 100.0 + d * np.arange(n); yg = 20.0 + d * np.arange(n)
    rng = np.random.default_rng(seed)
    topo = gaussian_hills(xg, yg, nhill, rng)
    periods = np.geomspace(3.0, 30.0, nper)
    return xg, yg, topo, periods, 2.8 + 0.04 * periods


def topo_checkerboard(n=401, d=0.005, nsrc=200, nper=20, seed=3, pert=0.08,
                      nhill=8, ncb=4, dim=2):
    """Curved-surface topography workload: n x n grid at d degrees, `nhill`
    Gaussian hills entering the metric, nper periods log-spaced 3-30 s,
    checkerboard +-pert over a period-dependent reference velocity; nsrc
    stations, each a source with the others as receivers."""
    xg = 100.0 + d * np.arange(n); yg = 20.0 + d * np.arange(n)
    rng = np.random.default_rng(seed)
    topo = gaussian_hills(xg, yg, nhill, rng)
    m11, m12, m22 = topo_metric(topo, xg, yg)
    periods = np.geomspace(3.0, 30.0, nper)
    vref = 2.8 + 0.04 * periods                     # a smooth dispersion curve
    amp = checkerboard_2d(n, n, ncb, ncb, pert)
    svel = vref[:, None, None] * (1 + amp)[None, :, :]
    sx, sy = stations_uniform(rng, nsrc, xg, yg)
    paths = all_pairs_paths(sx, sy, periods)
    rep = lambda f: np.repeat(f[None, :, :], nper, 0)
    bt = make_batch(dim, xg, yg, periods, svel, rep(m11), rep(m12), rep(m22), paths)
    bi = with_model(bt, vref[:, None, None] * np.ones((nper, n, n)))
    return bt, bi


def random_receiver_batch(xg, yg, svel, nitems, rng, nrec=(50, 200), period=10.0,
                          m11=None, m12=None, m22=None):
    """One grid and one period, `nitems` solves with a random source and
    nrec[0]..nrec[1] random receivers each (no source-receiver table: every
    receiver is its own path).  svel is (nx, ny).  Returns the batch with unit
    observed times; set them with set_observed."""
    nx, ny = len(xg), len(yg)
    ones = np.ones((1, nx, ny)); zeros = np.zeros((1, nx, ny))
    sx, sy = stations_uniform(rng, nitems, xg, yg)
    counts = rng.integers(nrec[0], nrec[1] + 1, nitems)
    rec_off = np.zeros(nitems + 1, dtype=np.int32)
    rec_off[1:] = np.cumsum(counts)
    nr = int(rec_off[-1])
    rx, ry = stations_uniform(rng, nr, xg, yg)
    f = lambda a, dflt: np.asfortranarray(dflt if a is None else a[None], dtype=np.float64)
    return {
        "dim": 2, "nx": nx, "ny": ny, "nperiod": 1, "periods": np.array([period]),
        "xgrids": np.ascontiguousarray(xg, dtype=np.float64),
        "ygrids": np.ascontiguousarray(yg, dtype=np.float64),
        "svel": np.asfortranarray(svel[None], dtype=np.float64),
        "m11": f(m11, ones), "m12": f(m12, zeros), "m22": f(m22, ones),
        "item_pidx": np.zeros(nitems, dtype=np.int32),
        "item_srcx": np.ascontiguousarray(sx), "item_srcy": np.ascontiguousarray(sy),
        "rec_off": rec_off, "recx": np.ascontiguousarray(rx), "recy": np.ascontiguousarray(ry),
        "tt": np.ones(nr), "weight": np.ones(nr),
        "idx_path": np.arange(nr, dtype=np.int32), "npath": nr,
    }


UNEVEN_SIZES = ((101, 101), (201, 161), (321, 321), (81, 241))


def uneven_mix(nsolve=2000, seed=2, d=0.01, sizes=UNEVEN_SIZES, pert=0.05):
    """Uneven batching: nsolve fwd+adj solves over grids of the given sizes
    (uniform choice), each grid with a smooth Gaussian random phase-velocity
    field (correlation length 10 cells) scaled to 2.5-4.2 km/s as the current
    model, the same field times (1 + a 3x3 checkerboard of +-pert) as the
    true model, and 50-200 random receivers per solve.  Returns
    [(batch_true, batch_current)] per size."""
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, len(sizes), nsolve)
    out = []
    for k, (nx, ny) in enumerate(sizes):
        xg = 100.0 + d * np.arange(nx); yg = 20.0 + d * np.arange(ny)
        v = smooth_random_field(nx, ny, 10.0, rng, 2.5, 4.2)
        vt = v * (1 + checkerboard_2d(nx, ny, 3, 3, pert))
        bt = random_receiver_batch(xg, yg, vt, int(np.sum(pick == k)), rng)
        out.append((bt, with_model(bt, v[None])))
    return out


SERVE_SIZES = (101, 201, 401, 801)
SERVE_PROBS = (0.4, 0.3, 0.2, 0.1)


def serving_requests(n, seed=4, sizes=SERVE_SIZES, probs=SERVE_PROBS, d=0.005, nrec=100):
    """The serving mix: n fwd+adj solve requests, grid size drawn from
    `sizes` (square, d degrees) with `probs`, period uniform on 5-50 s
    (setting the mean velocity 2.9 + 0.02*period km/s), a smooth random
    velocity field (+-10%, correlation 10 cells) per request, a random source
    and `nrec` random receivers whose observed times are the homogeneous
    great-circle times with 3% noise.  Returns a list of request dicts."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        m = int(rng.choice(sizes, p=probs))
        xg = 100.0 + d * np.arange(m); yg = 20.0 + d * np.arange(m)
        period = rng.uniform(5.0, 50.0)
        v0 = 2.9 + 0.02 * period
        v = smooth_random_field(m, m, 10.0, rng, 0.9 * v0, 1.1 * v0)
        sx, sy = stations_uniform(rng, 1, xg, yg)
        rx, ry = stations_uniform(rng, nrec, xg, yg)
        lon0, lat0 = np.radians(sx[0]), np.radians(sy[0])
        lon, lat = np.radians(rx), np.radians(ry)
        cosd = np.sin(lat) * np.sin(lat0) + np.cos(lat) * np.cos(lat0) * np.cos(lon - lon0)
        dist = RADIUS * np.arccos(np.clip(cosd, -1, 1))
        tt = dist / v0 * (1 + 0.03 * rng.standard_normal(nrec))
        out.append({"nx": m, "ny": m, "xgrids": xg, "ygrids": yg, "svel": v,
                    "srcx": sx[0], "srcy": sy[0], "recx": rx, "recy": ry,
                    "tt": np.maximum(tt, 0.1), "weight": np.ones(nrec)})
    return out


def request_batch(req):
    """The one-item batch of a serving request (for the batched plan)."""
    nx, ny = req["nx"], req["ny"]
    nr = len(req["recx"])
    ones = np.ones((1, nx, ny)); zeros = np.zeros((1, nx, ny))
    g = lambda k, dflt: np.asfortranarray(dflt if req.get(k) is None else req[k][None])
    return {
        "dim": 2, "nx": nx, "ny": ny, "nperiod": 1, "periods": np.array([10.0]),
        "xgrids": np.ascontiguousarray(req["xgrids"], dtype=np.float64),
        "ygrids": np.ascontiguousarray(req["ygrids"], dtype=np.float64),
        "svel": np.asfortranarray(req["svel"][None], dtype=np.float64),
        "m11": g("m11", ones), "m12": g("m12", zeros), "m22": g("m22", ones),
        "item_pidx": np.zeros(1, dtype=np.int32),
        "item_srcx": np.array([req["srcx"]], dtype=np.float64),
        "item_srcy": np.array([req["srcy"]], dtype=np.float64),
        "rec_off": np.array([0, nr], dtype=np.int32),
        "recx": np.ascontiguousarray(req["recx"], dtype=np.float64),
        "recy": np.ascontiguousarray(req["recy"], dtype=np.float64),
        "tt": np.ascontiguousarray(req["tt"], dtype=np.float64),
        "weight": np.ascontiguousarray(req["weight"], dtype=np.float64),
        "idx_path": np.arange(nr, dtype=np.int32), "npath": nr,
    }


def phase_velocity_maps(vs3d, depths, periods):
    """Synthetic phase-velocity maps from a Vs model, per column: 0.92 times
    the Vs average weighted by exp(-z / (c*T/3)) (a Rayleigh-like depth
    sensitivity, iterated for c), shape (nperiod, nx, ny).  A stand-in for the
    reference's surface-wave dispersion code (src/surfker), which is outside
    the path this build accelerates."""
    nz = len(depths)
    dz = np.gradient(depths)
    out = np.empty((len(periods),) + vs3d.shape[:2])
    for p, T in enumerate(periods):
        c = 0.92 * vs3d.mean(axis=2)
        for _ in range(4):
            w = np.exp(-depths[None, None, :] / np.maximum(c[..., None] * T / 3.0, 1e-3))
            w = w * dz[None, None, :]
            c = 0.92 * (vs3d * w).sum(axis=2) / w.sum(axis=2)
        out[p] = c
    return out


def acqui3d_model(n=41, nz=31, zmax=60.0, nper=16, pert=0.06, ncb=(2, 3, 2)):
    """The 3-D stage's true model: Vs (n, n, nz) on depths 0..zmax km rising
    linearly from 3.0 to 4.5 km/s times (1 + the reference's 3-D checkerboard,
    src/model.f90:276-300), and nper periods log-spaced 4-40 s."""
    depths = np.linspace(0.0, zmax, nz)
    vs1 = 3.0 + 1.5 * depths / zmax
    xp = np.sin(ncb[0] * PI_F * np.arange(1, n + 1) / n)
    yp = np.sin(ncb[1] * PI_F * np.arange(1, n + 1) / n)
    zp = np.sin(ncb[2] * PI_F * np.arange(1, nz + 1) / nz)
    vs = vs1[None, None, :] * (1 + pert * xp[:, None, None] * yp[None, :, None] * zp[None, None, :])
    return vs, depths, np.geomspace(4.0, 40.0, nper)


def acqui3d_checkerboard(n=41, d=0.05, nz=31, zmax=60.0, nsta=100, nper=16, seed=1,
                         pert=0.06, ncb=(2, 3, 2)):
    """The 3-D inversion's 2-D stage (att_acqui%forward_simulate, dim 3):
    n x n grid at d degrees, Vs on nz depths 0..zmax km rising linearly from
    3.0 to 4.5 km/s, times (1 + the reference's 3-D checkerboard of +-pert
    with ncb = (x, y, z) half-periods, src/model.f90:276-300); nper periods
    log-spaced 4-40 s with phase-velocity maps from phase_velocity_maps; nsta
    random stations (seed), each a source with the others as receivers
    (nsta*nper items).  The current model is the unperturbed 1-D Vs.
    Returns (batch_true, batch_current)."""
    xg = 100.0 + d * np.arange(n); yg = 20.0 + d * np.arange(n)
    vs_true, depths, periods = acqui3d_model(n, nz, zmax, nper, pert, ncb)
    vs1 = 3.0 + 1.5 * depths / zmax
    vs_cur = np.broadcast_to(vs1[None, None, :], (n, n, nz))
    ct = phase_velocity_maps(vs_true, depths, periods)
    cc = phase_velocity_maps(np.ascontiguousarray(vs_cur), depths, periods)
    rng = np.random.default_rng(seed)
    sx, sy = stations_uniform(rng, nsta, xg, yg)
    paths = all_pairs_paths(sx, sy, periods)
    ones = np.ones((nper, n, n)); zeros = np.zeros((nper, n, n))
    bt = make_batch(3, xg, yg, periods, ct, ones, zeros, ones, paths)
    return bt, with_model(bt, cc)


def linspace_ref(first, last, n):
    """linspace_r8r8 (src/shared/utils.f90:592-604): first + step*(i-1), the
    last node set to `last`."""
    step = (last - first) / (n - 1)
    out = first + step * np.arange(n, dtype=np.float64)
    out[-1] = last
    return out


def inv_grids(vx, vy, vz, n_inv, nset):
    """The inversion grids of att_model%get_inv_grids (src/model.f90:127-168)
    for model axes vx, vy, vz and n_inv = (ninvx, ninvy, ninvz): (xinv, yinv,
    zinv), each (ninv, nset) in Fortran order, the sets shifted as there."""
    nxi, nyi, nzi = n_inv
    x1, xn, y1, yn, z1, zn = vx[0], vx[-1], vy[0], vy[-1], vz[0], vz[-1]
    ix = linspace_ref(x1 - (xn - x1) / (nxi - 2) / 2.0, xn + (xn - x1) / (nxi - 2) / 2.0, nxi)
    iy = linspace_ref(y1 - (yn - y1) / (nyi - 2), yn, nyi)
    iz = linspace_ref(z1 - 0.5 * (zn - z1) / (nzi - 2), zn + 0.5 * (zn - z1) / (nzi - 2), nzi)
    zref = np.empty(nzi)
    zref[:-1] = iz[1:]
    zref[-1] = iz[-1]
    zadd = (zref - iz) / (nset + 1)
    zadd[0] = 0
    xinv = np.zeros((nxi, nset), order="F")
    yinv = np.zeros((nyi, nset), order="F")
    zinv = np.zeros((nzi, nset), order="F")
    polar = 1
    for i in range(nset):
        if i == 0:
            xinv[:, 0] = ix
        else:
            xinv[:, i] = xinv[:, i - 1] + polar * i * (ix[1] - ix[0]) / (nset + 1)
            polar = -polar
        yinv[:, i] = iy + i * (iy[1] - iy[0]) / (nset + 1)
        zinv[:, i] = iz + i * zadd
    return xinv, yinv, zinv
