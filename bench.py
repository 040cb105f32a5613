#!/usr/bin/env python3
"""Throughput of the SurfATT forward + adjoint hot path on one MI355X.

A step is one forward_simulate(chi, istotable=.true., isadj=.true.) over the
whole synthetic batch with the model, the gathers and every field resident
in device memory: for each (source, period) item a factored-eikonal forward
solve, receiver interpolation and misfit, an adjoint solve, the source mask
and the kernel stacking.  Default workload: the 401 x 401 curved-surface
topography grid at 0.005 deg, 200 sources x 20 periods = 4000 fwd+adj solves.

Prints one JSON line (rank 0).  `value` is fwd+adj solves per second of the
batched path (team sizes from the autotuner unless --no-tune); the plain FP64
one-solve-per-launch rate, the roofline of the dominant kernel and a host-core
baseline are reported beside it, and a model-side leg times the topographic
metric and the steepest-descent update on the same grid.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "surfatt-iso_amd"), os.path.join(ROOT, "oracle")]

import satt  # noqa: E402
import satt_workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
# sweep stages with their own team size (satt_tune_result.stage_waves order)
STAGES = ("fwd_fp32", "fwd_fp64", "adj_fp32", "adj_fp64")

WORKLOADS = {
    # name: (builder, kwargs, description); a builder returns (batch_true,
    # batch_current) for one grid or a list of them for several
    "topo401": (W.topo_checkerboard, dict(n=401, d=0.005, nsrc=200, nper=20, seed=3),
                "401x401 @0.005deg, 8-hill topographic metric, +-8% checkerboard, "
                "200 sources x 20 periods (3-30 s), 199 receivers each"),
    "tomo81": (W.tomo2d_checkerboard, dict(n=81, d=0.025, nsta=25, seed=0),
               "81x81 @0.025deg, 2x2 checkerboard +-5%, 25 stations, 1 period"),
    "acqui41": (W.acqui3d_checkerboard, dict(),
                "3-D stage (density adjoint): 41x41 @0.05deg, Vs 3.0-4.5 km/s with a 2x3x2 "
                "checkerboard +-6%, 16 periods 4-40 s, 100 stations -> 1600 solves"),
    "uneven": (W.uneven_mix, dict(nsolve=2000, seed=2),
               "2000 solves over 101x101, 201x161, 321x321, 81x241 @0.01deg in one plan, "
               "random fields 2.5-4.2 km/s, 50-200 receivers each"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes(N, out, nitems, mixed):
    """Algorithmic HBM bytes per kernel, summed over the run's items.

    FP64 forward: 72 B per cell-sweep (7 coefficient fields read, tau read and
    written) + 24 B per cell-iteration (convergence) + 48 B per cell once
    (T0 build, T = tau*T0).  FP64 adjoint: 64 B per cell-sweep + 16 B per
    cell-iteration + 80 B per cell once (coefficient build).  FP32 phase:
    half the per-sweep and per-iteration bytes (36 / 12 forward, 32 / 8
    adjoint), T0 build 24 + 12 B (FP64 and float), coefficient build 80 B.
    Polish: the FP64 figures plus 12 B per cell for the float-to-double
    conversion."""
    i64, i32 = out["iters"].astype(np.float64), out["iters_fp32"].astype(np.float64)
    f64, a64 = i64[:, 0].sum(), i64[:, 1].sum() + i64[:, 2].sum()
    nadj = nitems * (1 + int(np.any(i64[:, 2] > 0)))
    if not mixed:
        return {"k_fwd_strip": N * (72 * 4 * f64 + 24 * f64 + 48 * nitems),
                "k_adj_strip": N * (64 * 4 * a64 + 16 * a64 + 80 * nadj)}
    f32, a32 = i32[:, 0].sum(), i32[:, 1].sum() + i32[:, 2].sum()
    return {"k_fwd_mixed32": N * (36 * 4 * f32 + 12 * f32 + 36 * nitems),
            "k_fwd_polish": N * (72 * 4 * f64 + 24 * f64 + 36 * nitems),
            "k_adj_mixed32": N * (32 * 4 * a32 + 8 * a32 + 80 * nadj),
            "k_adj_polish": N * (64 * 4 * a64 + 92 * nadj)}


STAGE_OF = {"k_fwd_strip": "ms_fwd_fp64", "k_adj_strip": "ms_adj_fp64",
            "k_fwd_mixed32": "ms_fwd_fp32", "k_fwd_polish": "ms_fwd_fp64",
            "k_adj_mixed32": "ms_adj_fp32", "k_adj_polish": "ms_adj_fp64"}


def host_info():
    """What the host offers this process: its CPU model, the cores the OS
    reports, the cores this process may run on, and OMP_NUM_THREADS."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"cpu_model": model, "nproc_os": os.cpu_count(), "cores_this_process": affinity,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def scatter_block(count, size, rank):
    """The reference's static partition of `count` items over `size` ranks
    (scatter_all_i, src/shared/parallel.f90:1877-1903), 0-based [start, end)."""
    if count < size:
        return (rank, rank + 1) if rank < count else (0, 0)
    per, rem = divmod(count, size)
    if rank < rem:
        start = rank * (per + 1)
        return start, start + per + 1
    start = rank * per + rem
    return start, start + per


def cpu_baseline(parts, nits, nsample, threads, min_seconds=5.0):
    """The reference's own solver file (oracle/_ref, built from
    src/eikon_solver_2d.f90 at the -O3 of its release build) driving the loop
    body over evenly spaced items of the workload (spread over the parts in
    proportion) on `threads` host threads, each thread a contiguous block of
    them as the reference's ranks take theirs; interpolation, misfit and
    stacking by the oracle.  A first pass of `nsample` items sets the sample
    size so that the timed pass runs at least `min_seconds`.  Falls back to
    the oracle alone when the reference build is absent."""
    import oracle as O
    nit = sum(nits)

    def pick(nsamp):
        out = []
        for q, n in enumerate(nits):
            k = max(1, min(n, round(nsamp * n / nit)))
            out += [(q, int(i)) for i in np.unique(np.linspace(0, n - 1, k).round().astype(int))]
        return out

    if not O.RefSolver.available():
        picks = pick(nsample)
        t0 = time.perf_counter()
        for q in range(len(parts)):
            sel = [i for qq, i in picks if qq == q]
            O.forward_simulate(_subset(parts[q], np.array(sel)), nthreads=threads)
        dt = time.perf_counter() - t0
        return dict(value=len(picks) / dt, unit="solves/s", cores=threads, kind="port",
                    sample=f"{len(picks)} items of the workload, oracle loop, {dt:.1f} s wall")
    path = O.REF_O3_PATH if O.RefSolver.available(O.REF_O3_PATH) else O.REF_PATH
    build = "-O3" if path == O.REF_O3_PATH else "-O2"
    ref = O.RefSolver(path)
    err = []

    def work(items):
        try:
            acc = {}
            for q, it in items:
                b = parts[q]
                xg, yg = b["xgrids"], b["ygrids"]
                p = b["item_pidx"][it]
                r0, r1 = b["rec_off"][it], b["rec_off"][it + 1]
                v = b["svel"][p]
                a, c, bb = b["m11"][p], -b["m12"][p], b["m22"][p]
                x0, y0 = b["item_srcx"][it], b["item_srcy"][it]
                rx, ry = b["recx"][r0:r1], b["recy"][r0:r1]
                T = _ref_fwd(ref, xg, yg, a, bb, c, 1.0 / v, x0, y0)
                tek = O.interp2(xg, yg, T, rx, ry)
                tt, w = b["tt"][r0:r1], b["weight"][r0:r1].copy()
                w[np.abs(tek - tt) / tt > 0.5] = 0
                Ta = ref.mask(xg, yg, _ref_adj(ref, xg, yg, a, bb, c, T, rx, ry, w * (tek - tt)),
                              x0, y0)
                if b["dim"] == 3:   # density adjoint, RHS -w (src/measadj.f90:92-104)
                    Td = ref.mask(xg, yg, _ref_adj(ref, xg, yg, a, bb, c, T, rx, ry, -1.0 * w),
                                  x0, y0)
                    acc[(q, p)] = acc.get((q, p), 0) + Ta / (v * v * v) + Td / (v * v * v)
                else:
                    acc[(q, p)] = acc.get((q, p), 0) + Ta / (v * v)
        except BaseException as e:  # reported below
            err.append(e)

    def timed(picks):
        blocks = [picks[slice(*scatter_block(len(picks), threads, r))] for r in range(threads)]
        old = threading.stack_size()
        threading.stack_size(ref.STACK)
        try:
            ths = [threading.Thread(target=work, args=(blk,)) for blk in blocks]
            t0 = time.perf_counter()
            for t in ths:
                t.start()
            for t in ths:
                t.join()
            dt = time.perf_counter() - t0
        finally:
            threading.stack_size(old)
        if err:
            raise err[0]
        return dt

    picks = pick(nsample)
    dt = timed(picks)
    if dt < min_seconds and len(picks) < nit:      # scale the sample up to the time floor
        picks = pick(min(nit, int(len(picks) * 1.2 * min_seconds / dt) + threads))
        dt = timed(picks)
    passes = 1
    if dt < min_seconds:                           # every item: repeat the passes
        passes = int(np.ceil(1.2 * min_seconds / dt))
        dt = timed(picks * passes)
    return dict(value=passes * len(picks) / dt, unit="solves/s", cores=threads, kind="reference",
                sample=f"{len(picks)} of the workload's {nit} items (evenly spaced)"
                       f"{f', {passes} passes' if passes > 1 else ''}, reference "
                       f"src/eikon_solver_2d.f90 solvers built {build} (amdflang), {threads} "
                       f"threads each running a contiguous block of the items as the "
                       f"reference's ranks do, {dt:.2f} s wall, {dt * threads:.1f} thread-s",
                host=host_info())


def _ref_fwd(ref, xg, yg, a, b, c, fun, x0, y0):
    import ctypes
    import oracle as O
    nx, ny = len(xg), len(yg)
    a, b, c, fun = (np.asfortranarray(x) for x in (a, b, c, fun))
    T = np.zeros((nx, ny), order="F"); u = np.zeros((nx, ny), order="F")
    ref.lib.fsm_uw_ps_lonlat_2d_(O._d(xg), O._d(yg), ctypes.byref(ctypes.c_int(nx)),
                                 ctypes.byref(ctypes.c_int(ny)), O._d(a), O._d(b), O._d(c),
                                 O._d(T), O._d(fun), ctypes.byref(ctypes.c_double(x0)),
                                 ctypes.byref(ctypes.c_double(y0)), O._d(u))
    return T


def _ref_adj(ref, xg, yg, a, b, c, T, xr, yr, sa):
    import ctypes
    import oracle as O
    nx, ny = len(xg), len(yg)
    a, b, c, T = (np.asfortranarray(x) for x in (a, b, c, T))
    xr, yr, sa = (np.ascontiguousarray(x) for x in (xr, yr, sa))
    Ta = np.zeros((nx, ny), order="F")
    ref.lib.fsm_o1_jse_lonlat_2d_(O._d(xg), O._d(yg), ctypes.byref(ctypes.c_int(nx)),
                                  ctypes.byref(ctypes.c_int(ny)), O._d(a), O._d(b), O._d(c),
                                  O._d(T), O._d(Ta), O._d(xr), O._d(yr), O._d(sa),
                                  ctypes.byref(ctypes.c_int(len(xr))))
    return Ta


def _subset(b, items):
    """A batch with only `items` (same grid, model and path numbering)."""
    s = dict(b)
    s["item_pidx"] = np.ascontiguousarray(b["item_pidx"][items])
    s["item_srcx"] = np.ascontiguousarray(b["item_srcx"][items])
    s["item_srcy"] = np.ascontiguousarray(b["item_srcy"][items])
    offs = [0]
    sel = []
    for it in items:
        r0, r1 = b["rec_off"][it], b["rec_off"][it + 1]
        sel.append(np.arange(r0, r1))
        offs.append(offs[-1] + r1 - r0)
    sel = np.concatenate(sel)
    s["rec_off"] = np.array(offs, dtype=np.int32)
    for k in ("recx", "recy", "tt", "weight", "idx_path"):
        s[k] = np.ascontiguousarray(b[k][sel])
    return s


def copy_calibration(nbytes):
    """One hipMemcpy device-to-device of nbytes (the runtime's copy kernel):
    under rocprofv3 --pmc its FETCH_SIZE and WRITE_SIZE give the counters'
    scale for a transfer of known size."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes}), flush=True)


def _stream(srv, reqs, gaps):
    """Submit reqs at the given inter-arrival gaps; wait for every result."""
    tickets, t0 = [], time.perf_counter()
    due = t0
    for r, gap in zip(reqs, gaps):
        due += gap
        wait = due - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        tickets.append(srv.submit(r))
    outs = [srv.wait(t) for t in tickets]
    return outs, time.perf_counter() - t0


def serve(ctx, prec, nreq, qps, seed=4, plain_sample=6, sweep=(), sweep_requests=120):
    """The serving mix (grid sizes 101/201/401/801 squared with probabilities
    0.4/0.3/0.2/0.1, 100 receivers) as a Poisson stream of `nreq` requests at
    `qps`: latency percentiles from submit to results, throughput, and the
    plain FP64 one-solve-per-launch time of a sample of the same requests.
    `sweep`: further arrival rates, `sweep_requests` requests each, on the same
    server, for the latency/throughput curve up to saturation."""
    reqs = W.serving_requests(nreq, seed=seed)
    rng = np.random.default_rng(seed + 100)
    gaps = rng.exponential(1.0 / qps, nreq)
    # team size per grid: the autotuner's choice for one request of each size
    tuned = {}
    for m in sorted(set(int(r["nx"]) for r in reqs)):
        t = ctx.tune_waves(m, m, 1, prec)
        tuned[m] = {"waves": t["waves"], "stage_waves": t["stage_waves"],
                    "stage_prefetch_depth": t["stage_depth"],
                    "alone_ms_by_waves": [round(v, 2) for v in t["ms"]]}
    log(f"[bench] tuned launch shapes (one request): {tuned}")
    curve = []
    with satt.Server(ctx, prec) as srv:
        for r in reqs[:4]:                      # warm the kernels and pools
            srv.wait(srv.submit(r))
        outs, wall = _stream(srv, reqs, gaps)
        st = srv.stats()
        for k, q in enumerate(sweep):
            sreqs = W.serving_requests(sweep_requests, seed=seed + 11 + k)
            sg = np.random.default_rng(seed + 200 + k).exponential(1.0 / q, sweep_requests)
            so, sw = _stream(srv, sreqs, sg)
            sl = np.array([o["latency_ms"] for o in so])
            sq = np.array([o["queue_ms"] for o in so])
            curve.append({"qps": q, "requests": sweep_requests,
                          "p50_ms": round(float(np.percentile(sl, 50)), 2),
                          "p99_ms": round(float(np.percentile(sl, 99)), 2),
                          "queue_p99_ms": round(float(np.percentile(sq, 99)), 2),
                          "solves_per_s": round(sweep_requests / sw, 2),
                          "all_ok": bool(all(o["status"] == 0 for o in so))})
            log(f"[bench] serving at {q} requests/s: {curve[-1]}")
    lat = np.array([o["latency_ms"] for o in outs])
    queue = np.array([o["queue_ms"] for o in outs])
    sizes = np.array([r["nx"] for r in reqs])
    by_size = {int(m): {"n": int(np.sum(sizes == m)),
                        "p50_ms": round(float(np.percentile(lat[sizes == m], 50)), 2)}
               for m in sorted(set(sizes.tolist()))}
    plain = None
    if plain_sample > 0:
        tot, tplain = 0.0, 0.0
        for r in reqs[:plain_sample]:
            with ctx.plan(W.request_batch(r), satt.PREC_FP64) as pp:
                t1 = time.perf_counter()
                pp.run(True, True, satt.LAUNCH_PER_SOLVE)
                pp.wait()
                tplain += time.perf_counter() - t1
            tot += 1
        plain = tot / tplain
    return {"qps": qps, "requests": nreq,
            "mix": "squared grids 101/201/401/801 (p .4/.3/.2/.1) at 0.005 deg, 100 receivers",
            "p50_ms": round(float(np.percentile(lat, 50)), 2),
            "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "mean_ms": round(float(lat.mean()), 2), "by_size": by_size,
            "queue_p50_ms": round(float(np.percentile(queue, 50)), 2),
            "queue_p99_ms": round(float(np.percentile(queue, 99)), 2),
            "solves_per_s": round(nreq / wall, 2), "wall_s": round(wall, 2),
            "launch_groups": int(st["launches"]), "max_in_flight": int(st["max_in_flight"]),
            "launch_shapes": tuned,
            "all_ok": bool(all(o["status"] == 0 for o in outs)),
            "plain_fp64_per_solve_solves_per_s": None if plain is None else round(plain, 3),
            "rate_sweep": curve}


def dropin_leg(ctx, part, prec, calls=3):
    """satt_forward_simulate as the reference's caller uses it (the Fortran host
    calls it once per forward_simulate): host arrays in, results back to the
    host, PCIe both ways.  The first call creates the plan (device memory,
    uploads, team tuning on its own inputs); later calls with the same
    gathers and a new model upload only the model (src/tomo2d.f90:104,272)."""
    n = len(part["item_pidx"])
    t0 = time.perf_counter()
    ctx.forward_simulate_dropin(part, precision=prec)
    first = time.perf_counter() - t0
    times = []
    for k in range(calls):
        b = dict(part, svel=np.asfortranarray(part["svel"] * (1 + 1e-3 * (k + 1))))
        t1 = time.perf_counter()
        ctx.forward_simulate_dropin(b, precision=prec)
        times.append(time.perf_counter() - t1)
    ctx.forward_simulate_release()
    steady = float(np.median(times))
    model_mb = 4 * part["svel"].size * 8 / 1e6
    out_mb = (part["npath"] + part["svel"].size * (2 if part["dim"] == 3 else 1)) * 8 / 1e6
    return {"first_call_s": round(first, 3), "steady_s_per_call": round(steady, 4),
            "steady_solves_per_s": round(n / steady, 1), "calls": calls,
            "host_to_device_MB_per_call": round(model_mb, 1),
            "device_to_host_MB_per_call": round(out_mb, 1)}


def window_nodes(xg, yg, sigma):
    """Window nodes summed over the cells of one smoothing
    (src/shared/utils.f90:1204-1219): the window is a rectangle per cell, so
    the sum is (sum of x extents) x (sum of y extents)."""
    nx, ny = len(xg), len(yg)
    hx = int(sigma * 3 / (xg[1] - xg[0]))
    i = np.arange(nx)
    cx = np.sum(np.minimum(i + hx, nx - 1) - np.maximum(i - hx, 0) + 1)
    cy = 0
    for j in range(ny):
        hy = int(sigma * 3 / np.cos(np.deg2rad(yg[j])) / (yg[1] - yg[0]))
        cy += min(j + hy, ny - 1) - max(j - hy, 0) + 1
    return float(cx) * float(cy)


def model_side(ctx, cpu, sigma_2d=0.2, wavelen_factor=2.5, updatemax=0.02):
    """The model-side steps around the solver loop on the 401^2 workload grid:
    get_topo's metric for all 20 periods (smoothing of the topography with
    each period's sigma, src/grid.f90:114-123) and one optimize + steepest-
    descent update of a stacked kernel (src/tomo2d.f90:203-251).  Device
    time from HIP events around the kernels; the oracle times a few rows of
    the smoothing on one host thread for the CPU rate."""
    xg, yg, topo, periods, vref = W.topo_hills()
    sig = [satt.topo_sigma(v, p, wavelen_factor) for v, p in zip(vref, periods)]
    nodes = sum(window_nodes(xg, yg, s) for s in sig)
    ctx.topo_metric(xg, yg, topo, sig)
    t0 = time.perf_counter()
    m11, m12, m22, ang = ctx.topo_metric(xg, yg, topo, sig)
    wall = time.perf_counter() - t0
    dev = ctx.last_device_ms()
    n, npd = len(xg), len(periods)
    rng = np.random.default_rng(5)
    adj = np.asfortranarray(W.checkerboard_2d(n, n, 4, 4, 1e-3)[None] * rng.uniform(0.5, 1.5, (npd, 1, 1)))
    svel = np.asfortranarray(vref[:, None, None] * np.ones((npd, n, n)))
    ctx.steepest_descent(xg, yg, adj, sigma_2d, updatemax, svel)
    t0 = time.perf_counter()
    ctx.steepest_descent(xg, yg, adj, sigma_2d, updatemax, svel)
    sd_wall = time.perf_counter() - t0
    sd_dev = ctx.last_device_ms()
    sd_nodes = npd * window_nodes(xg, yg, sigma_2d)
    res = {
        "workload": f"{n}x{n} @{xg[1] - xg[0]:.3f}deg, topography of 8 hills, {npd} periods "
                    f"3-30 s, sigma = vel*T*{wavelen_factor}/(2 pi) "
                    f"({min(sig):.3f}-{max(sig):.3f} deg); steepest descent sigma_2d {sigma_2d}",
        "topo_metric": {"device_ms": round(dev, 3), "wall_ms": round(wall * 1e3, 3),
                        "window_nodes": nodes, "gnodes_per_s": round(nodes / dev / 1e6, 1)},
        "steepest_descent": {"device_ms": round(sd_dev, 3), "wall_ms": round(sd_wall * 1e3, 3),
                             "window_nodes": sd_nodes,
                             "gnodes_per_s": round(sd_nodes / sd_dev / 1e6, 1)},
        "metric_range": {"m11": [float(m11.min()), float(m11.max())],
                         "max_dip_deg": float(ang.max())},
    }
    # surface-wave dispersion of the 3-D stage's model (fwdsurf, before the loop)
    vs3, depz, per3 = W.acqui3d_model()
    ncol = vs3.shape[0] * vs3.shape[1]
    disp = {}
    for iw, igr, name in ((2, 0, "rayleigh_phase"), (1, 1, "love_group")):
        ctx.fwdsurf3d(vs3, depz, per3, iw, igr)
        t0 = time.perf_counter()
        sv = ctx.fwdsurf3d(vs3, depz, per3, iw, igr)
        disp[name] = {"device_ms": round(ctx.last_device_ms(), 3),
                      "wall_ms": round((time.perf_counter() - t0) * 1e3, 3),
                      "columns_per_s": round(ncol / (ctx.last_device_ms() * 1e-3), 1)}
        if cpu:
            import oracle as O
            if O.RefSolver.available():
                ref = O.RefSolver()
                cols = np.linspace(0, ncol - 1, 200).round().astype(int)
                t0 = time.perf_counter()
                worst = 0.0
                for c in cols:
                    i, j = c % vs3.shape[0], c // vs3.shape[0]
                    want = O.fwdsurf1d(ref, vs3[i, j], iw, igr, per3, depz)
                    worst = max(worst, float(np.max(np.abs(sv[:, i, j] - want) / want)))
                cdt = time.perf_counter() - t0
                disp[name]["cpu_baseline"] = {
                    "value": round(len(cols) / cdt, 1), "unit": "columns/s", "cores": 1,
                    "kind": "reference",
                    "sample": f"{len(cols)} of the {ncol} columns, the reference's surfdisp96 "
                              f"(oracle/_ref) through the oracle's fwdsurf1d, {cdt:.2f} s"}
                disp[name]["max_rel_diff_vs_reference"] = worst
                disp[name]["vs_cpu_core"] = round(disp[name]["columns_per_s"] / (len(cols) / cdt), 1)
    # Rayleigh depth kernels of the same model (depthkernel, before combine_kernels)
    ctx.depthkernel(vs3, depz, per3)
    t0 = time.perf_counter()
    dsv, _, _, dcp = ctx.depthkernel(vs3, depz, per3)
    dwall = time.perf_counter() - t0
    ddev = ctx.last_device_ms()
    disp["rayleigh_depth_kernels"] = {"device_ms": round(ddev, 3), "wall_ms": round(dwall * 1e3, 3),
                                      "columns_per_s": round(ncol / (ddev * 1e-3), 1)}
    if cpu:
        import oracle as O
        if O.RefSolver.available():
            ref = O.RefSolver()
            cols = np.linspace(0, ncol - 1, 40).round().astype(int)
            t0 = time.perf_counter()
            worst = 0.0
            for c in cols:
                i, j = c % vs3.shape[0], c // vs3.shape[0]
                thk, vp, vsr, rho = O.refinegrid(vs3[i, j], depz)
                cpr = ref.surfdisp96(thk, vp, vsr, rho, 2, 0, per3)
                osv, _, _ = O.depthkernel_rayleigh(vs3[i, j], depz, per3, cpr)
                worst = max(worst, float(np.max(np.abs(dsv[:, i, j, :] - osv))
                                         / np.max(np.abs(osv))))
            cdt = time.perf_counter() - t0
            disp["rayleigh_depth_kernels"]["cpu_baseline"] = {
                "value": round(len(cols) / cdt, 1), "unit": "columns/s", "cores": 1, "kind": "port",
                "sample": f"{len(cols)} of the {ncol} columns: the reference's surfdisp96 "
                          f"(oracle/_ref) on refinegrid's layers, then the oracle's sregn96, "
                          f"{cdt:.2f} s"}
            disp["rayleigh_depth_kernels"]["max_rel_diff_vs_cpu"] = worst
            disp["rayleigh_depth_kernels"]["vs_cpu_core"] = round(
                disp["rayleigh_depth_kernels"]["columns_per_s"] / (len(cols) / cdt), 1)
    res["dispersion"] = {"workload": f"{vs3.shape[0]}x{vs3.shape[1]} columns x {vs3.shape[2]} "
                                     f"depths 0-60 km, Vs 3.0-4.5 km/s +-6% checkerboard, "
                                     f"{len(per3)} periods 4-40 s, spherical earth, fundamental mode",
                         **disp}
    res["ker3d"] = ker3d_leg(ctx, cpu)
    if cpu:
        import oracle as O
        sample = [(sig[0], j) for j in (0, 200, 400)] + [(sig[npd // 2], 200), (sig[15], 200)]
        snodes = 0.0
        t0 = time.perf_counter()
        for s_, j in sample:
            O.smooth_geo2(xg, yg, topo, s_, rows=(j, j + 1))
            hx = int(s_ * 3 / (xg[1] - xg[0]))
            hy = int(s_ * 3 / np.cos(np.deg2rad(yg[j])) / (yg[1] - yg[0]))
            i = np.arange(n)
            snodes += float(np.sum(np.minimum(i + hx, n - 1) - np.maximum(i - hx, 0) + 1)) * \
                (min(j + hy, n - 1) - max(j - hy, 0) + 1)
        cdt = time.perf_counter() - t0
        rate = snodes / cdt
        res["cpu_baseline"] = {"value": round(rate / 1e6, 3), "unit": "Mnodes/s", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle gaussian_smooth_geo_2 on {len(sample)} rows "
                                         f"({snodes:.3g} window nodes, {cdt:.1f} s)",
                               "est_full_topo_metric_s": round(nodes / rate, 1)}
        res["vs_cpu_core"] = round(nodes / (dev * 1e-3) / rate, 1)
    return res


def ker3d_leg(ctx, cpu, n_inv=(8, 9, 8), nset=5, chi0=2.5, weight=1.0):
    """The 3-D stage's kernel steps after the loop on its 41x41x31 model with
    16 periods: combine_kernels through the Vs chain rule (src/acqui.f90:
    224-238), post_processing_for_kernel onto 5 inversion-grid sets of 8x9x8
    nodes (the reference's example 00 settings) and correct_depth
    (src/surfker/surfker.f90:260-283), with synthetic stacked kernels, depth
    sensitivities and dip angles of the right shapes.  The CPU
    leg runs the oracle's combination and the reference's own inv_grid_iso /
    inv2fwd_iso (oracle/_ref) on one host thread."""
    vs3, depz, per3 = W.acqui3d_model()
    nx, ny, nz = vs3.shape
    npd = len(per3)
    vx, vy = np.linspace(100.0, 102.0, nx), np.linspace(30.0, 32.0, ny)
    xi, yi, zi = W.inv_grids(vx, vy, depz, n_inv, nset)
    rng = np.random.default_rng(6)
    f = lambda *sh: np.asfortranarray(rng.standard_normal(sh))
    adj, adj_d = f(npd, nx, ny), f(npd, nx, ny)
    svs, svp, srho = f(npd, nx, ny, nz), f(npd, nx, ny, nz), f(npd, nx, ny, nz)
    vs = np.asfortranarray(vs3)
    vp = np.asfortranarray(0.9409 + 2.0947 * vs - 0.8206 * vs ** 2 + 0.2683 * vs ** 3 - 0.0251 * vs ** 4)
    grad0 = np.zeros((1, nx, ny, nz), order="F")
    out = {"workload": f"{nx}x{ny}x{nz} model, {npd} periods; {nset} inversion-grid sets of "
                       f"{n_inv[0]}x{n_inv[1]}x{n_inv[2]} nodes; combine through the Vs chain "
                       f"rule, no density preconditioning"}
    for rep in range(2):
        t0 = time.perf_counter()
        ker, kd = ctx.combine_kernels(adj, adj_d, svs, svp, srho, vs, vp, False, True)
        c_dev, c_wall = ctx.last_device_ms(), time.perf_counter() - t0
        t0 = time.perf_counter()
        _, grad = ctx.kernel_postproc(vx, vy, depz, xi, yi, zi, ker, kd, 0.0, chi0, weight, grad0)
        p_dev, p_wall = ctx.last_device_ms(), time.perf_counter() - t0
    out["combine"] = {"device_ms": round(c_dev, 4), "wall_ms": round(c_wall * 1e3, 3)}
    out["postproc"] = {"device_ms": round(p_dev, 4), "wall_ms": round(p_wall * 1e3, 3)}
    # correct_depth (depthkernel with topography) on the depth kernels, dip angles up to 30 deg
    dip = np.asfortranarray(rng.uniform(0.0, 30.0, (npd, nx, ny)))
    for rep in range(2):
        cd = ctx.correct_depth(svs, svp, srho, dip, depz)
        d_dev = ctx.last_device_ms()
    out["correct_depth"] = {"device_ms": round(d_dev, 4)}
    if cpu:
        import oracle as O
        t0 = time.perf_counter()
        wcd = O.correct_depth(svs, svp, srho, dip, depz)
        d_cpu = time.perf_counter() - t0
        out["correct_depth"]["max_rel_diff_vs_oracle"] = float(
            max(np.max(np.abs(a - b)) / np.max(np.abs(b)) for a, b in zip(cd, wcd)))
        out["correct_depth"]["cpu_baseline"] = {"value_ms": round(d_cpu * 1e3, 3), "cores": 1,
                                                "kind": "port", "sample": "the whole step"}
        out["correct_depth"]["vs_cpu_core"] = round(d_cpu * 1e3 / d_dev, 1)
    if cpu:
        import oracle as O
        t0 = time.perf_counter()
        wk, wkd = O.combine_kernels(adj, adj_d, svs, svp, srho, vs, vp, False, True)
        c_cpu = time.perf_counter() - t0
        out["combine"]["bitwise_vs_oracle"] = bool(np.array_equal(wk, ker))
        out["combine"]["cpu_baseline"] = {"value_ms": round(c_cpu * 1e3, 3), "cores": 1,
                                          "kind": "port", "sample": "the whole step"}
        if O.RefSolver.available():
            ref = O.RefSolver()
            t0 = time.perf_counter()
            gk = ref.inv_grid_iso(xi, yi, zi, wk[0], vx, vy, depz)
            up = ref.inv2fwd_iso(xi, yi, zi, gk, vx, vy, depz)
            p_cpu = time.perf_counter() - t0
            want = grad0[0] + up / nset / chi0 * weight
            out["postproc"]["bitwise_vs_reference"] = bool(np.array_equal(grad[0], want))
            out["postproc"]["cpu_baseline"] = {
                "value_ms": round(p_cpu * 1e3, 3), "cores": 1, "kind": "reference",
                "sample": "the whole step: the reference's inv_grid_iso + inv2fwd_iso (oracle/_ref)"}
            out["postproc"]["vs_cpu_core"] = round(p_cpu * 1e3 / p_dev, 1)
        out["combine"]["vs_cpu_core"] = round(c_cpu * 1e3 / c_dev, 1)
    return out


def host_threads():
    """The CPU threads this job is given: the cores it may run on, capped by
    OMP_NUM_THREADS where the host sets it (the share of a shared host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the HIP-graph replay timing of the same step")
    ap.add_argument("--no-tune", action="store_true",
                    help="one wave per item, no team-size tuning (and no comparison run)")
    ap.add_argument("--stage-depth", default=None, metavar="D,D,D,D",
                    help="with --stage-waves: prefetch depth per stage (2 or 4; default 2)")
    ap.add_argument("--stage-waves", default=None, metavar="W,W,W,W",
                    help="team sizes of the four sweep stages instead of tuning them "
                         "(forward FP32, forward FP64, adjoint FP32, adjoint FP64)")
    ap.add_argument("--workload", default="topo401", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--oracle-items", type=int, default=4,
                    help="items of the workload checked against the FP64 oracle (mixed)")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests of the serving stream (0: skip the serving measurement)")
    ap.add_argument("--serve-qps", type=float, default=40.0)
    ap.add_argument("--serve-sweep", default="20,40,80,160,320",
                    help="further arrival rates (requests/s) for the serving curve; '' for none")
    ap.add_argument("--serve-sweep-requests", type=int, default=120)
    ap.add_argument("--no-fp64-compare", action="store_true",
                    help="mixed mode: skip the FP64 batched run and the error report")
    ap.add_argument("--plain-team-items", type=int, default=4,
                    help="items timed one at a time through a one-item tuned strip plan")
    ap.add_argument("--plain-items", type=int, default=48,
                    help="items timed in the one-solve-per-launch FP64 form")
    ap.add_argument("--no-model-side", action="store_true",
                    help="skip the model-side leg (topographic metric, steepest descent)")
    ap.add_argument("--model-side-only", action="store_true",
                    help="run only the model-side leg and print its JSON")
    ap.add_argument("--no-hbm-probe", action="store_true",
                    help="skip the streaming kernels that measure achievable HBM bandwidth")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the end-to-end satt_forward_simulate calls (host arrays, PCIe)")
    ap.add_argument("--cpu-sample", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (a known-size "
                         "reference for the FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the (source, period) items shard across GPUs with one kernel all-reduce, "
              "which is not built here)", file=sys.stderr)
        sys.exit(2)

    if args.model_side_only:
        ctx = satt.Context(0)
        print(json.dumps(model_side(ctx, cpu=not args.no_cpu)), flush=True)
        return

    build, kw, desc = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    built = build(**kw)
    pairs = built if isinstance(built, list) else [built]
    nits = [len(bi["item_pidx"]) for _, bi in pairs]
    nit = sum(nits)
    grids = [[bi["nx"], bi["ny"]] for _, bi in pairs]
    log(f"[bench] workload {args.workload}: {nit} items, grids {grids} "
        f"(built in {time.perf_counter() - t0:.1f}s)")
    prec = satt.PREC_MIXED if args.precision == "mixed" else satt.PREC_FP64
    mixed = prec == satt.PREC_MIXED
    ctx = satt.Context(0)
    # serving first, on an otherwise idle device
    serving = None
    if args.serve_requests > 0:
        sweep = [float(v) for v in args.serve_sweep.split(",")] if args.serve_sweep else []
        serving = serve(ctx, prec, args.serve_requests, args.serve_qps, sweep=sweep,
                        sweep_requests=args.serve_sweep_requests)
        log(f"[bench] serving: {serving}")
    with ctx.plan_multi([bt for bt, _ in pairs]) as p:   # observed: FP64 forward, true model
        p.run(True, False)
        obs = [p.results(k)["tt_fwd"] for k in range(len(pairs))]
    parts = [W.set_observed(bi, o) for (_, bi), o in zip(pairs, obs)]
    log("[bench] observed times computed")

    with ctx.plan_multi(parts, prec) as plan:
        # team size per launch group from the autotuner (it times one wave
        # per item against teams of 2-4 waves on a synthetic batch of the
        # group's shape); the one-wave form is timed after, for comparison
        ttune, tuned_w = None, None
        if args.stage_waves:
            sw = [int(v) for v in args.stage_waves.split(",")]
            sd = [int(v) for v in args.stage_depth.split(",")] if args.stage_depth else [2] * 4
            plan.set_stage_waves(sw)
            plan.set_stage_depth(sd)
            tuned_w = {"stage_waves": sw, "stage_depth": sd, "stage_ms": [[0.0] * 4] * 8,
                       "pinned": True}
            ttune = 0.0
            log(f"[bench] launch shape pinned: waves {sw}, prefetch depth {sd}")
        elif not args.no_tune:
            t1 = time.perf_counter()
            tuned_w = plan.tune_waves()     # on the plan's own inputs
            ttune = time.perf_counter() - t1
            log(f"[bench] team sizes tuned in {ttune:.1f}s: {tuned_w}")
        for _ in range(args.warmup):
            plan.run(True, True)
            plan.wait()
        log(f"[bench] warmup done ({args.warmup})")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.run(True, True)
            plan.wait()
        dt = time.perf_counter() - t0
        outs = [plan.results(k) for k in range(len(parts))]
        st = plan.stats()
        # replay form: the same step captured once into a HIP graph
        graph = None
        if not args.no_graph:
            plan.set_graph(True)
            plan.run(True, True)
            plan.wait()
            t1 = time.perf_counter()
            for _ in range(args.steps):
                plan.run(True, True)
                plan.wait()
            gdt = time.perf_counter() - t1
            same = all(np.array_equal(plan.results(k)[key], outs[k][key])
                       for k in range(len(parts)) for key in ("tt_fwd", "adj_s"))
            plan.set_graph(False)
            graph = {"ms_per_step": round(gdt / args.steps * 1e3, 3),
                     "solves_per_s": round(args.steps * sum(len(b["item_pidx"]) for b in parts) / gdt, 3),
                     "results_same_as_eager": bool(same)}
            log(f"[bench] graph replay: {graph}")
        one_wave = None
        if not args.no_tune and not args.stage_waves:
            plan.set_waves(1)
            plan.set_stage_depth([2, 2, 2, 2])
            plan.run(True, True)
            plan.wait()
            t1 = time.perf_counter()
            plan.run(True, True)
            plan.wait()
            tdt = time.perf_counter() - t1
            same = all(np.array_equal(plan.results(k)[key], outs[k][key])
                       for k in range(len(parts)) for key in ("tt_fwd", "adj_s"))
            one_wave = {"tuning_s": round(ttune, 2), "one_wave_ms_per_step": round(tdt * 1e3, 3),
                        "results_same_as_one_wave": bool(same)}
            if tuned_w is not None:
                one_wave["stage_waves"] = dict(zip(STAGES, tuned_w["stage_waves"]))
                one_wave["stage_prefetch_depth"] = dict(zip(STAGES, tuned_w["stage_depth"]))
                one_wave["tune_stage_ms_by_waves"] = {
                    st: [round(tuned_w["stage_ms"][w][k], 2) for w in range(len(tuned_w["stage_ms"]))]
                    for k, st in enumerate(STAGES)}
                one_wave["tune_stage_ms_deep_prefetch"] = {
                    st: {"at_chosen_waves": round(tuned_w["stage_ms_deep"][0][k], 2),
                         "at_one_wave": round(tuned_w["stage_ms_deep"][1][k], 2)}
                    for k, st in enumerate(STAGES)}
            log(f"[bench] one wave per item: {one_wave}")
    log(f"[bench] timed {args.steps} steps in {dt:.3f}s")
    value = args.steps * nit / dt
    abytes = {}
    for b, o, n in zip(parts, outs, nits):
        for k, v in alg_bytes(b["nx"] * b["ny"], o, n, mixed).items():
            abytes[k] = abytes.get(k, 0.0) + v
    launches = max(1, st["launches"])
    kname = max(abytes, key=lambda k: st[STAGE_OF[k]])      # the dominant kernel
    kbytes, kms = abytes[kname], st[STAGE_OF[kname]]
    achieved = kbytes / (kms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{kname}_{args.workload}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    kernels = {k: {"ms": round(st[STAGE_OF[k]] / launches, 3),
                   "alg_GBs": round(abytes[k] / (st[STAGE_OF[k]] * 1e-3) / 1e9, 1)}
               for k in abytes if st[STAGE_OF[k]] > 0}
    kernels["fwd_setup"] = {"ms": round(st["ms_fwd_setup"] / launches, 3)}
    kernels["adj_setup"] = {"ms": round(st["ms_adj_setup"] / launches, 3)}
    kernels["step_device"] = {"ms": round(st["ms_total"], 3)}

    # mixed mode: the same step in FP64 (throughput and the error of the mixed
    # results against it), and a sample of items against the FP64 oracle
    fp64_value, mixed_err, fp64_shape, stage_cmp = None, None, None, None
    if mixed and not args.no_fp64_compare:
        with ctx.plan_multi(parts, satt.PREC_FP64) as p64:
            # the FP64 form gets the same launch-shape tuning as the mixed one
            t1 = time.perf_counter()
            tw64 = p64.tune_waves() if not args.no_tune else None
            t64 = time.perf_counter() - t1
            p64.run(True, True)
            p64.wait()
            t1 = time.perf_counter()
            p64.run(True, True)
            p64.wait()
            fp64_value = nit / (time.perf_counter() - t1)
            o64 = [p64.results(k) for k in range(len(parts))]
            s64 = p64.stats()
        if tw64 is not None:
            fp64_shape = {"tuning_s": round(t64, 2),
                          "stage_waves": dict(zip(STAGES, tw64["stage_waves"])),
                          "stage_prefetch_depth": dict(zip(STAGES, tw64["stage_depth"]))}
        # per-stage device times of one step, mixed against FP64, both tuned
        stage_cmp = {
            "forward_ms": {"mixed_fp32_phase": round(st["ms_fwd_fp32"], 3),
                           "mixed_fp64_polish": round(st["ms_fwd_fp64"], 3),
                           "mixed_total": round(st["ms_fwd_fp32"] + st["ms_fwd_fp64"], 3),
                           "fp64": round(s64["ms_fwd_fp64"], 3)},
            "adjoint_ms": {"mixed_fp32_phase": round(st["ms_adj_fp32"], 3),
                           "mixed_fp64_polish": round(st["ms_adj_fp64"], 3),
                           "mixed_total": round(st["ms_adj_fp32"] + st["ms_adj_fp64"], 3),
                           "fp64": round(s64["ms_adj_fp64"], 3)},
            "setup_ms": {"mixed": round(st["ms_fwd_setup"] + st["ms_adj_setup"], 3),
                         "fp64": round(s64["ms_fwd_setup"] + s64["ms_adj_setup"], 3)},
        }
        log(f"[bench] batched FP64 (tuned {fp64_shape}): {fp64_value:.1f} solves/s; "
            f"stages {stage_cmp}")
        cat = lambda key, ol: np.concatenate([o[key].ravel() for o in ol])
        d = np.abs(cat("tt_fwd", outs) - cat("tt_fwd", o64))
        mixed_err = {
            "vs": "batched FP64 on the device (parity-checked against the oracle)",
            "tt_max_ms": float(d.max() * 1e3), "tt_mean_ms": float(d.mean() * 1e3),
            "kernel_rel_l2": float(np.linalg.norm(cat("adj_s", outs) - cat("adj_s", o64))
                                   / np.linalg.norm(cat("adj_s", o64))),
            "chi_rel": float(abs(sum(o["chi"] for o in outs) - sum(o["chi"] for o in o64))
                             / abs(sum(o["chi"] for o in o64))),
        }
        if parts[0]["dim"] == 3:
            mixed_err["density_rel_l2"] = float(
                np.linalg.norm(cat("adj_density", outs) - cat("adj_density", o64))
                / np.linalg.norm(cat("adj_density", o64)))
        if args.oracle_items > 0:
            import oracle as O
            worst, worst64 = 0.0, 0.0
            for b, o, o6, n in zip(parts, outs, o64, nits):
                k = max(1, args.oracle_items // len(parts))
                sub = _subset(b, np.linspace(0, n - 1, k).round().astype(int))
                ref = O.forward_simulate(sub, istotable=True, isadj=False, nthreads=k)
                ps = sub["idx_path"]
                worst = max(worst, float(np.abs(o["tt_fwd"][ps] - ref["tt_fwd"][ps]).max()))
                worst64 = max(worst64, float(np.abs(o6["tt_fwd"][ps] - ref["tt_fwd"][ps]).max()
                                             / np.abs(ref["tt_fwd"][ps]).max()))
            mixed_err["oracle_sample"] = {"items": max(1, args.oracle_items // len(parts)) * len(parts),
                                          "tt_max_ms": worst * 1e3, "fp64_tt_max_rel": worst64}
        log(f"[bench] mixed error: {mixed_err}")

    # the drop-in call, host arrays in and out (one grid only)
    dropin = None
    if len(parts) == 1 and not args.no_dropin:
        dropin = dropin_leg(ctx, parts[0], prec)
        log(f"[bench] drop-in satt_forward_simulate: {dropin}")

    # plain FP64 form: one solve per launch, resident inputs, a sample of
    # items spread over the parts in proportion
    plain = None
    if args.plain_items > 0:
        tot, pdt = 0, 0.0
        for b, n in zip(parts, nits):
            k = max(1, round(args.plain_items * n / nit))
            sub = _subset(b, np.linspace(0, n - 1, k).round().astype(int))
            with ctx.plan(sub, satt.PREC_FP64) as pp:
                pp.run(True, True, satt.LAUNCH_PER_SOLVE)
                pp.wait()
                t1 = time.perf_counter()
                pp.run(True, True, satt.LAUNCH_PER_SOLVE)
                pp.wait()
                pdt += time.perf_counter() - t1
            tot += k
        plain = tot / pdt
        log(f"[bench] plain per-solve FP64: {plain:.1f} solves/s")
    # the same one-solve-at-a-time use through the strip kernels: a plan of
    # one item, its team size and prefetch depth tuned, one launch per stage
    plain_team = None
    if args.plain_team_items > 0:
        tot, pdt, shapes = 0, 0.0, []
        for b, n in zip(parts, nits):
            k = max(1, round(args.plain_team_items * n / nit))
            for it in np.unique(np.linspace(0, n - 1, k).round().astype(int)):
                with ctx.plan(_subset(b, np.array([it])), satt.PREC_FP64) as pp:
                    if not args.no_tune:
                        tw = pp.tune_waves()
                        shapes.append(tw["stage_waves"][1:4:2])
                    pp.run(True, True)
                    pp.wait()
                    t1 = time.perf_counter()
                    pp.run(True, True)
                    pp.wait()
                    pdt += time.perf_counter() - t1
                tot += 1
        plain_team = {"solves_per_s": round(tot / pdt, 3), "items": tot,
                      "fwd_adj_team_waves": shapes}
        log(f"[bench] one-item team-strip FP64: {plain_team}")

    cpu = None
    if not args.no_cpu:
        th = host_threads()
        cpu = cpu_baseline(parts, nits, max(args.cpu_sample, th), th)
        log(f"[bench] cpu baseline: {cpu['value']:.2f} solves/s on {th} threads")

    hbm = None if args.no_hbm_probe else ctx.hbm_probe(1 << 31)
    lds = None if args.no_hbm_probe else ctx.lds_probe()
    if hbm is not None:
        log(f"[bench] streaming kernels: read {hbm[0]:.0f} GB/s, copy {hbm[1]:.0f} GB/s; "
            f"LDS read {lds[0]:.0f} GB/s of {lds[1]:.0f}")
    side = None
    if not args.no_model_side:
        side = model_side(ctx, cpu=not args.no_cpu)
        log(f"[bench] model side: {side}")

    out = outs[0]
    iters = out["iters"]
    line = {
        "metric": "fwd+adj solves/s at 401^2 grid, batched mixed vs plain FP64 one-solve-per-launch",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == satt.PREC_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": desc, "items": nit,
                   "grid": grids[0] if len(grids) == 1 else grids, "precision": args.precision,
                   "launch": "batched", "chunk": st["chunk"],
                   "launch_shape": ("one wave per item, two-step prefetch" if args.no_tune else
                                    f"pinned per stage: waves {args.stage_waves}, prefetch depth "
                                    f"{args.stage_depth or '2,2,2,2'}" if args.stage_waves else
                                    "team size and prefetch depth autotuned per stage"),
                   "stage_waves": dict(zip(STAGES, tuned_w["stage_waves"])) if tuned_w else None,
                   "stage_prefetch_depth": (dict(zip(STAGES, tuned_w["stage_depth"]))
                                            if tuned_w else None),
                   "mean_iters_fwd": float(np.concatenate([o["iters"][:, 0] for o in outs]).mean()),
                   "mean_iters_adj": float(np.concatenate([o["iters"][:, 1] for o in outs]).mean()),
                   "mean_iters_fwd_fp32": float(np.concatenate([o["iters_fp32"][:, 0] for o in outs]).mean()),
                   "mean_iters_adj_fp32": float(np.concatenate([o["iters_fp32"][:, 1] for o in outs]).mean())},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "achievable_GBs": None if hbm is None else
                     {"stream_read": round(hbm[0], 1), "stream_copy": round(hbm[1], 1)},
                     # the strip kernels keep their operands in registers and pass
                     # neighbours by DPP; LDS carries only a strip's edge column
                     # (one value and its T0 per row and sweep), so it bounds nothing
                     "lds": None if lds is None else
                     {"probe_read_GBs": round(lds[0], 1), "peak_GBs": round(lds[1], 1),
                      "kernel_lds_bytes_per_strip_row_sweep": 16 if mixed else 32},
                     "measured_traffic_GBs": None if traffic is None else
                     round(traffic / (kms / launches * 1e-3) / 1e9, 1),
                     "alg_bytes_per_launch": kbytes / launches,
                     "avg_launch_ms": kms / launches},
        "kernels": kernels,
        "graph_replay": graph,
        "team_tuning": one_wave,
        "batched_fp64": None if fp64_value is None else round(fp64_value, 3),
        "batched_fp64_launch_shape": fp64_shape,
        "mixed_vs_fp64_stages": stage_cmp,
        "mixed_error": mixed_err,
        "dropin_end_to_end": dropin,
        "plain_fp64_per_solve": None if plain is None else round(plain, 3),
        "plain_fp64_team_strip_per_solve": plain_team,
        "batched_vs_plain": None if plain is None else round(value / plain, 2),
        "serving": serving,
        "cpu_baseline": cpu,
        "vs_cpu": None if cpu is None else round(value / cpu["value"], 1),
        "model_side": side,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
