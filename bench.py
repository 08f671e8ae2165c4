#!/usr/bin/env python3
"""Benchmark of the MI355X LV-ELVA time step.

Workload: the 64-member ensemble (256 x 256 ice grid at 16 km, N = 256, random
log10-viscosity field, lithosphere thickness and dome per member, seeds
1000 + m).  One bench step = one coupling step (isos_update to t + 10 yr) of the
whole batch with the batched mixed-precision path (FP32 Hartley passes, FP64
state, error norm and sea level); inputs are resident in HBM.

Reported beside it:
  plain_fp64    the single-domain FP64 drop-in path on a subset of members
  max_err_m     max |w_mixed - w_fp64| over all members after the timed steps
  roofline      the dominant kernel, timed with HIP events on the library
                stream over the timed region
  cpu_baseline  the FP64 oracle (a plain-C restatement of the reference path),
                single thread, on one member for a bounded sample

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNEL_NAMES = {0: "k_stencil", 1: "k_row_fwd", 2: "k_col", 3: "k_row_inv"}
FUSED_NAME = "kf_srow_fwd (stencil + forward row pass)"
PROF_ALL = 99


def fast_kernel_names(shape):
    """Template names (as rocprofv3 prints them) of the register-FFT attempt
    kernels the last timed attempt ran, from fi_prof_shape: {N, fp64, fused,
    fast, threads_row, C, row_tiles, srow_rows, threads_inv, dt_method,
    column-pass exchange buffers}."""
    n, fp64, fused, fast, th, c, rt, rows, inv, meth, nbuf = shape
    if not (fast and fused):
        return {}
    t = "double" if fp64 else "float"
    tpl = n // 8
    sth = (n if n % 3 else 384) if rows else th
    lf = rows // 2 if rows else sth // tpl
    return {0: f"kf_srow_fwd<{t}, {n}, {sth}, {lf}>", 2: f"kf_col<{t}, {n}, {c}, true, {'true' if nbuf == 2 else 'false'}>",
            3: f"kf_row_inv<{t}, {n}, {inv or th}, {rt}, {meth}>"}


# Algorithmic HBM bytes per cell of the once-per-coupling-step kernels of an
# LV-ELVA update without convolutions (uniform spacing, elastic and sea-surface
# convolutions off: the ensemble workload), all FP64 but the FP32 copy of
# canom_full and the 1-byte mask.  The per-domain scalar kernels (begin, bsl,
# decide, corner, check, gather) move a few hundred bytes per domain.  The
# FETCH_SIZE / WRITE_SIZE counters of the 64-member ensemble read 40.2 + 16.0,
# 65.1 + 45.0 and 40.1 + 24.1 B/cell for these three.
STEP_BYTES_PER_CELL = {
    "k_pre_haf": 40 + 16,        # read H_ice (ice grid, embedded each update), z_ss, z_bed, Haf_ref,
                                 # H_ice_ref; write H_ice, Haf
    "k_pre_fields": 65 + 45,     # read mask, z_ss_ref, Haf, z_bed, H_ice, Haf_ref, dz_ss, we_ref,
                                 # canom_mantle; write canom_load, we, dz_ss, z_ss, mask, canom_full (f64, f32)
    "k_finalize": 40 + 24,       # read w, w_ref, we, z_bed_ref, dz_ss; write now%w, canom_mantle, z_bed
}


def step_roofline(cells, attempts_per_step, ms_per_step, s=4, peak=8000.0):
    """The coupling step as run (concurrent sub-batches, bookkeeping kernels
    included): algorithmic bytes of every kernel of a step over its wall time."""
    attempt_b = sum(kernel_bytes_per_cell(k, s, True) for k in (0, 2, 3))
    per_cell = sum(STEP_BYTES_PER_CELL.values()) + attempt_b * attempts_per_step
    b = cells * per_cell
    gbs = b / (ms_per_step * 1e-3) / 1e9
    return {"bytes_per_step": int(b), "ms_per_step": round(ms_per_step, 4),
            "attempts_per_step": round(attempts_per_step, 2),
            "bytes_per_cell": {"attempt": attempt_b, **STEP_BYTES_PER_CELL},
            "achieved": round(gbs, 1), "peak": peak, "unit": "GB/s", "frac": round(gbs / peak, 4)}


def measured_traffic(kernel_name, cells):
    """HBM bytes per launch of a kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE measurement (scaled by the copy-kernel calibration), if it
    covers this kernel at this size."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_ensemble256.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"].get(kernel_name)
    except (OSError, ValueError, KeyError):
        return None
    if not k or k["cells"] != cells:
        return None
    return k["read_bytes"] + k["write_bytes"]


def kernel_bytes_per_cell(kernel, s, fused):
    """Algorithmic HBM bytes per cell per launch (s = bytes per RHS real).

    dwdt is written only by the attempt that reaches the coupling time, so the
    inverse row pass is priced without it."""
    if kernel == 0 and fused:
        return 8 + 3 * s + s       # read x (f64), D, 1/(2 eta), canom_full; write row-transformed
    return {0: 8 + 4 * s,          # read x (f64), D, 1/(2 eta), canom_full; write f
            1: 2 * s,              # read f, write row-transformed
            2: 3 * s,              # read, read kappa*R, write
            3: s + 8 + 8}[kernel]  # read, read x (f64), write x2 (f64)


def build_members(nmem, precision):
    from fastisostasy_amd import isostasy as G
    from fastisostasy_amd import workloads as W
    hs, bsls, wls = [], [], []
    for m in range(nmem):
        wl = W.ensemble_member(m)
        p = G.make_params(precision=precision, **wl["params"])
        h = G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"], eta_field=wl["eta_field"])
        b = G.bsl_init("const", 0.0)
        h.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
        hs.append(h)
        bsls.append(b)
        wls.append(wl)
    return hs, bsls, wls


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if xs else None


def autotune(n, precision, nmem, reps=3):
    """fi_autotune: launch shapes for `nmem` domains of N x N picked from timed
    trial updates; returns the choice (0 = the built-in one)."""
    import ctypes
    from fastisostasy_amd._lib import check, load
    choice = (ctypes.c_int * 8)()
    t0 = time.perf_counter()
    check(load().fi_autotune(n, 0 if precision == "fp64" else 1, nmem, reps, choice))
    keys = ("threads_row", "col_lines", "row_tiles", "subbatches", "generic_row_lines", "generic_col_lines",
            "srow_rows", "inv_threads")
    out = {k: int(v) for k, v in zip(keys, choice)}
    out["built_in"] = all(int(v) == 0 for v in choice)   # the built-in shapes were the fastest
    out["seconds"] = round(time.perf_counter() - t0, 2)
    return out


def run_batched(nmem, steps, warmup, precision, prof_kernel=None, timed=True, extra_steps=0):
    from fastisostasy_amd import isostasy as G
    from fastisostasy_amd._lib import check, load
    import ctypes
    hs, bsls, wls = build_members(nmem, precision)
    batch = G.Batch(hs)
    t = 0.0
    # first step uploads the loads (step-on at t = 1e-9 as in the reference driver)
    t += 10.0
    batch.update(t, bsls, H_ice=[wl["H_ice"] for wl in wls])
    for _ in range(max(0, warmup - 1)):
        t += 10.0
        batch.update(t, bsls)
    check(load().fi_sync())
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps if timed else 0):
        t += 10.0
        s0 = time.perf_counter()
        batch.update(t, bsls)
        lat.append(time.perf_counter() - s0)
    check(load().fi_sync())
    wall = time.perf_counter() - t0
    prof = None
    attempts_timed = sum(h.scalar("attempts") for h in hs)
    accepts_timed = sum(h.scalar("accepts") for h in hs)
    if prof_kernel is not None:
        # Kernel timing: the same number of steps again, continuing the run,
        # with HIP events on every launch of one kernel (the events slow the
        # stream, so they stay out of the region timed above), and the whole
        # batch in one launch per kernel (no concurrent sub-batches), so a
        # launch covers every member and runs alone on the GPU.
        saved = os.environ.get("FI_SUBBATCHES")
        os.environ["FI_SUBBATCHES"] = "1"
        check(load().fi_prof_select(prof_kernel))
        for _ in range(steps):
            t += 10.0
            batch.update(t, bsls)
        check(load().fi_sync())
        if saved is None:
            del os.environ["FI_SUBBATCHES"]
        else:
            os.environ["FI_SUBBATCHES"] = saved
        prof = {}
        for k in ((0, 2, 3) if prof_kernel == PROF_ALL else (prof_kernel,)):
            ms, cnt = ctypes.c_double(), ctypes.c_longlong()
            check(load().fi_prof_collect_kernel(k, ctypes.byref(ms), ctypes.byref(cnt)))
            prof[k] = (ms.value, cnt.value)
        prof["fused"] = bool(load().fi_prof_fused())
        shape = (ctypes.c_int * 11)()
        check(load().fi_prof_shape(shape))
        prof["names"] = fast_kernel_names(list(shape))
        check(load().fi_prof_select(-1))
    for _ in range(extra_steps):
        # untimed steps that bring a run to the end time of another
        t += 10.0
        batch.update(t, bsls)
    attempts, accepts = attempts_timed, accepts_timed
    w = [h.full("w") for h in hs]
    n = hs[0].n
    batch.close()
    for h in hs:
        h.close()
    return dict(wall=wall, lat=lat, prof=prof, attempts=attempts, accepts=accepts, w=w, n=n, t_end=t)


def run_plain(members, steps, warmup):
    """Single-domain FP64 drop-in path, one member after the other."""
    from fastisostasy_amd import isostasy as G
    from fastisostasy_amd import workloads as W
    total = 0.0
    n = None
    for m in range(members):
        wl = W.ensemble_member(m)
        p = G.make_params(precision="fp64", **wl["params"])
        h = G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"], eta_field=wl["eta_field"])
        b = G.bsl_init("const", 0.0)
        h.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
        t = 0.0
        for _ in range(warmup):
            t += 10.0
            G.bsl_update(b, t)
            h.update(wl["H_ice"], t, b)
        t0 = time.perf_counter()
        for _ in range(steps):
            t += 10.0
            G.bsl_update(b, t)
            h.update(wl["H_ice"], t, b)
        total += time.perf_counter() - t0
        n = h.n
        h.close()
    return dict(wall=total, n=n)


def run_single(precision, updates=40, warmup=5, member=0, tune=True):
    """The single-domain drop-in path: one 256^2 ensemble member (N = 256),
    isos_update by isos_update, after the autotuner picked its launch shape
    for one domain; per-update wall times and attempts."""
    from fastisostasy_amd import isostasy as G
    from fastisostasy_amd import workloads as W
    shape = autotune(256, precision, 1) if tune else None
    wl = W.ensemble_member(member)
    h = G.IsosClass(G.make_params(precision=precision, **wl["params"]), wl["nx"], wl["nx"], wl["dx"],
                    he_field=wl["he_field"], eta_field=wl["eta_field"])
    b = G.bsl_init("const", 0.0)
    h.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
    t = 0.0
    for _ in range(warmup):
        t += 10.0
        h.update(wl["H_ice"], t, b)
    a0 = h.scalar("attempts")
    lat = []
    for _ in range(updates):
        t += 10.0
        s0 = time.perf_counter()
        h.update(wl["H_ice"], t, b)
        lat.append(time.perf_counter() - s0)
    att = (h.scalar("attempts") - a0) / updates
    h.close()
    return {"p50_ms": round(1e3 * percentile(lat, 50), 3), "p99_ms": round(1e3 * percentile(lat, 99), 3),
            "attempts_per_update": round(att, 2), "us_per_attempt": round(1e6 * percentile(lat, 50) / att, 2),
            "launch_shape": shape}


def kernels_only(name, updates):
    """Untimed updates of one workload, for a profiler's kernel statistics:
    lv_rheology (N = 512, mixed, elastic convolution), sealevel (384^2 ->
    N = 512, mixed, elastic and sea-surface convolutions every update), n384
    (one 384^2 serving domain, mixed-radix kernels), single256 (one 256^2
    ensemble member, the single-domain path)."""
    from fastisostasy_amd import isostasy as G
    from fastisostasy_amd import workloads as W
    if name == "lv_rheology":
        wl, dt, H_of_t = W.lv_rheology(), 0.5, None
    elif name == "sealevel":
        wl, dt = W.sealevel_cycle(), 10.0
        H_of_t = lambda t: wl["H_max"] * float(W.sawtooth(t))
    elif name == "n384":
        wl, dt, H_of_t = W.ensemble_member(0, nx=384, seed=5555), 10.0, None
    elif name == "single256":
        wl, dt, H_of_t = W.ensemble_member(0), 10.0, None
    else:
        raise SystemExit(f"unknown workload {name}")
    H = (lambda t: wl["H_ice"]) if H_of_t is None else H_of_t
    h = G.IsosClass(G.make_params(precision="mixed", **wl["params"]), wl["nx"], wl["nx"], wl["dx"],
                    he_field=wl.get("he_field"), eta_field=wl.get("eta_field"))
    b = G.bsl_init(wl.get("bsl_method", "const"), 0.0)
    t0 = 89.9e3 if name == "sealevel" else 0.0
    if name == "sealevel":
        h.init_ref(wl["z_bed"], np.zeros_like(wl["z_bed"]))
    h.init_state(wl["z_bed"], H(t0) if name == "sealevel" else np.zeros_like(wl["z_bed"]), t0, b)
    t = t0
    s0 = time.perf_counter()
    for _ in range(updates):
        t += dt
        G.bsl_update(b, t)
        h.update(H(t), t, b)
    wall = time.perf_counter() - s0
    out = {"workload": name, "n": h.n, "updates": updates, "wall_s": round(wall, 4),
           "attempts": h.scalar("attempts")}
    h.close()
    return out


def run_serving(requests, n_domains, capacity_seconds):
    """Serving workload: resident domains of 128/256/384/512 cells (4:3:2:1,
    drawn like ensemble members, seeds 5555 + k), requests that each advance one
    domain by 10 yr, Poisson arrivals at 1.2 x the unbatched capacity (the
    closed-loop request rate when every request is launched on its own)."""
    from fastisostasy_amd import isostasy as G
    from fastisostasy_amd import serving as S
    from fastisostasy_amd import workloads as W
    doms = W.serving_domains(n_domains)
    hs, bs = [], []
    for wl in doms:
        p = G.make_params(precision="mixed", **wl["params"])
        h = G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"], eta_field=wl["eta_field"])
        b = G.bsl_init("const", 0.0)
        h.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
        hs.append(h)
        bs.append(b)
    foot = sum(h.scalar("hbm_bytes") for h in hs)
    srv = S.Server(hbm_budget_bytes=1.05 * foot, max_batch=64)
    slots = [srv.admit(h, b) for h, b in zip(hs, bs)]
    t_now = [10.0] * len(doms)
    for k, wl in enumerate(doms):   # load the domes (untimed)
        srv.submit(slots[k], t_now[k], k, H_ice=wl["H_ice"])
    while srv.pending():
        srv.run()
    srv.completed()
    one = S.Server(hbm_budget_bytes=1.05 * foot, max_batch=1)
    one_slots = [one.admit(h, b) for h, b in zip(hs, bs)]
    cap1 = S.closed_loop_rate(one, one_slots, t_now, capacity_seconds)
    one.close()
    capb = S.closed_loop_rate(srv, slots, t_now, capacity_seconds)
    st0 = srv.stats()
    rate = 1.2 * cap1
    arrivals, dom = W.serving_requests(len(doms), requests, rate)
    prog = 10000 if requests > 20000 else 0
    lat, wall = S.open_loop(srv, slots, t_now, arrivals, dom, progress=prog)
    st1 = srv.stats()
    cells = sum(int(doms[int(k)]["nx"]) ** 2 for k in dom)
    # the same stream at 0.8 x the batched capacity: latency under load
    rate_hi = 0.8 * capb
    arr_hi, dom_hi = W.serving_requests(len(doms), requests, rate_hi, seed=5556)
    lat_hi, wall_hi = S.open_loop(srv, slots, t_now, arr_hi, dom_hi, progress=prog)
    st2 = srv.stats()
    # continuous scheduling (fi_server_poll) on the same server: capacity, and
    # the same light-load and high-load streams
    capc = S.closed_loop_rate(srv, slots, t_now, capacity_seconds, continuous=True)
    st3 = srv.stats()
    lat_c, wall_c = S.open_loop(srv, slots, t_now, arrivals, dom, progress=prog, continuous=True)
    st4 = srv.stats()
    # the round-based high-load stream again (0.8 x the round-based capacity)
    rate_chi = rate_hi
    lat_chi, wall_chi = S.open_loop(srv, slots, t_now, arr_hi, dom_hi, progress=prog, continuous=True)
    st5 = srv.stats()
    srv.close()
    for h in hs:
        h.close()
    lat_ms = 1e3 * lat
    return {"requests": requests, "domains": len(doms), "sizes": "128/256/384/512 at 4:3:2:1",
            "rate_req_s": round(rate, 1), "unbatched_capacity_req_s": round(cap1, 1),
            "batched_capacity_req_s": round(capb, 1), "throughput_req_s": round(requests / wall, 1),
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 3), "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
            "max_ms": round(float(lat_ms.max()), 3),
            "mean_group": round((st1["served"] - st0["served"]) / max(1, st1["groups"] - st0["groups"]), 2),
            "grid_cell_years_s": round(cells * 10.0 / wall, 1), "hbm_budget_bytes": int(1.05 * foot),
            "high_load": {"rate_req_s": round(rate_hi, 1), "throughput_req_s": round(requests / wall_hi, 1),
                          "p50_ms": round(float(np.percentile(1e3 * lat_hi, 50)), 3),
                          "p99_ms": round(float(np.percentile(1e3 * lat_hi, 99)), 3),
                          "mean_group": round((st2["served"] - st1["served"]) / max(1, st2["groups"] - st1["groups"]), 2)},
            "continuous": {
                "batched_capacity_req_s": round(capc, 1), "rate_req_s": round(rate, 1),
                "throughput_req_s": round(requests / wall_c, 1),
                "p50_ms": round(float(np.percentile(1e3 * lat_c, 50)), 3),
                "p99_ms": round(float(np.percentile(1e3 * lat_c, 99)), 3),
                "max_ms": round(float(1e3 * lat_c.max()), 3),
                "mean_group": round((st4["served"] - st3["served"]) / max(1, st4["groups"] - st3["groups"]), 2),
                "high_load": {"rate_req_s": round(rate_chi, 1), "throughput_req_s": round(requests / wall_chi, 1),
                              "p50_ms": round(float(np.percentile(1e3 * lat_chi, 50)), 3),
                              "p99_ms": round(float(np.percentile(1e3 * lat_chi, 99)), 3),
                              "mean_group": round((st5["served"] - st4["served"]) /
                                                  max(1, st5["groups"] - st4["groups"]), 2)}}}


def cpu_model():
    """The host CPU's model name (first 'model name' line of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _cpu_worker(args):
    m, seconds = args
    return run_cpu_baseline(seconds, member=m)


def run_cpu_parallel(seconds, workers):
    """The oracle on `workers` ensemble members at once, one process each (the
    oracle is single-threaded): the host's aggregate FP64 CPU rate."""
    import multiprocessing as mp
    with mp.get_context("spawn").Pool(workers) as pool:
        res = pool.map(_cpu_worker, [(m, seconds) for m in range(workers)])
    return dict(value=sum(r["value"] for r in res), steps=sum(r["steps"] for r in res), workers=workers,
                n=res[0]["n"], wall=max(r["wall"] for r in res))


def run_cpu_baseline(seconds, member=0):
    """FP64 oracle, one thread, one ensemble member, as many coupling steps as fit."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from fastisostasy_amd import workloads as W
    wl = W.ensemble_member(member)
    o = O.OracleIsos(O.make_params(**wl["params"]), wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"],
                     eta_field=wl["eta_field"])
    b = O.OracleBSL("const", 0.0)
    o.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
    t = 0.0
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < seconds:
        t += 10.0
        b.update(t)
        o.update(wl["H_ice"], t, b)
        steps += 1
    wall = time.perf_counter() - t0
    n = o.n
    return dict(value=n * n * 10.0 * steps / wall, steps=steps, wall=wall, n=n)


def copy_bandwidth(mib=1024, reps=20):
    """Achieved device-to-device copy bandwidth (read + write bytes / s) of a
    plain copy of `mib` MiB on the library stream, beside the HBM peak."""
    import ctypes
    from fastisostasy_amd._lib import check, load
    v = ctypes.c_double()
    check(load().fi_bench_copy(float(mib), int(reps), ctypes.byref(v)))
    return v.value


def calibrate_copy(mib):
    """A device-to-device copy of known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters of rocprofv3 (not part of the benchmark)."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}))


def kernel_roofline(prof, k, cells, n, fused):
    """Algorithmic bytes per launch of attempt kernel k over its event-measured
    average launch time, and the calibrated HBM traffic where measured."""
    ms, cnt = prof[k]
    if cnt <= 0:
        return None
    avg_s = ms / cnt * 1e-3
    nbytes = kernel_bytes_per_cell(k, 4, fused) * cells
    achieved = nbytes / avg_s / 1e9
    full = prof.get("names", {}).get(k)
    name = (FUSED_NAME if k == 0 else full.split("<")[0]) if (fused and full) else KERNEL_NAMES[k]
    traffic = measured_traffic(full, cells) if full else None
    return {"kernel": name, "template": full, "bytes_per_launch": nbytes, "avg_launch_us": round(avg_s * 1e6, 2), "launches": cnt,
            "achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic}


def roofline(res, args, cells, n):
    """Roofline object: the dominant kernel (the fused stencil + forward row
    pass) at the top level, every attempt kernel in `kernels`, and the three
    transform passes together in `fft_passes` (their bytes over the sum of
    their average launch times)."""
    prof = res["prof"]
    if not prof:
        return None
    fused = prof["fused"]
    ks = {k: kernel_roofline(prof, k, cells, n, fused) for k in (0, 2, 3) if k in prof}
    ks = {k: v for k, v in ks.items() if v}
    if not ks:
        return None
    top = ks.get(0) or next(iter(ks.values()))
    out = {"kernel": top["kernel"], "template": top["template"], "bound": "hbm", "achieved": top["achieved"], "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": top["frac"], "traffic": top["traffic"], "avg_launch_us": top["avg_launch_us"],
           "launches": top["launches"], "bytes_per_launch": top["bytes_per_launch"]}
    if len(ks) > 1:
        out["kernels"] = [ks[k] for k in sorted(ks)]
        tot_b = sum(v["bytes_per_launch"] for v in ks.values())
        tot_us = sum(v["avg_launch_us"] for v in ks.values())
        agg = tot_b / (tot_us * 1e-6) / 1e9
        out["fft_passes"] = {"bytes_per_attempt": tot_b, "us_per_attempt": round(tot_us, 2),
                             "achieved": round(agg, 1), "frac": round(agg / HBM_PEAK_GBS, 4)}
    out["copy_GBs"] = round(copy_bandwidth(), 1)
    return out


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--calibrate-copy":
        calibrate_copy(int(sys.argv[2]) if len(sys.argv) > 2 else 256)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--plain-members", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=30)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-workers", type=int, default=16)
    ap.add_argument("--roofline-kernel", type=int, default=PROF_ALL,
                    help="attempt kernel timed with events: 0 fused stencil + row pass, 2 column pass, "
                         "3 inverse row pass, 99 all three (default), -1 none")
    ap.add_argument("--serve-requests", type=int, default=3000)
    ap.add_argument("--serve-domains", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in launch shapes instead of timing candidates first")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--kernels-only", default=None,
                    help="untimed updates of one workload (lv_rheology, sealevel, n384, single256) for profiling runs")
    ap.add_argument("--updates", type=int, default=10)
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the event-timed kernel pass (whole batch per launch), for profiling runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the ensemble "
              "would shard members across GPUs without communication, which is not built here.",
              file=sys.stderr)
        sys.exit(2)
    from fastisostasy_amd import _lib
    if args.kernels_only:
        _lib.load()
        print(json.dumps(kernels_only(args.kernels_only, args.updates)), flush=True)
        return
    if args.roofline_only:
        # every launch of the profiled kernel covers the whole batch, so a
        # profiler's per-kernel statistics describe the launch measured here
        os.environ["FI_SUBBATCHES"] = "1"
        _lib.load()
        if not args.no_autotune:
            tuned = autotune(256, "mixed", args.members)
            print(f"# autotune (N=256, mixed, {args.members} members): {tuned}", file=sys.stderr, flush=True)
        res = run_batched(args.members, args.steps, args.warmup, "mixed",
                          prof_kernel=args.roofline_kernel if args.roofline_kernel >= 0 else PROF_ALL,
                          timed=False)
        n = res["n"]
        print(json.dumps({"roofline": roofline(res, args, args.members * n * n, n),
                          "config": {"workload": f"ensemble {args.members} x 256^2 ice grid (N={n}), mixed",
                                     "steps": args.steps, "warmup": args.warmup}}), flush=True)
        return
    _lib.load()

    tuned = None
    if not args.no_autotune:
        tuned = autotune(256, "mixed", args.members)
        print(f"# autotune (N=256, mixed, {args.members} members): {tuned}", file=sys.stderr, flush=True)
    res = run_batched(args.members, args.steps, args.warmup, "mixed",
                      prof_kernel=args.roofline_kernel if args.roofline_kernel >= 0 else None)
    n = res["n"]
    cells = args.members * n * n
    years = 10.0 * args.steps
    value = cells * years / res["wall"]
    print(f"# batched mixed: {args.members} x {n}^2, {args.steps} steps in {res['wall']:.3f} s, "
          f"{res['attempts']:.0f} attempts total", file=sys.stderr, flush=True)

    max_err = rms_err = fp64_batched = None
    if not args.no_check:
        # the same coupling steps as the mixed run, its timing pass included
        ref = run_batched(args.members, args.steps, args.warmup, "fp64",
                          extra_steps=args.steps if res["prof"] else 0)
        assert ref["t_end"] == res["t_end"], (ref["t_end"], res["t_end"])
        max_err = max(float(np.abs(a - b).max()) for a, b in zip(res["w"], ref["w"]))
        rms_err = float(np.sqrt(np.mean([np.mean((a - b) ** 2) for a, b in zip(res["w"], ref["w"])])))
        print(f"# fp64 batch for the error check: {ref['wall']:.3f} s, max |dw| = {max_err:.3e} m",
              file=sys.stderr, flush=True)
        # the same batch in FP64 throughout (built-in launch shapes), timed the same way
        fp64_batched = {"value": round(cells * years / ref["wall"], 1),
                        "ms_per_step": round(1e3 * ref["wall"] / args.steps, 4)}

    plain = run_plain(args.plain_members, args.plain_steps, 2)
    plain_value = args.plain_members * plain["n"] ** 2 * 10.0 * args.plain_steps / plain["wall"]
    print(f"# plain fp64: {args.plain_members} members x {args.plain_steps} steps in {plain['wall']:.3f} s",
          file=sys.stderr, flush=True)

    roof = roofline(res, args, cells, n)
    roof["step"] = step_roofline(cells, res["attempts"] / args.members / (args.steps + args.warmup),
                                 1e3 * res["wall"] / args.steps)
    single = {"workload": "one 256^2 ensemble member (N=256), 10-yr isos_update",
              "fp64": run_single("fp64", tune=not args.no_autotune),
              "mixed": run_single("mixed", tune=not args.no_autotune)}
    print(f"# single domain: {single}", file=sys.stderr, flush=True)

    serve = None
    if args.serve_requests > 0:
        serve = run_serving(args.serve_requests, args.serve_domains, 1.5)
        print(f"# serving: {serve}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        c1 = run_cpu_baseline(args.cpu_seconds)
        workers = max(1, min(args.cpu_workers, os.cpu_count() or 1))
        cp = run_cpu_parallel(args.cpu_seconds, workers)
        cpu = {"value": round(cp["value"], 1), "unit": "grid-cell-years/s", "cores": workers, "kind": "port",
               "sample": f"oracle (plain-C FP64 restatement of the reference path), {workers} processes on "
                         f"{workers} ensemble members (N={cp['n']}) at once, {cp['steps']} coupling steps of "
                         f"10 yr in {cp['wall']:.1f} s",
               "cpu_model": cpu_model(), "host_threads": os.cpu_count(),
               "single_core": {"value": round(c1["value"], 1), "cores": 1,
                               "sample": f"1 member, {c1['steps']} coupling steps in {c1['wall']:.1f} s"}}

    out = {
        "metric": "grid-cell-years/s batched mixed vs plain FP64 at <=1 mm err; p99 step latency",
        "value": round(value, 1),
        "unit": "grid-cell-years/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * res["wall"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded ensemble members 1000+m)",
        "config": {"workload": f"ensemble {args.members} x 256^2 ice grid, dx 16 km, no pad (N={n}), "
                               "coupling step 10 yr, dt_max 0.2 yr",
                   "members": args.members, "n": n, "precision": "mixed", "launch_shape": tuned},
        "plain_fp64": {"value": round(plain_value, 1), "members": args.plain_members,
                       "steps": args.plain_steps},
        "speedup_vs_plain_fp64": round(value / plain_value, 2),
        "max_err_vs_fp64_m": max_err,
        "fp64_batched": fp64_batched,
        "rms_err_vs_fp64_m": rms_err,
        "step_latency_ms": ({"p50": serve["continuous"]["p50_ms"], "p99": serve["continuous"]["p99_ms"],
                             "of": "serving requests (continuous scheduling, 1.2 x unbatched capacity)"} if serve
                            else None),
        "single_update_ms": single,
        "batch_step_ms": {"p50": round(1e3 * percentile(res["lat"], 50), 4),
                          "p99": round(1e3 * percentile(res["lat"], 99), 4)},
        "serving": serve,
        "attempts_per_member_step": round(res["attempts"] / args.members / (args.steps + args.warmup), 2),
        "accepts_per_member_step": round(res["accepts"] / args.members / (args.steps + args.warmup), 2),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
