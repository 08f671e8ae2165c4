"""Long-horizon parity report: HIP library vs the FP64 oracle on the workloads.

Not a pytest module (long runs).  The two sides run separately so that the
oracle can run on CPU cores while the GPU side runs on the MI355X:

    python tests/parity_report.py oracle --out DIR [--only NAME ...] [--full] [--jobs J]
    python tests/parity_report.py gpu --out DIR [--precision fp64|mixed] [--full]
    python tests/parity_report.py compare ORACLE_DIR GPU_DIR

Each side writes one .npz per workload with out%w, out%z_bed and the three
masks of the ice grid every `dt_out` years.  `compare` prints max and RMS
|dw|, max |dz_bed| and the number of differing mask cells per workload.
"""
import argparse
import glob
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from fastisostasy_amd import workloads as W  # noqa: E402

MASKS = ("maskocean", "maskgrounded", "maskcontinent")


def cases(full=False, lv_years=100.0):
    """(name, workload dict, years, snapshot interval, H(t) callable, bsl method).

    full: the workloads' whole horizons (ensemble 10 kyr, sea-level cycle
    120 kyr) instead of the short default ones.  lv_years: horizon of the
    Gaussian-rheology workload (its dt_max of 0.02 yr makes the FP64 CPU
    oracle slow: about 12 min per 100 yr)."""
    out = []
    wl = W.plain_disc()
    out.append(("plain_disc", wl, 50e3, 1e3, lambda t, wl=wl: W.step_on_load(wl["H_ice"], t), "const"))
    wl = W.lv_rheology()
    out.append(("lv_rheology", wl, lv_years, max(20.0, lv_years / 10), lambda t, wl=wl: W.step_on_load(wl["H_ice"], t),
                "const"))
    for m in range(4):
        wl = W.ensemble_member(m)
        out.append((f"ensemble_{m}", wl, 10e3 if full else 1e3, 1e3 if full else 200.0,
                    lambda t, wl=wl: W.step_on_load(wl["H_ice"], t), "const"))
    wl = W.sealevel_cycle()
    out.append(("sealevel", wl, 120e3 if full else 3e3, 5e3 if full else 500.0,
                lambda t, wl=wl: wl["H_max"] * float(W.sawtooth(t)), "fastiso"))
    # the reference's test1b / test1c setups: LLRA and LV-ELRA under the disc
    for method, name in ((1, "llra_disc"), (2, "elra_disc")):
        wl = W.relax_disc(method)
        out.append((name, wl, 50e3 if full else 2e3, 1e3 if full else 200.0,
                    lambda t, wl=wl: W.step_on_load(wl["H_ice"], t), "const"))
    return out


def run_case(side, name, wl, years, every, H, bsl_method, precision="fp64"):
    if side == "oracle":
        import oracle as O
        iso = O.OracleIsos(O.make_params(**wl["params"]), wl["nx"], wl["nx"], wl["dx"],
                           he_field=wl.get("he_field"), eta_field=wl.get("eta_field"))
        bsl = O.OracleBSL(bsl_method, 0.0)
        get = iso.get
        upd_bsl = bsl.update
    else:
        from fastisostasy_amd import isostasy as G
        iso = G.IsosClass(G.make_params(precision=precision, **wl["params"]), wl["nx"], wl["nx"], wl["dx"],
                          he_field=wl.get("he_field"), eta_field=wl.get("eta_field"))
        bsl = G.bsl_init(bsl_method, 0.0)
        get = iso.out
        upd_bsl = lambda t: G.bsl_update(bsl, t)  # noqa: E731
    iso.init_state(wl["z_bed"], H(0.0), 0.0, bsl)
    snaps = {"t": [], "w": [], "z_bed": []}
    for m in MASKS:
        snaps[m] = []
    t0 = time.time()
    nsteps = int(round(years / 10.0))
    for k in range(1, nsteps + 1):
        t = 10.0 * k
        upd_bsl(t)
        iso.update(H(t), t, bsl)
        if abs(t / every - round(t / every)) < 1e-9 or k == nsteps:
            snaps["t"].append(t)
            snaps["w"].append(get("w"))
            snaps["z_bed"].append(get("z_bed"))
            for m in MASKS:
                snaps[m].append(np.asarray(get(m)).astype(np.uint8))
    wall = time.time() - t0
    return {k: np.asarray(v) for k, v in snaps.items()}, wall, iso.scalar("attempts")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("side", choices=["oracle", "gpu", "compare"])
    ap.add_argument("dirs", nargs="*")
    ap.add_argument("--out")
    ap.add_argument("--only", nargs="*")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--full", action="store_true", help="whole horizons (ensemble 10 kyr, sea level 120 kyr)")
    ap.add_argument("--lv-years", type=float, default=100.0)
    ap.add_argument("--jobs", type=int, default=1, help="oracle side: workloads in parallel processes")
    args = ap.parse_args()
    if args.side == "compare":
        a, b = args.dirs
        print("| workload | snapshots | max abs dw (m) | rms dw (m) | max abs dz_bed (m) | mask cells differing |")
        print("|---|---|---|---|---|---|")
        for fa in sorted(glob.glob(os.path.join(a, "*.npz"))):
            name = os.path.basename(fa)[:-4]
            fb = os.path.join(b, name + ".npz")
            if not os.path.exists(fb):
                continue
            A, B = np.load(fa), np.load(fb)
            dw = np.abs(A["w"] - B["w"])
            dz = np.abs(A["z_bed"] - B["z_bed"])
            dm = sum(int((A[m] != B[m]).sum()) for m in MASKS)
            print(f"| {name} | {len(A['t'])} (to {A['t'][-1]:.0f} yr) | {dw.max():.3e} | "
                  f"{np.sqrt((dw ** 2).mean()):.3e} | {dz.max():.3e} | {dm} |")
        return
    os.makedirs(args.out, exist_ok=True)
    todo = [c for c in cases(args.full, args.lv_years) if not args.only or c[0] in args.only]
    if args.side == "oracle" and args.jobs > 1:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(args.jobs) as pool:
            for line in pool.imap_unordered(_one, [(args, c[0]) for c in todo]):
                print(line, flush=True)
        return
    for c in todo:
        print(_one((args, c[0])), flush=True)


def _one(job):
    args, want = job
    for name, wl, years, every, H, bm in cases(args.full, args.lv_years):
        if name != want:
            continue
        snaps, wall, att = run_case(args.side, name, wl, years, every, H, bm, args.precision)
        np.savez_compressed(os.path.join(args.out, name + ".npz"), **snaps)
        return f"{args.side} {name}: {years:.0f} yr in {wall:.1f} s, {att:.0f} attempts"


if __name__ == "__main__":
    main()
