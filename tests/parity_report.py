"""Long-horizon parity report: HIP library vs the FP64 oracle on the workloads.

Not a pytest module (long runs).  The two sides run separately so that the
oracle can run on CPU cores while the GPU side runs on the MI355X:

    python tests/parity_report.py oracle --out DIR [--only NAME ...]
    python tests/parity_report.py gpu --out DIR [--precision fp64|mixed]
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


def cases():
    """(name, workload dict, years, snapshot interval, H(t) callable, bsl method)"""
    out = []
    wl = W.plain_disc()
    out.append(("plain_disc", wl, 50e3, 1e3, lambda t, wl=wl: W.step_on_load(wl["H_ice"], t), "const"))
    wl = W.lv_rheology()
    out.append(("lv_rheology", wl, 100.0, 20.0, lambda t, wl=wl: W.step_on_load(wl["H_ice"], t), "const"))
    for m in range(4):
        wl = W.ensemble_member(m)
        out.append((f"ensemble_{m}", wl, 1e3, 200.0, lambda t, wl=wl: W.step_on_load(wl["H_ice"], t), "const"))
    wl = W.sealevel_cycle()
    out.append(("sealevel", wl, 3e3, 500.0,
                lambda t, wl=wl: wl["H_max"] * float(W.sawtooth(t)), "fastiso"))
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
    for name, wl, years, every, H, bm in cases():
        if args.only and name not in args.only:
            continue
        snaps, wall, att = run_case(args.side, name, wl, years, every, H, bm, args.precision)
        np.savez_compressed(os.path.join(args.out, name + ".npz"), **snaps)
        print(f"{args.side} {name}: {years:.0f} yr in {wall:.1f} s, {att:.0f} attempts", flush=True)


if __name__ == "__main__":
    main()
