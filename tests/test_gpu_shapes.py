"""Launch shapes do not change results: the fused stencil pass with 4 or 8
rows per block (one column per thread), the inverse row pass with its own
block size, and the shapes the autotuner picks give the same bits as the
built-in shapes, since every line's arithmetic and every reduction order is
fixed by N alone."""
import ctypes

import numpy as np
import pytest

from fastisostasy_amd import isostasy as G
from fastisostasy_amd import workloads as W
from fastisostasy_amd._lib import check, load

pytestmark = pytest.mark.gpu

SHAPE_ENV = ("FI_SROW_ROWS", "FI_INV_THREADS", "FI_COL_LINES", "FI_ROW_THREADS")


def _run(nx, prec, steps=4, seed=7):
    wl = W.ensemble_member(0, nx=nx, seed=seed)
    h = G.IsosClass(G.make_params(precision=prec, **wl["params"]), wl["nx"], wl["nx"], wl["dx"],
                    he_field=wl["he_field"], eta_field=wl["eta_field"])
    b = G.bsl_init("const", 0.0)
    h.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
    for k in range(1, steps + 1):
        h.update(wl["H_ice"], 10.0 * k, b)
    out = {f: h.out(f) for f in ("w", "dwdt", "z_bed")}
    out["attempts"] = h.scalar("attempts")
    h.close()
    return out


# the built-in shapes of one domain are the small-batch ones (2-row fused
# blocks, whose march's edge windows reach into the halo rows; inverse blocks
# of 64 threads, 192 at N = 384); each case sets others
@pytest.mark.parametrize("nx,prec,env", [
    (256, "mixed", {"FI_SROW_ROWS": "0", "FI_INV_THREADS": "256", "FI_COL_LINES": "16"}),
    (256, "mixed", {"FI_SROW_ROWS": "8", "FI_INV_THREADS": "128"}),
    (256, "mixed", {"FI_SROW_ROWS": "4"}),
    (128, "fp64", {"FI_SROW_ROWS": "4", "FI_INV_THREADS": "128"}),
    (384, "mixed", {"FI_SROW_ROWS": "4"}),
    (256, "fp64", {"FI_SROW_ROWS": "0", "FI_INV_THREADS": "256", "FI_COL_LINES": "4"}),
    (128, "mixed", {"FI_SROW_ROWS": "0", "FI_INV_THREADS": "128"}),
    (384, "mixed", {"FI_SROW_ROWS": "0", "FI_INV_THREADS": "384", "FI_COL_LINES": "2"}),
    (384, "fp64", {"FI_SROW_ROWS": "8", "FI_INV_THREADS": "384"}),
    (512, "mixed", {"FI_SROW_ROWS": "0", "FI_INV_THREADS": "512"}),
    # column blocks that exchange through two LDS buffers (8 line pairs: 39 KiB
    # at N = 256; 78 KiB at N = 512, declared to the runtime above 64 KiB)
    (256, "mixed", {"FI_COL_LINES": "8"}),
    (512, "mixed", {"FI_COL_LINES": "8"}),
    (512, "fp64", {"FI_SROW_ROWS": "16", "FI_INV_THREADS": "256"}),
    (512, "fp64", {"FI_SROW_ROWS": "0", "FI_INV_THREADS": "256", "FI_ROW_THREADS": "256"}),
])
def test_small_batch_shapes_bitwise(nx, prec, env, monkeypatch):
    print(f"shapes: N={nx} {prec} {env}", flush=True)
    for k in SHAPE_ENV:
        monkeypatch.delenv(k, raising=False)
    ref = _run(nx, prec)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    alt = _run(nx, prec)
    assert alt["attempts"] == ref["attempts"] > 4
    for f in ("w", "dwdt", "z_bed"):
        assert np.array_equal(alt[f], ref[f]), (f, np.abs(alt[f] - ref[f]).max())


def test_autotuned_shape_bitwise(monkeypatch):
    """fi_autotune for one and for two 128^2 domains (mixed): the choice it
    reports is a shape the kernels take, and updates after tuning equal the
    built-in shape's bit for bit."""
    print("autotune: N=128 mixed, 1 and 2 domains", flush=True)
    for k in SHAPE_ENV:
        monkeypatch.delenv(k, raising=False)
    ref = _run(128, "mixed", seed=11)
    for nmem in (1, 2):
        choice = (ctypes.c_int * 8)()
        check(load().fi_autotune(128, 1, nmem, 1, choice))
        c = list(choice)
        assert c[0] in (0, 64, 128, 256) and c[1] in (0, 1, 2, 4, 8, 16) and c[3] in (0, 1, 2, 3)
        assert c[6] in (0, 2, 4, 8, 16) and c[7] in (0, 64, 128, 192, 256)
    tuned = _run(128, "mixed", seed=11)
    assert tuned["attempts"] == ref["attempts"]
    for f in ("w", "dwdt", "z_bed"):
        assert np.array_equal(tuned[f], ref[f]), f
