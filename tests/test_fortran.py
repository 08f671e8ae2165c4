"""The iso_c_binding drop-in module: a Fortran caller shaped like the
reference driver (tests/test_isostasy.f90) reads an &isostasy namelist and
runs through libfastiso.so."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from fastisostasy_amd import build as fb
from fastisostasy_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
NML = os.path.join(HERE, "golden", "disc.nml")
EXE = os.path.join(fb.FORTRAN_DIR, "drive_disc")

needs_flang = pytest.mark.skipif(not os.path.exists(fb.FLANG), reason="amdflang not installed")


@needs_flang
def test_fortran_module_builds_and_stops_without_gpu(tmp_path):
    exe = fb.build_fortran()
    assert os.path.exists(exe)
    from fastisostasy_amd import _lib
    if _lib.load().fi_sync() == 0:
        pytest.skip("a HIP device is present")
    r = subprocess.run([exe, NML, "32", "50e3", "20", str(tmp_path / "o.bin")], capture_output=True, text=True)
    assert r.returncode != 0
    assert "no HIP device" in r.stdout + r.stderr


@pytest.mark.gpu
def test_fortran_driver_matches_oracle(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("drive_disc not built")
    years = 500.0
    out = tmp_path / "w.bin"
    r = subprocess.run([EXE, NML, "128", "25e3", str(years), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    w_f, zb_f = raw[:128 * 128].reshape(128, 128), raw[128 * 128:].reshape(128, 128)
    wl = W.plain_disc()
    o = O.OracleIsos(O.make_params(**wl["params"]), wl["nx"], wl["nx"], wl["dx"])
    b = O.OracleBSL("const", 0.0)
    o.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
    for k in range(1, int(years / 10) + 1):
        t = 10.0 * k
        b.update(t)
        o.update(wl["H_ice"], t, b)
    assert np.abs(o.get("w") - w_f).max() < 1e-6
    assert np.abs(o.get("z_bed") - zb_f).max() < 1e-6


@pytest.mark.gpu
def test_fortran_driver_single_precision_matches_oracle(tmp_path):
    """The module built with wp = sp, the reference's default working
    precision (precision.f90:10): the caller's single-precision arrays are
    widened at the boundary and the results narrowed back, so out%w equals the
    FP64 oracle to single-precision rounding."""
    exe = EXE + "_sp"
    if not os.path.exists(exe):
        pytest.skip("drive_disc_sp not built")
    years = 500.0
    out = tmp_path / "w_sp.bin"
    r = subprocess.run([exe, NML, "128", "25e3", str(years), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, dtype=np.float32)
    assert raw.size == 2 * 128 * 128        # 4-byte reals: the sp build
    w_f, zb_f = raw[:128 * 128].reshape(128, 128), raw[128 * 128:].reshape(128, 128)
    wl = W.plain_disc()
    o = O.OracleIsos(O.make_params(**wl["params"]), wl["nx"], wl["nx"], wl["dx"])
    b = O.OracleBSL("const", 0.0)
    o.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
    for k in range(1, int(years / 10) + 1):
        t = 10.0 * k
        b.update(t)
        o.update(wl["H_ice"], t, b)
    w_o = o.get("w")
    assert np.abs(w_o).max() > 1.0
    # single-precision rounding of |w| <= ~30 m and of z_bed ~ 1e6 m
    assert np.abs(w_o - w_f).max() <= 4e-6 * np.abs(w_o).max()
    assert np.abs(o.get("z_bed") - zb_f).max() <= 0.07


@pytest.mark.gpu
@pytest.mark.parametrize("method", [1, 2])
def test_fortran_driver_relaxation_methods(tmp_path, method):
    """The namelist path of the module with method = 1 (LLRA) and 2 (LV-ELRA),
    tau and rho_uppermantle as in the reference's test1b/test1c parameter sets."""
    if not os.path.exists(EXE):
        pytest.skip("drive_disc not built")
    text = open(NML).read()
    assert "    method = 3\n" in text and "rho_uppermantle = 3300.0" in text
    nml = tmp_path / f"relax{method}.nml"
    nml.write_text(text.replace("    method = 3\n", f"    method = {method}\n", 1)
                   .replace("rho_uppermantle = 3300.0", "rho_uppermantle = 3400.0"))
    years = 300.0
    out = tmp_path / "w.bin"
    r = subprocess.run([EXE, str(nml), "128", "25e3", str(years), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    w_f = raw[:128 * 128].reshape(128, 128)
    wl = W.relax_disc(method)
    o = O.OracleIsos(O.make_params(**wl["params"]), wl["nx"], wl["nx"], wl["dx"])
    b = O.OracleBSL("const", 0.0)
    o.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
    for k in range(1, int(years / 10) + 1):
        t = 10.0 * k
        b.update(t)
        o.update(wl["H_ice"], t, b)
    assert np.abs(o.get("w")).max() > 1.0
    assert np.abs(o.get("w") - w_f).max() < 1e-6
