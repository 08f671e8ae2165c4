"""The output writers isos_write_init / isos_write_step (src/isostasy_io.f90:
197-405) and bsl_write_init / bsl_write_step (src/barysealevel.f90:453-494)
through the C ABI: the file holds the reference's variables, units and long
names on the ice grid (axes in km), the time axis starts at the initial time
with no data, a step at a new time appends a record, a step at a time already
on the axis overwrites it (nc_time_index), and the values are the out%
fields and ode%dt of the handle at the time of writing."""
import numpy as np
import pytest

from fastisostasy_amd import isostasy as G
from fastisostasy_amd import workloads as W
from fastisostasy_amd._lib import FiError

pytestmark = pytest.mark.gpu

VARS = {"H_ice": "Hice", "z_ss": "z_ss", "z_bed": "z_bed", "dwdt": "dwdt", "w_viscous": "w", "w_elastic": "we",
        "z_ss_perturbation": "dz_ss", "column_anomaly": "canom_full", "ocean_mask": "maskocean",
        "grounded_mask": "maskgrounded", "continent_mask": "maskcontinent"}


def test_isos_and_bsl_output_files(tmp_path):
    print("case: output writers, sea-level workload 48^2 (N = 64)", flush=True)
    wl = W.sealevel_cycle(nx=48, dx=50e3, pad=400e3)
    nx = wl["nx"]
    g = G.IsosClass(G.make_params(**wl["params"]), nx, nx, wl["dx"])
    b = G.bsl_init("fastiso", 0.0)
    H = lambda t: wl["H_max"] * min(1.0, t / 100.0)
    g.init_state(wl["z_bed"], H(0.0), 0.0, b)
    xc = W.ice_coords(nx, wl["dx"])
    path, bpath = tmp_path / "isos.out", tmp_path / "bsl.out"
    g.write_init(xc, xc, path, 0.0)
    b.write_init(bpath, 0.0)
    snap = {}
    for k in range(1, 4):
        t = 10.0 * k
        G.bsl_update(b, t)
        g.update(H(t), t, b)
        g.write_step(path, t, H_ice=H(t))
        b.write_step(bpath, t)
        snap[t] = ({v: g.out(f) for v, f in VARS.items()}, g.scalar("dt"), b.bsl_now, b.A_ocean_now)
    # rewrite t = 20 with the state at t = 30: overwritten in place, not appended
    g.write_step(path, 20.0)
    out = G.read_output(path)
    assert np.array_equal(out["time"], [0.0, 10.0, 20.0, 30.0])
    assert np.allclose(out["xc"], xc * 1e-3) and np.allclose(out["yc"], xc * 1e-3)
    assert out["units"]["dwdt"] == "m/yr" and out["units"]["column_anomaly"] == "N m^-2"
    assert out["long_name"]["w_viscous"] == "Displacement (viscous)"
    assert np.isnan(out["z_bed"][0]).all() and np.isnan(out["dt"][0])     # the initial time holds no data
    for i, t in enumerate((10.0, 20.0, 30.0), start=1):
        fields, dt, _, _ = snap[30.0 if t == 20.0 else t]
        for v in VARS:
            assert np.array_equal(out[v][i], fields[v].astype(np.float64)), (t, v)
        assert out["dt"][i] == dt
    bo = G.read_output(bpath)
    assert np.array_equal(bo["time"], [0.0, 10.0, 20.0, 30.0])
    for i, t in enumerate((10.0, 20.0, 30.0), start=1):
        assert bo["bsl"][i] == snap[t][2] and bo["A_ocean"][i] == snap[t][3]
    assert bo["bsl"][3] != 0.0


def test_output_with_benchmark_solution(tmp_path):
    print("case: output writer with z_bed_bench, plain disc 48^2", flush=True)
    nx, dx = 48, 50e3
    p = W.params(min_pad=400e3, rho_uppermantle=3300.0)
    g = G.IsosClass(G.make_params(**p), nx, nx, dx)
    b = G.bsl_init("const", 0.0)
    g.init_state(np.full((nx, nx), 1e6), np.zeros((nx, nx)), 0.0, b)
    xc = W.ice_coords(nx, dx)
    path = tmp_path / "bench.out"
    g.write_init(xc, xc, path, 0.0)
    g.update(W.disc_load(nx, dx), 10.0, b)
    g.write_step(path, 10.0)
    g.update(W.disc_load(nx, dx), 20.0, b)
    bench = np.full((nx, nx), -1.5)
    # the first benchmark solution adds its variables; earlier records hold NaN
    g.write_step(path, 20.0, z_bed_bench=bench)
    out = G.read_output(path)
    assert np.array_equal(out["time"], [0.0, 10.0, 20.0])
    assert np.isnan(out["z_bed_bench"][1]).all() and np.isnan(out["err_z_bed"][:2]).all()
    assert np.array_equal(out["z_bed_bench"][2], bench)
    assert np.array_equal(out["err_z_bed"][2], g.out("w") - bench)
    assert out["units"]["err_z_bed"] == "m"
    assert not np.isnan(out["w_viscous"][1]).any()
    # another grid's file is refused
    other = G.IsosClass(G.make_params(**p), 32, 32, dx)
    with pytest.raises(FiError):
        other.write_step(path, 10.0)
