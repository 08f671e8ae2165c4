"""The measurement's bookkeeping (bench.py, CPU only): algorithmic bytes per
cell of each attempt kernel, the kernel template names derived from the
launched shape (the names rocprofv3 prints, used to look up the measured HBM
traffic), and the roofline arithmetic over event timings."""
import json
import os

import bench


def test_bytes_per_cell():
    # fused stencil + row pass: x (8) + D, 1/(2 eta), canom_full (3 s) read, row spectrum (s) written
    assert bench.kernel_bytes_per_cell(0, 4, True) == 24
    assert bench.kernel_bytes_per_cell(0, 8, True) == 40
    # column pass: spectrum and kappa*R read, spectrum written
    assert bench.kernel_bytes_per_cell(2, 4, True) == 12
    # inverse row pass: spectrum (s) and x (8) read, x2 (8) written
    assert bench.kernel_bytes_per_cell(3, 4, True) == 20
    # the three passes of the mixed attempt: 24 + 8 s
    assert sum(bench.kernel_bytes_per_cell(k, 4, True) for k in (0, 2, 3)) == 56
    assert sum(bench.kernel_bytes_per_cell(k, 8, True) for k in (0, 2, 3)) == 88


def test_kernel_names_from_launch_shape():
    # {N, fp64, fused, fast, threads_row, C, row_tiles, srow_rows, threads_inv, dt_method,
    #  column-pass exchange buffers}
    big = bench.fast_kernel_names([256, 0, 1, 1, 256, 8, 1, 0, 0, 3, 2])
    assert big == {0: "kf_srow_fwd<float, 256, 256, 8>", 2: "kf_col<float, 256, 8, true, true>",
                   3: "kf_row_inv<float, 256, 256, 1, 3>"}
    small = bench.fast_kernel_names([256, 0, 1, 1, 256, 2, 1, 4, 64, 3, 1])
    assert small[0] == "kf_srow_fwd<float, 256, 256, 2>" and small[3] == "kf_row_inv<float, 256, 64, 1, 3>"
    assert small[2] == "kf_col<float, 256, 2, true, false>"
    r3 = bench.fast_kernel_names([384, 1, 1, 1, 384, 2, 1, 8, 192, -1, 1])
    assert r3 == {0: "kf_srow_fwd<double, 384, 384, 4>", 2: "kf_col<double, 384, 2, true, false>",
                  3: "kf_row_inv<double, 384, 192, 1, -1>"}
    assert bench.fast_kernel_names([96, 0, 0, 0, 96, 1, 1, 0, 0, 3, 1]) == {}   # generic kernels: no names


def test_traffic_lookup_uses_launched_template():
    path = os.path.join(os.path.dirname(bench.__file__), "profiles", "hbm_traffic_ensemble256.json")
    with open(path) as f:
        kern = json.load(f)["kernels"]
    cells = 64 * 256 * 256
    for name, k in kern.items():
        assert bench.measured_traffic(name, cells) == k["read_bytes"] + k["write_bytes"]
        # within 3% of the algorithmic bytes
        alg = (k["alg_read_per_cell"] + k["alg_write_per_cell"]) * k["cells"]
        assert abs(k["read_bytes"] + k["write_bytes"] - alg) <= 0.03 * alg, name
    assert bench.measured_traffic("kf_col<float, 256, 8, true, true>", cells // 2) is None   # other size
    assert bench.measured_traffic("kf_col<float, 256, 4, true, false>", cells) is None       # not measured


def test_roofline_arithmetic():
    cells = 64 * 256 * 256
    # 5100 launches of each kernel: 21.7, 12.9 and 16.6 us on average
    prof = {0: (5100 * 21.7e-3, 5100), 2: (5100 * 12.9e-3, 5100), 3: (5100 * 16.6e-3, 5100), "fused": True,
            "names": bench.fast_kernel_names([256, 0, 1, 1, 256, 8, 1, 0, 0, 3, 2])}
    ks = {k: bench.kernel_roofline(prof, k, cells, 256, True) for k in (0, 2, 3)}
    assert ks[0]["bytes_per_launch"] == 24 * cells and ks[2]["bytes_per_launch"] == 12 * cells
    assert abs(ks[0]["achieved"] - 24 * cells / 21.7e-6 / 1e9) < 0.1
    assert abs(ks[0]["frac"] - ks[0]["achieved"] / 8000.0) < 1e-4
    assert ks[2]["template"] == "kf_col<float, 256, 8, true, true>" and ks[2]["traffic"] is not None
    tot = sum(k["bytes_per_launch"] for k in ks.values()) / (sum(k["avg_launch_us"] for k in ks.values()) * 1e-6)
    assert abs(tot / 1e9 / 8000.0 - 0.5734) < 2e-3


def test_step_roofline_arithmetic():
    # 64 x 256^2 cells, 50.5 attempts per step of 2.585 ms: 56 B/cell per attempt
    # and 230 B/cell of once-per-step kernels
    cells = 64 * 256 * 256
    r = bench.step_roofline(cells, 50.5, 2.585)
    assert r["bytes_per_cell"]["attempt"] == 56 and sum(bench.STEP_BYTES_PER_CELL.values()) == 230
    assert r["bytes_per_step"] == int(cells * (230 + 56 * 50.5))
    assert abs(r["achieved"] - cells * (230 + 56 * 50.5) / 2.585e-3 / 1e9) < 0.1
    assert abs(r["frac"] - r["achieved"] / 8000.0) < 1e-4
