"""C ABI checks that need no GPU: the library loads, exports every entry point
declared in include/fastiso.h, sizes domains like the reference and refuses
to run without a HIP device (there is no CPU fallback)."""
import ctypes
import os
import re

import pytest

import oracle as O
from fastisostasy_amd import _lib
from fastisostasy_amd import isostasy as G
from fastisostasy_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "fastiso.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fi_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.EXPORTS)


def test_params_struct_layout_matches_header():
    """fi_params field order in the ctypes mirror equals the C declaration."""
    text = open(os.path.join(ROOT, "include", "fastiso.h")).read()
    body = text[text.index("typedef struct {"):text.index("} fi_params;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.replace("typedef struct {", "").strip()
        if not decl:
            continue
        decl = re.sub(r"^(int|double)\s+", "", decl)
        for part in decl.split(","):
            fields.append(re.sub(r"\[.*\]", "", part).strip())
    assert fields == [f[0] for f in _lib.FiParams._fields_]


@pytest.mark.parametrize("nx,ny,dx,pad", [(128, 128, 25e3, 1600e3), (121, 121, 50e3, 0.0),
                                          (256, 256, 16e3, 2048e3), (96, 80, 10e3, 100e3),
                                          (77, 90, 10e3, 35e3)])
def test_padded_size_matches_oracle(nx, ny, dx, pad):
    n = G.padded_size(nx, ny, dx, dx, pad)
    o = O.OracleIsos(O.make_params(**W.params(min_pad=pad)), nx, ny, dx)
    assert n == o.n
    if nx == ny:
        assert n == W.padded_size(nx, dx, pad)


def test_domain_errors_are_reported():
    with pytest.raises(_lib.FiError) as e:
        G.padded_size(64, 64, 10e3, 12e3, 0.0)
    assert e.value.code == _lib.FI_E_ARG and "square" in str(e.value)
    with pytest.raises(_lib.FiError) as e:
        G.padded_size(64, 64, 10e3, 10e3, -1.0)
    assert e.value.code == _lib.FI_E_ARG


def _has_gpu():
    return _lib.load().fi_sync() == 0


def test_compute_fails_loudly_without_gpu():
    if _has_gpu():
        pytest.skip("a HIP device is present")
    with pytest.raises(_lib.FiError) as e:
        G.bsl_init()
    assert e.value.code == _lib.FI_E_HIP and "no HIP device" in str(e.value)
    with pytest.raises(_lib.FiError) as e:
        G.IsosClass(W.params(), 32, 32, 10e3)
    assert e.value.code == _lib.FI_E_HIP


def test_server_fails_loudly_without_gpu():
    import ctypes
    from fastisostasy_amd._lib import load
    if _has_gpu():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = load().fi_server_create(1e9, 8, ctypes.byref(h))
    assert rc == -3 and b"no HIP device" in load().fi_last_error()


def test_series_file_reader(tmp_path):
    """read_series (barysealevel.f90:254-315) through the library's host-only
    entry point: header line skipped, comment and blank lines skipped, comma or
    blank separators."""
    from fastisostasy_amd import isostasy as G
    from fastisostasy_amd._lib import FiError
    f = tmp_path / "sl.dat"
    f.write_text("# time [yr]  bsl [m]\n-20000  -120.5\n# comment\n\n-10000, -60.25\n  0 0\n")
    t, v = G.read_series(f)
    assert t.tolist() == [-20000.0, -10000.0, 0.0] and v.tolist() == [-120.5, -60.25, 0.0]
    with pytest.raises(FiError):
        G.read_series(tmp_path / "missing.dat")


def test_bsl_restart_file_read_by_library(tmp_path):
    """The library's host-only restart reader understands the oracle's file:
    both implement the same container independently."""
    b = O.OracleBSL("fastiso", -3.5, A_ocean_pd=3.6e14)
    path = tmp_path / "bsl_restart.bin"
    b.restart_write(path, 42.0)
    assert G.bsl_restart_read(path) == (-3.5, 3.6e14)
    from fastisostasy_amd._lib import FiError
    with pytest.raises(FiError):
        G.bsl_restart_read(tmp_path / "nothing.bin")
