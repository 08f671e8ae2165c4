"""fi_isos_update_noalloc (the allocation-free update) against fi_isos_update:
bitwise equal results, called through the C ABI, on two calls in a row with
the second on a smaller domain (the static scratch keeps the first call's
contents), with and without convolutions, and a clear error beyond its limit."""
import ctypes

import numpy as np
import pytest

from fastisostasy_amd import isostasy as G
from fastisostasy_amd import workloads as W
from fastisostasy_amd._lib import FiError, check, load

pytestmark = pytest.mark.gpu

FI_E_CAPACITY = -7


def _pair(nx, dx, precision, pad, **over):
    p = G.make_params(precision=precision, **W.params(min_pad=pad, rho_uppermantle=3300.0, **over))
    out = []
    for _ in range(2):
        h = G.IsosClass(p, nx, nx, dx)
        b = G.bsl_init("const", 0.0)
        h.init_state(np.full((nx, nx), 1e6), np.zeros((nx, nx)), 0.0, b)
        out.append((h, b))
    return out


def _noalloc(h, H, t, b):
    a = np.ascontiguousarray(H, dtype=np.float64)
    check(load().fi_isos_update_noalloc(h.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), float(t), b.h))


def _same(a, b):
    for f in ("w", "z_bed", "dwdt", "we", "dz_ss", "canom_full"):
        assert np.array_equal(a.out(f), b.out(f)), f
    for s in ("t", "dt", "attempts", "accepts"):
        assert a.scalar(s) == b.scalar(s), s


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_update_equals_update_large_then_small(precision):
    # first a 512^2 domain with the elastic convolution (the largest the scratch
    # holds in FP64), then a 64^2 domain: the second call finds the first's data
    print(f"case: noalloc vs update, {precision}, N = 512 (elastic) then N = 64", flush=True)
    big = _pair(256, 16e3, precision, 2048e3, include_elastic=1, dt_max=0.5)
    small = _pair(48, 50e3, precision, 400e3)
    assert big[0][0].n == 512 and small[0][0].n == 64
    Hb = W.parabolic_dome(256, 16e3, 1200e3, 3000.0)
    Hs = W.disc_load(48, 50e3)
    for k in (1, 2):
        (r, rb), (q, qb) = big
        r.update(Hb, 2.0 * k, rb)
        _noalloc(q, Hb, 2.0 * k, qb)
        _same(r, q)
        (r, rb), (q, qb) = small
        r.update(Hs, 10.0 * k, rb)
        _noalloc(q, Hs, 10.0 * k, qb)
        _same(r, q)


def test_noalloc_update_refuses_beyond_its_limit():
    print("case: noalloc beyond the limit, N = 1024 fp64 with elastic", flush=True)
    (h, b), _ = _pair(512, 16e3, "fp64", 4096e3, include_elastic=1)
    assert h.n == 1024
    with pytest.raises(FiError) as e:
        _noalloc(h, np.zeros((512, 512)), 10.0, b)
    assert e.value.code == FI_E_CAPACITY and "N <= 512" in str(e.value)
    assert h.scalar("t") == 0.0
