"""Distinct handles driven from different host threads (the boundary's
threading contract: calls on distinct handles may come from different
threads, calls on one handle are serialised by the caller): two threads each
advance their own domain through the C ABI at the same time, and every
result equals that of the same updates run one thread after the other."""
import threading

import numpy as np
import pytest

from fastisostasy_amd import isostasy as G
from fastisostasy_amd import workloads as W

pytestmark = pytest.mark.gpu


def _domain(k, nx, precision):
    wl = W.ensemble_member(k, nx=nx, dx=50e3, seed=9100 + k)
    h = G.IsosClass(G.make_params(precision=precision, **wl["params"]), wl["nx"], wl["nx"], wl["dx"],
                    he_field=wl["he_field"], eta_field=wl["eta_field"])
    b = G.bsl_init("const", 0.0)
    h.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
    return h, b, wl


def _run(dom, steps, errors):
    h, b, wl = dom
    try:
        for k in range(1, steps + 1):
            h.update(wl["H_ice"], 10.0 * k, b)
    except Exception as e:   # reported by the main thread
        errors.append(e)


def test_two_threads_two_handles():
    print("case: two host threads, N = 64 (fp64) and N = 128 (mixed)", flush=True)
    spec = [(0, 48, "fp64"), (1, 96, "mixed")]
    threaded = [_domain(*s) for s in spec]
    serial = [_domain(*s) for s in spec]
    steps = 8
    errors = []
    th = [threading.Thread(target=_run, args=(d, steps, errors)) for d in threaded]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in th)
    assert not errors, errors
    for d in serial:
        _run(d, steps, errors)
    assert not errors, errors
    for (a, _, _), (s, _, _) in zip(threaded, serial):
        assert a.scalar("t") == s.scalar("t") == 10.0 * steps
        for f in ("w", "z_bed", "dwdt"):
            assert np.array_equal(a.out(f), s.out(f)), f


def test_two_threads_two_polled_servers():
    """Two request servers, each polled from its own host thread (their
    attempt graphs share the library's cache): every domain ends bitwise where
    the same updates run one thread after the other leave it."""
    from fastisostasy_amd import serving as S
    print("case: two host threads, two polled servers (N = 64 and 128, mixed)", flush=True)
    spec = [(0, 48, "mixed"), (1, 48, "mixed"), (2, 96, "mixed"), (3, 96, "mixed")]
    steps = 6

    def serve(doms, errors):
        try:
            srv = S.Server(hbm_budget_bytes=8e9, max_batch=4)
            slots = [srv.admit(h, b) for h, b, _ in doms]
            for k in range(1, steps + 1):
                for (h, b, wl), sl in zip(doms, slots):
                    srv.submit(sl, 10.0 * k, k, H_ice=wl["H_ice"] if k == 1 else None)
                while srv.pending():
                    srv.poll(wait=True)
            srv.completed()
            srv.close()
        except Exception as e:   # reported by the main thread
            errors.append(e)

    threaded = [_domain(*s) for s in spec]
    serial = [_domain(*s) for s in spec]
    errors = []
    th = [threading.Thread(target=serve, args=(threaded[2 * i:2 * i + 2], errors)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in th)
    assert not errors, errors
    _run(serial[0], steps, errors)
    _run(serial[1], steps, errors)
    _run(serial[2], steps, errors)
    _run(serial[3], steps, errors)
    assert not errors, errors
    for (a, _, _), (s, _, _) in zip(threaded, serial):
        assert a.scalar("t") == s.scalar("t") == 10.0 * steps
        for f in ("w", "z_bed", "dwdt"):
            assert np.array_equal(a.out(f), s.out(f)), f
