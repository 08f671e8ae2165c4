"""GPU tests of the continuous-batching server (fi_server_*): results of a
request served in a mixed batch are bitwise those of the same update run
alone, and admission honours the HBM budget."""
import numpy as np
import pytest

from fastisostasy_amd import isostasy as G
from fastisostasy_amd import serving as S
from fastisostasy_amd import workloads as W
from fastisostasy_amd._lib import FiError

pytestmark = pytest.mark.gpu

FI_E_CAPACITY = -7


def _domain(k, nx, prec, **over):
    wl = W.ensemble_member(k, nx=nx, dx=50e3, seed=7000 + k)
    p = G.make_params(precision=prec, **dict(wl["params"], **over))
    h = G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"], eta_field=wl["eta_field"])
    b = G.bsl_init("const", 0.0)
    h.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
    return h, b, wl


def test_server_mix_matches_single_updates():
    """Six domains (N = 64 and 128 on the power-of-two path, N = 96 on the
    generic path; FP64 and mixed) served through one server with random
    interleaving reproduce single-domain isos_update calls bit for bit."""
    spec = [(64, "fp64"), (64, "mixed"), (96, "fp64"), (96, "mixed"), (128, "mixed"), (128, "mixed")]
    alone = [_domain(k, nx, p) for k, (nx, p) in enumerate(spec)]
    served = [_domain(k, nx, p) for k, (nx, p) in enumerate(spec)]
    srv = S.Server(hbm_budget_bytes=8e9, max_batch=64)
    slots = [srv.admit(h, b) for h, b, _ in served]
    rng = np.random.default_rng(3)
    steps = 6
    order = np.concatenate([np.full(steps, k) for k in range(len(spec))])
    rng.shuffle(order)
    t_now = [0.0] * len(spec)
    tag = 0
    for k in order:
        t_now[k] += 10.0
        first = t_now[k] == 10.0
        srv.submit(slots[k], t_now[k], tag, H_ice=served[k][2]["H_ice"] if first else None)
        tag += 1
        if rng.random() < 0.4:
            srv.run()
    while srv.pending():
        srv.run()
    done = srv.completed()
    assert sorted(done) == list(range(tag))
    st = srv.stats()
    assert st["served"] == tag and st["groups"] < tag   # some requests shared launches
    for h, b, wl in alone:
        for k in range(1, steps + 1):
            h.update(wl["H_ice"], 10.0 * k, b)
    for (a, _, _), (s, _, _) in zip(alone, served):
        assert a.scalar("t") == s.scalar("t") == 10.0 * steps
        assert a.scalar("attempts") == s.scalar("attempts")
        for f in ("w", "z_bed", "dwdt"):
            assert np.array_equal(a.out(f), s.out(f)), f
    srv.close()


def test_server_integrators_in_one_round():
    """Two domains of one size and precision that differ only in the
    integrator (tsit54, bs32) are served as separate groups of the same round,
    concurrently, and each reproduces its own single-domain updates bit for
    bit (the launch shape, which compiles the integrator into the inverse row
    pass, is chosen per integrator)."""
    spec = [dict(dt_method="tsit54"), dict(dt_method="bs32")]
    alone = [_domain(k, 64, "mixed", **o) for k, o in enumerate(spec)]
    served = [_domain(k, 64, "mixed", **o) for k, o in enumerate(spec)]
    srv = S.Server(hbm_budget_bytes=8e9, max_batch=64)
    slots = [srv.admit(h, b) for h, b, _ in served]
    tag = 0
    for step in range(1, 5):
        for k in range(len(spec)):
            srv.submit(slots[k], 10.0 * step, tag, H_ice=served[k][2]["H_ice"] if step == 1 else None)
            tag += 1
        while srv.pending():
            srv.run()
    assert sorted(srv.completed()) == list(range(tag))
    for h, b, wl in alone:
        for step in range(1, 5):
            h.update(wl["H_ice"], 10.0 * step, b)
    for (a, _, _), (s, _, _) in zip(alone, served):
        assert a.scalar("attempts") == s.scalar("attempts")
        for f in ("w", "z_bed", "dwdt"):
            assert np.array_equal(a.out(f), s.out(f)), f
    srv.close()


def test_server_shared_bsl_across_groups():
    """Two domains of different sizes (separate launch groups) that share one
    fastiso bsl object are served in one round in request order, on one
    stream, and reproduce consecutive isos_update calls bit for bit, the
    barystatic sea level included."""
    def make(k, nx, b):
        wl = W.ensemble_member(k, nx=nx, dx=50e3, seed=7100 + k)
        p = G.make_params(precision="mixed", **wl["params"])
        h = G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"], eta_field=wl["eta_field"])
        h.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
        return h, wl

    ba, bs = G.bsl_init("fastiso", 0.0), G.bsl_init("fastiso", 0.0)
    alone = [make(0, 64, ba), make(1, 96, ba)]
    served = [make(0, 64, bs), make(1, 96, bs)]
    srv = S.Server(hbm_budget_bytes=8e9, max_batch=64)
    slots = [srv.admit(h, bs) for h, _ in served]
    tag = 0
    for step in range(1, 5):
        for k in range(2):
            srv.submit(slots[k], 10.0 * step, tag, H_ice=served[k][1]["H_ice"])
            tag += 1
        srv.run()
        assert srv.pending() == 0
        for h, wl in alone:
            h.update(wl["H_ice"], 10.0 * step, ba)
    assert sorted(srv.completed()) == list(range(tag))
    assert ba.bsl_now == bs.bsl_now != 0.0
    for (a, _), (s, _) in zip(alone, served):
        for f in ("w", "z_bed"):
            assert np.array_equal(a.out(f), s.out(f)), f
    srv.close()


def test_server_admission_budget():
    doms = [_domain(k, 64, "mixed") for k in range(3)]
    fp = doms[0][0].scalar("hbm_bytes")
    assert fp > 0
    srv = S.Server(hbm_budget_bytes=2.5 * fp, max_batch=8)
    s0 = srv.admit(doms[0][0], doms[0][1])
    s1 = srv.admit(doms[1][0], doms[1][1])
    with pytest.raises(FiError) as e:
        srv.admit(doms[2][0], doms[2][1])
    assert e.value.code == FI_E_CAPACITY
    srv.release(s0)
    s2 = srv.admit(doms[2][0], doms[2][1])
    assert srv.stats()["hbm_used"] == 2 * fp
    srv.submit(s1, 10.0, 0, H_ice=doms[1][2]["H_ice"])
    srv.submit(s2, 10.0, 1, H_ice=doms[2][2]["H_ice"])
    assert srv.run() == 2 and sorted(srv.completed()) == [0, 1]
    srv.close()


@pytest.mark.parametrize("max_batch", [1, 64])
def test_server_poll_matches_single_updates(max_batch):
    """Continuous scheduling (fi_server_poll): the same six-domain mix,
    requests submitted while earlier groups are still running, reproduces
    single-domain isos_update calls bit for bit (groups of one domain on
    concurrent streams with max_batch 1); fi_server_run refuses while polled
    work is in flight."""
    print(f"server poll: 6-domain mix, max_batch {max_batch}", flush=True)
    spec = [(64, "fp64"), (64, "mixed"), (96, "fp64"), (96, "mixed"), (128, "mixed"), (128, "mixed")]
    alone = [_domain(k, nx, p) for k, (nx, p) in enumerate(spec)]
    served = [_domain(k, nx, p) for k, (nx, p) in enumerate(spec)]
    srv = S.Server(hbm_budget_bytes=8e9, max_batch=max_batch)
    slots = [srv.admit(h, b) for h, b, _ in served]
    rng = np.random.default_rng(5)
    steps = 6
    order = np.concatenate([np.full(steps, k) for k in range(len(spec))])
    rng.shuffle(order)
    t_now = [0.0] * len(spec)
    tag = 0
    refused = False
    for k in order:
        t_now[k] += 10.0
        first = t_now[k] == 10.0
        srv.submit(slots[k], t_now[k], tag, H_ice=served[k][2]["H_ice"] if first else None)
        tag += 1
        if rng.random() < 0.5:
            srv.poll()
            if not refused:
                with pytest.raises(FiError):
                    srv.run()   # the group just launched is in flight
                refused = True
    while srv.pending():
        srv.poll(wait=True)
    done = srv.completed()
    assert sorted(done) == list(range(tag))
    assert srv.failed() == []
    st = srv.stats()
    assert st["served"] == tag and st["groups"] <= tag
    for h, b, wl in alone:
        for k in range(1, steps + 1):
            h.update(wl["H_ice"], 10.0 * k, b)
    for (a, _, _), (s, _, _) in zip(alone, served):
        assert a.scalar("t") == s.scalar("t") == 10.0 * steps
        assert a.scalar("attempts") == s.scalar("attempts")
        for f in ("w", "z_bed", "dwdt"):
            assert np.array_equal(a.out(f), s.out(f)), f
    srv.close()


def test_server_poll_shared_bsl_in_request_order():
    """Two domains of different sizes sharing one fastiso bsl object under
    continuous scheduling: the second waits while the first is in flight, so
    the pair reproduces consecutive isos_update calls bit for bit."""
    print("server poll: shared bsl", flush=True)

    def make(k, nx, b):
        wl = W.ensemble_member(k, nx=nx, dx=50e3, seed=7100 + k)
        p = G.make_params(precision="mixed", **wl["params"])
        h = G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"], eta_field=wl["eta_field"])
        h.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
        return h, wl

    ba, bs = G.bsl_init("fastiso", 0.0), G.bsl_init("fastiso", 0.0)
    alone = [make(0, 64, ba), make(1, 96, ba)]
    served = [make(0, 64, bs), make(1, 96, bs)]
    srv = S.Server(hbm_budget_bytes=8e9, max_batch=64)
    slots = [srv.admit(h, bs) for h, _ in served]
    tag = 0
    for step in range(1, 5):
        for k in range(2):
            srv.submit(slots[k], 10.0 * step, tag, H_ice=served[k][1]["H_ice"])
            tag += 1
        for h, wl in alone:
            h.update(wl["H_ice"], 10.0 * step, ba)
    while srv.pending():
        srv.poll(wait=True)
    assert sorted(srv.completed()) == list(range(tag))
    assert ba.bsl_now == bs.bsl_now != 0.0
    for (a, _), (s, _) in zip(alone, served):
        for f in ("w", "z_bed"):
            assert np.array_equal(a.out(f), s.out(f)), f
    srv.close()
