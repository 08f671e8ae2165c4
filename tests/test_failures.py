"""Failure detection: check_isos_instabilities and the NaN check of
isos_update (src/fastisostasy.f90:856-899) as status codes.

Every failure here is raised by the host from values the kernels reduce; no
kernel traps or faults, and each update ends normally on the device:
  * NaN in the initial bed: the update inside isos_init_state (no step) meets
    NaN in z_bed -> FI_E_NAN;
  * NaN in the ice thickness: the error norm of the first attempt is NaN.  The
    reference rejects every attempt and halves dt without end; the oracle and
    the library stop the update and report FI_E_NAN instead;
  * a load large enough that |w_equilibrium| > 1e5 m under LLRA, where w
    itself stays below its threshold -> FI_E_UNSTABLE naming w_equilibrium
    (the check the reference makes between we and dz_ss, :882);
  * an elastic response |we| > 1e5 m under LV-ELVA -> FI_E_UNSTABLE naming we;
  * a batch and a served round with one failing domain: the status names the
    member, every member's time advances, the others' results equal their
    single updates, and the server does not apply the failed request twice.
"""
import numpy as np
import pytest

import oracle as O
from fastisostasy_amd import workloads as W

FI_E_UNSTABLE, FI_E_NAN = -4, -5
NX, DX = 48, 50e3       # padded N = 64 (power-of-two kernels)


def _params(**over):
    return W.params(min_pad=400e3, rho_uppermantle=3300.0, **over)


def _oracle(**over):
    o = O.OracleIsos(O.make_params(**_params(**over)), NX, NX, DX)
    b = O.OracleBSL("const", 0.0)
    return o, b


def _expect_raise(fn, text):
    with pytest.raises(RuntimeError) as e:
        fn()
    assert text in str(e.value), str(e.value)
    return e.value


# ---------------------------------------------------------------- oracle (CPU)
def test_oracle_nan_bed_at_init_state():
    o, b = _oracle()
    z = np.full((NX, NX), 1e6)
    z[3, 4] = np.nan
    _expect_raise(lambda: o.init_state(z, np.zeros((NX, NX)), 0.0, b), "NaNs detected")


def test_oracle_nan_ice_stops_the_integrator():
    o, b = _oracle()
    o.init_state(np.full((NX, NX), 1e6), np.zeros((NX, NX)), 0.0, b)
    H = W.disc_load(NX, DX)
    H[5, 7] = np.nan
    _expect_raise(lambda: o.update(H, 10.0, b), "NaNs detected")
    assert o.scalar("attempts") == 1      # stopped at the first NaN error norm
    assert o.scalar("t") == 0.0


def test_oracle_w_equilibrium_threshold():
    o, b = _oracle(method=1)
    o.init_state(np.full((NX, NX), 1e6), np.zeros((NX, NX)), 0.0, b)
    _expect_raise(lambda: o.update(W.disc_load(NX, DX) * 1000.0, 10.0, b), "array w_equilibrium")
    assert np.abs(o.get_full("w")).max() < 1e5 < np.abs(o.get_full("w_equilibrium")).max()


# ---------------------------------------------------------------- GPU
def _gpu(**over):
    from fastisostasy_amd import isostasy as G
    g = G.IsosClass(G.make_params(**_params(**over)), NX, NX, DX)
    gb = G.bsl_init("const", 0.0)
    return g, gb


def _code(fn):
    from fastisostasy_amd._lib import FiError
    with pytest.raises(FiError) as e:
        fn()
    return e.value


@pytest.mark.gpu
def test_gpu_nan_bed_at_init_state():
    print("case: NaN bed at init_state, N = 64", flush=True)
    g, gb = _gpu()
    z = np.full((NX, NX), 1e6)
    z[3, 4] = np.nan
    e = _code(lambda: g.init_state(z, np.zeros((NX, NX)), 0.0, gb))
    assert e.code == FI_E_NAN and "NaNs detected" in str(e)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_nan_ice_stops_the_update(precision):
    print(f"case: NaN ice thickness, {precision}, N = 64", flush=True)
    g, gb = _gpu(precision=precision)
    g.init_state(np.full((NX, NX), 1e6), np.zeros((NX, NX)), 0.0, gb)
    H = W.disc_load(NX, DX)
    H[5, 7] = np.nan
    e = _code(lambda: g.update(H, 10.0, gb))
    assert e.code == FI_E_NAN and "NaNs detected" in str(e)
    assert g.scalar("attempts") == 1 and g.scalar("t") == 0.0
    assert g.scalar("status") == FI_E_NAN
    # the state is the last accepted one: finite; the domain can go on with a valid load
    assert np.isfinite(g.out("w")).all()
    g.update(W.disc_load(NX, DX), 10.0, gb)
    assert g.scalar("status") == 0 and g.scalar("t") == 10.0


@pytest.mark.gpu
def test_gpu_w_equilibrium_threshold_matches_oracle():
    print("case: LLRA huge load, N = 64", flush=True)
    o, ob = _oracle(method=1)
    g, gb = _gpu(method=1)
    z = np.full((NX, NX), 1e6)
    o.init_state(z, np.zeros((NX, NX)), 0.0, ob)
    g.init_state(z, np.zeros((NX, NX)), 0.0, gb)
    H = W.disc_load(NX, DX) * 1000.0
    _expect_raise(lambda: o.update(H, 10.0, ob), "array w_equilibrium")
    e = _code(lambda: g.update(H, 10.0, gb))
    assert e.code == FI_E_UNSTABLE and "array w_equilibrium" in str(e)
    # the update itself ran: same state as the oracle's
    assert np.abs(o.get("w") - g.out("w")).max() < 1e-8
    assert g.scalar("t") == 10.0


@pytest.mark.gpu
def test_gpu_elastic_threshold_names_we():
    print("case: LV-ELVA elastic huge load, N = 64", flush=True)
    over = dict(include_elastic=1)
    o, ob = _oracle(**over)
    g, gb = _gpu(**over)
    z = np.full((NX, NX), 1e6)
    o.init_state(z, np.zeros((NX, NX)), 0.0, ob)
    g.init_state(z, np.zeros((NX, NX)), 0.0, gb)
    H = W.disc_load(NX, DX) * 1e4       # we ~ 3.7e5 m, |w| ~ 7e3 m after 10 yr
    _expect_raise(lambda: o.update(H, 10.0, ob), "array we")
    e = _code(lambda: g.update(H, 10.0, gb))
    assert e.code == FI_E_UNSTABLE and "array we" in str(e)
    assert np.abs(o.get_full("w")).max() < 1e5
    assert np.abs(o.get("we") - g.out("we")).max() < 1e-9 * np.abs(o.get("we")).max()


@pytest.mark.gpu
def test_gpu_batch_reports_the_failing_member():
    """Three LLRA members; member 1 carries a load beyond the w_equilibrium
    threshold.  The batch names member 1, FI_S_STATUS flags only it, and
    members 0 and 2 equal their single updates bit for bit."""
    print("case: batch of 3, member 1 unstable, N = 64", flush=True)
    from fastisostasy_amd import isostasy as G
    z = np.full((NX, NX), 1e6)
    H = [W.disc_load(NX, DX) * s for s in (1.0, 1000.0, 2.0)]
    mem = [_gpu(method=1) for _ in range(3)]
    ref = [_gpu(method=1) for _ in range(3)]
    for g, gb in mem + ref:
        g.init_state(z, np.zeros((NX, NX)), 0.0, gb)
    batch = G.Batch([g for g, _ in mem])
    e = _code(lambda: batch.update(10.0, [gb for _, gb in mem], H_ice=H))
    assert e.code == FI_E_UNSTABLE and "member 1" in str(e) and "w_equilibrium" in str(e)
    assert [g.scalar("status") for g, _ in mem] == [0, FI_E_UNSTABLE, 0]
    assert all(g.scalar("t") == 10.0 for g, _ in mem)
    for k in (0, 2):
        ref[k][0].update(H[k], 10.0, ref[k][1])
        for f in ("w", "z_bed"):
            assert np.array_equal(ref[k][0].out(f), mem[k][0].out(f)), (k, f)
    # the next update of the batch goes on from the advanced state
    batch.update(20.0, [gb for _, gb in mem], H_ice=[H[0], H[0], H[2]])
    assert all(g.scalar("t") == 20.0 for g, _ in mem)
    batch.close()


@pytest.mark.gpu
@pytest.mark.parametrize("max_batch", [1, 8])
def test_gpu_server_fails_a_request_once(max_batch):
    """A served round in which one domain fails its checks: the failed tag is
    reported once, the round's other requests complete, the queue is empty
    afterwards (no retry re-applies an update), and every domain's state
    equals that of single updates.  max_batch 1 runs the groups one after
    another, 8 runs them concurrently (two sizes make two groups)."""
    print(f"case: server round with a failing domain, max_batch {max_batch}", flush=True)
    from fastisostasy_amd import isostasy as G
    from fastisostasy_amd import serving as S

    def dom(nx, scale):
        p = G.make_params(**_params(method=1))
        g = G.IsosClass(p, nx, nx, DX)
        gb = G.bsl_init("const", 0.0)
        g.init_state(np.full((nx, nx), 1e6), np.zeros((nx, nx)), 0.0, gb)
        return g, gb, W.disc_load(nx, DX) * scale

    spec = [(32, 1.0), (32, 1000.0), (48, 1.0)]
    served = [dom(nx, s) for nx, s in spec]
    alone = [dom(nx, s) for nx, s in spec]
    srv = S.Server(hbm_budget_bytes=8e9, max_batch=max_batch)
    slots = [srv.admit(g, gb) for g, gb, _ in served]
    for k, (g, gb, H) in enumerate(served):
        srv.submit(slots[k], 10.0, k, H_ice=H)
    e = _code(lambda: [srv.run() for _ in range(3)])
    assert e.code == FI_E_UNSTABLE
    while srv.pending():
        srv.run()
    assert srv.failed() == [1]
    assert sorted(srv.completed()) == [0, 2]
    assert srv.pending() == 0
    for k, (g, gb, H) in enumerate(alone):
        if k == 1:
            _code(lambda: g.update(H, 10.0, gb))
        else:
            g.update(H, 10.0, gb)
    for (a, _, _), (s, _, _) in zip(alone, served):
        assert a.scalar("t") == s.scalar("t") == 10.0
        assert np.array_equal(a.out("w"), s.out("w"))
    srv.close()


@pytest.mark.gpu
def test_gpu_server_poll_fails_a_request_once():
    """The same failing domain under continuous scheduling (fi_server_poll):
    the poll that retires its group returns FI_E_UNSTABLE, the failed tag is
    reported once, the other requests complete, nothing is retried, and every
    domain's state equals that of single updates."""
    print("case: server poll with a failing domain", flush=True)
    from fastisostasy_amd import isostasy as G
    from fastisostasy_amd import serving as S

    def dom(nx, scale):
        p = G.make_params(**_params(method=1))
        g = G.IsosClass(p, nx, nx, DX)
        gb = G.bsl_init("const", 0.0)
        g.init_state(np.full((nx, nx), 1e6), np.zeros((nx, nx)), 0.0, gb)
        return g, gb, W.disc_load(nx, DX) * scale

    spec = [(32, 1.0), (32, 1000.0), (48, 1.0)]
    served = [dom(nx, s) for nx, s in spec]
    alone = [dom(nx, s) for nx, s in spec]
    srv = S.Server(hbm_budget_bytes=8e9, max_batch=8)
    slots = [srv.admit(g, gb) for g, gb, _ in served]
    for k, (g, gb, H) in enumerate(served):
        srv.submit(slots[k], 10.0, k, H_ice=H)
    codes = []
    while srv.pending():
        try:
            srv.poll(wait=True)
        except Exception as ex:   # the retiring poll reports the failed checks
            codes.append(getattr(ex, "code", None))
    assert codes == [FI_E_UNSTABLE]
    assert srv.failed() == [1]
    assert sorted(srv.completed()) == [0, 2]
    for k, (g, gb, H) in enumerate(alone):
        if k == 1:
            _code(lambda: g.update(H, 10.0, gb))
        else:
            g.update(H, 10.0, gb)
    for (a, _, _), (s, _, _) in zip(alone, served):
        assert a.scalar("t") == s.scalar("t") == 10.0
        assert np.array_equal(a.out("w"), s.out("w"))
    srv.close()
