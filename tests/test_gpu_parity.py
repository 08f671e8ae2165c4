"""GPU parity: the HIP library (through its C ABI) against the FP64 oracle.

Tolerances: FP64 path -- displacement agrees to 1e-6 m over the short runs
here (only FFT rounding differs); mixed path -- max |dw| <= 1 mm.  Masks
must be bit-exact in both modes.
"""
import numpy as np
import pytest

import oracle as O
from fastisostasy_amd import isostasy as G
from fastisostasy_amd import workloads as W

pytestmark = pytest.mark.gpu

OUT = ["z_bed", "z_ss", "dz_ss", "w", "we", "dwdt", "canom_full", "Hice"]
MASKS = ["maskocean", "maskgrounded", "maskcontinent"]


def run_pair(wl, times, precision="fp64", bsl_method="const", H_of_t=None, bsl_kw=None):
    """Run the workload through oracle and GPU on the same inputs; yield after each update."""
    p = dict(wl["params"])
    op = O.make_params(**p)
    gp = G.make_params(precision=precision, **p)
    kw = dict(he_field=wl.get("he_field"), eta_field=wl.get("eta_field"))
    o = O.OracleIsos(op, wl["nx"], wl["nx"], wl["dx"], **kw)
    g = G.IsosClass(gp, wl["nx"], wl["nx"], wl["dx"], **kw)
    bkw = {"bsl_init": 0.0}
    bkw.update(bsl_kw or {})
    ob = O.OracleBSL(bsl_method, year_bp=times[0], **bkw)
    gb = G.bsl_init(bsl_method, year_bp=times[0], **bkw)
    H = (lambda t: W.step_on_load(wl["H_ice"], t)) if H_of_t is None else H_of_t
    o.init_state(wl["z_bed"], H(times[0]), times[0], ob)
    g.init_state(wl["z_bed"], H(times[0]), times[0], gb)
    yield o, g
    for t in times[1:]:
        ob.update(t)
        G.bsl_update(gb, t)
        o.update(H(t), t, ob)
        g.update(H(t), t, gb)
        yield o, g


def compare(o, g, wtol, rel=None):
    for f in OUT:
        a, b = o.get(f), g.out(f)
        scale = max(1.0, np.abs(a).max())
        tol = wtol if f in ("w", "z_bed") else (rel or 1e-9) * scale
        assert np.abs(a - b).max() <= tol, (f, np.abs(a - b).max(), tol)
    for f in MASKS:
        assert np.array_equal(o.get(f).astype(bool), g.out(f)), f


def small_disc(n_ice=48, dx=50e3, pad=400e3):
    wl = W.plain_disc()
    wl["nx"], wl["dx"] = n_ice, dx
    wl["params"] = dict(wl["params"], min_pad=pad)
    wl["H_ice"] = W.disc_load(n_ice, dx)
    wl["z_bed"] = np.full((n_ice, n_ice), 1e6)
    return wl


def test_init_fields_match_oracle():
    wl = W.lv_rheology()
    wl["params"] = dict(wl["params"], include_elastic=0)
    p = wl["params"]
    o = O.OracleIsos(O.make_params(**p), wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"],
                     eta_field=wl["eta_field"])
    g = G.IsosClass(G.make_params(**p), wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"],
                    eta_field=wl["eta_field"])
    assert g.n == o.n == 512
    for f in ("kappa", "R", "eta_eff", "D_lith", "He_lith"):
        a, b = o.get_full(f), g.full(f)
        assert np.allclose(a, b, rtol=1e-13, atol=0), (f, np.abs(a - b).max())


def test_plain_disc_fp64_short():
    wl = small_disc()
    times = [0.0] + [10.0 * k for k in range(1, 31)]
    for o, g in run_pair(wl, times):
        pass
    compare(o, g, wtol=1e-6)
    assert o.scalar("attempts") == g.scalar("attempts")
    assert abs(o.scalar("dt") - g.scalar("dt")) < 1e-9


def test_plain_disc_256_fp64():
    wl = W.plain_disc()
    times = [0.0] + [10.0 * k for k in range(1, 21)]
    for o, g in run_pair(wl, times):
        pass
    compare(o, g, wtol=1e-6)


def test_variable_rheology_fp64():
    wl = W.lv_rheology()
    wl["params"] = dict(wl["params"], include_elastic=0)
    times = [0.0, 0.5, 1.0]
    for o, g in run_pair(wl, times):
        pass
    compare(o, g, wtol=1e-6)


def test_elastic_convolution_fp64():
    wl = W.lv_rheology()
    times = [0.0, 0.5, 1.0]
    for o, g in run_pair(wl, times):
        pass
    assert np.abs(o.get("we")).max() > 0.1
    compare(o, g, wtol=1e-6)


def test_plain_disc_mixed_1mm():
    wl = small_disc()
    times = [0.0] + [10.0 * k for k in range(1, 31)]
    for o, g in run_pair(wl, times, precision="mixed"):
        pass
    compare(o, g, wtol=1e-3, rel=1e-4)


def test_batch_matches_single():
    members = [W.ensemble_member(m, nx=64, dx=50e3, pad=0.0) for m in range(3)]
    singles, gs, bsls = [], [], []
    for wl in members:
        p = G.make_params(**wl["params"])
        a = G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"], eta_field=wl["eta_field"])
        b = G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"], eta_field=wl["eta_field"])
        ba, bb = G.bsl_init(), G.bsl_init()
        a.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, ba)
        b.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, bb)
        singles.append((a, ba))
        gs.append(b)
        bsls.append(bb)
    batch = G.Batch(gs)
    for k in range(1, 11):
        t = 10.0 * k
        for (a, ba), wl in zip(singles, members):
            a.update(wl["H_ice"], t, ba)
        batch.update(t, bsls, H_ice=[wl["H_ice"] for wl in members])
    for (a, _), b in zip(singles, gs):
        for f in ("w", "z_bed", "dwdt"):
            assert np.array_equal(a.out(f), b.out(f)), f


def small_sealevel():
    wl = W.sealevel_cycle(nx=48, dx=50e3, pad=400e3)
    return wl


SERIES = (np.array([0.0, 100.0, 250.0]), np.array([0.0, -5.0, -2.0]))


@pytest.mark.parametrize("bsl_method", ["file", "mixed"])
def test_sealevel_series_bsl_fp64(bsl_method):
    """BSL from a sea-level series (file), and the series minus each domain's
    cumulative contribution (mixed), against the oracle: bsl bit-for-bit
    after every update, fields as in the fastiso case."""
    wl = small_sealevel()
    Hmax = wl["H_max"]
    times = [0.0] + [10.0 * k for k in range(1, 21)]
    H_of_t = lambda t: Hmax * min(1.0, t / 150.0)
    for o, g in run_pair(wl, times, bsl_method=bsl_method, H_of_t=H_of_t, bsl_kw=dict(bsl_init=1.0, series=SERIES)):
        assert abs(o.scalar("bsl") - g.scalar("bsl")) <= 1e-12 * max(1.0, abs(o.scalar("bsl"))), \
            (o.scalar("bsl"), g.scalar("bsl"))
    assert o.scalar("bsl") != 0.0
    compare(o, g, wtol=1e-6)


def test_sealevel_masks_bsl_fp64():
    wl = small_sealevel()
    Hmax = wl["H_max"]
    times = [0.0] + [10.0 * k for k in range(1, 21)]
    H_of_t = lambda t: Hmax * min(1.0, t / 150.0)
    for o, g in run_pair(wl, times, bsl_method="fastiso", H_of_t=H_of_t):
        pass
    assert o.get("maskocean").any() and o.get("maskcontinent").any() and o.get("maskgrounded").any()
    assert np.abs(o.get("dz_ss")).max() > 1e-3
    compare(o, g, wtol=1e-6)
    assert abs(o.scalar("bsl") - g.scalar("bsl")) < 1e-9



def _run_gpu(wl, precision, times, fused, threads, monkeypatch):
    from fastisostasy_amd._lib import load
    monkeypatch.setenv("FI_ROW_THREADS", str(threads))
    monkeypatch.setenv("FI_SROW_ROWS", "0")   # the fused pass in blocks of FI_ROW_THREADS threads
    if fused:
        monkeypatch.delenv("FI_NO_FUSED_STENCIL", raising=False)
    else:
        monkeypatch.setenv("FI_NO_FUSED_STENCIL", "1")
    g = G.IsosClass(G.make_params(precision=precision, **wl["params"]), wl["nx"], wl["nx"], wl["dx"],
                    he_field=wl.get("he_field"), eta_field=wl.get("eta_field"))
    b = G.bsl_init("const", 0.0)
    g.init_state(wl["z_bed"], W.step_on_load(wl["H_ice"], times[0]), times[0], b)
    for t in times[1:]:
        G.bsl_update(b, t)
        g.update(W.step_on_load(wl["H_ice"], t), t, b)
    out = {f: g.full(f) for f in ("w", "dwdt")}
    out["attempts"] = g.scalar("attempts")
    out["fused"] = bool(load().fi_prof_fused())
    g.close()
    return out


@pytest.mark.parametrize("case,prec,threads", [
    ("disc64", "fp64", 64),
    ("ens256", "fp64", 64), ("ens256", "fp64", 128), ("ens256", "fp64", 256),
    ("ens256", "mixed", 64), ("ens256", "mixed", 128), ("ens256", "mixed", 256),
    ("lv512", "mixed", 128), ("lv512", "mixed", 256), ("lv512", "mixed", 512)])
def test_fused_stencil_matches_separate_kernels(case, prec, threads, monkeypatch):
    """The stencil fused into the forward row pass (uniform spacing,
    power-of-two N) reproduces the separate stencil and row kernels bit for
    bit -- same arithmetic, same rounding -- on every row block, the top and
    bottom edge blocks included, for 1, 2 and 4 columns per thread (512-thread
    blocks at N = 512: one column per thread, 68 KiB of LDS)."""
    if case == "disc64":
        wl, times = small_disc(), [0.0, 10.0, 20.0, 30.0]
    elif case == "ens256":
        wl, times = W.ensemble_member(3), [0.0, 10.0, 20.0]
    else:
        wl, times = W.lv_rheology(), [0.0, 0.5, 1.0]
    print(f"fused vs separate: {case} {prec} threads={threads}", flush=True)
    a = _run_gpu(wl, prec, times, True, threads, monkeypatch)
    b = _run_gpu(wl, prec, times, False, threads, monkeypatch)
    assert a["fused"] and not b["fused"]
    assert a["attempts"] == b["attempts"]
    for f in ("w", "dwdt"):
        assert np.array_equal(a[f], b[f]), (case, f, np.abs(a[f] - b[f]).max())
    assert np.abs(a["w"]).max() > 0


@pytest.mark.parametrize("method", ["fastiso", "mixed"])
def test_batch_shared_bsl_matches_consecutive_updates(method):
    """Domains sharing one barystatic sea-level object: a batch update applies
    their volume changes to the shared bsl in member order, exactly as
    consecutive isos_update calls do (the last two members share one object,
    the first has its own)."""
    kw = dict(series=SERIES) if method == "mixed" else {}
    wl = small_sealevel()
    Hmax = wl["H_max"]
    scales = (1.0, 0.6, 0.8)
    p = G.make_params(**wl["params"])
    singles, gs = [], []
    for _ in scales:
        singles.append(G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"]))
        gs.append(G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"]))
    sb = [G.bsl_init(method, 0.0, **kw), G.bsl_init(method, 0.0, **kw)]
    bb = [G.bsl_init(method, 0.0, **kw), G.bsl_init(method, 0.0, **kw)]
    s_of = [sb[0], sb[1], sb[1]]
    b_of = [bb[0], bb[1], bb[1]]
    for a, b, sa, ba in zip(singles, gs, s_of, b_of):
        a.init_state(wl["z_bed"], np.zeros_like(wl["z_bed"]), 0.0, sa)
        b.init_state(wl["z_bed"], np.zeros_like(wl["z_bed"]), 0.0, ba)
    batch = G.Batch(gs)
    for k in range(1, 16):
        t = 10.0 * k
        Hs = [s * Hmax * min(1.0, t / 100.0) * np.ones_like(wl["z_bed"]) for s in scales]
        for o in sb + bb:
            G.bsl_update(o, t)
        for a, sa, H in zip(singles, s_of, Hs):
            a.update(H, t, sa)
        batch.update(t, b_of, H_ice=Hs)
    for a, b in zip(singles, gs):
        for f in ("w", "z_bed", "dz_ss", "z_ss"):
            assert np.array_equal(a.out(f), b.out(f)), f
        assert a.scalar("bsl") == b.scalar("bsl")
    assert sb[1].bsl_now == bb[1].bsl_now != 0.0
    assert sb[0].bsl_now == bb[0].bsl_now


@pytest.mark.parametrize("prec,nx,generic", [("fp64", 256, False), ("mixed", 256, False), ("fp64", 384, False),
                                             ("mixed", 384, False), ("fp64", 384, True), ("mixed", 384, True)])
def test_results_identical_alone_and_in_batch(prec, nx, generic, monkeypatch):
    """A domain stepped alone and the same domain stepped inside a batch of
    other domains give the same bits, whatever block shapes the two launches
    use: the corner and error-norm reductions follow an order fixed by N, on
    the power-of-two register kernels (N = 256), on the register kernels of
    N = 3 x 2^k (N = 384: one and 16 line pairs per column block) and on the
    mixed-radix LDS kernels (FI_GENERIC_FFT; one line pair per block alone,
    8 and 16 in the batch)."""
    target = 5
    others = [11, 12, 13]
    print(f"alone vs batch: {prec} N={nx} generic={generic}", flush=True)
    if generic:
        monkeypatch.setenv("FI_GENERIC_FFT", "1")

    def make(m):
        wl = W.ensemble_member(m, nx=nx)
        p = G.make_params(precision=prec, **wl["params"])
        h = G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"], eta_field=wl["eta_field"])
        b = G.bsl_init("const", 0.0)
        h.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
        return h, b, wl

    monkeypatch.setenv("FI_ROW_THREADS", "64")
    monkeypatch.setenv("FI_COL_LINES", "1")
    monkeypatch.setenv("FI_GEN_LINES", "1,1")
    alone, ab, awl = make(target)
    monkeypatch.setenv("FI_ROW_THREADS", "256")
    monkeypatch.setenv("FI_COL_LINES", "16")
    monkeypatch.setenv("FI_GEN_LINES", "8,16")
    mix = [make(m) for m in others[:2]] + [make(target)] + [make(others[2])]
    batch = G.Batch([h for h, _, _ in mix])
    for k in range(1, 6):
        t = 10.0 * k
        alone.update(awl["H_ice"], t, ab)
        batch.update(t, [b for _, b, _ in mix], H_ice=[wl["H_ice"] for _, _, wl in mix])
    h = mix[2][0]
    assert alone.scalar("attempts") == h.scalar("attempts") > 5
    for f in ("w", "dwdt", "z_bed"):
        assert np.array_equal(alone.out(f), h.out(f)), (f, np.abs(alone.out(f) - h.out(f)).max())


def test_sub_batches_match_one_launch(monkeypatch):
    """A batch split into concurrent sub-batches on several streams, and one
    launch with two row tiles per block in the inverse row pass, give the same
    bits as one plain launch over the whole batch."""
    members = [W.ensemble_member(m, nx=64, dx=50e3) for m in range(32)]

    def run(k, tiles=1):
        monkeypatch.setenv("FI_SUBBATCHES", str(k))
        monkeypatch.setenv("FI_ROW_TILES", str(tiles))
        hs, bs = [], []
        for wl in members:
            p = G.make_params(precision="mixed", **wl["params"])
            h = G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"], eta_field=wl["eta_field"])
            b = G.bsl_init()
            h.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
            hs.append(h)
            bs.append(b)
        batch = G.Batch(hs)
        for k_ in range(1, 4):
            batch.update(10.0 * k_, bs, H_ice=[wl["H_ice"] for wl in members] if k_ == 1 else None)
        out = [h.out("w") for h in hs]
        batch.close()
        for h in hs:
            h.close()
        return out

    one, four, tiled = run(1), run(4), run(1, tiles=2)
    for a, b in zip(one, tiled):
        assert np.array_equal(a, b)
    for a, b in zip(one, four):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_attempt_graphs_match_plain_launches(prec, monkeypatch):
    """Attempt blocks replayed from captured HIP graphs (kernel nodes only)
    give the same bits as plain launches, replay after replay, with the
    library's own copies and bookkeeping kernels between the replays."""
    members = [W.ensemble_member(m) for m in (21, 22, 23)]

    def run(graphs):
        if graphs:
            monkeypatch.setenv("FI_GRAPHS", "1")
        else:
            monkeypatch.delenv("FI_GRAPHS", raising=False)
        hs, bs = [], []
        for wl in members:
            p = G.make_params(precision=prec, **wl["params"])
            h = G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"], he_field=wl["he_field"], eta_field=wl["eta_field"])
            b = G.bsl_init()
            h.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, b)
            hs.append(h)
            bs.append(b)
        batch = G.Batch(hs)
        single = hs[0]
        for k in range(1, 6):
            batch.update(10.0 * k, bs, H_ice=[wl["H_ice"] for wl in members])
        out = [(h.out("w"), h.out("dwdt"), h.scalar("attempts")) for h in hs]
        batch.close()
        # the single-domain path too (one more update of member 0 alone)
        single.update(members[0]["H_ice"], 60.0, bs[0])
        out.append((single.out("w"), single.out("dwdt"), single.scalar("attempts")))
        for h in hs:
            h.close()
        return out

    a, b = run(True), run(False)
    assert a[0][2] > 3 * 8   # several graph replays per update
    for (wa, da, na), (wb, db, nb) in zip(a, b):
        assert na == nb
        assert np.array_equal(wa, wb) and np.array_equal(da, db)


# ---------------------------------------------------------------- methods 0-2
@pytest.mark.parametrize("method,prec", [(0, "fp64"), (1, "fp64"), (2, "fp64"), (1, "mixed"), (2, "mixed")])
def test_relaxation_methods_match_oracle(method, prec):
    """No displacement, LLRA and LV-ELRA (the reference's test1a-c setups) over
    400 yr of 10-yr coupling steps: displacement within 1e-8 m (FP64) or 1 mm
    (mixed: the viscous convolution runs in FP32), w_equilibrium alike, masks
    bit-exact."""
    wl = W.relax_disc(method)
    times = [0.0] + [10.0 * k for k in range(1, 41)]
    for o, g in run_pair(wl, times, precision=prec):
        pass
    print(f"method {method} {prec}: attempts {g.scalar('attempts'):.0f} / {o.scalar('attempts'):.0f}", flush=True)
    fp64 = prec == "fp64"
    compare(o, g, wtol=1e-8 if fp64 else 1e-3, rel=None if fp64 else 1e-5)
    a, b = o.get_full("w_equilibrium"), g.full("w_equilibrium")
    assert np.abs(a - b).max() <= (1e-8 if fp64 else 1e-3) * max(1.0, np.abs(a).max() / 100), np.abs(a - b).max()
    if method == 0:
        assert np.all(g.out("w") == 0) and np.all(b == 0)
    else:
        assert g.out("w").min() < -10.0   # the domain is relaxing under the load
    if method == 1:
        assert np.array_equal(a, b)       # pointwise equilibrium: same bits


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_relaxation_batch_matches_single(prec):
    """LV-ELRA domains with different loads: each member of a batch gives the
    same bits as the same domain updated alone."""
    wls = []
    for k, scale in enumerate((1.0, 0.6, 1.4)):
        wl = W.relax_disc(2)
        wl["H_ice"] = wl["H_ice"] * scale
        wls.append(wl)
    singles, gs, bsls = [], [], []
    for wl in wls:
        p = G.make_params(precision=prec, **wl["params"])
        a = G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"])
        b = G.IsosClass(p, wl["nx"], wl["nx"], wl["dx"])
        ba, bb = G.bsl_init(), G.bsl_init()
        a.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, ba)
        b.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, bb)
        singles.append((a, ba))
        gs.append(b)
        bsls.append(bb)
    batch = G.Batch(gs)
    for k in range(1, 11):
        t = 10.0 * k
        for (a, ba), wl in zip(singles, wls):
            a.update(wl["H_ice"], t, ba)
        batch.update(t, bsls, H_ice=[wl["H_ice"] for wl in wls])
    for (a, _), b in zip(singles, gs):
        for f in ("w", "z_bed", "dwdt"):
            assert np.array_equal(a.out(f), b.out(f)), f
        assert np.array_equal(a.full("w_equilibrium"), b.full("w_equilibrium"))


# ---------------------------------------------------------------- restart files
def test_restart_matches_oracle(tmp_path):
    """isos_restart_write / isos_restart_read through the library against the
    oracle's own implementation of the container: the files hold the same
    records (names, units; values within the FP64 difference of the two runs),
    the runs continued from them agree, and the library reads the oracle's
    file to the same result."""
    wl = small_sealevel()
    Hmax = wl["H_max"]
    H = lambda t: Hmax * min(1.0, t / 150.0)
    times = [0.0] + [10.0 * k for k in range(1, 11)]
    for o, g in run_pair(wl, times, bsl_method="fastiso", H_of_t=H):
        pass
    po, pg = tmp_path / "oracle.bin", tmp_path / "gpu.bin"
    o.restart_write(po, 100.0)
    g.restart_write(pg, 100.0)
    ro, to = O.read_restart_file(po)
    rg, tg = O.read_restart_file(pg)
    assert to == tg == 100.0 and list(ro) == list(rg)
    for k in ro:
        assert ro[k][0] == rg[k][0], k
        a, b = ro[k][1], rg[k][1]
        assert np.abs(a - b).max() <= 1e-9 * max(1.0, np.abs(a).max()), (k, np.abs(a - b).max())
    p = wl["params"]
    bsl0 = o.scalar("bsl")
    o2 = O.OracleIsos(O.make_params(**p), wl["nx"], wl["nx"], wl["dx"])
    g2 = G.IsosClass(G.make_params(**p), wl["nx"], wl["nx"], wl["dx"])
    g3 = G.IsosClass(G.make_params(**p), wl["nx"], wl["nx"], wl["dx"])
    o2.restart_read(po, 100.0)
    g2.restart_read(pg, 100.0)
    g3.restart_read(po, 100.0)
    ob = O.OracleBSL("fastiso", bsl0)
    gb2, gb3 = G.bsl_init("fastiso", bsl0), G.bsl_init("fastiso", bsl0)
    z_bed = o.get("z_bed")
    o2.init_state(z_bed, H(100.0), 100.0, ob)
    g2.init_state(z_bed, H(100.0), 100.0, gb2)
    g3.init_state(z_bed, H(100.0), 100.0, gb3)
    for k in range(11, 21):
        t = 10.0 * k
        ob.update(t)
        G.bsl_update(gb2, t)
        G.bsl_update(gb3, t)
        o2.update(H(t), t, ob)
        g2.update(H(t), t, gb2)
        g3.update(H(t), t, gb3)
    compare(o2, g2, wtol=1e-6)
    compare(o2, g3, wtol=1e-6)
    assert abs(o2.scalar("bsl") - g2.scalar("bsl")) < 1e-9


def test_bsl_restart_file_read_by_oracle(tmp_path):
    b = G.bsl_init("fastiso", -1.25, A_ocean_pd=3.5e14)
    G.bsl_update(b, 10.0)
    path = tmp_path / "bsl.bin"
    b.restart_write(path, 10.0)
    assert O.bsl_restart_read(path) == (b.bsl_now, b.A_ocean_now) == (-1.25, 3.5e14)
    recs, t = O.read_restart_file(path)
    assert t == 10.0 and {k: v[0] for k, v in recs.items()} == {"bsl": "m", "A_ocean": "m^2"}


# ---------------------------------------------------------------- rheology / mask files
def rheology_file_case():
    """A two-layer domain (24^2 ice cells at 100 km, N = 32) whose lithosphere,
    mantle viscosity and active mask come from gridded files on their own,
    coarser grids, smoothed over 150 km."""
    nx, dx = 24, 100e3
    p = dict(W.relax_disc(3)["params"], min_pad=400e3, nl=2, zl=[88e3, 250e3], viscosities=[1e21, 1e21],
             layering="equalized", viscosity_scaling_method="stddev", viscosity_scaling=0.5,
             lithosphere="rheology_file", mantle="rheology_file", interactive_sealevel=1)
    xr = np.linspace(-2500e3, 2500e3, 37)
    X, Y = np.meshgrid(xr, xr)
    T = 80e3 + 30e3 * np.exp(-(X * X + (Y - 300e3) ** 2) / (2 * 700e3 ** 2))
    zc = np.array([-400e3, -300e3, -200e3, -100e3, 0.0])
    lev = np.stack([19.5 + 0.6 * k + 0.4 * np.sin(X / 500e3) * np.cos(Y / 700e3) for k in range(len(zc))])
    sd = np.stack([0.1 + 0.05 * np.cos(X / 900e3) for _ in zc])
    xm = np.linspace(-1900e3, 1900e3, 29)
    XM, YM = np.meshgrid(xm, xm)
    mask = ((np.abs(XM) < 1000e3) & (YM > -800e3)).astype(float)
    grid = dict(xc=xr, yc=xr, litho_thickness=T, litho_thickness_scaling=1.05, zc=zc, log10_eta=lev,
                stddev_log10_eta=sd, rheo_smoothing_radius=150e3, mask_xc=xm, mask_yc=xm, mask=mask)
    return p, nx, dx, grid


def test_rheology_file_init_matches_oracle():
    p, nx, dx, grid = rheology_file_case()
    o = O.OracleIsos(O.make_params(**p), nx, nx, dx, grid=grid)
    g = G.IsosClass(G.make_params(**p), nx, nx, dx, grid=grid)
    assert g.n == o.n == 32
    for f in ("He_lith", "D_lith", "kappa", "R", "eta_eff", "maskactive"):
        a, b = o.get_full(f), g.full(f)
        assert np.allclose(a, b, rtol=1e-13, atol=0), (f, np.abs(a - b).max())
    act = g.full("maskactive")
    assert 0 < act.sum() < act.size
    H = W.disc_load(nx, dx, r0=600e3)
    z_bed = np.full((nx, nx), 1e6)
    ob, gb = O.OracleBSL("const", 0.0), G.bsl_init("const", 0.0)
    o.init_state(z_bed, np.zeros_like(H), 0.0, ob)
    g.init_state(z_bed, np.zeros_like(H), 0.0, gb)
    for k in range(1, 21):
        t = 10.0 * k
        o.update(H, t, ob)
        g.update(H, t, gb)
    compare(o, g, wtol=1e-6)
    assert np.abs(g.out("w")).max() > 1e-3


# ---------------------------------------------------------------- integrators
@pytest.mark.parametrize("case", ["llra", "elva"])
@pytest.mark.parametrize("dt_method", ["euler", "rk4", "bs32"])
def test_integrators_match_oracle(case, dt_method):
    """euler, rk4 and bs32 (integrators.f90:12-129) through the library (the
    inverse row pass then takes the integrator at run time) against the
    oracle, on LLRA and on LV-ELVA."""
    wl = W.relax_disc(1) if case == "llra" else small_disc()
    wl["params"] = dict(wl["params"], dt_method=dt_method, dt_init=1.0)
    times = [0.0] + [10.0 * k for k in range(1, 21)]
    for o, g in run_pair(wl, times):
        pass
    assert g.scalar("attempts") == o.scalar("attempts")
    compare(o, g, wtol=1e-8 if case == "llra" else 1e-6)


# ---------------------------------------------------------------- init-time rheology options
RHEO_CASES = [
    dict(lumping="bottom"), dict(lumping="min"), dict(lumping="max"), dict(lumping="mean"),
    dict(lumping="logmean"), dict(lumping="freqdomain", layering="parallel", dl=60e3),
    dict(viscosity_scaling_method="min"), dict(viscosity_scaling_method="max"),
    dict(viscosity_scaling_method="scale", viscosity_scaling=2.5, correct_compression=1),
    dict(mantle="gaussian_plus", lithosphere="gaussian_minus"),
    dict(mantle="gaussian_minus", lithosphere="gaussian_plus"),
]


@pytest.mark.parametrize("opts", RHEO_CASES, ids=lambda d: "-".join(f"{k}={v}" for k, v in d.items()))
def test_rheology_options_match_oracle(opts):
    """Lumping (lv_elva.f90:207-306), layering, viscosity scaling and the
    rheology tail (fastisostasy.f90:413-442), Gaussian rheologies
    (isos_utils.f90:960-1035): the init fields of the library against the
    oracle's on a two-layer 96^2 domain (N = 128), then 10 updates."""
    nx, dx = 96, 25e3
    n = W.padded_size(nx, dx, 400e3)
    xc = W.padded_coords(n, dx)
    X, Y = np.meshgrid(xc, xc)
    g = np.exp(-(X * X + Y * Y) / (2 * 600e3 ** 2))
    he = 120e3 - 50e3 * g
    eta = np.stack([10.0 ** (21.0 - 1.5 * g), np.full((n, n), 3e21)])
    p = W.params(min_pad=400e3, nl=2, zl=[120e3, 300e3], viscosities=[1e21, 3e21], layering="equalized",
                 lumping="freqdomain", lithosphere="fields", mantle="fields", dt_max=0.05)
    p.update(opts)
    kw = {}
    if p["lithosphere"] == "fields":
        kw["he_field"] = he
    if p["mantle"] == "fields":
        kw["eta_field"] = eta
    o = O.OracleIsos(O.make_params(**p), nx, nx, dx, **kw)
    gg = G.IsosClass(G.make_params(**p), nx, nx, dx, **kw)
    assert gg.n == o.n == n
    for f in ("kappa", "R", "eta_eff", "D_lith", "He_lith"):
        a, b = o.get_full(f), gg.full(f)
        assert np.allclose(a, b, rtol=1e-13, atol=0), (f, np.abs(a - b).max())
    wl = dict(params=p, nx=nx, dx=dx, H_ice=W.parabolic_dome(nx, dx, 800e3, 2500.0),
              z_bed=np.full((nx, nx), 1e6), **kw)
    for o2, g2 in run_pair(wl, [0.0] + [10.0 * k for k in range(1, 11)]):
        pass
    compare(o2, g2, wtol=1e-6)


# ---------------------------------------------------------------- distorted grid
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_distorted_grid_matches_oracle(prec):
    """correct_distortion (fastisostasy.f90:99,131-155): K on the padded grid
    scales dx_matrix, dy_matrix and the cell areas, and K^2 the elastic load;
    the stencil then takes the non-uniform-spacing kernels."""
    wl = small_disc()
    n = W.padded_size(wl["nx"], wl["dx"], wl["params"]["min_pad"])
    xc = W.padded_coords(n, wl["dx"])
    X, Y = np.meshgrid(xc, xc)
    K = 1.0 + 0.04 * (X * X + Y * Y) / (1500e3 ** 2)
    wl["params"] = dict(wl["params"], include_elastic=1)
    p = wl["params"]
    o = O.OracleIsos(O.make_params(**p), wl["nx"], wl["nx"], wl["dx"], K=K)
    g = G.IsosClass(G.make_params(precision=prec, **p), wl["nx"], wl["nx"], wl["dx"], K=K)
    ob, gb = O.OracleBSL("const", 0.0), G.bsl_init("const", 0.0)
    o.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, ob)
    g.init_state(wl["z_bed"], np.zeros_like(wl["H_ice"]), 0.0, gb)
    for k in range(1, 31):
        t = 10.0 * k
        o.update(wl["H_ice"], t, ob)
        g.update(wl["H_ice"], t, gb)
    assert np.abs(o.get("we")).max() > 1e-3
    compare(o, g, wtol=1e-6 if prec == "fp64" else 1e-3, rel=None if prec == "fp64" else 1e-5)


# ---------------------------------------------------------------- generic transform sizes
@pytest.mark.parametrize("n_ice,pad_cells,expect_n", [(64, 16, 96), (80, 20, 120), (96, 32, 160), (160, 32, 224)])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_generic_sizes_match_oracle(n_ice, pad_cells, expect_n, prec):
    """Padded sizes with factors 3, 5 and 7 run the mixed-radix LDS kernels
    (and the non-fused stencil); against the oracle over 300 yr."""
    wl = small_disc(n_ice=n_ice, dx=50e3, pad=pad_cells * 50e3)
    times = [0.0] + [10.0 * k for k in range(1, 31)]
    for o, g in run_pair(wl, times, precision=prec):
        pass
    assert g.n == o.n == expect_n
    compare(o, g, wtol=1e-6 if prec == "fp64" else 1e-3, rel=None if prec == "fp64" else 1e-5)


@pytest.mark.parametrize("n_ice,pad_cells,expect_n", [(128, 32, 192), (320, 32, 384)])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_three_times_power_of_two_register_kernels(n_ice, pad_cells, expect_n, prec, monkeypatch):
    """N = 192 and 384 (3 x 2^k) run the register kernels: three N/3-point
    register FFTs and a radix-3 pass per line (the stencil fused into the
    forward row pass at N = 384); against the oracle over 300 yr, and against
    the mixed-radix LDS kernels (FI_GENERIC_FFT) to FP64 rounding."""
    print(f"3 x 2^k register kernels: N={expect_n} {prec}", flush=True)
    wl = small_disc(n_ice=n_ice, dx=50e3, pad=pad_cells * 50e3)
    times = [0.0] + [10.0 * k for k in range(1, 31)]
    for o, g in run_pair(wl, times, precision=prec):
        pass
    assert g.n == o.n == expect_n
    compare(o, g, wtol=1e-6 if prec == "fp64" else 1e-3, rel=None if prec == "fp64" else 1e-5)
    monkeypatch.setenv("FI_GENERIC_FFT", "1")
    for o2, gen in run_pair(wl, times, precision=prec):
        pass
    tol = 1e-9 if prec == "fp64" else 1e-3
    assert np.abs(g.out("w") - gen.out("w")).max() < tol


def test_generic_kernels_on_power_of_two(monkeypatch):
    """FI_GENERIC_FFT forces the mixed-radix kernels for N = 128: same results
    as the register-FFT path to FP64 rounding, and as the oracle."""
    wl = small_disc(n_ice=96, dx=50e3, pad=16 * 50e3)
    times = [0.0] + [10.0 * k for k in range(1, 21)]
    for o, fast in run_pair(wl, times):
        pass
    monkeypatch.setenv("FI_GENERIC_FFT", "1")
    for o2, gen in run_pair(wl, times):
        pass
    assert gen.n == 128
    compare(o2, gen, wtol=1e-6)
    assert np.abs(fast.out("w") - gen.out("w")).max() < 1e-9


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_asymmetric_kappa_r_matches_oracle(prec):
    """A two-layer freqdomain rheology whose upper-layer viscosity follows an
    off-centre Gaussian: R, hence kappa*R, is even in neither x nor y (the
    lumping mixes the spatial viscosity with the frequency kappa at the same
    array index, lv_elva.f90:207-306); against the oracle over 100 yr."""
    print(f"asymmetric kappa*R: {prec}", flush=True)
    nx, dx = 96, 25e3
    n = W.padded_size(nx, dx, 400e3)
    xc = W.padded_coords(n, dx)
    X, Y = np.meshgrid(xc, xc)
    g = np.exp(-((X - 300e3) ** 2 + (Y + 200e3) ** 2) / (2 * 500e3 ** 2))
    eta = np.stack([10.0 ** (21.0 - 1.5 * g), np.full((n, n), 3e21)])
    p = W.params(min_pad=400e3, nl=2, zl=[100e3, 300e3], viscosities=[1e21, 3e21], layering="equalized",
                 lumping="freqdomain", mantle="fields", dt_max=0.05)
    wl = dict(params=p, nx=nx, dx=dx, H_ice=W.parabolic_dome(nx, dx, 800e3, 2500.0),
              z_bed=np.full((nx, nx), 1e6), eta_field=eta)
    for o, gg in run_pair(wl, [0.0] + [10.0 * k for k in range(1, 11)], precision=prec):
        pass
    assert gg.n == n == 128
    R = gg.full("R")
    # mirror k -> N - k (mod N) in y and in x
    assert not np.array_equal(R[1:, :], R[:0:-1, :]) and not np.array_equal(R[:, 1:], R[:, :0:-1])
    compare(o, gg, wtol=1e-6 if prec == "fp64" else 1e-3, rel=None if prec == "fp64" else 1e-5)
