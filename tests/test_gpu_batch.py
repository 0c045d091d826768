"""Batched evaluation (FP64 and mixed methods) against the FP64 CPU oracle.

Tolerances: FP64 method -- objective equal to 1e-10 relative (the same
kernels as the drop-in); mixed method -- objective and every frame's chi2
within 1e-6 relative, model images within 1e-4 of each frame's peak.
"""
import numpy as np
import pytest

import helpers as H

sh = H.sh
pytestmark = pytest.mark.gpu

TRIALS = [[], [(0, 1.02)], [(1, 0.97), (2, 1.01)], [(3, 25.2)], [(4, 2.1)], [(6, 0.25)], [(7, 1.9)],
          [(8, 0.3)], [(9, -0.1)]]


def _oracle_eval(orc, sc, trial):
    saved = [sc.get_param(i) for i in range(sc.nparams())]
    for i, v in trial:
        sc.set_param(i, v)
    H.oracle_forward(orc, sc)
    obj = orc.oracle_objective_tail(sc.par, sc.dat)
    chi = orc.oracle_chi2(sc.par, sc.dat, 0)
    frames = []
    for s, f in sc.all_frames():
        frames.append((sc.scalars(s, f)["chi2"] if f >= 0 else None, sc.image(s, max(f, 0), 0)))
    for i, v in enumerate(saved):
        sc.set_param(i, v)
    return obj, chi, frames


SHAPEPAR = 2  # scene parameter type of a vertex deviation


def _trials(sc):
    """TRIALS that fit the scene, plus its vertex deviation (the last
    parameter) 5% out: a target addressed by a vertex, not the parameter
    block."""
    ntr = [t for t in TRIALS if all(i < sc.nparams() for i, _ in t)]
    iv = sc.nparams() - 1
    if iv >= 0 and sc.param_type(iv) == SHAPEPAR:
        ntr.append([(iv, sc.get_param(iv) * 1.05)])
    return ntr


def _check(config, method, over, tol_chi, tol_img, flags=sh.BATCH_KEEP_FIT, law=None, tol_frame=None):
    orc = H.load_oracle()
    sc = H.make_scene_with_obs(config, **over)
    if law is not None:
        H.set_radar_laws(sc, *law)
    ntr = _trials(sc)
    B = sh.Batch([sc], [(0, t) for t in ntr], method=method, flags=flags)
    obj, chi = B.run()
    cnt = B.counters()
    print("config", config, "method", method, "counters", cnt, "sizes", B.sizes())
    assert cnt[1] == 0 and cnt[3] == 0
    worst_chi, worst_img = 0.0, 0.0
    for e, t in enumerate(ntr):
        o_obj, o_chi, frames = _oracle_eval(orc, sc, t)
        assert abs(obj[e] - o_obj) <= tol_chi * abs(o_obj), (e, t, obj[e], o_obj)
        assert abs(chi[e] - o_chi) <= tol_chi * abs(o_chi), (e, t, chi[e], o_chi)
        worst_chi = max(worst_chi, abs(chi[e] - o_chi) / abs(o_chi))
        fchi, _ = B.frame_results(e)
        for q, (c, img) in enumerate(frames):
            if c is not None:
                assert abs(fchi[q] - c) <= (tol_frame or tol_chi) * abs(c), (e, q, fchi[q], c)
            if e == 0:
                got = B.fit(e, q, img.shape)
                err = np.max(np.abs(got - img)) / max(np.max(np.abs(img)), 1e-300)
                worst_img = max(worst_img, err)
    print("worst chi2 rel err", worst_chi, "worst image rel err", worst_img)
    assert worst_img <= tol_img


SMEAR = dict(radar_nviews=3, view_interval_days=0.02, n_deldop=2, n_doppler=2, nbin_min=30, nbin_max=40)
SMEAR_LC = dict(n_deldop=3, n_doppler=3, n_lightcurves=2, lc_points=30, lc_ncalc=10, lc_nviews=3,
                view_interval_days=0.05)


@pytest.mark.parametrize("config,over", [(1, {}), (2, dict(n_deldop=16)),
                                         (1, dict(n_deldop=0, n_doppler=2, nbin_min=30, nbin_max=40)),
                                         (3, dict(n_deldop=3, n_doppler=3, n_lightcurves=2, lc_points=30,
                                                  lc_ncalc=10)),
                                         (1, SMEAR), (3, SMEAR_LC)])
def test_fp64_batch_matches_oracle(config, over):
    _check(config, sh.METHOD_FP64, over, 1e-10, 1e-12)


@pytest.mark.parametrize("config,over", [(1, {}), (2, dict(n_deldop=16)),
                                         (1, dict(n_deldop=0, n_doppler=2, nbin_min=30, nbin_max=40)),
                                         # the model reaches past a 41x41 POS: posbnd
                                         (1, dict(pos_n=20, n_deldop=2)),
                                         (5, dict(n_deldop=6)),
                                         (3, dict(n_deldop=3, n_doppler=3, n_lightcurves=2, lc_points=30,
                                                  lc_ncalc=10)),
                                         (1, SMEAR), (3, SMEAR_LC)])
def test_mixed_batch_matches_oracle(config, over):
    # Per frame, a chi2 near its noise floor amplifies the mixed method's
    # relative fit error eps (~1e-7) to ~2 eps sqrt(o2 / chi2): ~4e-6 at SNR
    # 20.  Smeared Doppler frames at the true model reach 1.7e-6 - 2.5e-6 (as
    # one unsmeared Doppler frame of SMEAR's scene does), so their frames are
    # held to that bound; every evaluation's chi2 and objective keep 1e-6.
    _check(config, sh.METHOD_MIXED, over, 1e-6, 1e-4, tol_frame=1e-5 if over is SMEAR else None)


@pytest.mark.parametrize("method,tol_chi,tol_img", [(sh.METHOD_FP64, 1e-10, 1e-12), (sh.METHOD_MIXED, 1e-6, 1e-4)])
@pytest.mark.parametrize("case", sorted(H.FOOTPRINTS))
def test_batch_footprints(method, tol_chi, tol_img, case):
    """Several samples per baud (spb 2-4, strides 1 and 2, the long-code
    and short-code reductions) and sinc^2 means over nsinc2 x nsinc2
    subpixels, in both methods against the oracle; "wide" footprints exceed
    the mixed method's unrolled register window and take its general loop.
    Each scene's Doppler spectrum has a chi2 of ~35 over 40-60 bins, at its
    noise floor, where the mixed method's ~1e-7 relative fit error becomes
    ~2e-6 of the frame's chi2 (test_mixed_batch_matches_oracle): its frames
    are held to 1e-5, every evaluation's chi2 and objective to 1e-6."""
    _check(1, method, H.FOOTPRINTS[case], tol_chi, tol_img,
           tol_frame=1e-5 if method == sh.METHOD_MIXED else None)


OTHER_LAWS = [(sh.HAGFORS_COSINE, [0.1, 1.2, 0.25, 40.0, 0.3]), (sh.GAUSSIAN_COSINE, [0.1, 1.2, 0.3, 6.0, 0.4]),
              (sh.COSINELAW_QS, [0.2, 3.0, 0.5])]


@pytest.mark.parametrize("method,tol_chi,tol_img", [(sh.METHOD_FP64, 1e-10, 1e-12), (sh.METHOD_MIXED, 1e-6, 1e-4)])
@pytest.mark.parametrize("law", range(len(OTHER_LAWS)))
def test_batch_other_radar_laws(method, tol_chi, tol_img, law):
    """Delay-Doppler and Doppler frames under quasispecular and hybrid laws
    (cutoffs inside the disk; the trials move the laws' R and C)."""
    _check(1, method, dict(n_deldop=2, n_doppler=2, nbin_min=30, nbin_max=40), tol_chi, tol_img,
           law=OTHER_LAWS[law])


@pytest.mark.parametrize("method,tol_f", [(sh.METHOD_FP64, 1e-10), (sh.METHOD_MIXED, 1e-6)])
def test_batch_line_search_matches_oracle(method, tol_f):
    """shape_batch_line_search over the size scale, 8 trial values per
    replay, against the oracle's mnbrak + brent_abs on the oracle objective
    (bestfit.c:446-468): the same minimum to brent_abs's tolerance, and the
    batch's objective there equal to the oracle's to the method's contract."""
    import ctypes as C
    orc = H.load_oracle()
    sc = H.make_scene_with_obs(1)
    x0, step, tol, abstol = 1.02, 0.01, 1e-4, 1e-9
    B = sh.Batch([sc], [(0, [(0, x0)]) for _ in range(8)], method=method)
    rc, xb, fb, rounds = B.line_search(x0, step, tol, abstol)
    assert rc == 0
    saved, calls = sc.get_param(0), [0]

    def f(x):
        calls[0] += 1
        sc.set_param(0, x)
        H.oracle_forward(orc, sc)
        return orc.oracle_objective_tail(sc.par, sc.dat)
    xo = C.c_double()
    fo = orc.oracle_line_search(H.ORACLE_FN(f), x0, step, tol, abstol, C.byref(xo))
    fo_at_xb = f(xb)
    sc.set_param(0, saved)
    print("method", method, "batched: x", xb, "f", fb, "rounds", rounds, "| oracle: x", xo.value, "f", fo,
          "evaluations", calls[0] - 1)
    t1 = tol * abs(xo.value) + abstol
    assert abs(xb - xo.value) <= 4 * t1
    assert abs(fb - fo_at_xb) <= tol_f * abs(fo_at_xb)
    assert rounds < calls[0] - 1


@pytest.mark.parametrize("method", [sh.METHOD_FP64, sh.METHOD_MIXED])
def test_graph_replays_match_eager(method):
    """Replays of the captured graph, with new trial values between them and
    eager work in between, reproduce eager evaluations of the same values."""
    sc = H.make_scene_with_obs(2, n_deldop=8)
    trials = [[(0, 1.0)], [(0, 1.01)], [(0, 0.99)]]
    G = sh.Batch([sc], [(0, t) for t in trials], method=method, flags=sh.BATCH_GRAPH)
    Eg = sh.Batch([sc], [(0, t) for t in trials], method=method, flags=0)
    for k, scale in enumerate((1.0, 1.02, 0.98, 1.0)):
        vals = [v * scale for t in trials for _, v in t]
        G.set_values(vals)
        Eg.set_values(vals)
        og, cg = G.run()
        oe, ce = Eg.run()
        # FP64: same kernels and order of everything but atomic accumulations (1e-12);
        # mixed: FP32 atomic accumulation order varies from run to run (1e-6)
        tol = 1e-12 if method == sh.METHOD_FP64 else 1e-6
        assert np.allclose(og, oe, rtol=tol, atol=0), (k, og, oe)
        assert np.allclose(cg, ce, rtol=tol, atol=0), (k, cg, ce)


@pytest.mark.parametrize("method,tol", [(sh.METHOD_FP64, 1e-10), (sh.METHOD_MIXED, 1e-6)])
def test_multi_model_batch(method, tol):
    """Several independent models (own seeds, frame counts and sizes) and
    trial values in one replay, each evaluation checked against the oracle
    on its own model (the shape of the batched-fit workload)."""
    orc = H.load_oracle()
    scenes = [H.make_scene_with_obs(5, n_deldop=3, seed_geom=5000 + m, seed_frames=5500 + m) for m in range(4)]
    evals = [(m, t) for m in range(4) for t in ([], [(0, 1.015)])]
    B = sh.Batch(scenes, evals, method=method, flags=sh.BATCH_GRAPH)
    obj, chi = B.run()
    assert B.counters()[1] == 0
    for e, (m, t) in enumerate(evals):
        o_obj, o_chi, _ = _oracle_eval(orc, scenes[m], t)
        assert abs(obj[e] - o_obj) <= tol * abs(o_obj), (e, m, t, obj[e], o_obj)
        assert abs(chi[e] - o_chi) <= tol * abs(o_chi), (e, m, t, chi[e], o_chi)
    # distinct models really differ
    assert len({round(float(c), 6) for c in chi[::2]}) == 4


@pytest.mark.parametrize("config,over", [(2, dict(n_deldop=16)),
                                         (1, dict(n_deldop=0, n_doppler=2, nbin_min=30, nbin_max=40)),
                                         (3, dict(n_deldop=3, n_doppler=3, n_lightcurves=2, lc_points=30,
                                                  lc_ncalc=10)),
                                         # the model reaches past a 41x41 POS (posbnd, k_outside32)
                                         (1, dict(pos_n=20, n_deldop=2)),
                                         # 81,920 facets on a 63x63 POS: tile lists overflow and the
                                         # tiles test every block
                                         (3, dict(n_deldop=3, n_doppler=0, n_lightcurves=0, pos_n=31,
                                                  km_per_pixel=0.04))])
def test_mixed_binned_matches_global_zbuf(config, over):
    """The binned tile kernel (LDS z-buffer per 32x32 tile) and the global
    z-buffer path build the same packed keys, replay the same pixels and do
    the same FP32 work per pixel; only the order of FP64 accumulation
    differs.  Chi-square, every frame's terms and the fit images agree to
    1e-11, and the replayed and claimed pixel counts are equal."""
    sc = H.make_scene_with_obs(config, **over)
    ntr = [t for t in TRIALS[:5] if all(i < sc.nparams() for i, _ in t)]
    evals = [(0, t) for t in ntr]
    Bt = sh.Batch([sc], evals, method=sh.METHOD_MIXED, flags=sh.BATCH_KEEP_FIT)
    Bg = sh.Batch([sc], evals, method=sh.METHOD_MIXED, flags=sh.BATCH_KEEP_FIT | sh.BATCH_GLOBAL_ZBUF)
    ot, ct = Bt.run()
    og, cg = Bg.run()
    nt, ng = Bt.counters(), Bg.counters()
    ft, fg = Bt.footprint(), Bg.footprint()
    print("binned counters", nt, "global counters", ng, "claimed", ft["claimed"], fg["claimed"])
    assert nt[1] == 0 and ng[1] == 0 and nt[3] == 0 and ng[3] == 0
    if not any(sc.set_type(s) == sh.LGHTCRV for s in range(sc.nsets)):
        # (lightcurve replays reuse counter 0 after the radar replay)
        assert nt[0] == ng[0]              # the same pixels are replayed
        assert ng[2] == ng[0]              # global path: every candidate list from the raster pass
        assert nt[2] <= nt[0]              # binned: from the tiles' own events unless they overflowed
    assert ft["claimed"] == fg["claimed"]  # the same keys are claimed
    assert np.allclose(ot, og, rtol=1e-11, atol=0), (ot, og)
    assert np.allclose(ct, cg, rtol=1e-11, atol=0), (ct, cg)
    for e in range(len(evals)):
        tt, tg = Bt.frame_terms(e), Bg.frame_terms(e)
        assert np.allclose(tt, tg, rtol=1e-11, atol=1e-300), e
    for q, (s, f) in enumerate(sc.all_frames()):
        if f < 0 or sc.set_type(s) == sh.LGHTCRV:
            continue
        shape = sc.image(s, f, 0).shape
        it, ig = Bt.fit(0, q, shape), Bg.fit(0, q, shape)
        assert np.max(np.abs(it - ig)) <= 1e-11 * max(np.max(np.abs(ig)), 1e-300), q


def _edge_pixels(fm):
    """Pixels of an owner map whose 4-neighbourhood holds another owner."""
    e = np.zeros(fm.shape, dtype=bool)
    e[1:, :] |= fm[1:, :] != fm[:-1, :]
    e[:-1, :] |= fm[:-1, :] != fm[1:, :]
    e[:, 1:] |= fm[:, 1:] != fm[:, :-1]
    e[:, :-1] |= fm[:, :-1] != fm[:, 1:]
    return e


@pytest.mark.parametrize("config,over", [(1, {}), (2, dict(n_deldop=16)),
                                         (3, dict(n_deldop=4, n_doppler=4, n_lightcurves=1, lc_points=20,
                                                  lc_ncalc=8))])
def test_mixed_facet_map_matches_fp64(config, over):
    """Owner facet of every POS pixel: the mixed method against the FP64
    method on the same trial values.  Pixels near a facet edge are replayed
    in FP64, so owners may differ only where two surfaces cross within FP32
    depth resolution: at most 1e-4 of the owned pixels, all on facet edges of
    the FP64 map.  The binned and global-z-buffer rasters agree exactly."""
    sc = H.make_scene_with_obs(config, **over)
    evals = [(0, [])]
    flags = sh.BATCH_KEEP_FIT | sh.BATCH_FACET_MAP
    Bm = sh.Batch([sc], evals, method=sh.METHOD_MIXED, flags=flags)
    Bg = sh.Batch([sc], evals, method=sh.METHOD_MIXED, flags=flags | sh.BATCH_GLOBAL_ZBUF)
    Bf = sh.Batch([sc], evals, method=sh.METHOD_FP64, flags=flags)
    Bm.run(); Bg.run(); Bf.run()
    n = sc.spec.pos_n
    owned = mism = off_edge = 0
    for q, (s, f) in enumerate(sc.all_frames()):
        if f < 0 or sc.set_type(s) == sh.LGHTCRV:
            continue
        fm, fg, ff = Bm.facet_map(0, q, n), Bg.facet_map(0, q, n), Bf.facet_map(0, q, n)
        assert np.array_equal(fm, fg), q
        d = fm != ff
        owned += int(np.count_nonzero(ff >= 0))
        mism += int(np.count_nonzero(d))
        off_edge += int(np.count_nonzero(d & ~_edge_pixels(ff)))
    print("config", config, "owned", owned, "mismatched", mism, "off edges", off_edge)
    assert owned > 0
    assert mism <= 1e-4 * owned and off_edge == 0
