"""Configurations 3, 4 and 5 at their full workloads: the mixed method
against the FP64 CPU oracle.

* 3: 81,920-facet icosphere, 50 delay-Doppler frames (50-300 x 30-250), 40
  Doppler spectra (100-400 bins, 0.25-2 Hz) and 20 lightcurves of 100 points
  (2,000 points, 40 calculated epochs each).
* 4: 327,680-facet icosphere, POS 1023 x 1023 at 2 m, 256 delay-Doppler
  frames of 128-512 x 64-256.
* 5: 32 models of 20,480 facets with 24 frames each (rows 60-250), 8 trial
  values of a size parameter per model: 256 evaluations in one replay.

The oracle runs on the host's cores (frames dealt to threads, each frame
with its own POS).  Contract: every evaluation's chi-square and objective
within 1e-6 of the oracle's; every frame's chi-square within 1e-6, or 1e-5
for a Doppler spectrum (its chi-square at the noise floor magnifies the
mixed method's ~1e-7 relative fit error, test_gpu_batch.py); every model
image within 1e-4 of its peak, and each pixel above 1% of the frame's peak
within 1e-4 of its own value; the facet maps against the FP64 method's
(bit-exact with the oracle's z-buffer, test_gpu_dropin.py): at most 1e-4 of
the owned pixels differ, all on facet edges.

The per-pixel floor: a pixel's FP32 cos(scattering) carries the FP32
rounding of its facet's projected corners times the facet's conditioning
(longest edge over projected area) times its vertex-normal spread.  Facets
past a bound on that product are shaded in FP64; below it the cosine stays
within ~4e-5 of FP64.  On the 3%-noise model one such pixel, 10% of a model
image pixel worth 1.2e-3 of its frame's peak, put that pixel off by 1.7e-4
of itself (2e-7 of the peak).  Dim image pixels fed by few POS pixels carry
that; above 1% of the peak every pixel sums enough POS pixels to stay within
1e-4.
"""
import numpy as np
import pytest

import helpers as H

sh = H.sh
pytestmark = pytest.mark.gpu

PIX_FLOOR = 1e-2  # per-pixel relative error is held above this fraction of the frame's peak


def _oracle_truth_scene(config, **over):
    return H.truth_scene(config, **over)


def _oracle_forward(orc, sc):
    return H.oracle_forward_threads(orc, sc)


def _oracle_eval(orc, sc, trial, images=True):
    saved = [sc.get_param(i) for i in range(sc.nparams())]
    for i, v in trial:
        sc.set_param(i, v)
    chi = _oracle_forward(orc, sc)
    obj = orc.oracle_objective_tail(sc.par, sc.dat)
    frames = []
    for s, f in sc.all_frames():
        c = sc.scalars(s, f)["chi2"] if f >= 0 else None
        frames.append((c, sc.set_type(s), sc.image(s, max(f, 0), 0) if images else None))
    for i, v in enumerate(saved):
        sc.set_param(i, v)
    return obj, chi, frames


def _image_err(got, ref):
    peak = max(float(np.max(np.abs(ref))), 1e-300)
    d = np.abs(got - ref)
    big = np.abs(ref) >= PIX_FLOOR * peak
    pix = float(np.max(d[big] / np.abs(ref[big]))) if np.any(big) else 0.0
    return float(np.max(d)) / peak, pix


def _edge_pixels(fm):
    e = np.zeros(fm.shape, dtype=bool)
    e[1:, :] |= fm[1:, :] != fm[:-1, :]
    e[:-1, :] |= fm[:-1, :] != fm[1:, :]
    e[:, 1:] |= fm[:, 1:] != fm[:, :-1]
    e[:, :-1] |= fm[:, :-1] != fm[:, 1:]
    return e


def _facet_maps(sc, Bm, em, Bf, ef, frames=None):
    """Owned pixels, mismatches and mismatches off facet edges, mixed batch
    evaluation em against FP64 batch evaluation ef, over the radar frames."""
    n = sc.spec.pos_n
    owned = mism = off = 0
    for q, (s, f) in enumerate(sc.all_frames()):
        if f < 0 or sc.set_type(s) == sh.LGHTCRV or (frames is not None and q not in frames):
            continue
        fm, ff = Bm.facet_map(em, q, n), Bf.facet_map(ef, q, n)
        d = fm != ff
        owned += int(np.count_nonzero(ff >= 0))
        mism += int(np.count_nonzero(d))
        off += int(np.count_nonzero(d & ~_edge_pixels(ff)))
    return owned, mism, off


def _check_eval(B, e, sc, ref, tol=1e-6, tol_doppler=1e-5, images=True):
    """One batch evaluation against the oracle's (obj, chi2, frames); returns
    the worst frame chi2 and image errors."""
    obj, chi = B.last
    o_obj, o_chi, frames = ref
    assert abs(chi[e] - o_chi) <= tol * abs(o_chi), (e, chi[e], o_chi)
    assert abs(obj[e] - o_obj) <= tol * abs(o_obj), (e, obj[e], o_obj)
    fchi, _ = B.frame_results(e)
    wf = wp = wx = 0.0
    for q, (c, typ, img) in enumerate(frames):
        if c is not None:
            t = tol_doppler if typ == sh.DOPPLER else tol
            assert abs(fchi[q] - c) <= t * abs(c), (e, q, fchi[q], c)
            wf = max(wf, abs(fchi[q] - c) / abs(c))
        if images and img is not None:
            got = B.fit(e, q, img.shape)
            pk, px = _image_err(got, img)
            wp, wx = max(wp, pk), max(wx, px)
    return wf, wp, wx


def _run(B):
    B.last = B.run()
    return B.last


def test_config4_full_workload():
    """256 frames at a 1023 x 1023 POS: the truth and a trial value 1% out,
    every frame's chi2 and image against the oracle, and every frame's
    facet map against the FP64 method's."""
    orc = H.load_oracle()
    sc = _oracle_truth_scene(4)
    trials = [[], [(0, sc.get_param(0) * 1.01)]]
    flags = sh.BATCH_KEEP_FIT | sh.BATCH_FACET_MAP
    Bm = sh.Batch([sc], [(0, t) for t in trials], method=sh.METHOD_MIXED, flags=flags)
    _run(Bm)
    cnt = Bm.counters()
    assert cnt[1] == 0, cnt
    for e, t in enumerate(trials):
        wf, wp, wx = _check_eval(Bm, e, sc, _oracle_eval(orc, sc, t))
        print("config 4 eval", e, "frame chi2 err", wf, "image err (peak)", wp, "per pixel", wx)
        assert wp <= 1e-4 and wx <= 1e-4, (wp, wx)
    Bf = sh.Batch([sc], [(0, trials[1])], method=sh.METHOD_FP64, flags=sh.BATCH_FACET_MAP)
    Bf.run()
    owned, mism, off = _facet_maps(sc, Bm, 1, Bf, 0)
    print("config 4 facet map: owned", owned, "mismatched", mism, "off edges", off)
    assert owned > 1e8 and mism <= 1e-4 * owned and off == 0


def test_config3_full_workload():
    """50 delay-Doppler frames, 40 spectra and 20 lightcurves of 100 points:
    four trial values (size, spin rate, radar law R, delay polynomial)."""
    orc = H.load_oracle()
    sc = _oracle_truth_scene(3)
    trials = [[], [(0, 1.02)], [(3, sc.get_param(3) * 1.0005)], [(6, 0.25)], [(8, 0.3)]]
    flags = sh.BATCH_KEEP_FIT | sh.BATCH_FACET_MAP
    Bm = sh.Batch([sc], [(0, t) for t in trials], method=sh.METHOD_MIXED, flags=flags)
    _run(Bm)
    assert Bm.counters()[1] == 0
    for e, t in enumerate(trials):
        wf, wp, wx = _check_eval(Bm, e, sc, _oracle_eval(orc, sc, t, images=(e < 2)), images=(e < 2))
        print("config 3 eval", e, "frame chi2 err", wf, "image err (peak)", wp, "per pixel", wx)
        assert wp <= 1e-4 and wx <= 1e-4, (wp, wx)
    Bf = sh.Batch([sc], [(0, trials[1])], method=sh.METHOD_FP64, flags=sh.BATCH_FACET_MAP)
    Bf.run()
    owned, mism, off = _facet_maps(sc, Bm, 1, Bf, 0)
    print("config 3 facet map: owned", owned, "mismatched", mism, "off edges", off)
    assert owned > 0 and mism <= 1e-4 * owned and off == 0


def test_config5_full_workload():
    """32 models (own seeds) x 24 frames, 8 trial values of the first scale
    factor per model, as a line search evaluates them: 256 evaluations in
    one replay, each against the oracle on its own model; images and facet
    maps of every model's middle trial."""
    orc = H.load_oracle()
    scenes = [_oracle_truth_scene(5, seed_geom=5000 + m, seed_frames=5500 + m) for m in range(32)]
    xs = [0.985 + 0.005 * k for k in range(8)]
    evals = [(m, [(0, x)]) for m in range(32) for x in xs]
    flags = sh.BATCH_KEEP_FIT | sh.BATCH_FACET_MAP
    Bm = sh.Batch(scenes, evals, method=sh.METHOD_MIXED, flags=flags | sh.BATCH_GRAPH)
    _run(Bm)
    assert Bm.counters()[1] == 0
    worst = [0.0, 0.0, 0.0]
    for e, (m, t) in enumerate(evals):
        img = (e % 8) == 4
        r = _check_eval(Bm, e, scenes[m], _oracle_eval(orc, scenes[m], t, images=img), images=img)
        worst = [max(a, b) for a, b in zip(worst, r)]
    print("config 5: frame chi2 err", worst[0], "image err (peak)", worst[1], "per pixel", worst[2])
    assert worst[1] <= 1e-4 and worst[2] <= 1e-4, worst
    # distinct models differ
    assert len({round(float(c), 6) for c in Bm.last[1][4::8]}) == 32
    Bf = sh.Batch(scenes, [(m, [(0, xs[4])]) for m in range(32)], method=sh.METHOD_FP64, flags=sh.BATCH_FACET_MAP)
    Bf.run()
    tot = [0, 0, 0]
    for m in range(32):
        r = _facet_maps(scenes[m], Bm, 8 * m + 4, Bf, m)
        tot = [a + b for a, b in zip(tot, r)]
    print("config 5 facet map: owned", tot[0], "mismatched", tot[1], "off edges", tot[2])
    assert tot[0] > 0 and tot[1] <= 1e-4 * tot[0] and tot[2] == 0
