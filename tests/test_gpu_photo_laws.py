"""Hapke, harmonic, inhomogeneous and tabular scattering laws on the device
against the oracle.

Jobs read from model files carrying the laws (the test job of test_io: two
delay-Doppler frames, a smeared Doppler set and a lightcurve).  For each pair
of laws:
* the FP64 drop-in chain (realize_mod, realize_spin, realize_delcor,
  calc_fits, chi2; the device resolves the laws' local values itself) matches
  the oracle chain with realize_photo to 1e-12 relative in chi2 and 1e-11 in
  the fit images (Hapke's reflectance chains a dozen transcendental calls whose
  device and host roundings differ by an ulp or two);
* radlaw and apply_photo, given the caller's resolved values, match the
  oracle's to 1e-13;
* the batched objective, FP64 and mixed, matches the oracle's (chi-square over
  degrees of freedom, the file's penalties, the bad-photometry factor) for
  trials of the size, a law parameter and a law value out of its bounds, to
  1e-11 (FP64) and 1e-6 (mixed).
"""
import numpy as np
import pytest

import helpers as H
import test_io as TI
import test_oracle_pen as P
import test_oracle_photo_laws as L

pytestmark = pytest.mark.gpu
sh = H.sh
NF = L.NF


def _inho_cos():
    locs = [[("f", 0.2 + 0.05 * np.sin(f)), ("f", 1.5)] if f % 4 else [("=", 0.0), ("=", 0.0)] for f in range(NF)]
    return L.inho_text("inhocosine", [("f", 0.22), ("c", 1.4)], locs)


def _inho_kaas():
    locs = [[("f", 0.3 + 0.02 * np.cos(f)), ("c", 0.1), ("c", 0.05), ("c", 8.0), ("c", -0.01)] if f % 5
            else [("=", 0.0)] * 5 for f in range(NF)]
    return L.inho_text("inhokaas", [("f", 0.31), ("c", 0.12), ("c", 0.0), ("c", 5.0), ("c", 0.0)], locs)


def _harm_cos():
    return L.harm_text("harmcosine", [(2, [0.2, 0.02, -0.01, 0.015, 0.01, 0.0, 0.005, -0.004, 0.003]),
                                      (1, [1.4, 0.1, 0.05, -0.08])])


def _harm_hapke():
    return L.harm_text("harmhapke", [(1, [0.5, 0.05, 0.02, -0.03]), (0, [0.08]), (0, [0.9]), (1, [-0.25, 0.05, 0, 0]),
                                     (0, [20.0])])


# (radar law, optical law, law parameter trials: (address key, value))
CASES = {
    "tabular+hapke": ("tabular 6 f 0.5 f 0.42 = 0.0 f 0.2 f 0.1 f 0.05",
                      "hapke f 0.45 c 0.07 c 1.1 c -0.28 c 18.0",
                      [(("law", 1, 0, 1, -1), 0.3), (("law", 0, 0, 0, -1), 0.5), (("law", 1, 0, 4, -1), -0.05)]),
    "harmcosine+harmhapke": (_harm_cos(), _harm_hapke(),
                             [(("law", 1, 0, 1, -1), 0.05), (("law", 0, 0, 0, -1), 0.55),
                              (("law", 0, 0, 0, -1), 1.02)]),
    "inhocosine+inhokaas": (_inho_cos(), _inho_kaas(),
                            [(("law", 1, 0, 0, 101), 0.3), (("law", 1, 0, 0, -1), 0.25),
                             (("law", 0, 0, 0, 7), -0.2)]),
    "inhocosine+harmlommel": (_inho_cos(), L.harm_text("harmlommel", [(2, [0.3, 0.03, 0.01, -0.02, 0.01, 0, 0, 0, 0])]),
                              [(("law", 0, 0, 1, -1), -0.04)]),
}


def _files(tmp_path, monkeypatch, case):
    monkeypatch.chdir(tmp_path)
    rad, opt, _ = CASES[case]
    return TI.write_job(".", radar_law=rad, optical_law=opt)[0]


def _oracle_chain(orc, job, objective=False):
    orc.oracle_realize_mod(job.par, job.mod)
    orc.oracle_realize_spin(job.par, job.mod, job.dat)
    orc.oracle_realize_photo(job.par, job.mod)
    orc.oracle_realize_delcor(job.dat, 0.0, 0)
    orc.oracle_calc_fits(job.par, job.mod, job.dat)
    if objective:
        return orc.oracle_objective(job.par, job.mod, job.dat)
    return orc.oracle_chi2(job.par, job.dat, 0)


def _orc():
    orc = P._oracle()
    L._orc()
    return orc


@pytest.mark.parametrize("case", sorted(CASES))
def test_dropin_chain_matches_oracle(tmp_path, monkeypatch, case):
    files = _files(tmp_path, monkeypatch, case)
    hip, orc = H.load_hip(), _orc()
    a, b = sh.Job(*files), sh.Job(*files)
    ca = _oracle_chain(orc, a)
    assert not orc.oracle_last_error()
    hip.shape_hip_clear_error()
    hip.realize_mod(b.par, b.mod)
    hip.realize_spin(b.par, b.mod, b.dat)
    hip.realize_delcor(b.dat, 0.0, 0)
    hip.calc_fits(b.par, b.mod, b.dat)
    cb = hip.chi2(b.par, b.dat, 0)
    assert hip.shape_hip_last_error() == 0
    print(case, "chi2 oracle", ca, "device", cb)
    assert abs(cb - ca) <= 1e-12 * abs(ca), (ca, cb)
    worst = 0.0
    for s, f in a.all_frames():
        fa, fb = a.image(s, max(f, 0), 0), b.image(s, max(f, 0), 0)
        worst = max(worst, H.rel_err(fb, fa))
    assert worst < 1e-11, worst
    a.close()
    b.close()


@pytest.mark.parametrize("case", ["tabular+hapke", "inhocosine+inhokaas", "harmcosine+harmhapke"])
def test_radlaw_and_apply_photo_dropins(tmp_path, monkeypatch, case):
    files = _files(tmp_path, monkeypatch, case)
    hip, orc = H.load_hip(), _orc()
    a = sh.Job(*files)
    _oracle_chain(orc, a)
    for c in (0.05, 0.3, 0.77, 0.999):
        for f in (-1, 0, 17, NF - 1):
            va = orc.oracle_radlaw(a.photo, 0, c, 0, f)
            vb = hip.radlaw(a.photo, 0, c, 0, f)
            assert hip.shape_hip_last_error() == 0
            assert abs(va - vb) <= 1e-13 * abs(va), (c, f, va, vb)
    pa = a.pos(2, 1)
    for phase, inten in ((0.3, 1.7e-3), (1.1, 2.0)):
        ya = orc.oracle_apply_photo(a.mod, 0, phase, inten, pa, 0, 2, 1)
        ba = a.pos_array(pa, 0).copy()
        yb = hip.apply_photo(a.mod, 0, phase, inten, pa, 0, 2, 1)
        assert hip.shape_hip_last_error() == 0
        print(case, "apply_photo", ya, yb)
        assert ya > 1e-30
        assert abs(ya - yb) <= 1e-13 * abs(ya)
        assert np.abs(a.pos_array(pa, 0) - ba).max() <= 1e-13 * np.abs(ba).max()
    a.close()


@pytest.mark.parametrize("method", [sh.METHOD_FP64, sh.METHOD_MIXED])
@pytest.mark.parametrize("case", sorted(CASES))
def test_batch_objective_matches_oracle(tmp_path, monkeypatch, case, method):
    files = _files(tmp_path, monkeypatch, case)
    orc = _orc()
    ref = sh.Job(*files)
    trials = [[], [(0, 1.03)]] + [[t] for t in CASES[case][2]] + [[(0, 0.98), CASES[case][2][0]]]
    want, bad = [], []
    for t in trials:
        saved = [(i, ref.get_param(i)) for i, _ in t]
        for i, v in t:
            ref.set_param(i, v)
        want.append(_oracle_chain(orc, ref, objective=True))
        bad.append(ref.badphoto()[0])
        for i, v in saved:
            ref.set_param(i, v)
    assert not orc.oracle_last_error()
    job = sh.Job(*files)
    batch = sh.Batch([job], [(0, t) for t in trials], method=method)
    obj, _ = batch.run()
    print(case, method, "bad photometry", bad, "\n oracle", want, "\n device", list(obj))
    tol = 1e-11 if method == sh.METHOD_FP64 else 1e-6
    for w, g in zip(want, obj):
        assert abs(g - w) <= tol * abs(w), (w, g)
    batch.close()
    job.close()
    ref.close()


LAW_PENALTIES = {
    "inhocosine+harmlommel": [("optalbdel", 0.3), ("optalbvar", 0.2), ("radalbdel", 0.1), ("radalbvar", 0.15),
                              ("rad_c_del", 0.05), ("rad_c_var", 0.07), ("nonsmooth", 0.5)],
    "tabular+harmhapke": [("thetadel", 0.2), ("thetavar", 0.1), ("optalbdel", 0.1), ("noncosine", 0.3),
                          ("volume", 0.25)],
}


@pytest.mark.parametrize("method", [sh.METHOD_FP64, sh.METHOD_MIXED])
@pytest.mark.parametrize("case", sorted(LAW_PENALTIES))
def test_batch_law_penalties_match_oracle(tmp_path, monkeypatch, case, method):
    """the penalties that read the laws' per-facet values and tables, in the
    batched objective, against the oracle's objective"""
    monkeypatch.chdir(tmp_path)
    if case == "tabular+harmhapke":
        rad, opt = CASES["tabular+hapke"][0], _harm_hapke()
        trials = [[], [(0, 1.02)], [(("law", 1, 0, 3, -1), 0.35)], [(("law", 0, 0, 1, -1), 0.08)]]
    else:
        rad, opt = CASES[case][0], CASES[case][1]
        trials = [[], [(0, 1.02)], [(("law", 1, 0, 0, 101), 0.3)], [(("law", 0, 0, 1, -1), -0.04)]]
    files = TI.write_job(".", radar_law=rad, optical_law=opt, penalties=LAW_PENALTIES[case])[0]
    orc = _orc()
    ref = sh.Job(*files)
    want = []
    for t in trials:
        saved = [(i, ref.get_param(i)) for i, _ in t]
        for i, v in t:
            ref.set_param(i, v)
        want.append(_oracle_chain(orc, ref, objective=True))
        for i, v in saved:
            ref.set_param(i, v)
    assert not orc.oracle_last_error()
    job = sh.Job(*files)
    batch = sh.Batch([job], [(0, t) for t in trials], method=method)
    obj, _ = batch.run()
    print(case, method, "oracle", want, "device", list(obj))
    tol = 1e-11 if method == sh.METHOD_FP64 else 1e-6
    for w, g in zip(want, obj):
        assert abs(g - w) <= tol * abs(w), (w, g)
    batch.close()
    job.close()
    ref.close()


@pytest.mark.parametrize("case", sorted(LAW_PENALTIES))
def test_penalties_dropin_reads_the_laws(tmp_path, monkeypatch, case):
    """the drop-in penalties() with the law-reading types, after the caller's
    realize_mod and realize_photo, against the oracle's penalties"""
    monkeypatch.chdir(tmp_path)
    rad, opt = (CASES["tabular+hapke"][0], _harm_hapke()) if case == "tabular+harmhapke" else CASES[case][:2]
    files = TI.write_job(".", radar_law=rad, optical_law=opt, penalties=LAW_PENALTIES[case])[0]
    hip, orc = H.load_hip(), _orc()
    a, b = sh.Job(*files), sh.Job(*files)
    for job in (a, b):
        orc.oracle_realize_mod(job.par, job.mod)
        orc.oracle_realize_photo(job.par, job.mod)
    pa = orc.oracle_penalties(a.par, a.mod, a.dat)
    assert not orc.oracle_last_error()
    hip.shape_hip_clear_error()
    pb = hip.penalties(b.par, b.mod, b.dat)
    assert hip.shape_hip_last_error() == 0
    print(case, "penalties oracle", pa, "device", pb)
    assert abs(pb - pa) <= 1e-12 * abs(pa), (pa, pb)
    a.close()
    b.close()


@pytest.mark.parametrize("flags", [sh.BATCH_GRAPH | sh.BATCH_KEEP_FIT,
                                   sh.BATCH_GRAPH | sh.BATCH_KEEP_FIT | sh.BATCH_GLOBAL_ZBUF])
def test_mixed_cosine_law_below_one_half_at_the_limb(tmp_path, monkeypatch, flags):
    """a diffuse cosine law with C < 1/2 grows without bound at the limb
    (c^(2C-1)); the mixed method replays limb pixels in FP64, so its images
    stay within 1e-5 of the FP64 batch's peak (binned and global kernels)"""
    monkeypatch.chdir(tmp_path)
    files = TI.write_job(".", radar_law="cosine c 0.2 f 0.3")[0]
    fits = {}
    for m in (sh.METHOD_FP64, sh.METHOD_MIXED):
        job = sh.Job(*files)
        b = sh.Batch([job], [(0, []), (0, [(("law", 1, 0, 1, -1), 0.45)])], method=m, flags=flags)
        obj, _ = b.run()
        fits[m] = (obj, [[b.fit(e, 0, (100, 80)), b.fit(e, 1, (100, 80)), b.fit(e, 2, (100,))] for e in range(2)])
        b.close()
        job.close()
    (o1, f1), (o2, f2) = fits[sh.METHOD_FP64], fits[sh.METHOD_MIXED]
    for e in range(2):
        assert abs(o2[e] - o1[e]) <= 1e-6 * abs(o1[e]), (e, o1[e], o2[e])
        for a, b in zip(f1[e], f2[e]):
            err = np.abs(a - b).max() / np.abs(a).max()
            print("eval", e, "image error", err)
            assert err < 1e-5, err
