"""Parity of the FP64 drop-in entry points against the FP64 CPU oracle.

The FP64 path evaluates every expression in the reference's order without
contraction, so the facet map and z-buffer must match exactly, and fits
differ only by libm-vs-device ulps of sin/pow and by the order of the
atomic accumulation (tolerance 1e-12 relative to the image maximum).
"""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _gpu_forward(hip, sc, realize=True):
    if realize:
        hip.realize_mod(sc.par, sc.mod)
        hip.realize_spin(sc.par, sc.mod, sc.dat)
        hip.realize_delcor(sc.dat, 0.0, 0)
    hip.calc_fits(sc.par, sc.mod, sc.dat)
    c = hip.chi2(sc.par, sc.dat, 0)
    assert hip.shape_hip_last_error() == 0
    return c


@pytest.mark.parametrize("config", [1, 2])
def test_realize_geometry_matches(config):
    hip, orc = H.load_hip(), H.load_oracle()
    a, b = H.sh.Scene(H.sh.spec_for(config)), H.sh.Scene(H.sh.spec_for(config))
    a.set_param(0, 1.03)
    b.set_param(0, 1.03)
    orc.oracle_realize_mod(a.par, a.mod)
    hip.realize_mod(b.par, b.mod)
    assert hip.shape_hip_last_error() == 0
    ga, gb = a.geometry(), b.geometry()
    for x, y in zip(ga, gb):
        assert np.array_equal(x, y), H.rel_err(y, x)


@pytest.mark.parametrize("config", [1, 2])
def test_spin_and_offsets_match(config):
    hip, orc = H.load_hip(), H.load_oracle()
    over = dict(delcor_a0=0.7, delcor_a1=-0.3)
    a, b = H.sh.Scene(H.sh.spec_for(config, **over)), H.sh.Scene(H.sh.spec_for(config, **over))
    orc.oracle_realize_spin(a.par, a.mod, a.dat)
    orc.oracle_realize_delcor(a.dat, 0.0, 0)
    hip.realize_spin(b.par, b.mod, b.dat)
    hip.realize_delcor(b.dat, 0.0, 0)
    assert hip.shape_hip_last_error() == 0
    for s, f in a.all_frames():
        va, vb = a.view(s, max(f, 0)), b.view(s, max(f, 0))
        # ae/oe/spin: device cos/sin may differ from libm by an ulp
        assert np.allclose(va, vb, rtol=0, atol=1e-13), (s, f, np.max(np.abs(va - vb)))


SMALL3 = dict(n_deldop=4, n_doppler=4, n_lightcurves=3, lc_points=40, lc_ncalc=12)
CASES = [(1, {}), (2, {}), (3, SMALL3), (1, dict(codemethod=0, delcor_a0=0.4, delcor_a1=-0.2)),
         (1, dict(n_deldop=0, n_doppler=2, nbin_min=30, nbin_max=40)),
         # smearing: three views per radar frame (calc_fits.c's view loop)
         (1, dict(radar_nviews=3, view_interval_days=0.02, n_deldop=2, n_doppler=2, nbin_min=30, nbin_max=40)),
         (3, dict(SMALL3, lc_nviews=3, view_interval_days=0.05))]
# several samples per baud, strides, code methods and sinc^2 subpixels
CASES += [(1, H.FOOTPRINTS[k]) for k in sorted(H.FOOTPRINTS)]


def _trial(sc):
    sc.set_param(0, 1.02)                      # a trial value away from the truth
    sc.set_param(3, sc.get_param(3) * 1.001)


def _oracle_realize(orc, sc):
    orc.oracle_realize_mod(sc.par, sc.mod)
    orc.oracle_realize_spin(sc.par, sc.mod, sc.dat)
    orc.oracle_realize_delcor(sc.dat, 0.0, 0)


@pytest.mark.parametrize("config,over", CASES)
def test_calc_fits_exact(config, over):
    """Same realized inputs: the FP64 drop-in reproduces the CPU's z-buffer and
    facet map bit for bit; fits agree to libm-vs-device ulps of sin/pow."""
    hip, orc = H.load_hip(), H.load_oracle()
    a, b = H.scene_pair(config, **over)
    for sc in (a, b):
        _trial(sc)
        _oracle_realize(orc, sc)
    orc.oracle_calc_fits(a.par, a.mod, a.dat)
    ca = orc.oracle_chi2(a.par, a.dat, 0)
    hip.shape_hip_set_strict(1)
    hip.calc_fits(b.par, b.mod, b.dat)
    hip.shape_hip_set_strict(0)
    cb = hip.chi2(b.par, b.dat, 0)
    assert hip.shape_hip_last_error() == 0
    print("config", config, over, "replayed pixels", hip.shape_hip_last_replayed(), "chi2", ca, cb)
    assert abs(cb - ca) / abs(ca) < 1e-12, (ca, cb)
    worst = 0.0
    for s, f in a.all_frames():
        fa, fb = a.image(s, f, 0), b.image(s, f, 0)
        worst = max(worst, H.rel_err(fb, fa))
        sa, sb = a.scalars(s, f), b.scalars(s, f)
        for k in ("idellim0", "idellim1", "idoplim0", "idoplim1", "overflow_o2"):
            assert sa[k] == sb[k], (s, f, k)
        for k in ("dellim0", "dellim1", "doplim0", "doplim1", "km2Hz", "cal", "chi2", "overflow_m2",
                  "overflow_xsec", "badradar_logfactor"):
            assert abs(sa[k] - sb[k]) <= 1e-9 * max(1.0, abs(sa[k])), (s, f, k, sa[k], sb[k])
    assert worst < 1e-12, worst
    s, f = a.all_frames()[-1]
    pa, pb = a.pos(s, max(f, 0)), b.pos(s, max(f, 0))
    for which in (3, 10, 11):
        x, y = a.pos_array(pa, which), b.pos_array(pb, which)
        bad = np.argwhere(x != y)
        assert len(bad) == 0, (which, len(bad), bad[:5], x[tuple(bad[0])], y[tuple(bad[0])])
    for which in (1, 2):
        assert H.rel_err(b.pos_array(pb, which), a.pos_array(pa, which)) < 1e-15, which
    assert a.pos_limits(pa) == b.pos_limits(pb)
    assert np.array_equal(a.facet_seen(), b.facet_seen())
    assert a.par_flags() == b.par_flags()


@pytest.mark.parametrize("config,over", CASES[:3])
def test_gpu_chain_matches(config, over):
    """The whole chain on the GPU (realize_mod, realize_spin, realize_delcor,
    calc_fits, chi2) against the oracle's chain: equal to 1e-10 relative."""
    hip, orc = H.load_hip(), H.load_oracle()
    a, b = H.scene_pair(config, **over)
    for sc in (a, b):
        _trial(sc)
    ca = H.oracle_forward(orc, a)
    cb = _gpu_forward(hip, b)
    assert abs(cb - ca) / abs(ca) < 1e-10, (ca, cb)
    for s, f in a.all_frames():
        assert H.rel_err(b.image(s, f, 0), a.image(s, f, 0)) < 1e-9


def _compare_fits(a, b, tol=1e-12):
    worst = 0.0
    for s, f in a.all_frames():
        worst = max(worst, H.rel_err(b.image(s, f, 0), a.image(s, f, 0)))
        sa, sb = a.scalars(s, f), b.scalars(s, f)
        for k in ("idellim0", "idellim1", "idoplim0", "idoplim1"):
            assert sa[k] == sb[k], (s, f, k)
        for k in ("dellim0", "dellim1", "doplim0", "doplim1", "km2Hz", "overflow_m2", "badradar_logfactor"):
            assert abs(sa[k] - sb[k]) <= 1e-9 * max(1.0, abs(sa[k])), (s, f, k, sa[k], sb[k])
    assert worst < tol, worst


def test_calc_fits_keeps_arenas_across_calls():
    """calc_fits keeps its device arenas for the same structs and refreshes
    them from the structs on every call: successive trial values (the model
    growing, then shrinking well inside its last rectangle) each match the
    oracle, POS arrays included (no pixel of an earlier call survives
    outside the new rectangle); another scene in between gets its own
    arenas."""
    hip, orc = H.load_hip(), H.load_oracle()
    a, b = H.scene_pair(3, **SMALL3)
    other = H.sh.Scene(H.sh.spec_for(1))
    _oracle_realize(orc, other)
    for scale in (1.0, 1.04, 0.7):
        for sc in (a, b):
            sc.set_param(0, scale)
            _oracle_realize(orc, sc)
        orc.oracle_calc_fits(a.par, a.mod, a.dat)
        ca = orc.oracle_chi2(a.par, a.dat, 0)
        hip.shape_hip_clear_error()
        hip.shape_hip_set_strict(1)
        hip.calc_fits(b.par, b.mod, b.dat)
        hip.calc_fits(other.par, other.mod, other.dat)
        hip.shape_hip_set_strict(0)
        cb = hip.chi2(b.par, b.dat, 0)
        assert hip.shape_hip_last_error() == 0
        assert abs(cb - ca) <= 1e-12 * abs(ca), (scale, ca, cb)
        _compare_fits(a, b)
        s, f = a.all_frames()[-1]
        pa, pb = a.pos(s, max(f, 0)), b.pos(s, max(f, 0))
        for which in (0, 1, 2, 3, 4, 5, 10, 13):
            x, y = a.pos_array(pa, which), b.pos_array(pb, which)
            assert np.max(np.abs(x - y)) <= 1e-15 * max(1.0, np.max(np.abs(x))), (scale, which)
        assert a.pos_limits(pa) == b.pos_limits(pb)
        assert np.array_equal(a.facet_seen(), b.facet_seen())
        assert a.par_flags() == b.par_flags()
    hip.shape_hip_release_cache()


@pytest.mark.parametrize("threads", [1, 4])
@pytest.mark.parametrize("config,over", [(3, SMALL3), (1, dict(pos_n=20, n_deldop=2, n_doppler=2, nbin_min=30,
                                                               nbin_max=40))])
def test_one_dataset_entries_match_oracle(config, over, threads):
    """calc_deldop / calc_doppler / calc_lghtcrv called per dataset, as the
    HostMultiThreading driver calls them (datasets dealt round-robin to host
    threads, calc_fits_hmt.c:120-160): the fits equal the oracle's
    calc_fits, and the flags, log factors summed over datasets and
    normalized as calc_fits does (calc_fits.c:407-409), and seen flags equal
    its.  The second case reaches past a 41x41 POS (posbnd)."""
    from concurrent.futures import ThreadPoolExecutor
    hip, orc = H.load_hip(), H.load_oracle()
    a, b = H.scene_pair(config, **over)
    for sc in (a, b):
        _trial(sc)
        _oracle_realize(orc, sc)
    orc.oracle_calc_fits(a.par, a.mod, a.dat)
    scl = H.sh.load_scene()
    scl.shp_fits_bracket(b.par, b.mod, b.dat, 1)
    hip.shape_hip_clear_error()
    entry = {H.sh.DELAY: hip.calc_deldop, H.sh.DOPPLER: hip.calc_doppler, H.sh.LGHTCRV: hip.calc_lghtcrv}

    def work(t):
        for s in range(t, b.nsets, threads):
            entry[b.set_type(s)](b.par, b.mod, b.set_desc(s), s)
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, range(threads)))
    scl.shp_fits_bracket(b.par, b.mod, b.dat, 0)
    assert hip.shape_hip_last_error() == 0
    _compare_fits(a, b)
    fa, fb = a.par_flags(), b.par_flags()
    for k in fa:
        assert abs(fa[k] - fb[k]) <= 1e-12 * max(1.0, abs(fa[k])), (k, fa[k], fb[k])
    if config == 1:
        assert fa["posbnd"] == 1
    assert np.array_equal(a.facet_seen(), b.facet_seen())
