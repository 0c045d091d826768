"""Parity of the per-POS entry points with the oracle's, on identical inputs.

posclr, posvis (fresh, merged onto an occupied POS with an orbit offset, and
the source view of a bistatic POS), posmask, radlaw, pos2deldop, pos2doppler
and apply_photo each run once on the oracle and once through the HIP library,
starting from two scenes in the same state.  The FP64 kernels evaluate the
reference's expressions in its order, so POS arrays, rectangles and flags
must match exactly; fits may differ by the order of atomic accumulation and
device-vs-libm ulps of sin/pow (1e-12 of the image maximum).
"""
import ctypes as C

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

DBL = [0, 1, 2, 3]          # b, cose, cosi, z
DBL_ILL = [4, 5]            # cosill, zill
INT = [10, 11, 12]          # f, body, comp
INT_ILL = [13]              # fill
SMALL3 = dict(n_deldop=2, n_doppler=2, n_lightcurves=2, lc_points=20, lc_ncalc=6)


def pair(config, **over):
    """Two scenes in one state: realized, spun and fitted by the oracle."""
    orc = H.load_oracle()
    a = H.sh.Scene(H.sh.spec_for(config, **over))
    b = H.sh.Scene(H.sh.spec_for(config, **over))
    H.oracle_forward(orc, a)
    H.oracle_forward(orc, b)
    return orc, H.load_hip(), a, b


def state(sc, pos, ill=False):
    out = {w: sc.pos_array(pos, w) for w in DBL + INT + ((DBL_ILL + INT_ILL) if ill else [])}
    out["limits"] = np.array(sc.pos_limits(pos))
    out["posbnd_logfactor"] = sc.pos_scalars(pos)["posbnd_logfactor"]
    return out


def assert_same_state(sa, sb, what):
    for k in sa:
        assert np.array_equal(np.asarray(sa[k]), np.asarray(sb[k])), (what, k)


def offset(x):
    return (C.c_double * 3)(*x)


@pytest.mark.parametrize("config,over,s,f", [(1, {}, 0, 0), (3, SMALL3, 4, 2)])
def test_posclr_matches(config, over, s, f):
    orc, hip, a, b = pair(config, **over)
    pa, pb = a.pos(s, f), b.pos(s, f)
    ill = a.pos_scalars(pa)["bistatic"] != 0
    assert_same_state(state(a, pa, ill), state(b, pb, ill), "before")
    orc.oracle_posclr(pa)
    hip.posclr(pb)
    assert hip.shape_hip_last_error() == 0
    assert_same_state(state(a, pa, ill), state(b, pb, ill), "posclr")


@pytest.mark.parametrize("config", [1, 2])
def test_posvis_fresh_then_merged(config):
    orc, hip, a, b = pair(config)
    pa, pb = a.pos(0, 0), b.pos(0, 0)
    orc.oracle_posclr(pa)
    orc.oracle_posclr(pb)
    zero = offset([0.0, 0.0, 0.0])
    ra = orc.oracle_posvis(a.vertices, zero, pa, 1, 0, 0, 0)
    rb = hip.posvis(b.vertices, zero, pb, 1, 0, 0, 0)
    assert hip.shape_hip_last_error() == 0
    assert ra == rb
    sa, sb = state(a, pa), state(b, pb)
    assert_same_state(sa, sb, "fresh")
    assert np.sum(sa[10] >= 0) > 100
    # a second body, offset in the plane of sky and toward the observer, drawn
    # onto the occupied POS: only nearer facets take pixels over
    off = offset([0.07, -0.05, 0.02])
    ra = orc.oracle_posvis(a.vertices, off, pa, 0, 0, 1, 2)
    rb = hip.posvis(b.vertices, off, pb, 0, 0, 1, 2)
    assert hip.shape_hip_last_error() == 0
    assert ra == rb
    sa, sb = state(a, pa), state(b, pb)
    assert_same_state(sa, sb, "merged")
    assert np.sum(sa[11] == 1) > 100 and np.sum(sa[11] == 0) > 100  # both bodies visible


def test_posvis_outside_window_sets_posbnd():
    orc, hip, a, b = pair(1)
    pa, pb = a.pos(0, 0), b.pos(0, 0)
    orc.oracle_posclr(pa)
    orc.oracle_posclr(pb)
    off = offset([0.9, 0.3, 0.0])  # pushes part of the model past the POS edge
    ra = orc.oracle_posvis(a.vertices, off, pa, 1, 0, 0, 0)
    rb = hip.posvis(b.vertices, off, pb, 1, 0, 0, 0)
    assert ra == rb == 1
    sa, sb = state(a, pa), state(b, pb)
    assert sa["posbnd_logfactor"] > 0.0
    assert_same_state(sa, sb, "outbnd")


def test_bistatic_posvis_posmask_apply_photo():
    orc, hip, a, b = pair(3, **SMALL3)
    s = next(q for q in range(a.nsets) if a.set_type(q) == H.sh.LGHTCRV)
    pa, pb = a.pos(s, 3), b.pos(s, 3)
    assert a.pos_scalars(pa)["bistatic"] == 1
    for sc, p in ((a, pa), (b, pb)):
        orc.oracle_posclr(p)
    zero = offset([0.0, 0.0, 0.0])
    for src in (0, 1):
        ra = orc.oracle_posvis(a.vertices, zero, pa, 1, src, 0, 0)
        rb = hip.posvis(b.vertices, zero, pb, 1, src, 0, 0)
        assert hip.shape_hip_last_error() == 0
        assert ra == rb
        assert_same_state(state(a, pa, True), state(b, pb, True), "posvis src=%d" % src)
    lit_before = np.sum(a.pos_array(pa, 1) > 0)
    orc.oracle_posmask(pa, 0.0)
    hip.posmask(pb, 0.0)
    assert hip.shape_hip_last_error() == 0
    assert_same_state(state(a, pa, True), state(b, pb, True), "posmask")
    assert np.sum(a.pos_array(pa, 1) > 0) <= lit_before
    for phase, inten in ((0.3, 1.7e-3), (1.1, 2.0)):
        ya = orc.oracle_apply_photo(a.mod, 0, phase, inten, pa, 0, s, 3)
        yb = hip.apply_photo(b.mod, 0, phase, inten, pb, 0, s, 3)
        assert hip.shape_hip_last_error() == 0
        assert abs(ya - yb) <= 1e-13 * abs(ya)
        assert np.array_equal(a.pos_array(pa, 0), b.pos_array(pb, 0))


def test_posmask_plane_case_matches():
    """posmask on a constructed input whose stencils often hold blank sky
    (the one-sided interpolation): the GPU's mask equals the oracle's and
    the plane's answer."""
    orc, hip = H.load_oracle(), H.load_hip()
    a, b = H.sh.Scene(H.sh.spec_for(1)), H.sh.Scene(H.sh.spec_for(1))
    pa, pb = a.pos(0, 0), b.pos(0, 0)
    for tol, seed in ((0.0, 3), (0.05, 4)):
        want, _, nk = H.posmask_plane_case(a, pa, tol, seed)
        H.posmask_plane_case(b, pb, tol, seed)
        orc.oracle_posmask(pa, tol)
        hip.posmask(pb, tol)
        assert hip.shape_hip_last_error() == 0
        assert np.array_equal(a.pos_array(pa, 1), want)
        assert np.array_equal(b.pos_array(pb, 1), want)


RADAR_LAWS = [(H.sh.GAUSSIANLAW, [0.3, 6.0, 0.4]), (H.sh.HAGFORSLAW, [0.25, 40.0, 0.3]),
              (H.sh.COSINELAW_QS, [0.2, 3.0, 0.5]), (H.sh.GAUSSIAN_COSINE, [0.1, 1.2, 0.3, 6.0, 0.4]),
              (H.sh.HAGFORS_COSINE, [0.1, 1.2, 0.25, 40.0, 0.3]), (H.sh.COSINE_COSINE, [0.1, 1.2, 0.2, 3.0, 0.5])]


def test_radlaw_matches():
    orc, hip, a, b = pair(2, n_deldop=4)
    cs = (1e-3, 0.2, 0.3, 0.4, 0.5, 0.9, 1.0)
    for ilaw in range(4):   # the scenes' cosine laws
        for c in cs:
            va = orc.oracle_radlaw(a.photo, ilaw, c, 0, 0)
            vb = hip.radlaw(b.photo, ilaw, c, 0, 0)
            assert hip.shape_hip_last_error() == 0
            assert abs(va - vb) <= 1e-14 * abs(va), (ilaw, c, va, vb)
    for t, p in RADAR_LAWS:   # the other homogeneous laws, cutoffs included
        a.set_radar_law(0, t, p)
        b.set_radar_law(0, t, p)
        for c in cs:
            va = orc.oracle_radlaw(a.photo, 0, c, 0, 0)
            vb = hip.radlaw(b.photo, 0, c, 0, 0)
            assert hip.shape_hip_last_error() == 0
            assert abs(va - vb) <= 1e-14 * abs(va), (t, c, va, vb)


SCALARS = ["overflow_o2", "overflow_m2", "badradar_logfactor", "idellim0", "idellim1", "idoplim0", "idoplim1",
           "dellim0", "dellim1", "doplim0", "doplim1", "km2Hz", "overflow_xsec", "overflow_dopmean"]


def radar_case(config, over, kind, offsets):
    orc, hip, a, b = pair(config, **over)
    s = next(q for q in range(a.nsets) if a.set_type(q) == kind)
    f = min(1, a.nframes(s) - 1)
    for sc in (a, b):  # empty fits: the call accumulates into them
        sc.set_image(s, f, 0, np.zeros_like(sc.image(s, f, 0)))
    ofn = orc.oracle_pos2deldop if kind == H.sh.DELAY else orc.oracle_pos2doppler
    hfn = hip.pos2deldop if kind == H.sh.DELAY else hip.pos2doppler
    ra = ofn(a.par, a.photo, *offsets, a.set_desc(s), 0, s, f, 0)
    rb = hfn(b.par, b.photo, *offsets, b.set_desc(s), 0, s, f, 0)
    assert hip.shape_hip_last_error() == 0
    assert ra == rb
    fa, fb = a.image(s, f, 0), b.image(s, f, 0)
    assert np.max(fa) > 0.0
    assert np.max(np.abs(fa - fb)) <= 1e-12 * np.max(np.abs(fa))
    sa, sb = a.scalars(s, f), b.scalars(s, f)
    for k in SCALARS:
        if k in sa:
            assert abs(sa[k] - sb[k]) <= 1e-12 * max(1.0, abs(sa[k])), (k, sa[k], sb[k])
    return sa


def test_pos2deldop_matches():
    radar_case(2, dict(n_deldop=4), H.sh.DELAY, (0.4, -0.3, 0.25))
    radar_case(1, dict(codemethod=0), H.sh.DELAY, (0.0, 0.0, 0.0))


@pytest.mark.parametrize("law", range(len(RADAR_LAWS)))
def test_pos2deldop_matches_other_laws(law):
    """pos2deldop with each quasispecular and hybrid law (cutoffs inside the
    disk) against the oracle."""
    t, p = RADAR_LAWS[law]
    orc, hip, a, b = pair(1, n_deldop=2)
    s = next(q for q in range(a.nsets) if a.set_type(q) == H.sh.DELAY)
    for sc in (a, b):
        H.set_radar_laws(sc, t, p)
        sc.set_image(s, 0, 0, np.zeros_like(sc.image(s, 0, 0)))
    ra = orc.oracle_pos2deldop(a.par, a.photo, 0.1, -0.2, 0.3, a.set_desc(s), 0, s, 0, 0)
    rb = hip.pos2deldop(b.par, b.photo, 0.1, -0.2, 0.3, b.set_desc(s), 0, s, 0, 0)
    assert hip.shape_hip_last_error() == 0
    assert ra == rb
    fa, fb = a.image(s, 0, 0), b.image(s, 0, 0)
    assert np.max(fa) > 0.0
    assert np.max(np.abs(fa - fb)) <= 1e-12 * np.max(np.abs(fa))


def test_pos2doppler_matches_with_overflow():
    sa = radar_case(1, dict(n_deldop=0, n_doppler=2, nbin_min=12, nbin_max=14), H.sh.DOPPLER, (0.2, 0.1, -0.5))
    assert sa["overflow_o2"] > 0.0  # the echo is wider than the spectrum
