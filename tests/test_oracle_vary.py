"""Pins of the oracle's vary_params (CPU only).

  * radar cross section: a sphere of radius r under the cosine law, whose
    projected differential cross section is R (C+1) cos^(2C-1)
    (radlaw.c:51-53), has total cross section pi r^2 R (C+1) / (C+1/2) (the
    projected disk integral of (1 - s^2/r^2)^(C-1/2) is pi r^2 / (C+1/2));
    the polyhedral sphere and the pixel grid leave well under 1%;
  * sub-radar point: for a sphere centred on the origin its distance toward
    Earth is r, less the facets' sag (a fraction of a per cent at this
    subdivision);
  * cos(sub-radar latitude): from the spin pole the scene was built with
    (ecliptic longitude and latitude) and each frame's line of sight, with no
    use of the Euler matrices the oracle reads;
  * lightcurve brightness: the mean of calc_fits's fit over the observation
    epochs (one view), a second route through the oracle;
  * what the means count: the first dataset only, as the reference's loop
    runs, divided by the weight sums over every dataset; frames without the
    calibration or floating states left out; zero where a weight sum is zero.
"""
import ctypes as C
import math

import numpy as np
import pytest

import helpers as H
import shape_hip as sh

VP = C.c_void_p
D = C.c_double


def _oracle():
    orc = H.load_oracle()
    orc.oracle_vary_params.argtypes = [VP, VP, VP] + [C.POINTER(D)] * 4
    orc.oracle_vary_params.restype = None
    return orc


def vary(lib_fn, sc):
    out = [D(), D(), D(), D()]
    lib_fn(sc.par, sc.mod, sc.dat, *[C.byref(x) for x in out])
    return np.array([x.value for x in out])


def _sphere(r=1.0, **over):
    base = dict(shape_kind=sh.SHAPE_ICOSPHERE, ico_level=5, semi_axes=(r, r, r), rdev_sigma=0.0, n_deldop=1)
    base.update(over)
    return sh.Scene(sh.spec_for(1, **base))


def _realize(orc, sc):
    orc.oracle_realize_mod(sc.par, sc.mod)
    orc.oracle_realize_spin(sc.par, sc.mod, sc.dat)
    orc.oracle_realize_delcor(sc.dat, 0.0, 0)


def _radar_law(sc, s=0):
    out = np.zeros(2)
    sc.lib.shp_radar_law(sc.mod, sc.dat, s, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out  # R, C


def test_cross_section_and_subradar_point_of_a_sphere():
    orc = _oracle()
    r = 1.0
    sc = _sphere(r)
    sc.set_vary(delcor0=1, radalb=1)
    sc.set_cal(0, 0, "c", 1.0)
    sc.vary_weights()
    _realize(orc, sc)
    zmax, xsec, bright, cosd = vary(orc.oracle_vary_params, sc)
    R, Claw = _radar_law(sc)
    want = math.pi * r * r * R * (Claw + 1) / (Claw + 0.5)
    print("zmax", zmax, "xsec", xsec, "closed form", want, "C", Claw)
    assert xsec == pytest.approx(want, rel=5e-3)
    assert r * (1 - 2e-3) < zmax <= r
    assert bright == 0.0 and cosd == 0.0


def test_cross_section_follows_the_exponent_and_the_calibration():
    orc = _oracle()
    res = []
    for Claw, cal in ((0.7, 1.0), (3.0, 1.0), (3.0, 2.5)):
        sc = _sphere(0.8)
        R = _radar_law(sc)[0]
        sc.set_radar_law(0, sh.COSINELAW_DIFF, [R, Claw])
        sc.set_vary(radalb=1)
        sc.set_cal(0, 0, "c", cal)
        sc.vary_weights()
        _realize(orc, sc)
        res.append(vary(orc.oracle_vary_params, sc)[1] / (cal * math.pi * 0.64 * R * (Claw + 1) / (Claw + 0.5)))
    print(res)
    assert res[0] == pytest.approx(1.0, rel=5e-3) and res[1] == pytest.approx(1.0, rel=5e-3)
    assert res[2] == pytest.approx(res[1], rel=1e-13)


def test_cos_subradar_latitude_from_the_pole():
    orc = _oracle()
    spec = sh.spec_for(1, n_deldop=3, one_set_per_frame=0, obs_lat_min_deg=-20.0, obs_lat_max_deg=35.0)
    sc = sh.Scene(spec)
    sc.set_vary(dopscale=1)
    sc.set_dopscale(0, "f", 1.0)
    sc.vary_weights()
    _realize(orc, sc)
    got = vary(orc.oracle_vary_params, sc)[3]
    lam, beta = math.radians(spec.pole_lambda_deg), math.radians(spec.pole_beta_deg)
    pole = np.array([math.cos(beta) * math.cos(lam), math.cos(beta) * math.sin(lam), math.sin(beta)])
    want = []
    for f in range(sc.nframes(0)):
        oe = sc.view(0, f)[9:18].reshape(3, 3)
        los = oe[2]  # the observer's line of sight in ecliptic coordinates
        want.append(math.sqrt(1 - np.dot(los, pole) ** 2))
    w = [sc.scalars(0, f)["weight"] for f in range(sc.nframes(0))]
    assert got == pytest.approx(np.dot(want, w) / np.sum(w), rel=1e-12)


def test_lightcurve_brightness_is_the_fit_mean():
    """A bumpy model (cast shadows, so the sun's view and the mask matter):
    the brightness is set 0's weighted fit over the observation epochs,
    divided by both lightcurves' weight sums."""
    orc = _oracle()
    spec = sh.spec_for(3, shape_kind=sh.SHAPE_ICOSPHERE, ico_level=4, rdev_sigma=0.12, n_deldop=0, n_doppler=0,
                       n_lightcurves=2, lc_points=25, lc_ncalc=12, lc_nviews=1, phase_min_deg=40.0,
                       phase_max_deg=70.0)
    sc = sh.Scene(spec)
    sc.set_vary(optalb=1)
    sc.set_cal(0, 0, "c", 1.0)
    sc.set_cal(1, 0, "c", 1.0)
    sc.set_weight(0, 0, 0.7)
    sc.set_weight(1, 0, 1.9)
    sc.vary_weights()
    _realize(orc, sc)
    got = vary(orc.oracle_vary_params, sc)[2]
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    fit = sc.image(0, 0, 0)
    n1 = len(sc.image(1, 0, 0))
    want = 0.7 * np.sum(fit) / (0.7 * len(fit) + 1.9 * n1)
    print("brightness", got, "from the fits", want)
    assert got == pytest.approx(want, rel=1e-12, abs=0)


def test_overflow_counts_in_the_cross_section():
    """A Doppler spectrum too narrow for the echo: what falls outside it
    lands in the overflow spectrum (pos2doppler.c keeps one; this version's
    pos2deldop does not), and the total still meets the closed form."""
    orc = _oracle()
    sc = _sphere(1.0, n_deldop=0, n_doppler=1, nbin_min=12, nbin_max=12, dop_per_bin_min=1.0, dop_per_bin_max=1.0)
    sc.set_vary(radalb=1)
    sc.set_cal(0, 0, "c", 1.0)
    sc.vary_weights()
    _realize(orc, sc)
    xsec = vary(orc.oracle_vary_params, sc)[1]
    R, Claw = _radar_law(sc)
    ov = sc.scalars(0, 0)["overflow_xsec"]
    print("xsec", xsec, "overflow part", ov)
    assert ov > 0.2 * xsec
    assert xsec == pytest.approx(math.pi * R * (Claw + 1) / (Claw + 0.5), rel=5e-3)


def test_first_dataset_only_and_weights():
    orc = _oracle()
    sc = _sphere(1.0, n_deldop=2, one_set_per_frame=1)
    sc.set_vary(radalb=1)
    sc.set_cal(0, 0, "c", 1.0)
    sc.set_cal(1, 0, "c", 1.0)
    sc.vary_weights()
    _realize(orc, sc)
    both = vary(orc.oracle_vary_params, sc)[1]
    sc.set_cal(1, 0, "f", 1.0)  # set 1 no longer counts in the weight sum
    sc.vary_weights()
    first = vary(orc.oracle_vary_params, sc)[1]
    w0, w1 = sc.scalars(0, 0)["weight"], sc.scalars(1, 0)["weight"]
    assert both == pytest.approx(first * w0 / (w0 + w1), rel=1e-13)
    sc.set_cal(0, 0, "f", 1.0)
    sc.vary_weights()
    assert vary(orc.oracle_vary_params, sc)[1] == 0.0
