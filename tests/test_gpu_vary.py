"""vary_params on the device against the FP64 CPU oracle.

Each case realizes two identical scenes, one through the oracle and one
through the drop-ins (realize_mod, realize_spin, realize_delcor), and asks
both for the four means.  Tolerance: 1e-10 relative (FP64 throughout; the
device sums the fit and the POS rectangle in a different order).
"""
import ctypes as C

import numpy as np
import pytest

import helpers as H
import test_oracle_vary as V

sh = H.sh
pytestmark = pytest.mark.gpu


def _pair(spec_kw, setup):
    out = []
    for _ in range(2):
        sc = sh.Scene(sh.spec_for(**spec_kw))
        setup(sc)
        sc.vary_weights()
        out.append(sc)
    return out


def _compare(spec_kw, setup):
    orc = V._oracle()
    hip = H.load_hip()
    a, b = _pair(spec_kw, setup)
    V._realize(orc, a)
    want = V.vary(orc.oracle_vary_params, a)
    hip.shape_hip_clear_error()
    hip.realize_mod(b.par, b.mod)
    hip.realize_spin(b.par, b.mod, b.dat)
    hip.realize_delcor(b.dat, 0.0, 0)
    outs = [C.c_double() for _ in range(4)]
    hip.vary_params(b.par, b.mod, b.dat, 0, *[C.byref(x) for x in outs])
    got = np.array([x.value for x in outs])
    assert hip.shape_hip_last_error() == 0
    print("oracle", want, "device", got)
    for g, w in zip(got, want):
        assert abs(g - w) <= 1e-10 * abs(w), (got, want)
    return got


def test_deldop_first_set():
    def setup(sc):
        sc.set_vary(delcor0=1, radalb=1, dopscale=1)
        sc.set_cal(0, 0, "c", 1.3)
        sc.set_dopscale(0, "f", 1.0)
        sc.set_weight(0, 0, 0.8)
    got = _compare(dict(config=1, n_deldop=2, one_set_per_frame=0), setup)
    assert got[0] > 0 and got[1] > 0 and got[3] > 0


def test_doppler_first_set_with_overflow():
    def setup(sc):
        sc.set_vary(radalb=1, dopscale=1)
        sc.set_cal(0, 0, "c", 1.0)
        sc.set_dopscale(0, "f", 1.0)
    got = _compare(dict(config=1, n_deldop=0, n_doppler=2, one_set_per_frame=1, nbin_min=12, nbin_max=12,
                        dop_per_bin_min=1.0, dop_per_bin_max=1.0), setup)
    assert got[1] > 0


def test_lightcurve_first_set():
    def setup(sc):
        sc.set_vary(optalb=1)
        sc.set_cal(0, 0, "c", 1.0)
        sc.set_cal(1, 0, "c", 1.0)
        sc.set_weight(1, 0, 1.7)
    got = _compare(dict(config=3, shape_kind=sh.SHAPE_ICOSPHERE, ico_level=4, rdev_sigma=0.1, n_deldop=0,
                        n_doppler=0, n_lightcurves=2, lc_points=25, lc_ncalc=12, lc_nviews=1), setup)
    assert got[2] > 0


def test_smeared_first_set_is_refused():
    hip = H.load_hip()
    sc = sh.Scene(sh.spec_for(1, radar_nviews=3, view_interval_days=0.01))
    sc.set_vary(radalb=1)
    sc.set_cal(0, 0, "c", 1.0)
    sc.vary_weights()
    hip.realize_mod(sc.par, sc.mod)
    hip.realize_spin(sc.par, sc.mod, sc.dat)
    outs = [C.c_double() for _ in range(4)]
    hip.shape_hip_clear_error()
    hip.vary_params(sc.par, sc.mod, sc.dat, 0, *[C.byref(x) for x in outs])
    assert hip.shape_hip_last_error() == -19
    hip.shape_hip_clear_error()


def test_batch_refuses_vary():
    sc = H.make_scene_with_obs(1)
    sc.set_vary(radalb=1)
    with pytest.raises(RuntimeError):
        sh.Batch([sc], [(0, [])], method=sh.METHOD_FP64, flags=0)
