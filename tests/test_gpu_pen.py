"""Moments and penalties on the device against the FP64 CPU oracle.

Tolerances (FP64 throughout; only the summation order differs from the
CPU's): moments -- area and volume 1e-12 relative, centre of mass 1e-12 of
the model's size, inertia tensor 1e-11 of its largest element; penalty bases
1e-9 relative (1e-12 absolute for near-zero ones: comdev of a centred model,
inertiadev of a nearly diagonal tensor); objectives 1e-10 relative (FP64
method) and 1e-6 (mixed method, whose chi2 carries that tolerance).
"""
import ctypes as C

import numpy as np
import pytest

import helpers as H
import test_oracle_pen as P

sh = H.sh
pytestmark = pytest.mark.gpu

ALL = [sh.NONSMOOTH, sh.CONCAVITY, sh.RDEV, sh.VOLUME, sh.COMDEV, sh.INERTIADEV, sh.INERTIADEV_UNI, sh.PA3TILT,
       sh.NONPA, sh.NONPA_UNI, sh.EULEROFFS, sh.FLATTENING, sh.MAXRDEV, sh.MAXELLIPDEV, sh.BIFURCATION, sh.IMPULSE]
WEIGHTS = [0.5, 2.0, 1.0, -0.1, 3.0, 1.5, 0.7, 0.2, 1.0, 0.3, 0.01, 1.0, 0.4, 0.6, 2.0, 0.05]


def _close_base(got, want, types):
    for t, g, w in zip(types, got, want):
        assert abs(g - w) <= 1e-9 * abs(w) + 1e-12, (t, g, w)


def _dress(sc):
    """spin inertia, impulses, dataset offsets and a waist, so that every
    penalty is away from its trivial value"""
    sc.set_spin_inertia([0.95, 1.0, 1.1])
    # an impulse before the spin epoch: counted by the penalty, never applied to the spin
    sc.set_impulses([sc.spec.t0 - 10.0], [[0.002, -0.001, 0.003]])
    for s in range(sc.nsets):
        for j in range(3):
            sc.lib.shp_set_offset(sc.dat, s, 0, j, ord("c"), 0.003 * (j + 1))
            sc.lib.shp_set_offset(sc.dat, s, 1, j, ord("c"), -0.001 * (s + 1))


@pytest.mark.parametrize("over", [dict(), dict(shape_kind=sh.SHAPE_ICOSPHERE, ico_level=5, rdev_sigma=0.03)])
def test_realize_mod_moments(over):
    orc = P._oracle()
    hip = H.load_hip()
    a = sh.Scene(sh.spec_for(1, **over))
    b = sh.Scene(sh.spec_for(1, **over))
    for sc in (a, b):
        sc.set_comp_transform([0.2, 0.5, -0.3], [0.01, -0.02, 0.005])
    orc.oracle_realize_mod(a.par, a.mod)
    hip.shape_hip_clear_error()
    hip.realize_mod(b.par, b.mod)
    assert hip.shape_hip_last_error() == 0
    for comp in (None, 0):
        ao, vo, co, Io = a.moments(comp)
        ah, vh, ch, Ih = b.moments(comp)
        print("area", ao, ah, "volume", vo, vh, "com", co, ch)
        assert ah == pytest.approx(ao, rel=1e-12)
        assert vh == pytest.approx(vo, rel=1e-12)
        assert np.max(np.abs(ch - co)) <= 1e-12 * vo ** (1 / 3)
        assert np.max(np.abs(Ih - Io)) <= 1e-11 * np.max(np.abs(Io))


def test_penalties_dropin():
    """penalties() on the host structs realize_mod left (both sides use the
    oracle's realization, so only the penalties differ)."""
    orc = P._oracle()
    hip = H.load_hip()
    sc = sh.Scene(sh.spec_for(1, shape_kind=sh.SHAPE_ICOSPHERE, ico_level=4, semi_axes=(1.0, 0.7, 0.5),
                              rdev_sigma=0.01))
    u, r = sc.vertex_dirs()
    sc.set_vertex_r(r * (1 - 0.4 * np.exp(-((u[:, 0] - 0.1) / 0.3) ** 2)))
    for j, v in enumerate((1.05, 0.95, 1.02)):
        sc.set_param(j, v)
    _dress(sc)
    orc.oracle_realize_mod(sc.par, sc.mod)
    sc.set_penalties(ALL, WEIGHTS)
    want_total = orc.oracle_penalties(sc.par, sc.mod, sc.dat)
    want = sc.pen_base().copy()
    sc.set_penalties(ALL, WEIGHTS)  # fresh base array
    hip.shape_hip_clear_error()
    got_total = hip.penalties(sc.par, sc.mod, sc.dat)
    assert hip.shape_hip_last_error() == 0
    got = sc.pen_base()
    print("penalties", dict(zip(ALL, got)))
    _close_base(got, want, ALL)
    assert got_total == pytest.approx(want_total, rel=1e-10)
    # a type outside the path records an error and returns 0
    sc.set_penalties([sh.NONCOSINE], [1.0])
    hip.shape_hip_clear_error()
    assert hip.penalties(sc.par, sc.mod, sc.dat) == 0.0
    assert hip.shape_hip_last_error() == -17
    hip.shape_hip_clear_error()


def _batch_check(method, tol):
    orc = P._oracle()
    sc = H.make_scene_with_obs(1)
    _dress(sc)
    sc.set_penalties(ALL, WEIGHTS)
    trials = [[], [(0, 1.03)], [(1, 0.96), (2, 1.02)], [(3, sc.get_param(3) * 1.001)]]
    iv = sc.nparams() - 1
    if sc.param_type(iv) == 2:
        trials.append([(iv, sc.get_param(iv) * 1.05)])
    B = sh.Batch([sc], [(0, t) for t in trials], method=method, flags=0)
    obj, chi = B.run()
    for e, t in enumerate(trials):
        saved = [sc.get_param(i) for i in range(sc.nparams())]
        for i, v in t:
            sc.set_param(i, v)
        H.oracle_forward(orc, sc)
        want = orc.oracle_objective(sc.par, sc.mod, sc.dat)
        want_base = sc.pen_base().copy()
        for i, v in enumerate(saved):
            sc.set_param(i, v)
        base, total = B.penalties(e)
        print("eval", e, t, "objective", obj[e], want, "penalty sum", total)
        _close_base(base, want_base, ALL)
        assert abs(obj[e] - want) <= tol * abs(want), (e, obj[e], want)
    B.close()


def test_batch_objective_with_penalties_fp64():
    _batch_check(sh.METHOD_FP64, 1e-10)


def test_batch_objective_with_penalties_mixed():
    _batch_check(sh.METHOD_MIXED, 1e-6)


def test_batch_rejects_unknown_penalty():
    sc = H.make_scene_with_obs(1)
    sc.set_penalties([sh.NONCOSINE], [1.0])
    with pytest.raises(RuntimeError):
        sh.Batch([sc], [(0, [])], method=sh.METHOD_FP64, flags=0)
