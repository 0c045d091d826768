"""Analytic components (ELLIPSE, OVOID, HARMONIC) on the device against the
FP64 CPU oracle: the realize_mod drop-in (vertex positions, realized radii,
normals, moments, the baddiam flag and log factor) and the batched objective
with trial values of the components' own parameters.

Tolerances: positions 1e-12 of the model's size (ellipse, harmonic: the same
operations in the same order); the ovoid's cubic runs through pow / acos /
cos, whose last bits differ between libm and the device, and the
trigonometric solution amplifies them as its cubic coefficient shrinks:
1e-10.  Objectives: 1e-10 (FP64 method) and 1e-6 (mixed).
"""
import numpy as np
import pytest

import helpers as H

sh = H.sh
pytestmark = pytest.mark.gpu

CASES = [
    ("ellipse", dict(comp_type=sh.COMP_ELLIPSE, semi_axes=(1.0, 0.8, 0.6)), 1e-12),
    ("ovoid", dict(comp_type=sh.COMP_OVOID, semi_axes=(1.0, 0.8, 0.6), ovoid_k=0.35), 1e-10),
    ("ovoid_neg", dict(comp_type=sh.COMP_OVOID, semi_axes=(1.0, 0.7, 0.5), ovoid_k=-0.6), 1e-10),
    ("harmonic", dict(comp_type=sh.COMP_HARMONIC, nhar=4, har_sigma=0.08), 1e-12),
]


def _pair(over):
    return sh.Scene(sh.spec_for(1, **over)), sh.Scene(sh.spec_for(1, **over))


def _compare_realize(a, b, tol):
    orc = H.load_oracle()
    hip = H.load_hip()
    orc.oracle_realize_mod(a.par, a.mod)
    hip.shape_hip_clear_error()
    hip.realize_mod(b.par, b.mod)
    assert hip.shape_hip_last_error() == 0
    xa, na, fa, ara = a.geometry()
    xb, nb, fb, arb = b.geometry()
    size = np.max(np.abs(xa))
    print("max |dx|", np.max(np.abs(xa - xb)), "size", size)
    assert np.max(np.abs(xa - xb)) <= tol * size
    assert np.max(np.abs(na - nb)) <= 100 * tol
    assert np.max(np.abs(ara - arb)) <= 10 * tol * np.max(ara)
    ua, ra = a.vertex_dirs()
    ub, rb = b.vertex_dirs()
    assert np.max(np.abs(ra - rb)) <= tol * size
    ma, mb = a.moments(), b.moments()
    assert mb[1] == pytest.approx(ma[1], rel=100 * tol)
    fa_, fb_ = a.par_flags(), b.par_flags()
    assert fa_["baddiam"] == fb_["baddiam"]
    assert fb_["baddiam_logfactor"] == pytest.approx(fa_["baddiam_logfactor"], rel=1e-12, abs=0)


@pytest.mark.parametrize("name,over,tol", CASES, ids=[c[0] for c in CASES])
def test_realize_mod_analytic(name, over, tol):
    a, b = _pair(over)
    _compare_realize(a, b, tol)


def test_realize_mod_guards():
    """the baddiam guards: a negative ellipse diameter, a harmonic radius below
    the hair width at some vertices, an ovoid k beyond +-1"""
    a, b = _pair(dict(comp_type=sh.COMP_ELLIPSE))
    for sc in (a, b):
        sc.set_param(0, -0.4)
        sc.set_param(2, 0.005)
    _compare_realize(a, b, 1e-12)
    assert a.par_flags()["baddiam"] == 1
    a, b = _pair(dict(comp_type=sh.COMP_HARMONIC, nhar=1, theta_steps=16))
    for sc in (a, b):
        sc.harmonic(np.array([0.2, 0.25, 0.1]), np.array([0.0, 0.0, 0.05]))  # l = 1 swamps l = 0 on one side
    _compare_realize(a, b, 1e-12)
    assert a.par_flags()["baddiam"] == 1
    a, b = _pair(dict(comp_type=sh.COMP_OVOID, ovoid_k=1.3))
    _compare_realize(a, b, 1e-10)
    assert a.par_flags()["baddiam"] == 1


def _batch_check(over, method, tol):
    orc = H.load_oracle()
    sc = H.make_scene_with_obs(1, **over)
    trials = [[], [(0, sc.get_param(0) * 1.02)], [(1, sc.get_param(1) * 0.97), (2, sc.get_param(2) * 1.01)]]
    for i in range(6, sc.nparams()):  # the component's own shape parameters (k, harmonic coefficients)
        if sc.param_type(i) == 2:
            trials.append([(i, sc.get_param(i) * 1.1 + 0.01)])
    B = sh.Batch([sc], [(0, t) for t in trials], method=method, flags=0)
    obj, chi = B.run()
    for e, t in enumerate(trials):
        saved = [sc.get_param(i) for i in range(sc.nparams())]
        for i, v in t:
            sc.set_param(i, v)
        H.oracle_forward(orc, sc)
        want = orc.oracle_objective_tail(sc.par, sc.dat)
        for i, v in enumerate(saved):
            sc.set_param(i, v)
        print("eval", e, t, obj[e], want)
        assert abs(obj[e] - want) <= tol * abs(want), (e, t, obj[e], want)
    B.close()


@pytest.mark.parametrize("name,over,tol", CASES, ids=[c[0] for c in CASES])
def test_batch_fp64_analytic(name, over, tol):
    _batch_check(over, sh.METHOD_FP64, 1e-10 if tol < 1e-11 else 1e-8)


@pytest.mark.parametrize("name,over,tol", CASES, ids=[c[0] for c in CASES])
def test_batch_mixed_analytic(name, over, tol):
    _batch_check(over, sh.METHOD_MIXED, 1e-6)
