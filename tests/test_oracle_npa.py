"""Pins of the oracle's NPA spin state (oracle_npa.c, inteuler.c:114-392).

Torque-free rotation has invariants and special cases that fix it without
the reference: a spin along a principal axis is the principal-axis
kinematics in closed form; a symmetric top is a rotation about the fixed
angular momentum composed with a rotation about the body's symmetry axis,
both at constant rates; a triaxial body conserves its energy and its
angular momentum vector in space; w returns to itself after the period
4 K(k) sqrt(...) of Landau & Lifshitz sec. 37.  The integrator is checked
against scipy's DOP853 (an independent method), the two methods (brute and
periodic) against each other, and K against scipy.special.ellipk.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.special import ellipk

import helpers as H

M33 = (C.c_double * 3) * 3
V3 = C.c_double * 3


def _orc():
    orc = H.load_oracle()
    orc.oracle_npa_integrate.argtypes = [V3, C.c_double, C.c_double, V3, M33]
    orc.oracle_npa_integrate.restype = None
    orc.oracle_npa_period.argtypes = [V3, V3]
    orc.oracle_npa_period.restype = C.c_double
    orc.oracle_ellip_k.argtypes = [C.c_double]
    orc.oracle_ellip_k.restype = C.c_double
    return orc


def euler(phi, theta, psi):
    """z-x-z Euler matrix, ecliptic -> body (the reference's util/euler.c)"""
    cp, sp, ct, st, cs, ss = math.cos(phi), math.sin(phi), math.cos(theta), math.sin(theta), math.cos(psi), math.sin(psi)
    return np.array([[cs * cp - ct * sp * ss, cs * sp + ct * cp * ss, ss * st],
                     [-ss * cp - ct * sp * cs, -ss * sp + ct * cp * cs, cs * st],
                     [st * sp, -st * cp, ct]])


def integrate(I, w, m, t0, t, method="brute"):
    orc = _orc()
    Iv, wv = V3(*I), V3(*w)
    mm = M33(*[V3(*row) for row in m])
    if method == "brute":
        orc.oracle_npa_integrate(Iv, t0, t, wv, mm)
    else:
        raise ValueError(method)
    return np.array(list(wv)), np.array([list(r) for r in mm])


def rot(axis, ang):
    """rotation by ang about the unit vector axis (Rodrigues)"""
    a = np.asarray(axis) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


I3 = [1.0, 1.3, 1.7]


def test_principal_axis_spin_is_pa_kinematics():
    w3 = 25.13
    m0 = euler(1.1, 0.6, 0.3)
    for dt in (0.37, 3.7):
        w, m = integrate(I3, [0.0, 0.0, w3], m0, 0.0, dt)
        assert np.allclose(w, [0.0, 0.0, w3], atol=1e-12)
        assert np.allclose(m, euler(1.1, 0.6, 0.3 + w3 * dt), atol=2e-11), np.abs(m - euler(1.1, 0.6, 0.3 + w3 * dt)).max()


def test_symmetric_top_closed_form():
    I = [1.0, 1.0, 1.7]
    w0 = np.array([2.0, -0.7, 25.0])
    m0 = euler(0.4, 1.2, -0.9)
    lam = (I[2] - I[0]) / I[0] * w0[2]
    Lb = np.array(I) * w0
    R0 = m0.T                               # body -> ecliptic
    Ls = R0 @ Lb
    alpha = np.linalg.norm(Lb) / I[0]       # rotation rate about the fixed L
    beta = -lam                             # about the body's symmetry axis
    for t in (0.05, 0.8, 2.5):
        w, m = integrate(I, w0, m0, 0.0, t)
        c, s = math.cos(lam * t), math.sin(lam * t)
        wx = np.array([w0[0] * c - w0[1] * s, w0[0] * s + w0[1] * c, w0[2]])
        assert np.allclose(w, wx, atol=1e-10 * np.linalg.norm(w0)), (t, w - wx)
        R = rot(Ls, alpha * t) @ R0 @ rot([0, 0, 1], beta * t)
        assert np.allclose(m, R.T, atol=5e-11), (t, np.abs(m - R.T).max())


def test_triaxial_invariants_over_thirty_days():
    w0 = np.array([2.0, 0.3, 25.0])
    m0 = euler(0.7, 0.9, 0.1)
    L0 = m0.T @ (np.array(I3) * w0)
    E0 = 0.5 * np.sum(np.array(I3) * w0 ** 2)
    for t in (1.3, 30.0):
        w, m = integrate(I3, w0, m0, 0.0, t)
        assert np.allclose(m @ m.T, np.eye(3), atol=1e-11)
        L = m.T @ (np.array(I3) * w)
        assert np.linalg.norm(L - L0) < 1e-11 * np.linalg.norm(L0), (t, L - L0)
        assert abs(0.5 * np.sum(np.array(I3) * w ** 2) - E0) < 1e-11 * E0


def test_integrator_against_dop853():
    w0 = np.array([3.0, -1.5, 22.0])
    m0 = euler(2.0, 0.4, 1.0)
    Iv = np.array(I3)

    def f(t, y):
        w = y[:3]
        d = np.empty(12)
        d[0] = (Iv[1] - Iv[2]) / Iv[0] * w[1] * w[2]
        d[1] = (Iv[2] - Iv[0]) / Iv[1] * w[2] * w[0]
        d[2] = (Iv[0] - Iv[1]) / Iv[2] * w[0] * w[1]
        for j in range(3):
            d[3 + 3 * j:6 + 3 * j] = np.cross(y[3 + 3 * j:6 + 3 * j], w)
        return d

    y0 = np.concatenate([w0, m0.T.reshape(-1)])   # columns of m
    sol = solve_ivp(f, (0, 0.6), y0, method="DOP853", rtol=1e-13, atol=1e-15)
    w, m = integrate(I3, w0, m0, 0.0, 0.6)
    assert np.allclose(w, sol.y[:3, -1], atol=1e-10)
    assert np.allclose(m, sol.y[3:, -1].reshape(3, 3).T, atol=1e-11)


def test_period_returns_w_and_k_is_ellipk():
    orc = _orc()
    for kc in (0.1, 0.5, 0.9, 0.999):
        assert orc.oracle_ellip_k(kc) == pytest.approx(float(ellipk(1 - kc * kc)), rel=2e-15)
    for w0 in ([2.0, 0.3, 25.0], [9.0, 8.0, 1.0], [0.5, 10.0, 0.2]):   # short- and long-axis modes
        T = orc.oracle_npa_period(V3(*I3), V3(*w0))
        assert T > 0
        w, _ = integrate(I3, w0, np.eye(3), 0.0, T)
        assert np.linalg.norm(w - w0) < 1e-9 * np.linalg.norm(w0), (w0, T, w - w0)
        wh, _ = integrate(I3, w0, np.eye(3), 0.0, T / 2)
        assert np.linalg.norm(wh - w0) > 1e-3 * np.linalg.norm(w0)


def _npa_scene(method, omega=(2.0, 0.3, 25.0), impulses=None):
    sh = H.sh
    sc = sh.Scene(sh.spec_for(1, n_deldop=3, epoch_span_days=20.0))
    sc.set_npa(omega, I3, method)
    if impulses:
        t, imp = impulses
        sc.lib.shp_set_impulses(sc.mod, len(t), np.ascontiguousarray(t, dtype=float).ctypes.data_as(H.DP),
                                np.ascontiguousarray(imp, dtype=float).ctypes.data_as(H.DP))
    return sc


@pytest.mark.parametrize("method", ["brute", "periodic"])
def test_realize_spin_views_follow_the_state(method):
    """oracle_realize_spin's view matrices are the integrated state at each
    view's epoch (angle offsets zero here), its spins aeT (w + omegaoff) + orbspin"""
    orc = H.load_oracle()
    sc = _npa_scene(method)
    orc.oracle_realize_spin(sc.par, sc.mod, sc.dat)
    assert not orc.oracle_last_error()
    spin = sc.spin_state()
    for s, f in sc.all_frames()[:3]:
        v = sc.view(s, max(f, 0))
        ae, intspin = v[0:9].reshape(3, 3), v[21:24]
        t = sc.view_epoch(s, max(f, 0))[0]
        w, m = integrate(I3, spin["omega"], euler(*spin["angle"]), spin["t0"], t)
        assert np.allclose(ae, m, atol=1e-8), (s, f, np.abs(ae - m).max())
        assert np.allclose(intspin, m.T @ w, atol=1e-8 * np.linalg.norm(w))


def test_brute_and_periodic_agree():
    orc = H.load_oracle()
    a, b = _npa_scene("brute"), _npa_scene("periodic")
    orc.oracle_realize_spin(a.par, a.mod, a.dat)
    orc.oracle_realize_spin(b.par, b.mod, b.dat)
    worst = 0.0
    for s, f in a.all_frames():
        worst = max(worst, np.abs(a.view(s, max(f, 0))[:9] - b.view(s, max(f, 0))[:9]).max())
    assert worst < 1e-9, worst


def test_impulse_adds_to_w_between_intervals():
    orc = H.load_oracle()
    spin_t0 = H.sh.spec_for(1).t0
    imp_t, imp = [spin_t0 + 4.0], [[0.5, -0.4, 0.3]]
    sc = _npa_scene("brute", impulses=(imp_t, np.array(imp)))
    orc.oracle_realize_spin(sc.par, sc.mod, sc.dat)
    spin = sc.spin_state()
    for s, f in sc.all_frames()[:3]:
        t = sc.view_epoch(s, max(f, 0))[0]
        w, m = np.array(spin["omega"]), euler(*spin["angle"])
        if t > imp_t[0]:
            w, m = integrate(I3, w, m, spin["t0"], imp_t[0])
            w = w + np.array(imp[0])
            w, m = integrate(I3, w, m, imp_t[0], t)
        else:
            w, m = integrate(I3, w, m, spin["t0"], t)
        v = sc.view(s, max(f, 0))
        assert np.allclose(v[0:9].reshape(3, 3), m, atol=1e-9), (s, f)
