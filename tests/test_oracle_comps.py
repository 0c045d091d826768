"""Pins of the oracle's analytic components (ELLIPSE, OVOID, HARMONIC), CPU only.

  * ellipse: every vertex on (x/a)^2 + (y/b)^2 + (z/c)^2 = 1; scale factors
    1; a non-positive diameter made positive with the baddiam log factor in
    closed form;
  * ovoid: k = 0 is the ellipse; for k != 0 each vertex's r / a is the
    largest admissible root (non-negative, |x/a - alpha0| <= 1) of its cubic
    as numpy's companion-matrix solver finds them; the centre of mass at the
    origin (what x0 is for), to the tessellation's accuracy;
  * harmonic: the vertex factors against scipy's associated Legendre
    functions with the stated normalization (the integral of afactor^2 is
    4 pi for m = 0 and 2 pi otherwise); the radius as the coefficients
    times scipy's factors; a sphere for degree 0; the per-vertex guard's
    log factor in closed form.
"""
import math

import numpy as np
import pytest
from scipy.special import lpmv, gammaln

import helpers as H
import shape_hip as sh

HAIRWIDTH = 1.0e-7


def _scene(ct, **over):
    return sh.Scene(sh.spec_for(1, comp_type=ct, **over))


def _realize(sc):
    orc = H.load_oracle()
    orc.oracle_realize_mod(sc.par, sc.mod)
    assert orc.oracle_last_error() in (b"", None)
    vx, _, _, _ = sc.geometry()
    return vx


def test_ellipse_vertices_on_the_ellipsoid():
    sc = _scene(sh.COMP_ELLIPSE, semi_axes=(1.1, 0.7, 0.45))
    vx = _realize(sc)
    q = (vx[:, 0] / 1.1) ** 2 + (vx[:, 1] / 0.7) ** 2 + (vx[:, 2] / 0.45) ** 2
    assert np.max(np.abs(q - 1)) < 1e-14
    f = sc.par_flags()
    assert f["baddiam"] == 0 and f["baddiam_logfactor"] == 0.0


def test_ellipse_negative_diameter():
    sc = _scene(sh.COMP_ELLIPSE)
    sc.set_param(0, -0.5)  # 2a
    vx = _realize(sc)
    f = sc.par_flags()
    assert f["baddiam"] == 1
    assert f["baddiam_logfactor"] == pytest.approx(math.log(1 + HAIRWIDTH + 0.5), rel=1e-15)
    a = 1 / ((2.0 / HAIRWIDTH) * (1 + HAIRWIDTH + 0.5))  # the guarded radius
    u, r = sc.vertex_dirs()
    i = int(np.argmax(np.abs(u[:, 0])))
    assert abs(vx[i, 0]) == pytest.approx(a, rel=1e-6)


def test_ovoid_zero_k_is_the_ellipse():
    e = _scene(sh.COMP_ELLIPSE, semi_axes=(1.0, 0.8, 0.6))
    o = _scene(sh.COMP_OVOID, semi_axes=(1.0, 0.8, 0.6), ovoid_k=0.0)
    assert np.max(np.abs(_realize(e) - _realize(o))) < 1e-15


@pytest.mark.parametrize("k", [0.35, -0.6, 5e-3])
def test_ovoid_roots_and_centre(k):
    a_, b_, c_ = 1.0, 0.8, 0.6
    sc = _scene(sh.COMP_OVOID, semi_axes=(a_, b_, c_), ovoid_k=k, theta_steps=40)
    vx = _realize(sc)
    u, r = sc.vertex_dirs()
    ab, bc = a_ / b_, b_ / c_
    if abs(k) > 0.01:
        t = 3 * (1 - k * k) * math.log((1 + k) / (1 - k))
        x0 = (a_ / k) * ((2 * k * (3 - 2 * k * k) - t) / (2 * k * (3 - k * k) - t))
    else:
        x0 = 0.4 * k * a_
    al = x0 / a_
    worst = 0.0
    for i in range(0, len(r), 7):
        ux = u[i, 0]
        h = ab * ab * u[i, 1] ** 2 + (ab * bc) ** 2 * u[i, 2] ** 2
        c3 = (h - ux * ux) * k * ux
        c2 = (1 + 3 * k * al) * ux * ux + h * (1 - k * al)
        c1 = (k - (2 + 3 * k * al) * al) * ux
        c0 = -(1 - al * al) * (1 + k * al)
        roots = np.roots([c3, c2, c1, c0]) if abs(c3) > 1e-12 else np.roots([c2, c1, c0])
        good = [z.real for z in roots if abs(z.imag) < 1e-9 and z.real >= 0 and abs(z.real * ux - al) - 1 < 1e-6]
        assert good
        rho = r[i] / a_
        scale = max(abs(c3), abs(c2), abs(c1), abs(c0))
        # the oracle's root solves its cubic (to the trigonometric / Cardano
        # method's accuracy, which falls as the cubic coefficient shrinks)
        assert abs(((c3 * rho + c2) * rho + c1) * rho + c0) < 1e-15 * scale * max(1.0, (c2 / c3) ** 2)
        worst = max(worst, abs(rho - max(good)))
    # (numpy's eigenvalue roots lose digits when the cubic coefficient is small)
    assert worst < 1e-7
    _, vol, com, _ = sc.moments()
    assert np.max(np.abs(com)) < 3e-3 * a_, com


def _basis(u, L):
    """afactor / bfactor from scipy's associated Legendre functions (with the
    Condon-Shortley phase) and N_lm = sqrt((2l+1)(l-m)!/(l+m)!)"""
    ct = u[:, 2]
    phi = np.arctan2(u[:, 1], u[:, 0])
    A, Bf = {}, {}
    for l in range(L + 1):
        for m in range(l + 1):
            n = math.sqrt((2 * l + 1) * math.exp(gammaln(l - m + 1) - gammaln(l + m + 1)))
            p = n * lpmv(m, l, ct)
            A[l, m] = np.cos(m * phi) * p
            Bf[l, m] = np.sin(m * phi) * p
    return A, Bf


def test_harmonic_radius_from_scipy_factors():
    sc = _scene(sh.COMP_HARMONIC, nhar=4, har_sigma=0.08, theta_steps=24)
    _realize(sc)
    L, a, b = sc.harmonic()
    assert L == 4
    u, r = sc.vertex_dirs()
    A, Bf = _basis(u, L)
    want = np.zeros(len(r))
    q = 0
    for l in range(L + 1):
        for m in range(l + 1):
            want += a[q] * A[l, m] + (b[q] * Bf[l, m] if m > 0 else 0.0)
            q += 1
    assert np.max(np.abs(r - want)) < 1e-13 * np.max(np.abs(want))
    # the factors' normalization over the sphere, by quadrature on a fine grid
    th = (np.arange(400) + 0.5) * math.pi / 400
    ph = (np.arange(800) + 0.5) * 2 * math.pi / 800
    T, P = np.meshgrid(th, ph, indexing="ij")
    uu = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    A2, B2 = _basis(uu, 3)
    w = (np.sin(T) * (math.pi / 400) * (2 * math.pi / 800)).reshape(-1)
    for (l, m), v in A2.items():
        assert np.sum(w * v * v) == pytest.approx(4 * math.pi if m == 0 else 2 * math.pi, rel=1e-4)


def test_harmonic_degree_zero_sphere_and_guard():
    sc = _scene(sh.COMP_HARMONIC, nhar=0, theta_steps=12)
    L, a, b = sc.harmonic(np.array([0.37]), np.array([0.0]))
    vx = _realize(sc)
    assert np.max(np.abs(np.linalg.norm(vx, axis=1) - 0.37)) < 1e-15
    sc.harmonic(np.array([-0.1]), np.array([0.0]))
    _realize(sc)
    f = sc.par_flags()
    assert f["baddiam"] == 1
    assert f["baddiam_logfactor"] == pytest.approx(sc.nv * math.log(1 + HAIRWIDTH + 0.2), rel=1e-12)
