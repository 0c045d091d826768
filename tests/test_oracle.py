"""Pins for the FP64 CPU oracle (CPU only).

The reference ships no tests, fixtures or golden vectors for this path, so
each oracle function is pinned by what the mathematics fixes: closed forms,
conservation laws, brute force, physical definitions computed another way and
an independent library (scipy) -- never by re-typing the oracle's own code.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import helpers as H

sh = H.sh
AU = 1.4959787061e8
KM2US = 6.67128190396304          # 2/c, usec per km of range
KM2HZFACT = 7.72139109254982e-5   # 2e6/c, c in km/day


@pytest.fixture(scope="module")
def orc():
    return H.load_oracle()


def _realized(orc, config=1, **over):
    sc = sh.Scene(sh.spec_for(config, **over))
    orc.oracle_realize_mod(sc.par, sc.mod)
    orc.oracle_realize_spin(sc.par, sc.mod, sc.dat)
    orc.oracle_realize_delcor(sc.dat, 0.0, 0)
    return sc


def _rz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, s, 0], [-s, c, 0], [0, 0, 1.0]])


def _rx(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1.0, 0, 0], [0, c, s], [0, -s, c]])


def _oa(sc, s, f):
    v = sc.view(s, f)
    return v[9:18].reshape(3, 3) @ v[:9].reshape(3, 3).T


# ------------------------------------------------------------------ a1 realize
def test_realize_sphere_closed_form(orc):
    R = 0.7
    sc = sh.Scene(sh.spec_for(1, semi_axes=(R, R, R), rdev_sigma=0.0))
    orc.oracle_realize_mod(sc.par, sc.mod)
    x, vn, fn, area = sc.geometry()
    assert np.allclose(np.linalg.norm(x, axis=1), R, rtol=0, atol=1e-14)
    tot = area.sum()  # inscribed polyhedron: just below 4 pi R^2
    assert 0.99 * 4 * math.pi * R * R < tot < 4 * math.pi * R * R
    # sphere: vertex normals are nearly radial
    assert np.min(np.sum(vn * x / R, axis=1)) > 0.995


def test_realize_vertex_normals_plain_sum(orc):
    """Facet normals and areas from the corners' cross product, and each
    vertex normal the normalized plain sum of its attached facets' unit
    normals (realize_mod.c:532-563), all recomputed here from the realized
    positions (facets oriented outward, away from the center of the
    star-shaped mesh).  On this noisy mesh the facets around a vertex differ
    in area, so an area-weighted sum would miss by > 1e-4."""
    sc = sh.Scene(sh.spec_for(1))
    orc.oracle_realize_mod(sc.par, sc.mod)
    x, vn, fn, area = sc.geometry()
    af = _attached(sc)
    nf = len(fn)
    tri = -np.ones((nf, 3), dtype=int)
    cnt = np.zeros(nf, dtype=int)
    for v, fs in enumerate(af):
        for f in fs:
            tri[f, cnt[f]] = v
            cnt[f] += 1
    assert np.all(cnt == 3)
    a, b, c = x[tri[:, 0]], x[tri[:, 1]], x[tri[:, 2]]
    cr = np.cross(b - a, c - a)
    cr *= np.sign(np.sum(cr * (a + b + c), axis=1))[:, None]
    ar = 0.5 * np.linalg.norm(cr, axis=1)
    un = cr / (2 * ar[:, None])
    assert np.max(np.abs(fn - un)) < 1e-13
    assert np.max(np.abs(area - ar) / ar) < 1e-12
    plain = np.array([un[fs].sum(axis=0) for fs in af])
    plain /= np.linalg.norm(plain, axis=1)[:, None]
    wtd = np.array([(ar[fs, None] * un[fs]).sum(axis=0) for fs in af])
    wtd /= np.linalg.norm(wtd, axis=1)[:, None]
    assert np.max(np.abs(vn - plain)) < 1e-13
    assert np.max(np.abs(vn - wtd)) > 1e-4


def test_realize_divergence_theorem(orc):
    """Closed outward-oriented surface: sum(area * n) = 0 and
    (1/3) sum(area * n . centroid) = volume of the polyhedron > 0."""
    sc = sh.Scene(sh.spec_for(2))
    orc.oracle_realize_mod(sc.par, sc.mod)
    x, vn, fn, area = sc.geometry()
    assert np.allclose(np.sum(area[:, None] * fn, axis=0), 0.0, atol=1e-12)
    assert np.allclose(np.linalg.norm(fn, axis=1), 1.0, atol=1e-14)
    assert np.allclose(np.linalg.norm(vn, axis=1), 1.0, atol=1e-14)
    assert np.all(np.sum(vn * x, axis=1) > 0)  # outward


def test_realize_scale_factors_linear(orc):
    """x = s * (u r + a): positions scale componentwise with the scale factors."""
    a, b = sh.Scene(sh.spec_for(1)), sh.Scene(sh.spec_for(1))
    for i, v in enumerate((1.3, 0.7, 2.0)):
        b.set_param(i, v)
    orc.oracle_realize_mod(a.par, a.mod)
    orc.oracle_realize_mod(b.par, b.mod)
    assert np.allclose(b.geometry()[0], a.geometry()[0] * np.array([1.3, 0.7, 2.0]), rtol=1e-15, atol=0)


def test_realize_baddiam(orc):
    sc = sh.Scene(sh.spec_for(1))
    sc.set_param(1, 0.005)
    orc.oracle_realize_mod(sc.par, sc.mod)
    fl = sc.par_flags()
    assert fl["baddiam"] == 1
    assert fl["baddiam_logfactor"] == pytest.approx(math.log(1 + 0.01 - 0.005), rel=1e-15)


# ------------------------------------------------------------------ a2 spin, a3 offsets
def test_spin_state_closed_form(orc):
    """ae = Rz(psi) Rx(theta) Rz(phi), psi = psi0 + omega (t - t0);
    total spin = omega * pole (orbital spin is zero in the synthetic scenes)."""
    sc = sh.Scene(sh.spec_for(2))
    orc.oracle_realize_spin(sc.par, sc.mod, sc.dat)
    phi, theta, omega = sc.get_param(4), sc.get_param(5), sc.get_param(3)
    t0 = sc.spec.t0
    pole = np.array([math.sin(theta) * math.sin(phi), -math.sin(theta) * math.cos(phi), math.cos(theta)])
    for s, f in sc.all_frames():
        v = sc.view(s, f)
        ae = v[:9].reshape(3, 3)
        t = sc.consts(s, f)["t"]
        expect = _rz(omega * (t - t0)) @ _rx(theta) @ _rz(phi)
        assert np.allclose(ae, expect, atol=1e-12)  # psi up to ~750 rad: libm argument reduction
        assert np.allclose(v[18:21], omega * pole, rtol=0, atol=1e-12)


def test_delay_and_doppler_offsets(orc):
    """deloff = (a0 + a1 (t - t0))/del_per_pixel; dopoff = -a1 Ftx/(dop_per_pixel 86400)."""
    a0, a1 = 0.8, -0.25
    sc = sh.Scene(sh.spec_for(2, delcor_a0=a0, delcor_a1=a1))
    orc.oracle_realize_delcor(sc.dat, 0.0, 0)
    tmid = sc.spec.t0 + 0.5 * sc.spec.epoch_span_days
    for s, f in sc.all_frames():
        v, k = sc.view(s, f), sc.consts(s, f)
        assert v[24] == pytest.approx((a0 + a1 * (k["t"] - tmid)) / k["del_per_pixel"], rel=1e-13)
        assert v[25] == pytest.approx(-a1 * k["Ftx"] / (k["dop_per_pixel"] * 86400.0), rel=1e-15)


# ------------------------------------------------------------------ a4, a5 POS
def test_posclr_resets(orc):
    sc = _realized(orc, 1)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    pos = sc.pos(0, 0)
    orc.oracle_posclr(pos)
    assert np.all(sc.pos_array(pos, 10) == -1) and np.all(sc.pos_array(pos, 3) == -1e20)
    n = sc.spec.pos_n
    assert sc.pos_limits(pos)[:4] == [n, -n, n, -n]


def _tri_contains(px, py, a, b, c, tol=1e-9):
    def side(p, q, r):
        return (q[0] - p[0]) * (r[1] - p[1]) - (q[1] - p[1]) * (r[0] - p[0])
    d1, d2, d3 = side(a, b, (px, py)), side(b, c, (px, py)), side(c, a, (px, py))
    return (d1 >= -tol and d2 >= -tol and d3 >= -tol) or (d1 <= tol and d2 <= tol and d3 <= tol)


def test_posvis_zbuffer_brute_force(orc):
    """For a sample of pixels: the owner's triangle contains the pixel center, its
    plane gives z, no front-facing facet containing the center lies nearer, and
    equal-depth ties go to the lower facet index (brute force over all facets)."""
    sc = _realized(orc, 1)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    pos = sc.pos(0, 0)
    z, fac = sc.pos_array(pos, 3), sc.pos_array(pos, 10).astype(int)
    x, vn, fn, area = sc.geometry()
    tri = _triangles(sc)
    oa = _oa(sc, 0, 0)
    xo = x @ oa.T
    front = (fn @ oa.T)[:, 2] > 0
    n, kmpp = sc.spec.pos_n, sc.spec.km_per_pixel
    rng = np.random.default_rng(3)
    cov = np.argwhere(fac >= 0)
    checked = 0
    for i, j in cov[rng.choice(len(cov), 300, replace=False)]:
        px, py = (i - n) * kmpp, (j - n) * kmpp
        best_z, best_f = -np.inf, -1
        for f in np.nonzero(front)[0]:
            a, b, c = xo[tri[f]]
            if not _tri_contains(px, py, a, b, c):
                continue
            # plane depth at the pixel center
            nrm = np.cross(b - a, c - a)
            zz = a[2] - (nrm[0] * (px - a[0]) + nrm[1] * (py - a[1])) / nrm[2]
            if zz > best_z + 1e-12:
                best_z, best_f = zz, f
        assert best_f == fac[i, j] or abs(best_z - z[i, j]) < 1e-12
        assert abs(best_z - z[i, j]) < 1e-12
        checked += 1
    assert checked == 300
    assert np.all(z[fac < 0] == -1e20)


def _triangles(sc):
    """Facet corner indices recovered from the realized geometry: each facet's
    normal-weighted area equals half the cross product of its corners."""
    x, vn, fn, area = sc.geometry()
    nv, nf = len(x), len(fn)
    # rebuild from the vertex -> attached-facet lists exported as the seen-free
    # topology: every vertex lists its facets, so each facet appears under exactly
    # three vertices
    counts = np.zeros(nf, dtype=int)
    tri = -np.ones((nf, 3), dtype=int)
    af = _attached(sc)
    for v in range(nv):
        for f in af[v]:
            tri[f, counts[f]] = v
            counts[f] += 1
    assert np.all(counts == 3)
    # orient each triangle like its stored normal
    for f in range(nf):
        a, b, c = x[tri[f]]
        if np.dot(np.cross(b - a, c - b), fn[f]) < 0:
            tri[f, [1, 2]] = tri[f, [2, 1]]
    return tri


def _attached(sc):
    lib = sc.lib
    lib.shp_vertex_facets.restype = C.c_int
    lib.shp_vertex_facets.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_int]
    out = []
    buf = (C.c_int * 64)()
    for v in range(sc.nv):
        k = lib.shp_vertex_facets(sc.mod, v, buf, 64)
        out.append(list(buf[:k]))
    return out


def test_posvis_against_analytic_ellipsoid(orc):
    """Undeformed ellipsoid: z(x, y) at pixel centers lies on or just inside the
    analytic surface (inscribed facets), within the tessellation's chord sag."""
    ax = (1.0, 0.8, 0.6)
    sc = _realized(orc, 1, rdev_sigma=0.0, semi_axes=ax)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    pos = sc.pos(0, 0)
    z, fac, cose = sc.pos_array(pos, 3), sc.pos_array(pos, 10), sc.pos_array(pos, 1)
    oa = _oa(sc, 0, 0)
    Q = oa @ np.diag([1 / a ** 2 for a in ax]) @ oa.T
    n, kmpp = sc.spec.pos_n, sc.spec.km_per_pixel
    errs = []
    # away from the limb (cos e > 0.5) the depth error is at most twice the chord sag
    for i, j in np.argwhere((fac >= 0) & (cose > 0.5))[::3]:
        px, py = (i - n) * kmpp, (j - n) * kmpp
        a2 = Q[2, 2]
        b2 = 2 * (Q[0, 2] * px + Q[1, 2] * py)
        c2 = Q[0, 0] * px * px + 2 * Q[0, 1] * px * py + Q[1, 1] * py * py - 1
        disc = b2 * b2 - 4 * a2 * c2
        if disc < 0.05:
            continue
        # sag of an inscribed facet is at most edge^2/(8 rho_min): edges <= 0.13 km and
        # rho_min = c^2/a = 0.36 km give 6 m along the normal, /cos(e) along z
        errs.append(((-b2 + math.sqrt(disc)) / (2 * a2) - z[i, j]) * cose[i, j] / 0.0059)
    errs = np.array(errs)
    assert len(errs) > 1000
    assert np.all(errs > -1e-9)
    assert np.max(errs) < 1.0


def test_posvis_smoothed_cose_closed_form(orc):
    """Smoothed cos(scattering) (posvis.c:243-260) interpolates the corners'
    vertex normals linearly in the pixel center's barycentric coordinates,
    rotates the result into the observer frame and normalizes it.  With
    every vertex normal set to its own vertex position, linear interpolation
    over a facet reproduces the surface point under the pixel center, so on
    any mesh centered at the origin cose = z / |(x, y, z)| in closed form
    (x, y the pixel center, z its depth).  An interpolation that weights the
    corners otherwise (e.g. t (c - a) for t (c - b)) misses it by ~ the edge
    length over the radius."""
    sc = _realized(orc, 1)  # the 2%-deformed ellipsoid
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)  # POS frames of the view
    sc.set_vertex_normals(sc.geometry()[0])
    pos = sc.pos(0, 0)
    orc.oracle_posclr(pos)
    off = (C.c_double * 3)(0.0, 0.0, 0.0)
    orc.oracle_posvis(sc.vertices, off, pos, 1, 0, 0, 0)
    z, fac, cose = sc.pos_array(pos, 3), sc.pos_array(pos, 10), sc.pos_array(pos, 1)
    n, kmpp = sc.spec.pos_n, sc.spec.km_per_pixel
    ii, jj = np.nonzero((fac >= 0) & (z > 0))
    px, py, pz = (ii - n) * kmpp, (jj - n) * kmpp, z[ii, jj]
    expect = pz / np.sqrt(px * px + py * py + pz * pz)
    assert len(ii) > 5000
    assert np.max(np.abs(cose[ii, jj] - expect)) < 1e-13


def test_posbnd_when_model_exceeds_window(orc):
    """A model wider than the POS window sets posbnd; the log factor is the log of
    the area ratio of the enlarged frame (bounded by the projected vertex extent)."""
    sc = _realized(orc, 1, semi_axes=(1.4, 1.2, 0.6))
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    fl = sc.par_flags()
    assert fl["posbnd"] == 1
    n = sc.spec.pos_n
    xo = sc.geometry()[0] @ _oa(sc, 0, 0).T / sc.spec.km_per_pixel
    xf = (max(np.floor(xo[:, 0].max() + 0.5), n) - min(np.floor(xo[:, 0].min() + 0.5), -n) + 1) / (2 * n + 1)
    yf = (max(np.floor(xo[:, 1].max() + 0.5), n) - min(np.floor(xo[:, 1].min() + 0.5), -n) + 1) / (2 * n + 1)
    # one frame that is the whole data set: par logfactor = log(xf*yf) of the front half
    assert 0.5 * math.log(xf * yf) < fl["posbnd_logfactor"] <= math.log(xf * yf) + 1e-12


# ------------------------------------------------------------------ a7, a8 delay-Doppler
def _pixel_power(sc, s, f):
    pos = sc.pos(s, f)
    cose, z = sc.pos_array(pos, 1), sc.pos_array(pos, 3)
    law = sc.set_type(s)
    R, Cc = _law_rc(sc, s)
    ii, jj = np.nonzero(cose > 0)
    P = sc.spec.km_per_pixel ** 2 * R * (Cc + 1) * cose[ii, jj] ** (2 * Cc - 1)
    del law
    return ii, jj, z[ii, jj], P


def _law_rc(sc, s):
    out = np.zeros(2)
    lib = sc.lib
    lib.shp_radar_law.restype = None
    lib.shp_radar_law.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_double)]
    lib.shp_radar_law(sc.mod, sc.dat, s, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out[0], out[1]


def test_radlaw_cosine_closed_form(orc):
    sc = sh.Scene(sh.spec_for(1))
    R, Cc = _law_rc(sc, 0)
    for c in (0.1, 0.5, 0.999):
        got = orc.oracle_radlaw(sc.photo, 0, c, 0, 0)
        assert got == pytest.approx(R * (Cc + 1) * c ** (2 * Cc - 1), rel=1e-15)


def _law(orc, sc, law_type, p):
    sc.set_radar_law(0, law_type, p)
    return lambda mu: orc.oracle_radlaw(sc.photo, 0, mu, 0, 0)


def _quad(g, a, b):
    from scipy.integrate import quad
    return quad(g, a, b, epsabs=0.0, epsrel=1e-13, limit=200)[0]


def test_radlaw_weighted_integrals(orc):
    """Each law against an exact weighted integral over mu = cos(incidence),
    from the cutoff mu_c to 1, evaluated by quadrature of the oracle's values:
      cosine R(C+1)mu^(2C-1):     int 2mu law      = 2R(C+1)(1 - mu_c^(2C+1))/(2C+1)
                                  (a sphere's cross-section over pi a^2 at mu_c = 0)
      Gaussian RC e^(-C tan^2)/mu^5:  int 2mu^2 law = R(1 - e^(-C tan^2(theta_c)))
                                  (2 mu^2 law = R d/dmu e^(-C(1/mu^2 - 1)))
      Hagfors RC/2 q^(-3/2)/mu, q = mu^4 + C(1-mu^2):
                                  int 2mu(C - mu^4) law = RC(1 - mu_c/sqrt(q(mu_c)))
                                  (d/dmu [mu/sqrt(q)] = (C - mu^4) q^(-3/2))
    A wrong exponent, factor or term in any law moves its integral."""
    sc = sh.Scene(sh.spec_for(1))
    R, mc = 0.23, 0.35
    for Cc in (0.8, 1.7, 3.0):
        law = _law(orc, sc, sh.COSINELAW_DIFF, [R, Cc])
        assert _quad(lambda m: 2 * m * law(m), 0.0, 1.0) == pytest.approx(2 * R * (Cc + 1) / (2 * Cc + 1), rel=1e-10)
        law = _law(orc, sc, sh.COSINELAW_QS, [R, Cc, mc])
        want = 2 * R * (Cc + 1) * (1 - mc ** (2 * Cc + 1)) / (2 * Cc + 1)
        assert _quad(lambda m: 2 * m * law(m), mc, 1.0) == pytest.approx(want, rel=1e-10)
    for Cc in (0.5, 8.0, 40.0):
        law = _law(orc, sc, sh.GAUSSIANLAW, [R, Cc, mc])
        want = R * (1 - math.exp(-Cc * (1 / mc ** 2 - 1)))
        assert _quad(lambda m: 2 * m * m * law(m), mc, 1.0) == pytest.approx(want, rel=1e-10)
        law = _law(orc, sc, sh.HAGFORSLAW, [R, Cc, mc])
        want = R * Cc * (1 - mc / math.sqrt(mc ** 4 + Cc * (1 - mc * mc)))
        assert _quad(lambda m: 2 * m * (Cc - m ** 4) * law(m), mc, 1.0) == pytest.approx(want, rel=1e-10)


def test_radlaw_cutoffs_and_hybrids(orc):
    """The quasispecular laws switch on at cos >= cutoff (inclusive), and each
    hybrid is its diffuse cosine law plus its quasispecular law, each with its
    own parameters (so a swapped parameter set fails)."""
    sc = sh.Scene(sh.spec_for(1))
    mc = 0.6
    for t in (sh.GAUSSIANLAW, sh.HAGFORSLAW, sh.COSINELAW_QS):
        law = _law(orc, sc, t, [0.3, 2.5, mc])
        assert law(mc) > 0.0
        assert law(math.nextafter(mc, 0.0)) == 0.0
    diff, qs = [0.11, 1.3], [0.27, 6.0, mc]
    for hyb, part in ((sh.GAUSSIAN_COSINE, sh.GAUSSIANLAW), (sh.HAGFORS_COSINE, sh.HAGFORSLAW),
                      (sh.COSINE_COSINE, sh.COSINELAW_QS)):
        for mu in (0.05, 0.3, math.nextafter(mc, 0.0), mc, 0.8, 0.999, 1.0):
            h = _law(orc, sc, hyb, diff + qs)(mu)
            d = _law(orc, sc, sh.COSINELAW_DIFF, diff)(mu)
            q = _law(orc, sc, part, qs)(mu)
            assert h == pytest.approx(d + q, rel=1e-15, abs=0.0), (hyb, mu)


def test_smearing_identical_views_equal_one_view(orc):
    """calc_fits.c's smearing averages the fits of a frame's views: three views
    at the same epoch give the one-view fit (to rounding), and views spread
    over time give a different one."""
    base = dict(n_deldop=2, n_doppler=2, nbin_min=30, nbin_max=40)
    one = sh.Scene(sh.spec_for(1, **base))
    same = sh.Scene(sh.spec_for(1, radar_nviews=3, view_interval_days=0.0, **base))
    spread = sh.Scene(sh.spec_for(1, radar_nviews=3, view_interval_days=0.02, **base))
    for sc in (one, same, spread):
        H.oracle_forward(orc, sc)
        assert orc.oracle_last_error() in (None, b"")
    for s, f in one.all_frames():
        if f < 0:
            continue
        a, b, c = one.image(s, f, 0), same.image(s, f, 0), spread.image(s, f, 0)
        assert np.max(np.abs(b - a)) <= 1e-14 * np.max(np.abs(a)), (s, f)
        assert np.max(np.abs(c - a)) > 1e-6 * np.max(np.abs(a)), (s, f)


def test_lightcurve_smearing_averages_views(orc):
    """A lightcurve point smeared over views averages the spline at the views'
    epochs (calc_fits.c's lightcurve branch): views at one epoch give the
    one-view fit, views spread in time another."""
    base = dict(n_deldop=1, n_doppler=0, n_lightcurves=1, lc_points=20, lc_ncalc=8)
    one = sh.Scene(sh.spec_for(3, **base))
    same = sh.Scene(sh.spec_for(3, lc_nviews=3, view_interval_days=0.0, **base))
    spread = sh.Scene(sh.spec_for(3, lc_nviews=3, view_interval_days=0.05, **base))
    for sc in (one, same, spread):
        H.oracle_forward(orc, sc)
        assert orc.oracle_last_error() in (None, b"")
    s = next(q for q in range(one.nsets) if one.set_type(q) == sh.LGHTCRV)
    a, b, c = one.image(s, 0, 0), same.image(s, 0, 0), spread.image(s, 0, 0)
    assert np.max(np.abs(b - a)) <= 1e-14 * np.max(np.abs(a))
    assert np.max(np.abs(c - a)) > 1e-6 * np.max(np.abs(a))


def test_deldop_power_conservation(orc):
    """sum(fit) = kmpp^2 * sum over lit pixels of radlaw(cos) when every footprint
    lies inside the frame (long code: codefactor 1)."""
    sc = _realized(orc, 1)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    ii, jj, z, P = _pixel_power(sc, 0, 0)
    assert sc.image(0, 0, 0).sum() == pytest.approx(P.sum(), rel=1e-12)


def _dd_coords(sc, s, f, ii, jj, z):
    """Delay bin from the round-trip range (2/c per km toward the radar) and Doppler
    bin from the line-of-sight velocity of the rotating surface, (w x r)_z."""
    k, v = sc.consts(s, f), sc.view(s, f)
    n, kmpp = sc.spec.pos_n, sc.spec.km_per_pixel
    w = v[9:18].reshape(3, 3) @ v[18:21]           # apparent spin in observer coords
    x, y = (ii - n) * kmpp, (jj - n) * kmpp
    vlos = w[0] * y - w[1] * x                      # (w x r)_z in km/day
    delpos = -KM2US * z / k["del_per_pixel"] + k["delcom_vig"] + v[24]
    dop = KM2HZFACT * k["Ftx"] * vlos / k["dop_per_pixel"] + k["dopcom_vig"] + v[25]
    return delpos, dop


def test_deldop_marginals(orc):
    """Row sums: sum_pixels P TRI2(delPOS - i)/sum_i TRI2; column sums:
    sum_pixels P sinc^2(dopPOS - j)/sum_j sinc^2 over the six nearest bins --
    each marginal is independent of the other axis' response."""
    sc = _realized(orc, 2, pos_scope=sh.LOCAL, n_deldop=6)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    for s, f in sc.all_frames():
        ii, jj, z, P = _pixel_power(sc, s, f)
        delpos, dop = _dd_coords(sc, s, f, ii, jj, z)
        fit = sc.image(s, f, 0)
        ndel, ndop = fit.shape
        rows, cols = np.zeros(ndel), np.zeros(ndop)
        for p, d, q in zip(P, delpos, dop):
            ids = np.arange(int(math.floor(d - 1)) + 1, int(math.ceil(d + 1)))
            wd = np.clip(1 - np.abs(d - ids), 0, None) ** 2
            jd = np.arange(int(math.floor(q + 1 - 3)), int(math.floor(q + 3)) + 1)
            arg = np.pi * np.abs(q - jd)
            wj = np.where(arg > 1e-4, (np.sin(arg) / np.where(arg > 0, arg, 1)) ** 2, 1.0)
            if ids.min() < 1 or ids.max() > ndel or jd.min() < 1 or jd.max() > ndop:
                continue
            rows[ids - 1] += p * wd / wd.sum()
            cols[jd - 1] += p * wj / wj.sum()
        assert np.allclose(fit.sum(axis=1), rows, rtol=0, atol=1e-12 * rows.max())
        assert np.allclose(fit.sum(axis=0), cols, rtol=0, atol=1e-12 * cols.max())


def _delay_rows(d, spb, stride, code, ndel):
    """Response of every image row to a pixel at fractional row d, from its
    definition (pos2deldop.c:371-424).  Long code, original reduction: each
    sample per baud reduced on its own and the rows interleaved, a squared
    unit triangle of half-width spb/stride rows.  Otherwise: spb unit
    triangles of that half-width, the m-th centred m/stride - (spb-1)/(2
    stride) rows after the row centre, summed, divided by spb and squared.
    Every row of the frame is evaluated (zero outside the support)."""
    hw = spb // stride
    rows = np.arange(1, ndel + 1, dtype=float)

    def tri(u):
        return np.clip(1.0 - np.abs(u), 0.0, None)
    if code == sh.LONG_ORIG:
        return tri((d - rows) / hw) ** 2
    c2 = (spb - 1) / (2.0 * stride)
    acc = sum(tri((d - (rows - c2 + m / stride)) / hw) for m in range(spb))
    return (acc / spb) ** 2


def _sinc2(u):
    a = np.pi * np.abs(u)
    return np.where(a > 1e-4, (np.sin(a) / np.where(a > 0, a, 1.0)) ** 2, 1.0)


@pytest.mark.parametrize("spb,stride,code,ns", [(2, 1, sh.LONG_MOD, 1), (2, 1, sh.LONG_ORIG, 1),
                                                (3, 1, sh.LONG_MOD, 1), (4, 2, sh.LONG_MOD, 1),
                                                (2, 2, sh.LONG_ORIG, 1), (1, 1, sh.LONG_MOD, 2),
                                                (2, 1, sh.LONG_MOD, 3)])
def test_deldop_marginals_samples_and_subpixels(orc, spb, stride, code, ns):
    """Row and column sums of a delay-Doppler fit with several samples per
    baud and nsinc2 x nsinc2 sinc^2 points per pixel.  Rows: each lit
    pixel's power spread by the delay response written from its definition
    (_delay_rows).  Columns: the mean of sinc^2 over the pixel's ns x ns
    subpixel points, evenly spaced ((a + 1/2)/ns - 1/2 pixels from the
    center in x and y, their Doppler from (w x r) at the subpoint), over the
    sinc2width = 6 nearest bins widened by the subpoints' largest Doppler
    offset.  Delay and Doppler bins from the range and the line-of-sight
    velocity (_dd_coords)."""
    sc = _realized(orc, 1, spb=spb, stride=stride, codemethod=code, nsinc2=ns)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    ii, jj, z, P = _pixel_power(sc, 0, 0)
    delpos, dop = _dd_coords(sc, 0, 0, ii, jj, z)
    # Doppler bins per pixel step in x and y, from two neighbouring centers
    _, dx = _dd_coords(sc, 0, 0, ii + 1, jj, z)
    _, dy = _dd_coords(sc, 0, 0, ii, jj + 1, z)
    ax, ay = dx - dop, dy - dop
    fit = sc.image(0, 0, 0)
    ndel, ndop = fit.shape
    rows, cols = np.zeros(ndel), np.zeros(ndop)
    offs = [(a + 0.5) / ns - 0.5 for a in range(ns)]
    cbins = np.arange(1, ndop + 1, dtype=float)
    used = 0
    for p, d, q, gx, gy in zip(P, delpos, dop, ax, ay):
        wd = _delay_rows(d, spb, stride, code, ndel)
        sub = np.array([q + u * gx + v * gy for u in offs for v in offs])
        dmax = np.max(np.abs(sub - q))
        j0, j1 = int(math.floor(q - dmax + 1 - 3)), int(math.floor(q + dmax + 3))
        nz = np.nonzero(wd)[0]
        if j0 < 1 or j1 > ndop or nz.min() == 0 or nz.max() == ndel - 1:
            continue
        wp = np.where((cbins >= j0) & (cbins <= j1), np.mean(_sinc2(sub[:, None] - cbins[None, :]), axis=0), 0.0)
        rows += p * wd / wd.sum()
        cols += p * wp / wp.sum()
        used += 1
    assert used > 1000
    assert np.allclose(fit.sum(axis=1), rows, rtol=0, atol=1e-11 * rows.max())
    assert np.allclose(fit.sum(axis=0), cols, rtol=0, atol=1e-11 * cols.max())


def test_badradar_from_sinc_width(orc):
    """More sinc^2 evaluation bins than MAXBINS = 100 flags badradar with
    log((2 dopdiff_max + width + 1)/100) per frame."""
    sc = _realized(orc, 1)
    sc.par_set("sinc2width", 120)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    fl = sc.par_flags()
    assert fl["badradar"] == 1
    # nsinc2 = 1: dopdiff_max = 0; par factor = dof * log(121/100) / dof
    assert fl["badradar_logfactor"] == pytest.approx(math.log(121 / 100), rel=1e-14)


def test_short_code_factor(orc):
    """Short code: each pixel's power is scaled by sinc^2((dopPOS - dopDC)/fftlen)."""
    a = _realized(orc, 1)
    b = _realized(orc, 1, codemethod=sh.SHORT)
    orc.oracle_calc_fits(a.par, a.mod, a.dat)
    orc.oracle_calc_fits(b.par, b.mod, b.dat)
    ii, jj, z, P = _pixel_power(b, 0, 0)
    delpos, dop = _dd_coords(b, 0, 0, ii, jj, z)
    k = b.consts(0, 0)
    arg = np.pi * np.abs((dop - k["dopDC_vig"]) / k["dopfftlen"])
    cf = np.where(arg > 1e-4, (np.sin(arg) / arg) ** 2, 1.0)
    assert b.image(0, 0, 0).sum() == pytest.approx(np.sum(P * cf), rel=1e-12)
    assert a.image(0, 0, 0).sum() == pytest.approx(P.sum(), rel=1e-12)


# ------------------------------------------------------------------ a9 Doppler
def test_doppler_overflow_conservation(orc):
    """In-bounds pixels feed the spectrum, out-of-bounds pixels the overflow
    spectrum: together they hold the whole cross section."""
    sc = _realized(orc, 1, n_deldop=0, n_doppler=1, nbin_min=30, nbin_max=30,
                   dop_per_bin_min=0.5, dop_per_bin_max=0.5)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    ii, jj, z, P = _pixel_power(sc, 0, 0)
    sv = sc.scalars(0, 0)
    assert sv["overflow_xsec"] > 0
    assert sc.image(0, 0, 0).sum() + sv["overflow_xsec"] == pytest.approx(P.sum(), rel=1e-12)
    assert sv["overflow_o2"] == int(sv["overflow_o2"]) and sv["overflow_o2"] > 0
    assert sv["overflow_m2"] > 0


def test_doppler_spectrum_marginal(orc):
    """A Doppler spectrum equals the column sums of the delay-Doppler image of the
    same view (same pixels, same sinc^2 response)."""
    sc = _realized(orc, 1, n_doppler=1, nbin_min=80, nbin_max=80, dop_per_bin_min=0.5,
                   dop_per_bin_max=0.5)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    ii, jj, z, P = _pixel_power(sc, 1, 0)
    delpos, dop = _dd_coords(sc, 1, 0, ii, jj, z)
    spec = sc.image(1, 0, 0)
    cols = np.zeros(spec.size)
    for p, q in zip(P, dop):
        jd = np.arange(int(math.floor(q + 1 - 3)), int(math.floor(q + 3)) + 1)
        arg = np.pi * np.abs(q - jd)
        wj = np.where(arg > 1e-4, (np.sin(arg) / np.where(arg > 0, arg, 1)) ** 2, 1.0)
        if jd.min() >= 1 and jd.max() <= spec.size:
            cols[jd - 1] += p * wj / wj.sum()
    assert np.allclose(spec, cols, rtol=0, atol=1e-12 * cols.max())


# ------------------------------------------------------------------ a6, a10 photometry
def _lc_scene(orc, wt, phase_deg=0.0, r=0.6, **over):
    kw = dict(n_deldop=0, n_lightcurves=1, lc_points=20, lc_ncalc=8, rdev_sigma=0.0,
              semi_axes=(r, r, r), kaas_wt=wt, phase_min_deg=phase_deg, phase_max_deg=phase_deg,
              pos_n=100, km_per_pixel=0.008)
    kw.update(over)
    return _realized(orc, 1, **kw)


def test_lambert_disk_integral(orc):
    """Sphere at zero phase, Lambert (Kaasalainen wt = 1): the disk integral of
    cos(i) is (2/3) pi r^2, so sum b -> (2/3) R r^2 / AU^2."""
    r = 0.6
    sc = _lc_scene(orc, 1.0, 0.0, r)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    assert np.allclose(sc.image(0, 0, 3)[1:], (2.0 / 3.0) * r * r / AU ** 2, rtol=5e-3)


def test_lommel_seeliger_disk_integral(orc):
    """Zero phase, Lommel-Seeliger (wt = 0): b = I R/(4 pi) * 1/2 on every lit
    pixel, so sum b -> R r^2 / (8 AU^2)."""
    r = 0.6
    sc = _lc_scene(orc, 0.0, 0.0, r)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    assert np.allclose(sc.image(0, 0, 3)[1:], r * r / (8 * AU ** 2), rtol=5e-3)


def test_kaasalainen_is_linear_mix(orc):
    ys = []
    for wt in (0.0, 1.0, 0.3):
        sc = _lc_scene(orc, wt, 20.0)
        orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
        ys.append(sc.image(0, 0, 3)[1:])
    assert np.allclose(ys[2], 0.7 * ys[0] + 0.3 * ys[1], rtol=1e-12)


def test_posmask_zero_phase_shadows_nothing(orc):
    """Sun behind the observer: no pixel is shadowed (cos(emission) survives on
    every lit pixel), even on a bumpy model."""
    sc = _lc_scene(orc, 1.0, 0.0, rdev_sigma=0.08)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    pos = sc.pos(0, 7)
    cose, cosi, fac = sc.pos_array(pos, 1), sc.pos_array(pos, 2), sc.pos_array(pos, 10)
    lit = (fac >= 0) & (cosi > 1e-6)
    assert np.all(cose[lit] > 0)


def test_posmask_shadows_at_large_phase(orc):
    """A bumpy model at 60 deg phase casts shadows: lit, Earth-facing pixels with
    cos(emission) zeroed by the mask."""
    sc = _lc_scene(orc, 1.0, 60.0, rdev_sigma=0.08)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    pos = sc.pos(0, 7)
    cose, cosi, fac = sc.pos_array(pos, 1), sc.pos_array(pos, 2), sc.pos_array(pos, 10)
    assert np.sum((fac >= 0) & (cosi > 0.05) & (cose == 0)) > 0


def test_posmask_interpolation_exact_on_a_plane(orc):
    """Shadowing decided by the interpolated source depth, one-sided beside
    blank sky included: on a planar source z-buffer both interpolations are
    exact, so the mask is known pixel by pixel (a nearest-pixel depth, a
    transposed neighbour or a wrong sign on the fallback side flips pixels
    whose margin is 1e-7)."""
    sc = sh.Scene(sh.spec_for(1))
    pos = sc.pos(0, 0)
    for tol, seed in ((0.0, 1), (0.05, 2)):
        want, eligible, nk = H.posmask_plane_case(sc, pos, tol, seed)
        assert nk > 100 and eligible.sum() > 1000
        orc.oracle_posmask(pos, tol)
        got = sc.pos_array(pos, 1)
        assert np.array_equal(got, want), np.argwhere(got != want)[:5]


# ------------------------------------------------------------------ a11 spline
def test_spline_matches_scipy_natural(orc):
    rng = np.random.default_rng(7)
    n = 40
    x = np.sort(rng.uniform(0, 10, n))
    y = np.sin(x) + 0.1 * rng.normal(size=n)
    xs, ys = np.concatenate([[0.0], x]), np.concatenate([[0.0], y])
    y2 = np.zeros(n + 1)
    dp = C.POINTER(C.c_double)
    orc.oracle_spline(xs.ctypes.data_as(dp), ys.ctypes.data_as(dp), n, 2e30, 2e30, y2.ctypes.data_as(dp))
    ref = CubicSpline(x, y, bc_type="natural")
    assert np.allclose(y2[1:], ref(x, 2), rtol=1e-9, atol=1e-10)
    for t in np.linspace(x[0], x[-1], 97):
        out = C.c_double()
        orc.oracle_splint(xs.ctypes.data_as(dp), ys.ctypes.data_as(dp), y2.ctypes.data_as(dp), n, t, C.byref(out))
        assert out.value == pytest.approx(float(ref(t)), rel=1e-10, abs=1e-12)


# ------------------------------------------------------------------ a13 chi-square
def test_chi2_residual_form(orc):
    """chi2 of each frame equals sum w (obs - c fit)^2 at the least-squares
    calibration c = sum w obs fit / sum w fit^2 (a different algebraic form)."""
    sc = H.make_scene_with_obs(2)
    sc.set_param(0, 1.01)
    chi = H.oracle_forward(orc, sc)
    tot = 0.0
    for s, f in sc.all_frames():
        fit, obs, w = sc.image(s, f, 0), sc.image(s, f, 1), sc.image(s, f, 2)
        c = np.sum(w * obs * fit) / np.sum(w * fit * fit)
        r = np.sum(w * (obs - c * fit) ** 2)
        sv = sc.scalars(s, f)
        assert sv["cal"] == pytest.approx(c, rel=1e-12)
        assert sv["chi2"] == pytest.approx(r, rel=1e-9)
        tot += r
    assert chi == pytest.approx(tot, rel=1e-10)


def test_chi2_zero_for_noiseless(orc):
    sc = sh.Scene(sh.spec_for(1))
    H.oracle_forward(orc, sc)
    fit = sc.image(0, 0, 0)
    sc.set_image(0, 0, 1, 3.7 * fit)
    chi = orc.oracle_chi2(sc.par, sc.dat, 0)
    assert abs(chi) < 1e-12 * np.sum((3.7 * fit) ** 2)
    assert sc.scalars(0, 0)["cal"] == pytest.approx(3.7, rel=1e-14)


def test_chi2_negative_calibration_branch(orc):
    sc = sh.Scene(sh.spec_for(1))
    H.oracle_forward(orc, sc)
    fit = sc.image(0, 0, 0)
    sc.set_image(0, 0, 1, -fit)
    chi = orc.oracle_chi2(sc.par, sc.dat, 0)
    assert sc.scalars(0, 0)["cal"] == 1e-40
    assert chi == pytest.approx(np.sum(fit ** 2), rel=1e-12)


def test_lightcurve_chi2(orc):
    sc = H.make_scene_with_obs(1, n_deldop=0, n_lightcurves=2, lc_points=30, lc_ncalc=10)
    chi = H.oracle_forward(orc, sc)
    tot = 0.0
    for s in range(sc.nsets):
        fit, obs, w = sc.image(s, 0, 0), sc.image(s, 0, 1), sc.image(s, 0, 2)
        c = np.sum(w * obs * fit) / np.sum(w * fit * fit)
        tot += np.sum(w * (obs - c * fit) ** 2)
    assert chi == pytest.approx(tot, rel=1e-9)


# ------------------------------------------------------------------ a14 objective
def test_objective_bad_factor(orc):
    sc = H.make_scene_with_obs(1)
    sc2 = sh.Scene(sh.spec_for(1, semi_axes=(1.4, 1.2, 0.6)))
    H.copy_obs(sc, sc2)
    H.oracle_forward(orc, sc2)
    obj = orc.oracle_objective_tail(sc2.par, sc2.dat)
    fl = sc2.par_flags()
    chi = orc.oracle_chi2(sc2.par, sc2.dat, 0)
    assert fl["posbnd"] == 1
    assert obj == pytest.approx(chi / 8000.0 * 2.0 * math.exp(fl["posbnd_logfactor"]), rel=1e-14)


# ------------------------------------------------------------------ HMT baseline
def test_threaded_baseline_matches_serial(orc):
    a = H.make_scene_with_obs(2, pos_scope=sh.LOCAL)
    b = sh.Scene(sh.spec_for(2, pos_scope=sh.LOCAL))
    H.copy_obs(a, b)
    orc.oracle_realize_mod(b.par, b.mod)
    orc.oracle_realize_spin(b.par, b.mod, b.dat)
    orc.oracle_realize_delcor(b.dat, 0.0, 0)
    orc.oracle_calc_fits_threads(b.par, b.mod, b.dat, 4)
    assert orc.oracle_chi2(a.par, a.dat, 0) == orc.oracle_chi2(b.par, b.dat, 0)
    for s, f in a.all_frames():
        assert np.array_equal(a.image(s, f, 0), b.image(s, f, 0))
