"""Pins of the oracle's photometric laws beyond the cosine and Lambert
families: Hapke's reflectance (oracle_hapke, astro/hapke.c:41-176), the
tabular radar law (radlaw.c:56-66), and realize_photo's local values
(oracle_realize_photo, realize_photo.c:122-969).

Hapke's function is fixed here by what its physics and construction require,
not by its own formula: the reflectance is reciprocal (r(i, e) / cos i =
r(e, i) / cos e) with roughness on, continuous across its two roughness
branches at i = e and as the mean slope goes to zero; its smooth-surface
multiple-scattering factor is within Hapke 2002's stated 1% of Chandrasekhar's
H function solved from its integral equation; the Henyey-Greenstein phase
function it carries integrates to one over the sphere; the opposition surge
falls to half at tan(phase / 2) = h.  The harmonic laws' local values are
checked against scipy's spherical harmonics at the facet centers, the
associated Legendre functions against scipy.special.lpmv, the tabular law
against a linear table (whose interpolation is exact), and whole chains
against laws they reduce to (an inhomogeneous cosine law whose every facet
takes the global values, a harmonic Lambert law of degree 0).
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate
from scipy.special import lpmv, sph_harm_y

import helpers as H
import test_io as TI

sh = H.sh
D = C.c_double


def _orc():
    orc = H.load_oracle()
    orc.oracle_hapke.argtypes = [D] * 8
    orc.oracle_hapke.restype = D
    orc.oracle_plm.argtypes = [C.c_int, C.c_int, D]
    orc.oracle_plm.restype = D
    orc.oracle_realize_photo.argtypes = [C.c_void_p, C.c_void_p]
    orc.oracle_realize_photo.restype = None
    orc.oracle_radlaw.argtypes = [C.c_void_p, C.c_int, D, C.c_int, C.c_int]
    orc.oracle_radlaw.restype = D
    return orc


def hapke(ci, ce, phase, w, h=0.1, B0=0.0, g=0.0, theta=0.0):
    return _orc().oracle_hapke(ci, ce, phase, w, h, B0, g, theta)


def _phase(i, e, az):
    """phase angle of incidence i and emission e with azimuth az between them"""
    return math.acos(math.cos(i) * math.cos(e) + math.sin(i) * math.sin(e) * math.cos(az))


# ---- Hapke ------------------------------------------------------------------
def test_hapke_is_reciprocal_with_roughness():
    for i, e, az in [(0.3, 1.1, 0.4), (1.2, 0.2, 2.5), (0.7, 0.9, 1.6), (0.05, 1.4, 3.0)]:
        ph = _phase(i, e, az)
        kw = dict(w=0.6, h=0.08, B0=1.2, g=-0.3, theta=math.radians(25))
        a = hapke(math.cos(i), math.cos(e), ph, **kw)
        b = hapke(math.cos(e), math.cos(i), ph, **kw)
        assert a > 0
        assert a * math.cos(e) == pytest.approx(b * math.cos(i), rel=1e-12), (i, e)


def test_hapke_roughness_branches_meet_at_equal_angles():
    kw = dict(w=0.5, h=0.1, B0=0.5, g=0.2, theta=math.radians(30))
    for i, az in [(0.4, 0.5), (1.0, 2.0), (1.3, 1.0)]:
        d = 1e-8
        lo = hapke(math.cos(i), math.cos(i - d), _phase(i, i - d, az), **kw)
        hi = hapke(math.cos(i), math.cos(i + d), _phase(i, i + d, az), **kw)
        assert abs(hi - lo) <= 1e-6 * lo, (i, lo, hi)


def test_hapke_smooth_limit_and_invalid_parameters():
    ci, ce, ph = math.cos(0.5), math.cos(0.8), 0.9
    smooth = hapke(ci, ce, ph, 0.4, 0.1, 0.3, 0.1, 0.0)
    rough = hapke(ci, ce, ph, 0.4, 0.1, 0.3, 0.1, 1e-7)
    assert rough == pytest.approx(smooth, rel=1e-5)
    assert abs(hapke(ci, ce, ph, 0.4, 0.1, 0.3, 0.1, 0.3) - smooth) > 1e-3 * smooth   # roughness matters
    for bad in [dict(w=0.0), dict(w=1.0), dict(w=0.5, h=0.0), dict(w=0.5, B0=-0.1), dict(w=0.5, g=1.0),
                dict(w=0.5, theta=math.pi / 2), dict(w=0.5, theta=-0.1)]:
        assert hapke(ci, ce, ph, **bad) == 0.0, bad
    assert hapke(ci, ce, -0.1, 0.5) == 0.0 and hapke(ci, ce, 3.2, 0.5) == 0.0


def _chandrasekhar_h(w, mus, n=400):
    """H(mu) for isotropic scattering of albedo w, from
    1/H(mu) = 1 - (w/2) mu int_0^1 H(t) / (mu + t) dt, iterated on
    Gauss-Legendre nodes"""
    x, wt = np.polynomial.legendre.leggauss(n)
    t, wt = 0.5 * (x + 1), 0.5 * wt
    Ht = np.ones(n)
    for _ in range(500):
        new = 1 / (1 - 0.5 * w * t * np.sum(wt * Ht / (t[:, None] + t[None, :]), axis=1))
        if np.max(np.abs(new - Ht)) < 1e-14:
            break
        Ht = new
    mus = np.asarray(mus, dtype=float)
    return 1 / (1 - 0.5 * w * mus * np.sum(wt * Ht / (mus[:, None] + t[None, :]), axis=1))


def test_hapke_h_function_within_one_percent_of_chandrasekhar():
    for w in (0.3, 0.7, 0.95):
        mus = [0.1, 0.3, 0.6, 0.9]
        exact = _chandrasekhar_h(w, mus)
        for mu, hx in zip(mus, exact):
            # smooth, isotropic, no surge, mu0 = mu: r = (w / (8 pi)) H(mu)^2
            r = hapke(mu, mu, 0.0, w, 0.1, 0.0, 0.0, 0.0)
            happrox = math.sqrt(r * 8 * math.pi / w)
            assert abs(happrox / hx - 1) < 0.01, (w, mu, happrox, hx)
            assert happrox != hx


def test_hapke_phase_function_is_normalised():
    ci, ce = math.cos(0.6), math.cos(0.3)
    pref = 0.55 * ci / (4 * math.pi * (ci + ce))
    for g in (-0.4, 0.3):
        def P(a):
            return (hapke(ci, ce, a, 0.55, 0.1, 0.0, g, 0.0) - hapke(ci, ce, a, 0.55, 0.1, 0.0, 0.0, 0.0)) / pref + 1
        norm = 0.5 * integrate.quad(lambda a: P(a) * math.sin(a), 0, math.pi, epsabs=1e-12, limit=200)[0]
        assert norm == pytest.approx(1.0, abs=1e-8), g
        assert P(0.0) == pytest.approx((1 - g) / (1 + g) ** 2, rel=1e-10)


def test_hapke_opposition_surge_half_width():
    ci, ce, B0, h = math.cos(0.4), math.cos(0.5), 0.8, 0.06
    pref = 0.5 * ci / (4 * math.pi * (ci + ce))

    def B(a):
        return (hapke(ci, ce, a, 0.5, h, B0, 0.0, 0.0) - hapke(ci, ce, a, 0.5, h, 0.0, 0.0, 0.0)) / pref
    assert B(0.0) == pytest.approx(B0, rel=1e-10)
    assert B(2 * math.atan(h)) == pytest.approx(B0 / 2, rel=1e-10)
    assert B(2.5) < B0 / 10


# ---- Legendre functions and realize_photo -----------------------------------
def test_plm_matches_scipy():
    orc = _orc()
    for l in range(9):
        for m in range(l + 1):
            for x in (-0.93, -0.2, 0.0, 0.41, 0.99):
                assert orc.oracle_plm(l, m, x) == pytest.approx(float(lpmv(m, l, x)), rel=1e-12, abs=1e-12)


def harm_text(name, expansions):
    """a harmonic law's text: each expansion as (degree, values in the file's
    order a00 a10 a11 b11 ...), every coefficient free"""
    parts = [name]
    for L, vals in expansions:
        assert len(vals) == (L + 1) ** 2
        parts.append("%d " % L + " ".join("f %.17g" % v for v in vals))
    return "  ".join(parts)


def inho_text(name, glob, locs):
    """an inhomogeneous law's text: global (state, value) pairs, then one
    component with a list of (state, value) pairs per facet"""
    g = " ".join("%s %.17g" % sv for sv in glob)
    body = " ".join(" ".join("%s %.17g" % sv for sv in fv) for fv in locs)
    return "%s %s 1 %d %s" % (name, g, len(locs), body)


def _job(tmp_path, monkeypatch, **laws):
    monkeypatch.chdir(tmp_path)
    files = TI.write_job(".", **laws)[0]
    return files, sh.Job(*files)


NF = 2292  # facets of the test job's model


def _coef_to_lm(L):
    out = []
    for l in range(L + 1):
        out.append((l, 0, "a"))
        for m in range(1, l + 1):
            out += [(l, m, "a"), (l, m, "b")]
    return out


def test_realize_photo_harmonic_locals_are_real_spherical_harmonics(tmp_path, monkeypatch):
    rng = np.random.default_rng(7)
    L = 3
    coefs = rng.uniform(-0.05, 0.05, (L + 1) ** 2)
    coefs[0] = 0.3
    files, job = _job(tmp_path, monkeypatch, optical_law=harm_text("harmlambert", [(L, coefs)]))
    orc = _orc()
    orc.oracle_realize_mod(job.par, job.mod)
    orc.oracle_realize_photo(job.par, job.mod)
    assert not orc.oracle_last_error()
    for f in range(0, NF, 97):
        g = job.facet_geom(f)
        x = g[:3]
        th, ph = math.atan2(math.hypot(x[0], x[1]), x[2]), math.atan2(x[1], x[0])
        assert g[3] == pytest.approx(th, abs=1e-15) and g[4] == pytest.approx(ph, abs=1e-15)
        want = 0.0
        for c, (l, m, ab) in zip(coefs, _coef_to_lm(L)):
            y = complex(sph_harm_y(l, m, th, ph)) * math.sqrt(4 * math.pi)
            want += c * (y.real if ab == "a" else y.imag)
        assert job.law_local(0, 0, f)[0] == pytest.approx(want, rel=1e-12, abs=1e-14), f
    assert job.badphoto() == (0, 0.0)
    job.close()


def test_realize_photo_ties_and_bounds(tmp_path, monkeypatch):
    locs = [[("f", 0.5 + 1e-4 * f)] if f % 3 else [("=", 0.0)] for f in range(NF)]
    locs[5] = [("f", -0.1)]                          # R <= 0: bad (mode 3)
    rho = [("f", 0.9), ("=", 0.0), ("f", 0.5), ("=", 0.0), ("f", 0.2), ("f", 0.0)]   # rho 0 at the limit: bad
    files, job = _job(tmp_path, monkeypatch, optical_law=inho_text("inholommel", [("c", 0.37)], locs),
                      radar_law="tabular %d %s" % (len(rho), " ".join("%s %g" % sv for sv in rho)))
    orc = _orc()
    orc.oracle_realize_mod(job.par, job.mod)
    orc.oracle_realize_photo(job.par, job.mod)
    for f in (0, 3, 6, 1, 2, 5):
        want = 0.37 if f % 3 == 0 else (-0.1 if f == 5 else 0.5 + 1e-4 * f)
        assert job.law_local(0, 0, f)[0] == want, f
    assert list(job.law_table(0)) == [0.9, 0.9, 0.5, 0.5, 0.2, 0.0]
    bad, logf = job.badphoto()
    assert bad == 1
    assert logf == pytest.approx(math.log(1.1) + math.log(1.0), rel=1e-15)
    job.close()


def test_tabular_law_interpolates_a_linear_table_exactly(tmp_path, monkeypatch):
    n, a, b = 7, 0.8, -0.3
    step = (math.pi / 2) / (n - 1)
    rho = [a + b * i * step for i in range(n)]
    files, job = _job(tmp_path, monkeypatch, radar_law="tabular %d %s" % (n, " ".join("c %.17g" % r for r in rho)))
    orc = _orc()
    photo = job.lib.shp_mod_photo(job.mod)
    for inc in (0.0, 0.1, 0.5, step, 1.2, 1.5, math.pi / 2 - 1e-9):
        c = math.cos(inc)
        got = orc.oracle_radlaw(photo, 0, c, 0, 3)
        assert got == pytest.approx((a + b * math.acos(c)) / c, rel=1e-13), inc
    assert orc.oracle_radlaw(photo, 0, 0.5, 0, -1) == 0.0     # blank sky
    job.close()


def _chain(files, photo=True):
    orc = _orc()
    job = sh.Job(*files)
    orc.oracle_realize_mod(job.par, job.mod)
    orc.oracle_realize_spin(job.par, job.mod, job.dat)
    if photo:
        orc.oracle_realize_photo(job.par, job.mod)
    orc.oracle_realize_delcor(job.dat, 0.0, 0)
    orc.oracle_calc_fits(job.par, job.mod, job.dat)
    chi2 = orc.oracle_chi2(job.par, job.dat, 0)
    assert not orc.oracle_last_error()
    job.close()
    return chi2


def test_laws_reduce_to_their_homogeneous_forms(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    base = _chain(TI.write_job(".", radar_law="cosine c 0.2 c 1.5", optical_law="lambertian c 0.3")[0])
    inho = inho_text("inhocosine", [("c", 0.2), ("c", 1.5)], [[("=", 0.0), ("=", 0.0)]] * NF)
    harm = harm_text("harmlambert", [(1, [0.3, 0.0, 0.0, 0.0])])
    other = _chain(TI.write_job(".", radar_law=inho, optical_law=harm)[0])
    assert other == pytest.approx(base, rel=1e-13)
    # and a law that differs does move chi2
    moved = _chain(TI.write_job(".", radar_law=inho, optical_law=harm_text("harmlambert", [(1, [0.3, 0.05, 0, 0])]))[0])
    assert abs(moved - base) > 1e-6 * base


# ---- the penalties that read the laws (penalties.c:150-588) -----------------
def _pen(files, penalties):
    import test_oracle_pen as P
    orc = P._oracle()
    _orc()
    job = sh.Job(*files)
    orc.oracle_realize_mod(job.par, job.mod)
    orc.oracle_realize_photo(job.par, job.mod)
    pen = orc.oracle_penalties(job.par, job.mod, job.dat)
    err = orc.oracle_last_error()
    geo = job.geometry()
    fac = np.zeros(3 * NF, dtype=np.int32)
    job.lib.shp_facets(job.mod, fac.ctypes.data_as(C.POINTER(C.c_int)))
    job.close()
    return pen, err, geo[0], fac.reshape(NF, 3)


def _edges(x, fac):
    """each edge once: (length, facet a, facet b), from the facet list alone"""
    seen = {}
    for f, (a, b, c) in enumerate(fac):
        for u, v in ((a, b), (b, c), (c, a)):
            seen.setdefault((min(u, v), max(u, v)), []).append(f)
    return [(np.linalg.norm(x[u] - x[v]), fs[0], fs[1]) for (u, v), fs in seen.items()]


def test_albedo_variance_and_contrast_penalties(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    two = [[("f", 1.0 if f % 2 else 3.0)] for f in range(NF)]
    files = TI.write_job(".", optical_law=inho_text("inholambert", [("c", 2.0)], two), penalties=[("optalbvar", 1.0)])[0]
    pen, err, _, _ = _pen(files, None)
    assert not err
    assert pen == pytest.approx(0.25, rel=1e-12)          # variance 1 over mean 2, squared
    one = [[("f", 3.0 if f == 0 else 1.0)] for f in range(NF)]
    files = TI.write_job(".", optical_law=inho_text("inholommel", [("c", 1.0)], one), penalties=[("optalbdel", 1.0)])[0]
    pen, err, x, fac = _pen(files, None)
    assert not err
    val = np.where(np.arange(NF) == 0, 3.0, 1.0)
    e = _edges(x, fac)
    want = sum(l * abs(val[a] - val[b]) for l, a, b in e) / sum(l * abs(val[a] + val[b]) for l, a, b in e)
    assert pen == pytest.approx(want, rel=1e-12) and pen > 0
    flat = [[("=", 0.0), ("f", 1.5)] for f in range(NF)]
    files = TI.write_job(".", radar_law=inho_text("inhocosine", [("c", 0.2), ("c", 1.5)], flat),
                         penalties=[("radalbdel", 1.0), ("rad_c_var", 1.0), ("radalbvar", 1.0)])[0]
    pen, err, _, _ = _pen(files, None)
    assert not err and pen == pytest.approx(0.0, abs=1e-12)   # uniform R and C: no contrast, no variance
    #  (to the rounding of ntot sum(x^2) / (sum x)^2 - 1)


def test_noncosine_penalty(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    n, A, p = 7, 0.6, 2.4
    th = np.arange(n) * (math.pi / 2) / (n - 1)
    rho = A * np.cos(th[:-1]) ** p
    def run(vals):
        text = "tabular %d %s" % (n, " ".join("c %.17g" % v for v in vals))
        files = TI.write_job(".", radar_law=text, penalties=[("noncosine", 1.0)])[0]
        pen, err, _, _ = _pen(files, None)
        assert not err
        return pen
    assert run(list(rho) + [0.0]) == pytest.approx(0.0, abs=1e-28)       # a cosine power law fits exactly
    last = 0.5                                                            # the 90-degree point above its neighbour
    y = math.log10(last) - (p * math.log10(math.cos(th[-2])) + math.log10(A))
    assert run(list(rho) + [last]) == pytest.approx(y * y / n, rel=1e-10)
    bent = rho.copy()
    bent[2] *= 1.3
    lx, ly = np.log10(np.cos(th[:-1])), np.log10(bent)
    res = ly - np.polyval(np.polyfit(lx, ly, 1), lx)
    assert run(list(bent) + [0.0]) == pytest.approx(np.sum(res ** 2) / n, rel=1e-9)
