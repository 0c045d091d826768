"""The input-file readers (libshape_io.so): par, mod and obs files into the ABI
structs, as read_par.c, read_mod.c and read_dat.c read them.

Pins that do not come from the readers themselves:
* calendar dates against known Julian dates (J2000.0, the MJD epoch, the
  Gregorian reform, JD 0);
* the observing geometry against the ephemeris vectors computed here in numpy
  (line of sight, orthonormal frame, the obliquity rotation, the solar phase
  angle, the Sun's direction) and the plane-of-sky spin against a finite
  difference of the line of sight;
* the analytic components' tessellation: the vertex count of the ring formula
  (1,148 at theta_steps = 30, the count config 1 states), F = 2V - 4, a closed
  and consistently outward-oriented surface;
* the side and attached-facet lists against the scene builder's independent
  construction of the same ordering rule;
* the data layout end to end: observations written as the oracle's own fits
  of the read job, in every data format, read back to a chi-square at the
  FP32 rounding floor (a transposed, shifted or mis-vignetted image would not).
"""
import math
import os

import numpy as np
import pytest

import helpers as H
import shape_files as SF

sh = H.sh

T0 = 2458849.5 + 1.0 / 24  # 2020-01-01 01:00:00
RA, DEC, DIST = 30.0, 20.0, 0.05
SUN = (280.0, -23.0, 0.983)


def io():
    return sh.load_io()


# ---- calendar -----------------------------------------------------------------
@pytest.mark.parametrize("date,jd", [
    ((2000, 1, 1, 12, 0, 0), 2451545.0),          # J2000.0
    ((1858, 11, 17, 0, 0, 0), 2400000.5),         # MJD 0
    ((1582, 10, 15, 0, 0, 0), 2299160.5),         # first Gregorian day
    ((1582, 10, 4, 0, 0, 0), 2299159.5),          # last Julian-calendar day
    ((-4713, 1, 1, 12, 0, 0), 0.0),               # JD 0 (4713 BC, Julian calendar)
    ((2020, 2, 29, 6, 30, 15), 2458908.5 + (6 * 3600 + 30 * 60 + 15) / 86400.0),
])
def test_cal2jd_known_dates(date, jd):
    assert io().shape_io_cal2jd(*date) == pytest.approx(jd, abs=1e-9)


def test_cal2jd_inverts_writer():
    for jd in [T0, T0 + 17.3125, 2451545.0 + 0.123 * 86400 // 1 / 86400]:
        jd = round(jd * 86400) / 86400
        assert io().shape_io_cal2jd(*SF.jd_to_cal(jd)) == pytest.approx(jd, abs=1e-9)


# ---- ephemerides --------------------------------------------------------------
def _ephem(points):
    """an ephem_t in ctypes: (JD, RA deg, Dec deg, AU) points"""
    import ctypes as C

    class Pnt(C.Structure):
        _fields_ = [("t", C.c_double), ("ra", C.c_double), ("dec", C.c_double), ("dist", C.c_double)]

    class Eph(C.Structure):
        _fields_ = [("n", C.c_int), ("pnt", C.POINTER(Pnt))]

    arr = (Pnt * len(points))(*[Pnt(t, math.radians(ra), math.radians(dec), d) for t, ra, dec, d in points])
    e = Eph(len(points), C.cast(arr, C.POINTER(Pnt)))
    e._keep = arr
    return e


def _geom(ast, sun, t, bistatic):
    import ctypes as C
    oe, se, s = np.zeros((3, 3)), np.zeros((3, 3)), np.zeros(3)
    ph, az = C.c_double(), C.c_double()
    dp = C.POINTER(C.c_double)
    d = io().shape_io_ephem2mat(C.byref(ast), C.byref(sun), t, oe.ctypes.data_as(dp), se.ctypes.data_as(dp),
                                s.ctypes.data_as(dp), C.byref(ph), C.byref(az), bistatic)
    return d, oe, se, s, ph.value, az.value


EPS = math.radians(23 + 26 / 60 + 21.406 / 3600)   # J2000 obliquity (IAU 2006)


def _ecl(v):
    c, s = math.cos(EPS), math.sin(EPS)
    return np.array([v[0], v[1] * c + v[2] * s, v[2] * c - v[1] * s])


def _unit_radec(ra, dec):
    ra, dec = math.radians(ra), math.radians(dec)
    return np.array([math.cos(dec) * math.cos(ra), math.cos(dec) * math.sin(ra), math.sin(dec)])


def test_ephem2mat_fixed_direction():
    ast = _ephem([(T0 - 1, RA, DEC, DIST), (T0 + 1, RA, DEC, DIST)])
    d, oe, se, s, ph, az = _geom(ast, ast, T0 + 0.25, 0)
    assert d == pytest.approx(DIST, rel=1e-15)
    assert np.allclose(oe @ oe.T, np.eye(3), atol=1e-15)
    assert np.linalg.det(oe) == pytest.approx(1.0, abs=1e-14)
    # z: from the target toward the observer, in ecliptic coordinates
    assert np.allclose(oe[2], -_ecl(_unit_radec(RA, DEC)), atol=1e-15)
    # x lies in the celestial equator (its equatorial z is zero) and y points north
    c, sn = math.cos(EPS), math.sin(EPS)
    x_eq_z = oe[0][1] * sn + oe[0][2] * c
    assert abs(x_eq_z) < 1e-15
    y_eq_z = oe[1][1] * sn + oe[1][2] * c
    assert y_eq_z > 0
    assert np.all(s == 0.0) and ph == 0.0 and az == 0.0


def test_ephem2mat_spin_is_line_of_sight_rate():
    """orbital spin = (d e/dt) x e for the unit vector e target->observer"""
    ast = _ephem([(T0 - 1, RA - 3, DEC - 1, DIST), (T0 + 1, RA + 5, DEC + 2, DIST * 1.1)])
    t = T0 + 0.3
    _, oe, _, s, _, _ = _geom(ast, ast, t, 0)
    h = 1e-2  # days: epochs near JD 2.46e6 carry only ~5e-10 days of precision
    ep = _geom(ast, ast, t + h, 0)[1][2]
    em = _geom(ast, ast, t - h, 0)[1][2]
    edot = (ep - em) / (2 * h)
    assert np.linalg.norm(s - np.cross(edot, oe[2])) < 1e-6 * np.linalg.norm(s)
    assert np.linalg.norm(s) > 1e-3


def test_ephem2mat_ra_across_zero():
    """RA interpolated across 360/0 deg takes the short way"""
    ast = _ephem([(T0 - 1, 359.0, DEC, DIST), (T0 + 1, 1.0, DEC, DIST)])
    _, oe, _, _, _, _ = _geom(ast, ast, T0, 0)
    assert np.allclose(oe[2], -_ecl(_unit_radec(0.0, DEC)), atol=1e-12)


def test_ephem2mat_solar_geometry():
    ast = _ephem([(T0 - 1, RA, DEC, DIST), (T0 + 1, RA, DEC, DIST)])
    sun = _ephem([(T0 - 1,) + SUN, (T0 + 1,) + SUN])
    _, oe, se, _, ph, az = _geom(ast, sun, T0, 1)
    to_obs = -_unit_radec(RA, DEC) * DIST
    to_sun = _unit_radec(SUN[0], SUN[1]) * SUN[2] + to_obs
    phase = math.acos(np.dot(to_obs, to_sun) / np.linalg.norm(to_obs) / np.linalg.norm(to_sun))
    assert ph == pytest.approx(phase, abs=1e-13)
    # se's z row: the unit vector from the target to the Sun (ecliptic)
    assert np.allclose(se[2], _ecl(to_sun / np.linalg.norm(to_sun)), atol=1e-14)
    assert np.allclose(se @ se.T, np.eye(3), atol=1e-14)
    # azimuth: the Sun's direction in the plane of the sky, from north (+y) through east (-x)
    so = oe @ _ecl(to_sun)
    assert az == pytest.approx((math.atan2(so[1], so[0]) - math.pi / 2) % (2 * math.pi), abs=1e-12)


def test_ephem2mat_outside_ephemeris_is_an_error():
    ast = _ephem([(T0, RA, DEC, DIST), (T0 + 1, RA, DEC, DIST)])
    d = _geom(ast, ast, T0 - 0.5, 0)[0]
    assert d < 0
    assert b"before" in io().shape_io_error()


# ---- a complete job ------------------------------------------------------------
PAR = [("action", "fit"), ("pos_pixels", "151"), ("pos_width", "2.25"), ("term_prec", "1e-6"),
       ("length_step", "0.01"), ("length_tol", "0.001"), ("ratio_step", "0.01"), ("ratio_tol", "0.001"),
       ("angle_step", "1.0"), ("angle_tol", "0.001"), ("spin_step", "0.5"), ("spin_tol", "0.001"),
       ("perform_ltc", "no"), ("pos_scope", "global"), ("nsinc2", "1"), ("sinc2width", "6"),
       ("maskdir", "''"), ("delcor_step", "0.1 0.2 {usec, usec/day}"), ("dd_maxsides", "left-bottom")]


def config1_mesh():
    sc = sh.Scene(sh.spec_for(1))
    nv, nf = sc.nv, sc.nf
    import ctypes as C
    u, r = np.zeros(3 * nv), np.zeros(nv)
    sc.lib.shp_vertex_dirs(sc.mod, u.ctypes.data_as(C.POINTER(C.c_double)), r.ctypes.data_as(C.POINTER(C.c_double)))
    fac = np.zeros(3 * nf, dtype=np.int32)
    sc.lib.shp_facets(sc.mod, fac.ctypes.data_as(C.POINTER(C.c_int)))
    return sc, u.reshape(nv, 3), r, fac.reshape(nf, 3)


def write_job(d, data=None, fmts=("binary", "rdf", "ascii.gz"), ncalc=-1, mask_rows=(40, 61), mask_cols=(25, 56),
              radar_law="cosine c 0.2 f 1.5", optical_law="kaasalainen f 0.3 c 0.1 c 0.0 c 0.1 c 0.0",
              penalties=(("nonsmooth", 0.5), ("volume", 0.25))):
    """A job with two delay-Doppler frames (the second masked), one Doppler
    set of two smeared views and one lightcurve, as files under d (a path
    relative to the working directory: the structs hold file names in fields
    of 40 and 80 characters).  data maps (set, frame) to the values to write
    (FP32, in sdev units); the default is a ramp.  Returns the file names and
    the frame records."""
    _, u, r, fac = config1_mesh()
    spin = {"t0": T0 - 1.0 / 24, "angles": [("f", 120.0), ("f", 30.0), ("c", 10.0)],
            "omega": [("c", 0.0), ("c", 0.0), ("f", 1440.0)]}
    SF.write_mod(os.path.join(d, "job.mod"), [SF.vertex_component(u, r, fac)],
                 [radar_law], [optical_law], spin)
    SF.write_par(os.path.join(d, "job.par"), PAR, list(penalties))
    eph = [(T0 - 1, RA, DEC, DIST), (T0 + 2, RA, DEC, DIST)]
    os.makedirs(os.path.join(d, "dd"), exist_ok=True)
    os.makedirs(os.path.join(d, "dop"), exist_ok=True)
    ndel, ndop = 100, 80
    frames = []

    def values(key, n):
        if data is not None and key in data:
            return np.asarray(data[key], dtype=np.float32)
        return (np.arange(n, dtype=np.float32) % 97 + 1) / 50.0

    names = []
    for i, fmt in enumerate(fmts[:2]):
        name = "frame%d.%s" % (i, fmt)
        SF.write_values(os.path.join(d, "dd", name), values((0, i), ndel * ndop), fmt.replace(".gz", ""))
        names.append(name)
    mask = np.zeros((ndel, ndop))
    mask[mask_rows[0]:mask_rows[1], mask_cols[0]:mask_cols[1]] = 0.5
    mask[45, 30] = 0.0
    SF.write_mask(os.path.join(d, "dd", "mask_00_01.txt"), mask)
    dd = {"iradlaw": 0, "ephem": eph, "Ftx": 8560.0, "ndel": ndel, "del_per_pixel": 0.5, "spb": 1, "stride": 1,
          "code": "long_orig", "ndop": ndop, "dop_per_pixel": 0.5, "dopcom": 40.5, "dopDC": 40.5, "dopfftlen": 256,
          "delcor_t0": T0, "delcor": [("f", 0.25), ("c", 0.5)], "dopscale": ("c", 1.0), "nviews": 1,
          "view_interval_s": 0.0, "smearing": "center", "dir": os.path.join(d, "dd"),
          "frames": [{"name": names[0], "t0": T0 + 1.0 / 24, "sdev": 2.0, "cal": ("f", 1.0), "nlooks": 0,
                      "delcom": 50.5, "weight": 1.0, "mask": 0},
                     {"name": names[1], "t0": T0 + 3.5 / 24, "sdev": 1.5, "cal": ("f", 1.0), "nlooks": 0,
                      "delcom": 50.5, "weight": 0.75, "mask": 1}]}
    nbin = 100
    dname = "spec0." + fmts[2]
    SF.write_values(os.path.join(d, "dop", dname), values((1, 0), nbin), fmts[2].replace(".gz", ""))
    dp = {"iradlaw": 0, "ephem": eph, "Ftx": 8560.0, "ndop": nbin, "dop_per_bin": 0.5, "dopcom": 50.5,
          "delcor_t0": T0, "delcor": [("=", 0.25), ("=", 0.5)], "dopscale": ("=", 1.0), "nviews": 2,
          "view_interval_s": 600.0, "smearing": "center", "dir": os.path.join(d, "dop"),
          "frames": [{"name": dname, "t0": T0 + 2.0 / 24, "sdev": 1.0, "cal": ("f", 1.0), "nlooks": 0,
                      "weight": 1.0, "mask": 0}]}
    lc_t = T0 + np.array([0.1, 0.2, 0.2, 0.35, 0.5]) + 0.0
    lc_t = np.round(lc_t * 86400) / 86400
    lcv = values((2, 0), 5).astype(np.float64)
    mags = SF.mag_of(lcv * 1e-12)
    SF.write_lightcurve(os.path.join(d, "lc0.dat"), lc_t, mags, np.full(5, 0.02))
    lc = {"ioptlaw": 0, "ephem": eph, "sun_ephem": [(T0 - 1,) + SUN, (T0 + 2,) + SUN], "ncalc": ncalc,
          "calc_epochs": list(np.round((T0 + np.linspace(0.05, 0.55, 6)) * 86400) / 86400),
          "calc_grid": (T0, T0 + 0.6, 0.1), "nviews": 1, "view_interval_s": 0.0, "smearing": "center",
          "n": 5, "file": os.path.join(d, "lc0.dat"), "cal": ("f", 1.0), "weight": 1.0}
    SF.write_obs(os.path.join(d, "job.obs"), [SF.deldop_set(dd), SF.doppler_set(dp), SF.lightcurve_set(lc)])
    return (os.path.join(d, "job.par"), os.path.join(d, "job.mod"), os.path.join(d, "job.obs")), dd, dp, lc


def test_job_reads(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    files, dd, dp, lc = write_job(".")
    job = sh.Job(*files)
    assert job.nsets == 3
    assert [job.set_type(s) for s in range(3)] == [sh.DELAY, sh.DOPPLER, sh.LGHTCRV]
    # free parameters counted by the readers: the scale factors, the vertex
    # deviations, C of the cosine law, R of the Kaasalainen law, two Euler
    # angles and the spin rate; in the obs file the delay polynomial's a0 (the
    # frames' calibration factors float but are not counted, read_dat.c:617)
    assert job.nfpar_mod == 3 + 1148 + 1 + 1 + 3
    assert job.nfpar_dat == 1
    # frame 0: whole image; frame 1: vignetted to the mask's positive window
    assert job.dims(0, 0) == (100, 80)
    assert job.dims(0, 1) == (21, 31)
    ramp = (np.arange(8000, dtype=np.float32) % 97 + 1) / 50.0
    img0 = ramp.reshape(100, 80).astype(np.float64) * 2.0
    assert np.array_equal(job.image(0, 0, 1), img0)
    assert np.array_equal(job.image(0, 0, 2), np.full((100, 80), 0.25))
    img1 = ramp.reshape(100, 80).astype(np.float64)[40:61, 25:56] * 1.5
    assert np.array_equal(job.image(0, 1, 1), img1)
    w1 = np.full((21, 31), 0.5 / 1.5 ** 2)
    w1[45 - 40, 30 - 25] = 0.0
    assert np.array_equal(job.image(0, 1, 2), w1)
    s0, s1 = job.scalars(0, 0), job.scalars(0, 1)
    assert s0["dof"] == 8000 * 1.0
    assert s1["dof"] == (21 * 31 - 1) * 0.75
    spec = ((np.arange(100, dtype=np.float32) % 97 + 1) / 50.0).astype(np.float64)
    assert np.array_equal(job.image(1, 0, 1), spec)
    # view epochs: frame epoch, and the smeared Doppler views +-5 min about it
    t, spin = job.view_epoch(0, 1)
    assert t == pytest.approx(dd["frames"][1]["t0"], abs=1e-9)
    assert np.all(spin == 0.0)
    t0d = dp["frames"][0]["t0"]
    assert job.view_epoch(1, 0, 0)[0] == pytest.approx(t0d - 600 / 86400, abs=1e-9)
    assert job.view_epoch(1, 0, 1)[0] == pytest.approx(t0d, abs=1e-9)
    # the lightcurve: 4 distinct observation epochs become the calculated points
    assert job.dims(2, 0) == (5, 4)
    lcv = ((np.arange(5, dtype=np.float32) % 97 + 1) / 50.0).astype(np.float64) * 1e-12
    assert np.allclose(job.image(2, 0, 1), lcv, rtol=1e-13)
    # oneovervar = 1 / (0.4 ln10 magerr I)^2
    assert np.allclose(job.image(2, 0, 2), 1 / (0.4 * math.log(10) * 0.02 * lcv) ** 2, rtol=1e-12)
    job.close()


def test_job_par_values(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    files = write_job(".")[0]
    job = sh.Job(*files)
    import ctypes as C
    par = job.par
    flags = np.zeros(6)
    job.lib.shp_par_flags(par, flags.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.all(flags == 0.0)
    base = np.zeros(2)
    job.lib.shp_pen_base(par, base.ctypes.data_as(C.POINTER(C.c_double)))
    job.close()


def test_tessellation_of_an_ellipse_component(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    d = "."
    spin = {"t0": T0, "angles": [("c", 0.0)] * 3, "omega": [("c", 0.0), ("c", 0.0), ("c", 360.0)]}
    for nt in (3, 5, 8, 30):
        SF.write_mod(os.path.join(d, "e.mod"), [SF.ellipse_component(2.0, 1.25, 1.3333333333333333, nt)],
                     ["cosine c 0.2 c 1.0"], [], spin)
        SF.write_par(os.path.join(d, "e.par"), [("action", "write"), ("pos_pixels", "51"), ("pos_width", "3")])
        lib = io()
        import ctypes as C
        bufs = [C.create_string_buffer(int(sh.load_scene().shp_struct_size(w))) for w in range(2)]
        par, mod = (C.cast(b, C.c_void_p).value for b in bufs)
        assert lib.shape_read_par(os.path.join(d, "e.par").encode(), par) == 0, lib.shape_io_error()
        sh.load_scene().shp_set_name(1, mod, os.path.join(d, "e.mod").encode())
        assert lib.shape_read_mod(par, mod) == 3, lib.shape_io_error()
        scn = sh.load_scene()
        nv, nf = scn.shp_mod_nv(mod), scn.shp_mod_nf(mod)
        want = sum(max(1, int(math.floor(2 * nt * math.sin(t * math.pi / nt) + 0.5))) for t in range(nt + 1))
        assert nv == want
        if nt == 30:
            assert (nv, nf) == (1148, 2292)
        assert nf == 2 * nv - 4
        u, r = np.zeros(3 * nv), np.zeros(nv)
        scn.shp_vertex_dirs(mod, u.ctypes.data_as(C.POINTER(C.c_double)), r.ctypes.data_as(C.POINTER(C.c_double)))
        u = u.reshape(nv, 3)
        assert np.allclose(np.linalg.norm(u, axis=1), 1.0, atol=1e-15)
        fac = np.zeros(3 * nf, dtype=np.int32)
        scn.shp_facets(mod, fac.ctypes.data_as(C.POINTER(C.c_int)))
        fac = fac.reshape(nf, 3)
        # every directed edge once, its reverse once: closed and consistently oriented
        edges = {}
        for a, b, c in fac:
            for e in ((a, b), (b, c), (c, a)):
                assert e not in edges
                edges[e] = 1
        assert all((b, a) in edges for (a, b) in edges)
        # outward: each facet's normal points away from the centre
        x0, x1, x2 = u[fac[:, 0]], u[fac[:, 1]], u[fac[:, 2]]
        n = np.cross(x1 - x0, x2 - x0)
        assert np.all(np.einsum("ij,ij->i", n, x0 + x1 + x2) > 0)
        lib.shape_io_free_mod(mod)
        lib.shape_io_free_par(par)


def test_sides_and_attached_facets_match_the_scene_builder(tmp_path, monkeypatch):
    """the readers' side numbering and vertex lists against the scene builder's
    own construction of the same rule, on the config-1 mesh"""
    monkeypatch.chdir(tmp_path)
    files = write_job(".")[0]
    job = sh.Job(*files)
    sc = config1_mesh()[0]
    import ctypes as C
    ip = C.POINTER(C.c_int)
    cap = 3 * sc.nf
    a, b = np.zeros(4 * cap, dtype=np.int32), np.zeros(4 * cap, dtype=np.int32)
    na = sc.lib.shp_sides(sc.mod, a.ctypes.data_as(ip), cap)
    nb = job.lib.shp_sides(job.mod, b.ctypes.data_as(ip), cap)
    assert na == nb == 3 * sc.nf // 2
    assert np.array_equal(a[:4 * na], b[:4 * nb])
    for v in (0, 1, 17, 573, sc.nv - 1):
        fa, fb = np.zeros(16, dtype=np.int32), np.zeros(16, dtype=np.int32)
        ka = sc.lib.shp_vertex_facets(sc.mod, v, fa.ctypes.data_as(ip), 16)
        kb = job.lib.shp_vertex_facets(job.mod, v, fb.ctypes.data_as(ip), 16)
        assert ka == kb and np.array_equal(fa[:ka], fb[:kb])
    # each facet's sides are the sides of its corner pairs, ascending
    fs = np.zeros(3 * sc.nf, dtype=np.int32)
    job.lib.shp_facet_sides(job.mod, fs.ctypes.data_as(ip))
    fs = fs.reshape(-1, 3)
    assert np.all(np.diff(fs, axis=1) > 0)
    job.close()


def test_observations_of_the_model_give_zero_chi2(tmp_path, monkeypatch):
    """obs = the oracle's fit of the read job, written in each data format,
    read back to a chi-square at the FP32 rounding floor"""
    orc = H.load_oracle()
    monkeypatch.chdir(tmp_path)
    d = "."
    files, dd, dp, lc = write_job(d, fmts=("ascii", "ascii", "ascii"))
    job = sh.Job(*files)
    H.oracle_forward(orc, job)
    fits = {(0, 0): job.image(0, 0, 0), (0, 1): job.image(0, 1, 0), (1, 0): job.image(1, 0, 0),
            (2, 0): job.image(2, 0, 0)}
    assert fits[(0, 0)].max() > 0 and fits[(1, 0)].max() > 0 and fits[(2, 0)].max() > 0
    job.close()
    data = {}
    full0 = fits[(0, 0)] / 2.0
    data[(0, 0)] = full0.reshape(-1)
    full1 = np.full((100, 80), -1.0)
    full1[40:61, 25:56] = fits[(0, 1)]
    data[(0, 1)] = (full1 / 1.5).reshape(-1)
    data[(1, 0)] = fits[(1, 0)] / 1.0
    data[(2, 0)] = fits[(2, 0)] / 1e-12
    # a raw binary file is told by a byte above 127 in its first 128, so the
    # frames whose images start with 32 zero pixels go in the text formats
    for fmts in (("ascii", "binary", "rdf"), ("rdf.gz", "binary.gz", "binary")):
        sub = fmts[0].replace(".", "_")
        os.makedirs(sub, exist_ok=True)
        files, _, _, _ = write_job(sub, data=data, fmts=fmts)
        job = sh.Job(*files)
        chi2 = H.oracle_forward(orc, job)
        scale = sum(float(np.sum(fits[k] ** 2 * job.image(k[0], k[1], 2))) for k in fits)
        assert chi2 <= 1e-12 * scale, (fmts, chi2, scale)
        job.close()


@pytest.mark.parametrize("change,code,text", [
    ("nofile", -1, b"can't open"),
    ("keyword", -2, b"don't recognize parameter"),
    ("twice", -2, b"more than once"),
    ("fits", -3, b"FITS"),
    ("pos", -3, b"plane-of-sky"),
])
def test_reader_errors(tmp_path, monkeypatch, change, code, text):
    monkeypatch.chdir(tmp_path)
    d = "."
    files, dd, dp, lc = write_job(d)
    par, mod, obs = files
    if change == "nofile":
        par = os.path.join(d, "missing.par")
    elif change == "keyword":
        txt = open(par).read().replace("PENALTIES", "no_such_keyword 3\nPENALTIES")
        open(par, "w").write(txt)
    elif change == "twice":
        txt = open(par).read().replace("PENALTIES", "pos_width 2.0\nPENALTIES")
        open(par, "w").write(txt)
    elif change == "fits":
        name = dd["frames"][0]["name"]
        with open(os.path.join(dd["dir"], name), "w") as f:
            f.write("SIMPLE  =                    T / FITS" + " " * 200)
    elif change == "pos":
        txt = open(obs).read().replace("\ndoppler\n", "\nplane-of-sky\n", 1)
        open(obs, "w").write(txt)
    with pytest.raises(sh.ShapeIOError) as e:
        sh.Job(par, mod, obs)
    assert e.value.code == code
    assert text in str(e.value).encode()


def test_library_exports_reference_reader_names():
    import ctypes as C
    lib = C.CDLL(os.path.join(sh.LIBDIR, "libshape_io.so"))
    for name in ("read_par", "read_mod", "read_dat", "setupsides", "setupvertices"):
        assert hasattr(lib, name)
    # internals stay local
    assert not hasattr(lib, "io_word")


@pytest.mark.skipif(not os.path.isdir("/root/reference/src"), reason="reference tree not present")
def test_every_host_caller_of_the_reader_objects_resolves():
    """Routines that read_par.c, read_mod.c and read_dat.c define and that any
    other C file of the reference calls (shape-cuda-v1.0.c, convex_hull.c,
    merge_comps.c, split_mod.c) are exported by libshape_io.so, so those three
    objects can leave a host link.  (read_dat_gpu.cu, the CUDA build's own
    reader, calls read_dat.c's file-format helpers; it is not part of a build
    for this device.)"""
    import ctypes as C
    import glob
    import re
    import test_abi as TA
    lib = C.CDLL(os.path.join(sh.LIBDIR, "libshape_io.so"))
    called = {}
    for obj in ("read_par", "read_mod", "read_dat"):
        src = os.path.join(TA.REF_SRC, "shape", obj + ".c")
        defs = TA._external_definitions(src)
        for path in glob.glob(TA.REF_SRC + "/**/*.c", recursive=True):
            if path == src:
                continue
            t = TA._strip_c(open(path, errors="replace").read())
            for name in defs:
                if re.search(r"(?<![\w.>])%s\s*\(" % name, t):
                    called.setdefault(name, path)
    assert {"read_par", "read_mod", "read_dat", "setupsides", "setupvertices"} <= set(called)
    missing = sorted(n for n in called if not hasattr(lib, n))
    assert not missing, [(n, called[n]) for n in missing]
