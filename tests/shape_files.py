"""Writers of SHAPE input files for the reader tests.

The formats are those the reference's readers parse (read_par.c, read_mod.c,
read_dat.c): whitespace-separated words with {comments}; parameters as
"state value" pairs; calendar dates as six integers.  The writers take plain
Python values (degrees, the file's own units) and produce a small but complete
job: a parameter file, a model file, an obs file and its data files in each
data format the readers accept (ASCII, raw binary, RDF, and gzip of any).
"""
import gzip
import math
import os
import struct

import numpy as np

SUN_VMAG = -26.762


def jd_to_cal(jd):
    """Julian date -> (y, mo, d, h, mi, s) for a date whose seconds are integral
    (the inverse of the readers' calendar conversion, Gregorian dates)."""
    z = jd + 0.5
    day = int(math.floor(z))
    secs = int(round((z - day) * 86400))
    if secs == 86400:
        day += 1
        secs = 0
    a = day + 32044
    b = (4 * a + 3) // 146097
    c = a - 146097 * b // 4
    d = (4 * c + 3) // 1461
    e = c - 1461 * d // 4
    m = (5 * e + 2) // 153
    dd = e - (153 * m + 2) // 5 + 1
    mo = m + 3 - 12 * (m // 10)
    yy = 100 * b + d - 4800 + m // 10
    return yy, mo, dd, secs // 3600, (secs % 3600) // 60, secs % 60


def cal(jd):
    return "%d %d %d %d %d %d" % jd_to_cal(jd)


def p(state, val):
    return "%s %.17g" % (state, val)


# ---- parameter file ---------------------------------------------------------
def write_par(path, params, penalties=()):
    lines = ["{ parameter file written by the reader tests }", "PARAMETERS"]
    for k, v in params:
        lines.append("%s %s" % (k, v))
    lines.append("PENALTIES %d" % len(penalties))
    for name, w in penalties:
        lines.append("  %s %.17g" % (name, w))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


# ---- model file --------------------------------------------------------------
def vertex_component(u, r, facets, scale=(1.0, 1.0, 1.0), r_state="f", a=None):
    """A VERTEX component: directions u (nv x 3), deviations r, facets (nf x 3)."""
    nv = len(r)
    a = np.zeros((nv, 3)) if a is None else a
    out = ["  c 0.0 c 0.0 c 0.0 { linear offsets }", "  c 0.0 c 0.0 c 0.0 { angular offsets (deg) }",
           "  vertex", "  %d { vertices }" % nv]
    out.append("  %s %s %s { scale factors }" % tuple(p("f", s) for s in scale))
    for i in range(nv):
        out.append("  %s %.17g %.17g %.17g %.17g %.17g %.17g" % (
            p(r_state, r[i]), u[i][0], u[i][1], u[i][2], a[i][0], a[i][1], a[i][2]))
    out.append("  %d { facets }" % len(facets))
    for t in facets:
        out.append("  %d %d %d" % tuple(t))
    return "\n".join(out)


def ellipse_component(two_a, a_over_b, b_over_c, ntheta):
    return "\n".join(["  c 0.0 c 0.0 c 0.0", "  c 0.0 c 0.0 c 0.0", "  ellipse",
                      "  %s %s %s" % (p("f", two_a), p("f", a_over_b), p("f", b_over_c)),
                      "  %d { theta steps }" % ntheta])


def write_mod(path, components, radar_laws, optical_laws, spin):
    """radar_laws / optical_laws: lists of preformatted law texts (name and
    parameters); spin: dict with t0 (JD), angles (deg), omega (deg/day),
    inertia, omegadot (deg/day^2), lib (amp, freq, phase), impulses."""
    lines = ["{ model file written by the reader tests }", "%d { components }" % len(components)]
    lines += components
    lines.append("%d { radar laws }" % len(radar_laws))
    lines += ["  " + t for t in radar_laws]
    lines.append("%d { optical laws }" % len(optical_laws))
    lines += ["  " + t for t in optical_laws]
    lines.append("%s { spin epoch }" % cal(spin["t0"]))
    lines.append("  ".join(p(s, v) for s, v in spin["angles"]) + " { Euler angles }")
    lines.append("  ".join(p(s, v) for s, v in spin["omega"]) + " { spin vector }")
    lines.append("  ".join(p(s, v) for s, v in spin.get("inertia", [("c", 1.0)] * 3)) + " { moments }")
    lines.append("  ".join(p(s, v) for s, v in spin.get("omegadot", [("c", 0.0)] * 3)) + " { YORP }")
    lines.append("  ".join(p(s, v) for s, v in spin.get("lib", [("c", 0.0)] * 3)) + " { libration }")
    imps = spin.get("impulses", [])
    lines.append("%d { spin impulses }" % len(imps))
    for t, imp in imps:
        lines.append("%s %s" % (cal(t), "  ".join(p(s, v) for s, v in imp)))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


# ---- data files --------------------------------------------------------------
def write_values(path, values, fmt, big_endian=True):
    """values in file order (rows of an image, or a spectrum) as ASCII, raw
    FP32 binary or RDF; a path ending in .gz is compressed."""
    v = np.asarray(values, dtype=np.float32).reshape(-1)
    if fmt == "ascii":
        body = ("\n".join("%.9g" % x for x in v) + "\n").encode()
    else:
        raw = v.astype(">f4" if big_endian else "<f4").tobytes()
        if fmt == "rdf":
            # the readers take a file as RDF only if its first 128 bytes are
            # text, as real RDF headers (hundreds of bytes of keywords) are
            head = ("# RDF header written by the reader tests\ntype = float\nbyte order = big endian\n"
                    "rows = 1\ncolumns = %d\nsamples per baud = 1\ntarget = synthetic asteroid\n"
                    "observer = none, the test writer\n.\n" % v.size).encode()
            assert len(head) > 128
            body = head + raw
        else:
            # the readers tell a raw binary file by a byte above 127 among its
            # first 128, so such a file must not start with 32 zero values
            body = raw
    if path.endswith(".gz"):
        with gzip.open(path, "wb") as f:
            f.write(body)
    else:
        with open(path, "wb") as f:
            f.write(body)


def write_mask(path, weights):
    with open(path, "w") as f:
        f.write("{ pixel weights }\n")
        for w in np.asarray(weights).reshape(-1):
            f.write("%.17g\n" % w)


def ephem_lines(points):
    """points: (JD, RA deg, Dec deg, dist AU)"""
    out = ["%d { ephemeris points }" % len(points)]
    for t, ra, dec, d in points:
        out.append("  %s %.17g %.17g %.17g" % (cal(t), ra, dec, d))
    return out


def deldop_set(s):
    """s: dict with the set's fields; returns the obs-file text of the set."""
    out = ["%d" % s.get("node", 0), "c 0.0 c 0.0 c 0.0 { angle offsets }", "c 0.0 c 0.0 c 0.0 { spin offsets }",
           "delay-doppler", "%d { radar law }" % s["iradlaw"]]
    out += ephem_lines(s["ephem"])
    out.append("%.17g { Ftx MHz }" % s["Ftx"])
    out.append("%d %.17g %d %d %s { ndel, del_per_pixel, spb, stride, code }" % (
        s["ndel"], s["del_per_pixel"], s["spb"], s["stride"], s["code"]))
    out.append("%d %.17g %.17g %.17g %d { ndop, dop_per_pixel, dopcom, dopDC, fft length }" % (
        s["ndop"], s["dop_per_pixel"], s["dopcom"], s["dopDC"], s["dopfftlen"]))
    out.append("%s %d %s { delay polynomial }" % (cal(s["delcor_t0"]), len(s["delcor"]) - 1,
                                                   " ".join(p(st, v) for st, v in s["delcor"])))
    out.append("%s { Doppler scale }" % p(*s["dopscale"]))
    out.append("%d %.17g %s { smearing }" % (s["nviews"], s["view_interval_s"], s["smearing"]))
    out.append("%s %d { data directory, frames }" % (s["dir"], len(s["frames"])))
    for fr in s["frames"]:
        out.append("  %s %s %.17g %s %.17g %.17g %.17g %d" % (
            fr["name"], cal(fr["t0"]), fr["sdev"], p(*fr["cal"]), fr["nlooks"], fr["delcom"], fr["weight"],
            fr["mask"]))
    return "\n".join(out)


def doppler_set(s):
    out = ["%d" % s.get("node", 0), "c 0.0 c 0.0 c 0.0", "c 0.0 c 0.0 c 0.0", "doppler",
           "%d { radar law }" % s["iradlaw"]]
    out += ephem_lines(s["ephem"])
    out.append("%.17g" % s["Ftx"])
    out.append("%d %.17g %.17g { ndop, dop_per_bin, dopcom }" % (s["ndop"], s["dop_per_bin"], s["dopcom"]))
    out.append("%s %d %s" % (cal(s["delcor_t0"]), len(s["delcor"]) - 1,
                             " ".join(p(st, v) for st, v in s["delcor"])))
    out.append(p(*s["dopscale"]))
    out.append("%d %.17g %s" % (s["nviews"], s["view_interval_s"], s["smearing"]))
    out.append("%s %d" % (s["dir"], len(s["frames"])))
    for fr in s["frames"]:
        out.append("  %s %s %.17g %s %.17g %.17g %d" % (
            fr["name"], cal(fr["t0"]), fr["sdev"], p(*fr["cal"]), fr["nlooks"], fr["weight"], fr["mask"]))
    return "\n".join(out)


def lightcurve_set(s):
    out = ["%d" % s.get("node", 0), "c 0.0 c 0.0 c 0.0", "c 0.0 c 0.0 c 0.0",
           "heliocurve" if s.get("helio") else "lightcurve", "%d { optical law }" % s["ioptlaw"]]
    out += ephem_lines(s["ephem"])
    out += ephem_lines(s["sun_ephem"])
    np_calc = s["ncalc"]
    if np_calc > 0:
        out.append("%d %s" % (np_calc, " ".join("%.17g" % t for t in s["calc_epochs"])))
    elif np_calc == 0:
        out.append("0 %.17g %.17g %.17g" % tuple(s["calc_grid"]))
    else:
        out.append("%d { calculation epochs = observation epochs }" % np_calc)
    out.append("%d %.17g %s" % (s["nviews"], s["view_interval_s"], s["smearing"]))
    out.append("%d %s %s %.17g { points, file, calibration, weight }" % (
        s["n"], s["file"], p(*s["cal"]), s["weight"]))
    return "\n".join(out)


def write_obs(path, sets_text):
    with open(path, "w") as f:
        f.write("{ obs file written by the reader tests }\n%d { datasets }\n" % len(sets_text))
        for t in sets_text:
            f.write(t + "\n")


def write_lightcurve(path, epochs, mags, magerrs):
    with open(path, "w") as f:
        for t, m, e in zip(epochs, mags, magerrs):
            f.write("%.17g %.17g %.17g\n" % (t, m, e))


def mag_of(intensity, sunmag=SUN_VMAG):
    """the magnitude whose intensity relative to the Sun's is the given one"""
    return sunmag - 2.5 * np.log10(intensity)
