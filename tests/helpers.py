"""Shared test helpers: library loading and scene pairs for parity checks.

The oracle (oracle/_build/liboracle.so) is the FP64 CPU checker; the HIP
library (shape-cuda-v1.2_amd/lib/libshape_hip.so) is what is being tested.
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "shape-cuda-v1.2_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import shape_hip as sh  # noqa: E402

VP = C.c_void_p
_oracle = None
_hip = None


DP = C.POINTER(C.c_double)
ORACLE_FN = C.CFUNCTYPE(C.c_double, C.c_double)


def load_oracle():
    global _oracle
    if _oracle is None:
        path = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built (make -C oracle)")
        lib = C.CDLL(path)
        sigs = {
            "oracle_realize_mod": (None, [VP, VP]),
            "oracle_realize_spin": (None, [VP, VP, VP]),
            "oracle_realize_photo": (None, [VP, VP]),
            "oracle_hapke": (C.c_double, [C.c_double] * 8),
            "oracle_plm": (C.c_double, [C.c_int, C.c_int, C.c_double]),
            "oracle_realize_delcor": (None, [VP, C.c_double, C.c_int]),
            "oracle_posclr": (None, [VP]),
            "oracle_posvis": (C.c_int, [VP, C.POINTER(C.c_double), VP, C.c_int, C.c_int, C.c_int, C.c_int]),
            "oracle_posmask": (None, [VP, C.c_double]),
            "oracle_radlaw": (C.c_double, [VP, C.c_int, C.c_double, C.c_int, C.c_int]),
            "oracle_pos2deldop": (C.c_int, [VP, VP, C.c_double, C.c_double, C.c_double, VP, C.c_int,
                                            C.c_int, C.c_int, C.c_int]),
            "oracle_pos2doppler": (C.c_int, [VP, VP, C.c_double, C.c_double, C.c_double, VP, C.c_int,
                                             C.c_int, C.c_int, C.c_int]),
            "oracle_apply_photo": (C.c_double, [VP, C.c_int, C.c_double, C.c_double, VP, C.c_int,
                                                C.c_int, C.c_int]),
            "oracle_spline": (None, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_double,
                                     C.c_double, C.POINTER(C.c_double)]),
            "oracle_splint": (None, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.c_int, C.c_double, C.POINTER(C.c_double)]),
            "oracle_calc_fits": (None, [VP, VP, VP]),
            "oracle_calc_fits_threads": (None, [VP, VP, VP, C.c_int]),
            "oracle_chi2": (C.c_double, [VP, VP, C.c_int]),
            "oracle_objective_tail": (C.c_double, [VP, VP]),
            "oracle_mnbrak": (None, [DP, DP, DP, DP, DP, DP, ORACLE_FN]),
            "oracle_brent_abs": (C.c_double, [C.c_double, C.c_double, C.c_double, ORACLE_FN, C.c_double,
                                              C.c_double, DP]),
            "oracle_line_search": (C.c_double, [ORACLE_FN, C.c_double, C.c_double, C.c_double, C.c_double, DP]),
            "oracle_last_error": (C.c_char_p, []),
            "oracle_clear_error": (None, []),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _oracle = lib
    return _oracle


def load_hip():
    global _hip
    if _hip is None:
        _hip = sh.load_hip()
    return _hip


def oracle_forward(orc, sc, realize=True):
    """realize (optional) + calc_fits + chi2 on the oracle; returns chi2."""
    if realize:
        orc.oracle_realize_mod(sc.par, sc.mod)
        orc.oracle_realize_spin(sc.par, sc.mod, sc.dat)
        orc.oracle_realize_photo(sc.par, sc.mod)
        orc.oracle_realize_delcor(sc.dat, 0.0, 0)
    orc.oracle_calc_fits(sc.par, sc.mod, sc.dat)
    return orc.oracle_chi2(sc.par, sc.dat, 0)


def make_scene_with_obs(config, snr=20.0, **over):
    """Scene whose observations are the oracle's truth fits plus noise."""
    orc = load_oracle()
    sc = sh.Scene(sh.spec_for(config, **over))
    oracle_forward(orc, sc)
    sc.synthesize_obs(snr=snr)
    return sc


def oracle_threads():
    """Host threads for the oracle: the process's CPU share, at most 16."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def oracle_forward_threads(orc, sc):
    """The oracle's chain with frames dealt to host threads (scenes with one
    POS per frame, pos_scope LOCAL); returns chi2."""
    orc.oracle_realize_mod(sc.par, sc.mod)
    orc.oracle_realize_spin(sc.par, sc.mod, sc.dat)
    orc.oracle_realize_photo(sc.par, sc.mod)
    orc.oracle_realize_delcor(sc.dat, 0.0, 0)
    orc.oracle_calc_fits_threads(sc.par, sc.mod, sc.dat, oracle_threads())
    return orc.oracle_chi2(sc.par, sc.dat, 0)


def truth_scene(config, **over):
    """Scene with one POS per frame whose observations are the oracle's truth
    fits plus noise, the fits computed on the host's threads."""
    orc = load_oracle()
    sc = sh.Scene(sh.spec_for(config, pos_scope=sh.LOCAL, **over))
    oracle_forward_threads(orc, sc)
    sc.synthesize_obs(snr=20.0)
    return sc


def oracle_objective(orc, sc, values):
    """The oracle's objective with parameters {index: value} set, restored after."""
    saved = {i: sc.get_param(i) for i in values}
    for i, v in values.items():
        sc.set_param(i, v)
    oracle_forward_threads(orc, sc)
    f = orc.oracle_objective_tail(sc.par, sc.dat)
    for i, v in saved.items():
        sc.set_param(i, v)
    return f


def copy_obs(src, dst):
    for s, f in src.all_frames():
        ff = max(f, 0)
        dst.set_image(s, ff, 1, src.image(s, ff, 1))
        dst.set_image(s, ff, 2, src.image(s, ff, 2))


def scene_pair(config, **over):
    """Two identical scenes (same spec) with identical observations."""
    a = make_scene_with_obs(config, **over)
    b = sh.Scene(sh.spec_for(config, **over))
    copy_obs(a, b)
    return a, b


def set_radar_laws(sc, law_type, p):
    """Give every radar law of the scene (one per radar set) type law_type."""
    for l in range(sc.nsets):
        try:
            sc.set_radar_law(l, law_type, p)
        except ValueError:
            break


def posmask_plane_case(sc, pos, tol, seed, theta=0.37, a=0.37, b=-0.21, c0=3.0):
    """A posmask input whose answer a plane fixes.  The source's z-buffer is
    the plane zill = a i + b j + c0, blank on a lattice of isolated pixels
    (i and j multiples of 3), so every valid pixel keeps a valid neighbour on
    each axis; the source frame is the observer's turned by theta about the
    line of sight, so each POS pixel lands at a fractional source position
    (i0, j0).  Bilinear interpolation and its one-sided fallback beside blank
    pixels (posmask.c:112-166) both reproduce a plane exactly, so the pixel
    is shadowed iff plane(i0, j0) - z > tol.  z is set 1e-7 above or below
    plane - tol at random.  Returns (expected cose, eligible, stencils with
    a blank corner among the eligible pixels)."""
    n = sc.spec.pos_n
    I, J = np.meshgrid(np.arange(-n, n + 1), np.arange(-n, n + 1), indexing="ij")
    blank = (I % 3 == 0) & (J % 3 == 0)
    zill = np.where(blank, -1.0e20, a * I + b * J + c0)
    c, s = np.cos(theta), np.sin(theta)
    sc.pos_set_frames(pos, np.eye(3), np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]]))
    i0, j0 = c * I + s * J, -s * I + c * J
    sign = np.random.default_rng(seed).choice([-1.0, 1.0], size=I.shape)
    z = a * i0 + b * j0 + c0 - tol - sign * 1e-7
    cose = np.full(I.shape, 0.5)
    for which, arr in ((1, cose), (3, z), (5, zill), (10, np.ones_like(z)), (13, np.full(I.shape, 2.0)),
                       (11, np.zeros_like(z)), (12, np.zeros_like(z))):
        sc.pos_set_array(pos, which, arr)
    inside = (np.abs(i0) < n) & (np.abs(j0) < n)
    im = np.clip(np.floor(i0 + 0.5).astype(int), -n, n)
    jm = np.clip(np.floor(j0 + 0.5).astype(int), -n, n)
    eligible = inside & ~blank[im + n, jm + n]
    i1 = np.clip(np.floor(i0).astype(int), -n, n - 1)
    j1 = np.clip(np.floor(j0).astype(int), -n, n - 1)
    kludge = eligible & (blank[i1 + n, j1 + n] | blank[i1 + 1 + n, j1 + n] | blank[i1 + n, j1 + 1 + n]
                         | blank[i1 + 1 + n, j1 + 1 + n])
    return np.where(eligible & (sign > 0), 0.0, cose), eligible, int(np.sum(kludge))


def rel_err(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)


# Delay and Doppler response cases beyond one sample per baud and one sinc^2
# sample per pixel (pos2deldop.c:225-236, :371-456; pos2doppler.c), on a
# one-frame delay-Doppler plus one-spectrum scene of config 1.  Rows per
# pixel footprint: LONG_MOD / SHORT 2*ceil(const1) with const1 = (3 spb -
# 1) / (2 stride), LONG_ORIG 2*spb/stride.  "wide" has pixels of 0.12 km
# at 0.05 Hz per bin with nsinc2 = 3 and spb = 3: ~20 Doppler bins and 8
# delay rows per pixel, beyond the mixed method's unrolled register window
# (4 rows x 16 bins), so the general loop synthesizes them.
_DD1 = dict(n_deldop=1, n_doppler=1, nbin_min=40, nbin_max=60)
FOOTPRINTS = {
    "spb2_longmod": dict(_DD1, spb=2, stride=1, codemethod=sh.LONG_MOD),
    "spb2_longorig": dict(_DD1, spb=2, stride=1, codemethod=sh.LONG_ORIG),
    "spb3_short": dict(_DD1, spb=3, stride=1, codemethod=sh.SHORT),
    "spb4_stride2": dict(_DD1, spb=4, stride=2, codemethod=sh.LONG_MOD),
    "spb3_longorig": dict(_DD1, spb=3, stride=1, codemethod=sh.LONG_ORIG),
    "nsinc2_2": dict(_DD1, nsinc2=2),
    "nsinc2_3_spb2": dict(_DD1, nsinc2=3, spb=2, stride=1),
    "wide": dict(n_deldop=1, n_doppler=1, pos_n=12, km_per_pixel=0.12, dop_per_pixel=0.05, ndop_min=800,
                 ndop_max=800, nbin_min=800, nbin_max=800, dop_per_bin_min=0.05, dop_per_bin_max=0.05,
                 nsinc2=3, spb=3, stride=1),
}
