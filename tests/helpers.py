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


def rel_err(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)
