"""The C-ABI boundary, checked without a GPU.

* libshape_hip.so loads and exports every function include/shape_hip.h
  declares (no compute call is made: there may be no device here).
* include/shape_abi.h declares the reference's structs with the same members,
  in the same order, with the same types and array sizes as the reference's
  src/shape/shape2.h (read as text; skipped where the reference is absent).
* The structs' sizes as the HIP library and the C scene builder see them agree
  (both are compiled from include/shape_abi.h by different compilers).
"""
import ctypes as C
import os
import re

import pytest

import abi_parse
import helpers as H

sh = H.sh

ROOT = H.ROOT
HDR = os.path.join(ROOT, "include", "shape_hip.h")
ABI = os.path.join(ROOT, "include", "shape_abi.h")
REF_SHAPE2 = "/root/reference/src/shape/shape2.h"
REF_CONST = "/root/reference/src/shape/const.h"

# structs on the hot path (the drop-in reads or writes every one of them)
HOT_STRUCTS = ["pos_t", "deldopfrm_t", "deldopview_t", "deldop_t", "dopfrm_t", "dopview_t", "doppler_t",
               "lghtcrv_t", "crvrend_t", "vertices_t", "vertex_t", "facet_t", "photo_t", "RC_t", "kaas_t",
               "spin_t", "mod_t", "set_t", "dat_t", "par_t", "param_t", "delcor_t", "component_t"]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"^\s*#.*$", " ", text, flags=re.M)
    text = re.sub(r"\btypedef\b[^;]*;", " ", text)  # (function-pointer types are not functions)
    names = re.findall(r"\b(\w+)\s*\([^;{}]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "while", "sizeof")))


def test_header_declares_reference_entry_points():
    names = declared_functions(HDR)
    for n in ["realize_mod", "realize_spin", "realize_delcor", "calc_fits", "chi2", "posclr", "posvis",
              "posmask", "radlaw", "pos2deldop", "pos2doppler", "apply_photo"]:
        assert n in names, n


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(os.path.join(H.PKG, "lib", "libshape_hip.so"))
    missing = [n for n in declared_functions(HDR) if not hasattr(lib, n)]
    assert not missing, missing


def test_no_compute_without_device_is_loud():
    """The package exposes no CPU fallback: the only compute path is the HIP library."""
    src = open(os.path.join(H.PKG, "shape_hip.py")).read()
    assert "oracle" not in src


def test_batch_create_rejects_bad_arguments():
    """Bad batch arguments are refused before any device call (NULL)."""
    lib = sh.load_hip()
    sc = sh.Scene(sh.spec_for(1))
    mods = (C.c_void_p * 1)(sc.mod)
    dats = (C.c_void_p * 1)(sc.dat)
    addr = C.cast(sc.lib.shp_param_ptr(sc.h, 0), C.c_void_p).value
    tg = (C.c_void_p * 1)(addr)
    vals = (C.c_double * 1)(1.0)

    def create(nmodels, nevals, model, ntarg, targets=tg, method=sh.METHOD_FP64):
        em = (C.c_int * max(1, nevals))(*([model] * max(1, nevals)))
        nt = (C.c_int * max(1, nevals))(*([ntarg] * max(1, nevals)))
        return lib.shape_batch_create(sc.par, mods, dats, nmodels, nevals, em, nt, targets, vals, method, 0)
    assert not create(1, 0, 0, 1)                    # no evaluations
    assert not create(0, 1, 0, 1)                    # no models
    assert not create(1, 1, 1, 1)                    # model index out of range
    assert not create(1, 1, 0, -1)                   # negative target count
    assert not create(1, 1, 0, 1, method=7)          # unknown method
    assert not create(1, 1, 0, 1, targets=(C.c_void_p * 1)(None))  # null target


@pytest.mark.skipif(not os.path.exists(REF_SHAPE2), reason="reference tree not present")
def test_struct_members_match_reference():
    ours = abi_parse.parse_structs(open(ABI).read())
    # const.h supplies the array-size macros shape2.h uses (MAXOVERFLOW, ...)
    ref_text = (open(REF_CONST).read() if os.path.exists(REF_CONST) else "") + open(REF_SHAPE2).read()
    ref = abi_parse.parse_structs(ref_text)
    checked = 0
    for tag in HOT_STRUCTS:
        if tag not in ref:
            continue
        assert tag in ours, tag
        assert ours[tag] == ref[tag], (tag, [a for a, b in zip(ours[tag], ref[tag]) if a != b][:3])
        checked += 1
    assert checked >= 15


def test_struct_sizes_agree_between_compilers():
    """sizeof() of the ABI structs as seen by the HIP library and by the gcc-built scene builder."""
    hip = C.CDLL(os.path.join(H.PKG, "lib", "libshape_hip.so"))
    scn = H.sh.load_scene()
    if not hasattr(hip, "shape_hip_abi_sizes") or not hasattr(scn, "shp_abi_sizes"):
        pytest.skip("size probes not exported")
    a = (C.c_longlong * 32)()
    b = (C.c_longlong * 32)()
    na = hip.shape_hip_abi_sizes(a, 32)
    nb = scn.shp_abi_sizes(b, 32)
    assert na == nb and na > 10
    assert list(a)[:na] == list(b)[:nb]


# Objects of the reference's link whose routines libshape_hip.so provides
# (the link recipe drops them), and the routines of those objects the
# recipe keeps instead: calc_fits.c also defines the plane-of-sky dataset
# (calc_poset) and the write action's POS images (write_pos_*), which this
# library does not compute; the recipe keeps them from a copy of
# calc_fits.o whose other routines objcopy renames.
REPLACED_OBJECTS = ["realize_mod", "realize_spin", "realize_delcor", "posclr", "posvis", "posmask", "radlaw",
                    "pos2deldop", "pos2doppler", "apply_photo", "calc_fits", "chi2", "penalties", "vary_params"]
KEPT_BY_RECIPE = {"calc_poset", "write_pos_deldop", "write_pos_doppler", "write_pos_poset", "write_pos_lghtcrv"}
REF_SRC = "/root/reference/src"


def _strip_c(t):
    t = re.sub(r"/\*.*?\*/", " ", t, flags=re.S)
    return re.sub(r"//[^\n]*", " ", t)


def _external_definitions(path):
    """Names of the functions a C file defines without `static`."""
    t = _strip_c(open(path, errors="replace").read())
    out = set()
    for m in re.finditer(r"^([A-Za-z_][\w \t\*]*?)\b(\w+)\s*\(([^;{)]*(?:\([^)]*\)[^;{)]*)*)\)\s*\{", t, flags=re.M):
        if "static" not in m.group(1) and m.group(2) not in ("if", "while", "for", "switch"):
            out.add(m.group(2))
    return out


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference tree not present")
def test_every_cross_file_callee_of_the_replaced_objects_resolves():
    """Every routine that the dropped objects define and that any other source
    file of the reference (C or CUDA, every directory) calls must be exported
    by libshape_hip.so or kept by the recipe, so the documented link leaves no
    dangling symbol (e.g. write_dat.c:123-124, calc_fits_hmt.c:146-155)."""
    import glob
    defs = {}
    for obj in REPLACED_OBJECTS:
        for name in _external_definitions(os.path.join(REF_SRC, "shape", obj + ".c")):
            defs[name] = obj
    replaced_paths = {os.path.join(REF_SRC, "shape", o + ".c") for o in REPLACED_OBJECTS}
    called = {}
    for path in glob.glob(REF_SRC + "/**/*.c", recursive=True) + glob.glob(REF_SRC + "/**/*.cu", recursive=True):
        if path in replaced_paths:
            continue
        t = _strip_c(open(path, errors="replace").read())
        for name in defs:
            if re.search(r"(?<![\w.>])%s\s*\(" % name, t):
                called.setdefault(name, path)
    assert {"realize_angleoff", "calc_deldop", "calc_lghtcrv", "penalties", "vary_params"} <= set(called)
    lib = C.CDLL(os.path.join(H.PKG, "lib", "libshape_hip.so"))
    missing = sorted(n for n in called if n not in KEPT_BY_RECIPE and not hasattr(lib, n))
    assert not missing, [(n, called[n]) for n in missing]
    recipe = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    for n in REPLACED_OBJECTS:
        assert "./src/shape/%s.o" % n in recipe, n
    # the recipe renames the library's routines in its copy of calc_fits.o
    for n in ("calc_fits", "calc_deldop", "calc_doppler", "calc_lghtcrv"):
        assert "--redefine-sym %s=" % n in recipe, n


def test_offset_equal_states():
    """realize_angleoff / realize_omegaoff (realize_spin.c:286-342): an '='
    component takes the value of the last earlier dataset whose state is not
    '='; '=' on the first dataset is an error.  Host bookkeeping: no device
    call."""
    lib = sh.load_hip()
    scl = sh.load_scene()
    sc = sh.Scene(sh.spec_for(2, n_deldop=4))
    dat = sc.dat
    for which, fn in ((0, lib.realize_angleoff), (1, lib.realize_omegaoff)):
        for s in range(4):
            for j in range(3):
                scl.shp_set_offset(dat, s, which, j, ord("c"), 10.0 * s + j + 0.5 * which)
        scl.shp_set_offset(dat, 1, which, 0, ord("="), -1.0)
        scl.shp_set_offset(dat, 2, which, 0, ord("="), -2.0)
        scl.shp_set_offset(dat, 3, which, 2, ord("="), -3.0)
        lib.shape_hip_clear_error()
        fn(dat)
        assert lib.shape_hip_last_error() == 0
        got = [[scl.shp_get_offset(dat, s, which, j) for j in range(3)] for s in range(4)]
        b = 0.5 * which
        assert got == [[0 + b, 1 + b, 2 + b], [0 + b, 11 + b, 12 + b], [0 + b, 21 + b, 22 + b],
                       [30 + b, 31 + b, 22 + b]]
        scl.shp_set_offset(dat, 0, which, 1, ord("="), 5.0)
        fn(dat)
        assert lib.shape_hip_last_error() != 0
        lib.shape_hip_clear_error()
