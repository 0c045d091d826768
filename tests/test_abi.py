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


@pytest.mark.skipif(not os.path.exists(REF_SHAPE2), reason="reference tree not present")
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
