"""ctypes bindings for the SHAPE scene builder and the HIP model-evaluation library.

The libraries are C-ABI shared objects built in-tree (``lib/``):

* ``libshapescene.so`` -- synthetic scenes as SHAPE ABI structs (par_t/mod_t/dat_t),
  plus accessors that copy NR-offset arrays into numpy.
* ``libshape_hip.so`` -- the HIP drop-ins for the reference's entry points
  (realize_mod, realize_spin, realize_delcor, posclr, posvis, posmask, radlaw,
  pos2deldop, pos2doppler, apply_photo, calc_fits, chi2) and the batched
  evaluation engine.

Loading the HIP library never falls back to anything: if it is missing,
``load_hip()`` raises.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

GLOBAL, LOCAL = 1, 2
BRUTE, PERIODIC = 1, 2  # par_t int_method
SHORT, LONG_ORIG, LONG_MOD = 0, 1, 2
DELAY, DOPPLER, POS, LGHTCRV = 0, 1, 2, 3
# radar scattering law types (shape2.h's numbering)
(COSINELAW_DIFF, GAUSSIANLAW, HAGFORSLAW, COSINELAW_QS, GAUSSIAN_COSINE, HAGFORS_COSINE,
 COSINE_COSINE) = 1, 2, 3, 4, 5, 6, 7
SHAPE_ELLIPSOID_TESS, SHAPE_ICOSPHERE = 0, 1
# scene component types (shp_spec.comp_type)
COMP_VERTEX, COMP_ELLIPSE, COMP_OVOID, COMP_HARMONIC = 0, 1, 2, 3
# penalty types (const.h's numbering)
(NONSMOOTH, CONCAVITY, RDEV, VOLUME, COMDEV, INERTIADEV, INERTIADEV_UNI, PA3TILT, NONPA, NONPA_UNI,
 EULEROFFS, FLATTENING, MAXRDEV, MAXELLIPDEV, NONCOSINE, BIFURCATION, IMPULSE) = range(9, 26)


class Spec(C.Structure):
    """Mirror of shp_spec in csrc/scene/scene.h."""
    _fields_ = [
        ("shape_kind", C.c_int), ("theta_steps", C.c_int), ("ico_level", C.c_int),
        ("semi_axes", C.c_double * 3), ("rdev_sigma", C.c_double),
        ("seed_geom", C.c_ulonglong), ("seed_frames", C.c_ulonglong),
        ("spin_period_h", C.c_double), ("pole_lambda_deg", C.c_double),
        ("pole_beta_deg", C.c_double), ("phase0_deg", C.c_double),
        ("t0", C.c_double), ("epoch_span_days", C.c_double),
        ("pos_n", C.c_int), ("km_per_pixel", C.c_double), ("pos_scope", C.c_int),
        ("pos_smooth", C.c_int),
        ("Ftx", C.c_double), ("codemethod", C.c_int), ("spb", C.c_int), ("stride", C.c_int),
        ("dopfftlen", C.c_int),
        ("R_min", C.c_double), ("R_max", C.c_double), ("C_min", C.c_double), ("C_max", C.c_double),
        ("one_set_per_frame", C.c_int),
        ("n_deldop", C.c_int), ("ndel_min", C.c_int), ("ndel_max", C.c_int),
        ("ndop_min", C.c_int), ("ndop_max", C.c_int),
        ("del_per_pixel", C.c_double), ("dop_per_pixel", C.c_double),
        ("n_doppler", C.c_int), ("nbin_min", C.c_int), ("nbin_max", C.c_int),
        ("dop_per_bin_min", C.c_double), ("dop_per_bin_max", C.c_double),
        ("n_lightcurves", C.c_int), ("lc_points", C.c_int), ("lc_ncalc", C.c_int),
        ("phase_min_deg", C.c_double), ("phase_max_deg", C.c_double), ("kaas_wt", C.c_double),
        ("delcor_a0", C.c_double), ("delcor_a1", C.c_double),
        ("obs_lat_min_deg", C.c_double), ("obs_lat_max_deg", C.c_double),
        ("radar_nviews", C.c_int), ("view_interval_days", C.c_double), ("lc_nviews", C.c_int),
        ("nsinc2", C.c_int), ("sinc2width", C.c_int),
        ("comp_type", C.c_int), ("ovoid_k", C.c_double), ("nhar", C.c_int), ("har_sigma", C.c_double),
    ]


_scene_lib = None


def load_scene():
    global _scene_lib
    if _scene_lib is None:
        path = os.path.join(LIBDIR, "libshapescene.so")
        if not os.path.exists(path):
            raise RuntimeError("libshapescene.so is not built (run __graft_entry__.build())")
        lib = C.CDLL(path)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        sigs = {
            "shp_spec_defaults": (None, [C.POINTER(Spec)]),
            "shp_scene_create": (vp, [C.POINTER(Spec)]),
            "shp_scene_destroy": (None, [vp]),
            "shp_scene_par": (vp, [vp]), "shp_scene_mod": (vp, [vp]), "shp_scene_dat": (vp, [vp]),
            "shp_scene_nv": (C.c_int, [vp]), "shp_scene_nf": (C.c_int, [vp]),
            "shp_scene_synthesize_obs": (None, [vp, C.c_double, C.c_double, C.c_ulonglong]),
            "shp_set_type": (C.c_int, [vp, C.c_int]),
            "shp_dat_nsets": (C.c_int, [vp]),
            "shp_set_nframes": (C.c_int, [vp, C.c_int]),
            "shp_frame_dims": (None, [vp, C.c_int, C.c_int, ip]),
            "shp_frame_image": (None, [vp, C.c_int, C.c_int, C.c_int, dp]),
            "shp_frame_set_image": (None, [vp, C.c_int, C.c_int, C.c_int, dp]),
            "shp_frame_scalars": (None, [vp, C.c_int, C.c_int, dp]),
            "shp_frame_pos": (vp, [vp, C.c_int, C.c_int]),
            "shp_pos_array": (None, [vp, C.c_int, dp]),
            "shp_pos_limits": (None, [vp, ip]),
            "shp_pos_copy": (None, [vp, vp]),
            "shp_pos_scalars": (None, [vp, dp]),
            "shp_mod_vertices": (vp, [vp]),
            "shp_set_desc": (vp, [vp, C.c_int]),
            "shp_view_data": (None, [vp, C.c_int, C.c_int, C.c_int, dp]),
            "shp_set_view_data": (None, [vp, C.c_int, C.c_int, C.c_int, dp]),
            "shp_geometry": (None, [vp, dp, dp, dp, dp]),
            "shp_facet_seen": (None, [vp, C.POINTER(C.c_ubyte)]),
            "shp_set_vertex_normals": (None, [vp, dp]),
            "shp_fits_bracket": (None, [vp, vp, vp, C.c_int]),
            "shp_set_offset": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]),
            "shp_get_offset": (C.c_double, [vp, C.c_int, C.c_int, C.c_int]),
            "shp_par_flags": (None, [vp, dp]),
            "shp_frame_consts": (None, [vp, C.c_int, C.c_int, dp]),
            "shp_par_set": (None, [vp, C.c_int, C.c_double]),
            "shp_mod_photo": (vp, [vp]),
            "shp_set_radar_law": (C.c_int, [vp, C.c_int, C.c_int, dp]),
            "shp_pos_set_array": (None, [vp, C.c_int, dp]),
            "shp_pos_set_frames": (None, [vp, dp, dp]),
            "shp_param_ptr": (dp, [vp, C.c_int]),
            "shp_param_count": (C.c_int, [vp]),
            "shp_param_type": (C.c_int, [vp, C.c_int]),
            "shp_scene_set_penalties": (C.c_int, [vp, C.c_int, ip, dp]),
            "shp_pen_base": (None, [vp, dp]),
            "shp_set_spin_inertia": (None, [vp, dp]),
            "shp_set_impulses": (C.c_int, [vp, C.c_int, dp, dp]),
            "shp_moments": (None, [vp, C.c_int, dp]),
            "shp_sides": (C.c_int, [vp, ip, C.c_int]),
            "shp_vertex_facets": (C.c_int, [vp, C.c_int, ip, C.c_int]),
            "shp_facets": (None, [vp, ip]),
            "shp_harmonic": (C.c_int, [vp, C.c_int, dp, dp]),
            "shp_set_vertex_r": (None, [vp, dp]),
            "shp_set_comp_transform": (None, [vp, dp, dp]),
            "shp_set_vary": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
            "shp_radar_law": (None, [vp, vp, C.c_int, dp]),
            "shp_set_state": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]),
            "shp_dat_vary_weights": (None, [vp]),
            "shp_vertex_dirs": (None, [vp, dp, dp]),
            "shp_struct_size": (C.c_size_t, [C.c_int]),
            "shp_set_name": (None, [C.c_int, vp, C.c_char_p]),
            "shp_mod_nv": (C.c_int, [vp]), "shp_mod_nf": (C.c_int, [vp]),
            "shp_vertex_sides": (C.c_int, [vp, ip, ip, C.c_int]),
            "shp_facet_sides": (None, [vp, ip]),
            "shp_view_epoch": (None, [vp, C.c_int, C.c_int, C.c_int, dp]),
            "shp_mod_param_ptr": (dp, [vp, C.c_int]),
            "shp_set_npa": (None, [vp, vp, dp, dp, C.c_int]),
            "shp_spin_state": (None, [vp, dp]),
            "shp_law_local": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp]),
            "shp_law_table": (C.c_int, [vp, C.c_int, dp, C.c_int]),
            "shp_facet_geom": (None, [vp, C.c_int, dp]),
            "shp_badphoto": (None, [vp, dp]),
            "shp_law_param_ptr": (dp, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _scene_lib = lib
    return _scene_lib


_io_lib = None


def load_io():
    """libshape_io.so: the par / mod / obs file readers (host C)."""
    global _io_lib
    if _io_lib is None:
        path = os.path.join(LIBDIR, "libshape_io.so")
        if not os.path.exists(path):
            raise RuntimeError("libshape_io.so is not built (run __graft_entry__.build())")
        lib = C.CDLL(path)
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        sigs = {
            "shape_read_par": (C.c_int, [C.c_char_p, vp]),
            "shape_read_mod": (C.c_int, [vp, vp]),
            "shape_read_dat": (C.c_int, [vp, vp, vp]),
            "shape_io_error": (C.c_char_p, []),
            "shape_io_free_par": (None, [vp]), "shape_io_free_mod": (None, [vp]),
            "shape_io_free_dat": (None, [vp]),
            "shape_io_cal2jd": (C.c_double, [C.c_int] * 6),
            "shape_io_ephem2mat": (C.c_double, [vp, vp, C.c_double, dp, dp, dp, dp, dp, C.c_int]),
            "setupsides": (None, [vp]), "setupvertices": (None, [vp]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _io_lib = lib
    return _io_lib


class ShapeIOError(RuntimeError):
    """A reader refused its input; code is SHAPE_IO_E* (-1 file, -2 parse, -3 unsupported)."""

    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


# shape_eval_fn: evaluate the objective at x[0..k) into f; 0 on success
SEARCH_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double))

_hip_lib = None

HIP_EXPORTS = ["realize_mod", "realize_spin", "realize_delcor", "calc_fits", "chi2", "penalties", "vary_params",
               "shape_hip_last_error", "shape_hip_clear_error", "shape_hip_set_strict",
               "shape_hip_device_count", "shape_hip_realize_delcor"]


def load_hip():
    """Load the HIP library.  Raises if it is missing: there is no fallback."""
    global _hip_lib
    if _hip_lib is None:
        path = os.path.join(LIBDIR, "libshape_hip.so")
        if not os.path.exists(path):
            raise RuntimeError("libshape_hip.so is not built (run __graft_entry__.build())")
        lib = C.CDLL(path)
        vp = C.c_void_p
        sigs = {
            "posclr": (None, [vp]),
            "posvis": (C.c_int, [vp, C.POINTER(C.c_double), vp, C.c_int, C.c_int, C.c_int, C.c_int]),
            "posmask": (None, [vp, C.c_double]),
            "radlaw": (C.c_double, [vp, C.c_int, C.c_double, C.c_int, C.c_int]),
            "pos2deldop": (C.c_int, [vp, vp, C.c_double, C.c_double, C.c_double, vp, C.c_int, C.c_int,
                                     C.c_int, C.c_int]),
            "pos2doppler": (C.c_int, [vp, vp, C.c_double, C.c_double, C.c_double, vp, C.c_int, C.c_int,
                                      C.c_int, C.c_int]),
            "apply_photo": (C.c_double, [vp, C.c_int, C.c_double, C.c_double, vp, C.c_int, C.c_int, C.c_int]),
            "realize_mod": (None, [vp, vp]),
            "realize_spin": (None, [vp, vp, vp]),
            "realize_delcor": (None, [vp, C.c_double, C.c_int]),
            "calc_fits": (None, [vp, vp, vp]),
            "calc_deldop": (None, [vp, vp, vp, C.c_int]),
            "calc_doppler": (None, [vp, vp, vp, C.c_int]),
            "calc_lghtcrv": (None, [vp, vp, vp, C.c_int]),
            "realize_angleoff": (None, [vp]),
            "realize_omegaoff": (None, [vp]),
            "shape_hip_release_cache": (None, []),
            "chi2": (C.c_double, [vp, vp, C.c_int]),
            "penalties": (C.c_double, [vp, vp, vp]),
            "vary_params": (None, [vp, vp, vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "shape_chi2_noalloc": (C.c_double, [vp, vp, vp, vp, vp]),
            "shape_hip_last_error": (C.c_int, []),
            "shape_hip_clear_error": (None, []),
            "shape_hip_set_strict": (None, [C.c_int]),
            "shape_hip_device_count": (C.c_int, []),
            "shape_hip_wave_selftest": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "shape_hip_realize_delcor": (None, [vp, vp, vp, C.c_double, C.c_int]),
            "shape_hip_last_replayed": (C.c_int, []),
            "shape_batch_create": (vp, [vp, C.POINTER(vp), C.POINTER(vp), C.c_int, C.c_int,
                                        C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(vp),
                                        C.POINTER(C.c_double), C.c_int, C.c_int]),
            "shape_batch_set_values": (C.c_int, [vp, C.POINTER(C.c_double)]),
            "shape_batch_launch": (C.c_int, [vp]),
            "shape_batch_wait": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "shape_batch_run": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "shape_line_search": (C.c_int, [SEARCH_FN, vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                            C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                            C.POINTER(C.c_int)]),
            "shape_batch_line_search": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                                  C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                  C.POINTER(C.c_int)]),
            "shape_line_searches": (C.c_int, [SEARCH_FN, vp, C.c_int, C.c_int, C.POINTER(C.c_double),
                                              C.POINTER(C.c_double), C.c_double, C.c_double, C.c_int,
                                              C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int),
                                              C.POINTER(C.c_int)]),
            "shape_batch_line_searches": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                    C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double),
                                                    C.POINTER(C.c_double), C.POINTER(C.c_int),
                                                    C.POINTER(C.c_int)]),
            "shape_batch_fit_pass": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                               C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double),
                                               C.POINTER(C.c_double), C.POINTER(C.c_int)]),
            "shape_batch_tune": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
            "shape_batch_get_tune": (C.c_int, [vp, C.POINTER(C.c_int)]),
            "shape_batch_set_tune": (C.c_int, [vp, C.POINTER(C.c_int)]),
            "shape_batch_last_ms": (C.c_float, [vp]),
            "shape_batch_nframes": (C.c_int, [vp, C.c_int]),
            "shape_batch_frame_results": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double),
                                                    C.POINTER(C.c_double)]),
            "shape_batch_fit": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]),
            "shape_batch_facet_map": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
            "shape_batch_frame_terms": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double)]),
            "shape_batch_penalties": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "shape_batch_counters": (C.c_int, [vp, C.POINTER(C.c_int)]),
            "shape_batch_sizes": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
            "shape_batch_destroy": (None, [vp]),
            "shape_batch_stream": (vp, [vp]),
            "shape_batch_time_kernels": (C.c_int, [vp, C.c_int]),
            "shape_batch_kernel_stat": (C.c_int, [vp, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_double),
                                                  C.POINTER(C.c_longlong)]),
            "shape_batch_footprint": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _hip_lib = lib
    return _hip_lib


# ---------------------------------------------------------------------------
# configurations (the five benchmark / parity workloads)
# ---------------------------------------------------------------------------
def spec_for(config, **over):
    """Return a Spec for workload `config` (1..5); keyword overrides win.

    1: triaxial ellipsoid 1.0x0.8x0.6 km, theta_steps=30 (1148 V / 2292 F), 2% radial noise,
       one 100x80 delay-Doppler frame, POS 151^2 at 15 m.
    2: ellipsoid tessellation theta_steps=80 (8150 V / 16296 F; 16382 F is not reachable
       by a closed tessellation of that family), 3% noise, 64 frames rows 60-200 x cols
       40-160, POS 255^2 at 10 m, epochs over 30 days, R~U(0.1,0.3), C~U(1,3).
    3: icosphere level 6 (40962 V / 81920 F), 50 delay-Doppler (50-300 x 30-250),
       40 Doppler (100-400 bins, 0.25-2 Hz), 20 lightcurves x 100 points (2000 points),
       Kaasalainen (L-S 0.9 + Lambert 0.1), solar phase 5-60 deg.
    4: icosphere level 7 (163842 V / 327680 F) scaled to 0.9x0.72x0.54 km, 0.2% radial
       noise, POS 1023^2 at 2 m, 256 frames 128-512 x 64-256.
    5: icosphere level 5 (10242 V / 20480 F), 24 frames rows 60-250 x cols 40-160,
       POS 255^2 at 10 m (one model; the batch driver makes 32 with per-model seeds).
    """
    lib = load_scene()
    s = Spec()
    lib.shp_spec_defaults(C.byref(s))
    if config == 1:
        pass
    elif config == 2:
        s.theta_steps = 80
        s.rdev_sigma = 0.03
        s.seed_geom, s.seed_frames = 2001, 2002
        s.pos_n, s.km_per_pixel = 127, 0.010
        s.n_deldop = 64
        s.ndel_min, s.ndel_max, s.ndop_min, s.ndop_max = 60, 200, 40, 160
        s.dop_per_pixel = 0.0
        s.epoch_span_days = 30.0
        s.R_min, s.R_max, s.C_min, s.C_max = 0.1, 0.3, 1.0, 3.0
        s.obs_lat_min_deg, s.obs_lat_max_deg = -20.0, 40.0
    elif config == 3:
        s.shape_kind, s.ico_level = SHAPE_ICOSPHERE, 6
        s.rdev_sigma = 0.03
        s.seed_geom, s.seed_frames = 3001, 3002
        s.pos_n, s.km_per_pixel = 127, 0.010
        s.n_deldop = 50
        s.ndel_min, s.ndel_max, s.ndop_min, s.ndop_max = 50, 300, 30, 250
        s.dop_per_pixel = 0.0
        s.n_doppler = 40
        s.nbin_min, s.nbin_max = 100, 400
        s.dop_per_bin_min, s.dop_per_bin_max = 0.25, 2.0
        s.n_lightcurves, s.lc_points, s.lc_ncalc = 20, 100, 40
        s.epoch_span_days = 30.0
        s.R_min, s.R_max, s.C_min, s.C_max = 0.1, 0.3, 1.0, 3.0
        s.obs_lat_min_deg, s.obs_lat_max_deg = -20.0, 40.0
    elif config == 4:
        s.shape_kind, s.ico_level = SHAPE_ICOSPHERE, 7
        s.semi_axes[0], s.semi_axes[1], s.semi_axes[2] = 0.9, 0.72, 0.54
        # 0.2% radial noise: facets are ~6.7 m across, so this keeps surface
        # slopes near config 1's (noise / facet size ~0.2); 3% here would be
        # 27 m spikes on 6.7 m facets, not a surface
        s.rdev_sigma = 0.002
        s.seed_geom, s.seed_frames = 4001, 4002
        s.pos_n, s.km_per_pixel = 511, 0.002
        s.n_deldop = 256
        s.ndel_min, s.ndel_max, s.ndop_min, s.ndop_max = 128, 512, 64, 256
        s.dop_per_pixel = 0.0
        s.epoch_span_days = 30.0
        s.R_min, s.R_max, s.C_min, s.C_max = 0.1, 0.3, 1.0, 3.0
        s.obs_lat_min_deg, s.obs_lat_max_deg = -20.0, 40.0
    elif config == 5:
        s.shape_kind, s.ico_level = SHAPE_ICOSPHERE, 5
        s.rdev_sigma = 0.03
        s.seed_geom, s.seed_frames = 5000, 5500
        s.pos_n, s.km_per_pixel = 127, 0.010
        s.n_deldop = 24
        s.ndel_min, s.ndel_max, s.ndop_min, s.ndop_max = 60, 250, 40, 160
        s.dop_per_pixel = 0.0
        s.epoch_span_days = 30.0
        s.R_min, s.R_max, s.C_min, s.C_max = 0.1, 0.3, 1.0, 3.0
        s.obs_lat_min_deg, s.obs_lat_max_deg = -20.0, 40.0
    else:
        raise ValueError("config must be 1..5")
    for k, v in over.items():
        if k == "semi_axes":
            for i in range(3):
                s.semi_axes[i] = v[i]
        else:
            setattr(s, k, v)
    return s


class Scene:
    """Owns one synthetic scene (par/mod/dat ABI structs)."""

    def __init__(self, spec):
        self.lib = load_scene()
        self.spec = spec
        self.h = self.lib.shp_scene_create(C.byref(spec))
        if not self.h:
            raise RuntimeError("scene creation failed")
        self.par = self.lib.shp_scene_par(self.h)
        self.mod = self.lib.shp_scene_mod(self.h)
        self.dat = self.lib.shp_scene_dat(self.h)

    def close(self):
        if self.h:
            self.lib.shp_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- shape ----
    @property
    def nv(self):
        return self.lib.shp_scene_nv(self.h)

    @property
    def nf(self):
        return self.lib.shp_scene_nf(self.h)

    def geometry(self):
        nv, nf = self.nv, self.nf
        vx, vn = np.zeros(3 * nv), np.zeros(3 * nv)
        fn, area = np.zeros(3 * nf), np.zeros(nf)
        dp = C.POINTER(C.c_double)
        self.lib.shp_geometry(self.mod, vx.ctypes.data_as(dp), vn.ctypes.data_as(dp),
                              fn.ctypes.data_as(dp), area.ctypes.data_as(dp))
        return vx.reshape(nv, 3), vn.reshape(nv, 3), fn.reshape(nf, 3), area

    def set_vertex_normals(self, vn):
        vn = np.ascontiguousarray(vn, dtype=np.float64).reshape(-1)
        self.lib.shp_set_vertex_normals(self.mod, vn.ctypes.data_as(C.POINTER(C.c_double)))

    def facet_seen(self):
        out = np.zeros(self.nf, dtype=np.uint8)
        self.lib.shp_facet_seen(self.mod, out.ctypes.data_as(C.POINTER(C.c_ubyte)))
        return out

    # ---- data sets ----
    @property
    def nsets(self):
        return self.lib.shp_dat_nsets(self.dat)

    def set_type(self, s):
        return self.lib.shp_set_type(self.dat, s)

    def nframes(self, s):
        return self.lib.shp_set_nframes(self.dat, s)

    def dims(self, s, f):
        out = (C.c_int * 2)()
        self.lib.shp_frame_dims(self.dat, s, f, out)
        return out[0], out[1]

    def image(self, s, f, which=0):
        """which: 0 fit, 1 obs, 2 oneovervar, 3 lightcurve y (calculated points)."""
        t = self.set_type(s)
        a, b = self.dims(s, f)
        if t == DELAY:
            out = np.zeros((a, b))
        elif t == DOPPLER:
            out = np.zeros(b)
        else:
            out = np.zeros(b if which == 3 else a)
        self.lib.shp_frame_image(self.dat, s, f, which, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def set_image(self, s, f, which, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self.lib.shp_frame_set_image(self.dat, s, f, which, arr.ctypes.data_as(C.POINTER(C.c_double)))

    def scalars(self, s, f):
        out = np.zeros(18)
        self.lib.shp_frame_scalars(self.dat, s, f, out.ctypes.data_as(C.POINTER(C.c_double)))
        keys = ["chi2", "cal", "overflow_o2", "overflow_m2", "badradar_logfactor", "dof",
                "idellim0", "idellim1", "idoplim0", "idoplim1", "dellim0", "dellim1",
                "doplim0", "doplim1", "km2Hz", "overflow_xsec", "overflow_dopmean", "weight"]
        return dict(zip(keys, out))

    def pos(self, s, f):
        return self.lib.shp_frame_pos(self.dat, s, f)

    def pos_array(self, pos, which):
        n = int(self.pos_scalars(pos)["n"])
        out = np.zeros((2 * n + 1, 2 * n + 1))
        self.lib.shp_pos_array(pos, which, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def pos_set_array(self, pos, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.lib.shp_pos_set_array(pos, which, a.ctypes.data_as(C.POINTER(C.c_double)))

    def pos_set_frames(self, pos, oe, se):
        oe = np.ascontiguousarray(oe, dtype=np.float64)
        se = np.ascontiguousarray(se, dtype=np.float64)
        self.lib.shp_pos_set_frames(pos, oe.ctypes.data_as(C.POINTER(C.c_double)),
                                    se.ctypes.data_as(C.POINTER(C.c_double)))

    def pos_copy(self, dst, src):
        self.lib.shp_pos_copy(dst, src)

    def pos_scalars(self, pos):
        out = np.zeros(4)
        self.lib.shp_pos_scalars(pos, out.ctypes.data_as(C.POINTER(C.c_double)))
        return dict(zip(["posbnd_logfactor", "bistatic", "km_per_pixel", "n"], out))

    @property
    def vertices(self):
        return self.lib.shp_mod_vertices(self.mod)

    def set_desc(self, s):
        return self.lib.shp_set_desc(self.dat, s)

    def pos_limits(self, pos):
        out = (C.c_int * 8)()
        self.lib.shp_pos_limits(pos, out)
        return list(out)

    def view(self, s, f, v=0):
        out = np.zeros(29)
        self.lib.shp_view_data(self.dat, s, f, v, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def view_epoch(self, s, f, v=0):
        """a view's epoch and its plane-of-sky spin contribution (orbspin)"""
        out = np.zeros(4)
        self.lib.shp_view_epoch(self.dat, s, f, v, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[0], out[1:]

    def set_npa(self, omega, inertia, method="periodic"):
        """make the model a non-principal-axis rotator: body-frame spin vector
        omega (rad/day), principal moments, and the integration method"""
        w = np.ascontiguousarray(omega, dtype=np.float64)
        I = np.ascontiguousarray(inertia, dtype=np.float64)
        dp = C.POINTER(C.c_double)
        self.lib.shp_set_npa(self.par, self.mod, w.ctypes.data_as(dp), I.ctypes.data_as(dp),
                             {"brute": BRUTE, "periodic": PERIODIC}[method])

    def spin_state(self):
        out = np.zeros(7)
        self.lib.shp_spin_state(self.mod, out.ctypes.data_as(C.POINTER(C.c_double)))
        return {"t0": out[0], "angle": out[1:4], "omega": out[4:7]}

    def law_local(self, radar, ilaw, f):
        """a per-facet law's resolved values at facet f (empty for other laws)"""
        out = np.zeros(5)
        k = self.lib.shp_law_local(self.mod, int(radar), ilaw, f, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[:k]

    def law_table(self, ilaw):
        out = np.zeros(512)
        n = self.lib.shp_law_table(self.mod, ilaw, out.ctypes.data_as(C.POINTER(C.c_double)), 512)
        return out[:n]

    def facet_geom(self, f):
        """facet f's center, theta and phi"""
        out = np.zeros(5)
        self.lib.shp_facet_geom(self.mod, f, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def badphoto(self):
        out = np.zeros(2)
        self.lib.shp_badphoto(self.par, out.ctypes.data_as(C.POINTER(C.c_double)))
        return int(out[0]), out[1]

    def set_view(self, s, f, v, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self.lib.shp_set_view_data(self.dat, s, f, v, arr.ctypes.data_as(C.POINTER(C.c_double)))

    def consts(self, s, f):
        out = np.zeros(9)
        self.lib.shp_frame_consts(self.dat, s, f, out.ctypes.data_as(C.POINTER(C.c_double)))
        return dict(zip(["delcom_vig", "dopcom_vig", "dopDC_vig", "t", "del_per_pixel",
                         "dop_per_pixel", "Ftx", "codemethod", "dopfftlen"], out))

    def par_set(self, which, value):
        """which: nsinc2, sinc2width, pos_smooth, exclude_seen, mask_tol, speckle, dd_gamma"""
        idx = ["nsinc2", "sinc2width", "pos_smooth", "exclude_seen", "mask_tol", "speckle",
               "dd_gamma"].index(which)
        self.lib.shp_par_set(self.par, idx, float(value))

    @property
    def photo(self):
        return self.lib.shp_mod_photo(self.mod)

    def set_radar_law(self, l, law_type, p):
        """Make radar law l a homogeneous law of type 1-7 (radlaw.c:51-117):
        p = [R, C] (cosine), [R, C, cos_cutoff] (quasispecular) or
        [diffuse R, C, quasispecular R, C, cos_cutoff] (hybrids)."""
        a = np.zeros(6)
        a[:len(p)] = p
        if self.lib.shp_set_radar_law(self.mod, l, law_type, a.ctypes.data_as(C.POINTER(C.c_double))):
            raise ValueError("radar law %d: type %d not settable" % (l, law_type))

    def par_flags(self):
        out = np.zeros(6)
        self.lib.shp_par_flags(self.par, out.ctypes.data_as(C.POINTER(C.c_double)))
        return dict(zip(["posbnd", "badradar", "baddiam", "posbnd_logfactor",
                         "badradar_logfactor", "baddiam_logfactor"], out))

    # ---- free parameters ----
    def nparams(self):
        return self.lib.shp_param_count(self.h)

    def param_address(self, i):
        """address of free parameter i's value (the batch's trial targets)"""
        return C.cast(self.lib.shp_param_ptr(self.h, i), C.c_void_p).value

    def get_param(self, i):
        return self.lib.shp_param_ptr(self.h, i)[0]

    def set_param(self, i, val):
        self.lib.shp_param_ptr(self.h, i)[0] = val

    def param_type(self, i):
        return self.lib.shp_param_type(self.h, i)

    # ---- penalties (penalties.c) and the quantities they read ----
    def set_penalties(self, types, weights):
        """par->pen: penalty types (NONSMOOTH ... IMPULSE) and weights."""
        t = np.ascontiguousarray(types, dtype=np.int32)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        if len(t) != len(w):
            raise ValueError("one weight per penalty")
        self.lib.shp_scene_set_penalties(self.h, len(t), t.ctypes.data_as(C.POINTER(C.c_int)),
                                         w.ctypes.data_as(C.POINTER(C.c_double)))
        self.npen = len(t)

    def pen_base(self):
        """par->pen.base[1..n] as penalties() left them."""
        n = getattr(self, "npen", 0)
        out = np.zeros(max(n, 1))
        self.lib.shp_pen_base(self.par, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[:n]

    def set_spin_inertia(self, moments):
        a = np.ascontiguousarray(moments, dtype=np.float64)
        self.lib.shp_set_spin_inertia(self.mod, a.ctypes.data_as(C.POINTER(C.c_double)))

    def set_impulses(self, t, imp):
        t = np.ascontiguousarray(t, dtype=np.float64)
        imp = np.ascontiguousarray(imp, dtype=np.float64).reshape(-1)
        if self.lib.shp_set_impulses(self.mod, len(t), t.ctypes.data_as(C.POINTER(C.c_double)),
                                     imp.ctypes.data_as(C.POINTER(C.c_double))):
            raise ValueError("at most 10 impulses")

    def moments(self, comp=None):
        """(area, volume, com[3], inertia[3x3]) of the model, or of component comp."""
        out = np.zeros(14)
        self.lib.shp_moments(self.mod, 0 if comp is None else 1 + comp, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[0], out[1], out[2:5].copy(), out[5:14].reshape(3, 3).copy()

    def harmonic(self, a=None, b=None):
        """the harmonic component's coefficients (a, b), (l, m) lower-triangular in
        order; with a and b given, writes them first"""
        out_a, out_b = np.zeros(1024), np.zeros(1024)
        dp = C.POINTER(C.c_double)
        if a is not None:
            wa = np.ascontiguousarray(a, dtype=np.float64)
            wb = np.ascontiguousarray(b, dtype=np.float64)
            self.lib.shp_harmonic(self.mod, 1, wa.ctypes.data_as(dp), wb.ctypes.data_as(dp))
        L = self.lib.shp_harmonic(self.mod, 0, out_a.ctypes.data_as(dp), out_b.ctypes.data_as(dp))
        if L < 0:
            raise ValueError("not a harmonic component")
        n = (L + 1) * (L + 2) // 2
        return L, out_a[:n].copy(), out_b[:n].copy()

    def facets(self):
        """facet corners, (nf, 3)"""
        out = np.zeros((self.nf, 3), dtype=np.int32)
        self.lib.shp_facets(self.mod, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def vertex_dirs(self):
        """(u (nv, 3), r (nv,)): vertex directions and deviations"""
        u, r = np.zeros((self.nv, 3)), np.zeros(self.nv)
        self.lib.shp_vertex_dirs(self.mod, u.ctypes.data_as(C.POINTER(C.c_double)),
                                 r.ctypes.data_as(C.POINTER(C.c_double)))
        return u, r

    # ---- joint adjustments (vary_params.c) ----
    def set_vary(self, delcor0=0, radalb=0, optalb=0, dopscale=0):
        """par's vary_* flags (VARY_NONE 0, VARY_SIZE 1, VARY_SPIN 2, VARY_ALL 3)"""
        self.lib.shp_set_vary(self.par, delcor0, radalb, optalb, dopscale)

    def set_cal(self, s, f, state, val):
        """calibration of frame f of set s (a lightcurve's own when set s is one)"""
        self.lib.shp_set_state(self.dat, s, f, 0, ord(state), float(val))

    def set_dopscale(self, s, state, val):
        self.lib.shp_set_state(self.dat, s, 0, 1, ord(state), float(val))

    def set_delcor0_state(self, s, state):
        self.lib.shp_set_state(self.dat, s, 0, 2, ord(state), 0.0)

    def set_weight(self, s, f, w):
        """relative weight of frame f of set s (a lightcurve's own when set s is one)"""
        self.lib.shp_set_state(self.dat, s, f, 3, 0, float(w))

    def vary_weights(self):
        """recompute dat's vary weight sums as read_dat does"""
        self.lib.shp_dat_vary_weights(self.dat)

    def set_comp_transform(self, rot, off):
        """Euler angles (rad) and offset (km) of the model's component"""
        a = np.ascontiguousarray(rot, dtype=np.float64)
        b = np.ascontiguousarray(off, dtype=np.float64)
        self.lib.shp_set_comp_transform(self.mod, a.ctypes.data_as(C.POINTER(C.c_double)),
                                        b.ctypes.data_as(C.POINTER(C.c_double)))

    def set_vertex_r(self, r):
        a = np.ascontiguousarray(r, dtype=np.float64)
        self.lib.shp_set_vertex_r(self.mod, a.ctypes.data_as(C.POINTER(C.c_double)))

    def sides(self):
        """side list of the model: rows (v0, v1, f0, f1)."""
        ns = self.lib.shp_sides(self.mod, None, 0)
        out = np.zeros((ns, 4), dtype=np.int32)
        self.lib.shp_sides(self.mod, out.ctypes.data_as(C.POINTER(C.c_int)), ns)
        return out

    def synthesize_obs(self, snr=20.0, lc_rel_sigma=0.01, seed=1002):
        self.lib.shp_scene_synthesize_obs(self.h, snr, lc_rel_sigma, seed)

    def all_frames(self):
        """(s, f) pairs for every radar frame and every lightcurve set (f = -1)."""
        out = []
        for s in range(self.nsets):
            t = self.set_type(s)
            if t == LGHTCRV:
                out.append((s, -1))
            else:
                for f in range(self.nframes(s)):
                    out.append((s, f))
        return out


METHOD_FP64, METHOD_MIXED = 0, 1
BATCH_GRAPH, BATCH_KEEP_FIT, BATCH_GLOBAL_ZBUF, BATCH_FACET_MAP = 1, 2, 4, 8


def line_search(f, k, x0, step, tol, abstol, max_rounds=40):
    """shape_line_search over a Python objective f(x) (k trial values per
    round, evaluated here one by one); returns (status, xmin, fmin, rounds,
    the x values of every round)."""
    lib = load_hip()
    seen = []

    def cb(ctx, kk, x, out):
        xs = [x[q] for q in range(kk)]
        seen.append(xs)
        for q in range(kk):
            out[q] = f(xs[q])
        return 0

    fn = SEARCH_FN(cb)
    xmin, fmin, rounds = C.c_double(), C.c_double(), C.c_int()
    rc = lib.shape_line_search(fn, None, k, x0, step, tol, abstol, max_rounds, C.byref(xmin), C.byref(fmin),
                               C.byref(rounds))
    return rc, xmin.value, fmin.value, rounds.value, seen


class Batch:
    """Many objective evaluations in one device-resident replay.

    scenes: list of Scene (one model each; all share the first scene's par_t).
    evals:  list of (scene_index, [(param_index, value), ...]).
    """

    def __init__(self, scenes, evals, method=METHOD_MIXED, flags=BATCH_GRAPH):
        self.lib = load_hip()
        self.scenes = scenes
        nm = len(scenes)
        mods = (C.c_void_p * nm)(*[s.mod for s in scenes])
        dats = (C.c_void_p * nm)(*[s.dat for s in scenes])
        self.nevals = len(evals)
        self.eval_model = (C.c_int * self.nevals)(*[m for m, _ in evals])
        self.ntarg = (C.c_int * self.nevals)(*[len(t) for _, t in evals])
        addrs, vals = [], []
        for m, tv in evals:
            for pi, v in tv:
                addrs.append(scenes[m].param_address(pi))
                vals.append(v)
        self.nvals = len(vals)
        tarr = (C.c_void_p * max(1, len(addrs)))(*addrs)
        varr = (C.c_double * max(1, len(vals)))(*vals)
        self.h = self.lib.shape_batch_create(scenes[0].par, mods, dats, nm, self.nevals, self.eval_model,
                                             self.ntarg, tarr, varr, method, flags)
        if not self.h:
            raise RuntimeError("shape_batch_create failed")

    def set_values(self, values):
        arr = (C.c_double * max(1, self.nvals))(*values)
        rc = self.lib.shape_batch_set_values(self.h, arr)
        if rc:
            raise RuntimeError("shape_batch_set_values failed (%d)" % rc)

    def run(self):
        obj = np.zeros(self.nevals)
        chi = np.zeros(self.nevals)
        dp = C.POINTER(C.c_double)
        rc = self.lib.shape_batch_run(self.h, obj.ctypes.data_as(dp), chi.ctypes.data_as(dp))
        if rc:
            raise RuntimeError("shape_batch_run failed (%d)" % rc)
        return obj, chi

    def line_search(self, x0, step, tol, abstol, max_rounds=40):
        """Minimize the objective over the one parameter every evaluation sets
        (shape_batch_line_search); returns (status, xmin, fmin, rounds)."""
        xmin, fmin, rounds = C.c_double(), C.c_double(), C.c_int()
        rc = self.lib.shape_batch_line_search(self.h, x0, step, tol, abstol, max_rounds, C.byref(xmin),
                                              C.byref(fmin), C.byref(rounds))
        if rc < 0:
            raise RuntimeError("shape_batch_line_search failed (%d)" % rc)
        return rc, xmin.value, fmin.value, rounds.value

    def line_searches(self, x0, step, tol, abstol, max_rounds=40):
        """One search per group of nevals / len(x0) consecutive evaluations
        (shape_batch_line_searches); returns (status, xmin[], fmin[],
        status[], rounds)."""
        ns = len(x0)
        dp = C.POINTER(C.c_double)
        a0, st = np.ascontiguousarray(x0, dtype=float), np.ascontiguousarray(step, dtype=float)
        xm, fm, ss = np.zeros(ns), np.zeros(ns), np.zeros(ns, dtype=np.int32)
        r = C.c_int()
        rc = self.lib.shape_batch_line_searches(self.h, ns, a0.ctypes.data_as(dp), st.ctypes.data_as(dp), tol, abstol,
                                                max_rounds, xm.ctypes.data_as(dp), fm.ctypes.data_as(dp),
                                                ss.ctypes.data_as(C.POINTER(C.c_int)), C.byref(r))
        if rc < 0:
            raise RuntimeError("shape_batch_line_searches failed (%d)" % rc)
        return rc, xm, fm, ss, r.value

    def fit_pass(self, x, step, tol, abstol, max_rounds=40):
        """bestfit's pass over the parameters every evaluation sets
        (shape_batch_fit_pass); returns (status, x after the pass, fmin, rounds)."""
        dp = C.POINTER(C.c_double)
        xx = np.array(x, dtype=float)
        st, tl, ab = (np.ascontiguousarray(v, dtype=float) for v in (step, tol, abstol))
        fm, r = C.c_double(), C.c_int()
        rc = self.lib.shape_batch_fit_pass(self.h, st.ctypes.data_as(dp), tl.ctypes.data_as(dp), ab.ctypes.data_as(dp),
                                           max_rounds, xx.ctypes.data_as(dp), C.byref(fm), C.byref(r))
        if rc < 0:
            raise RuntimeError("shape_batch_fit_pass failed (%d)" % rc)
        return rc, xx, fm.value, r.value

    def tune(self, replays=3, force=False):
        """Time the launch-setting candidates (shape_batch_tune) and keep the
        fastest; returns ([tile_grid, replay_grid, collect_grid], best ms)."""
        out, ms = (C.c_int * 3)(), C.c_double()
        rc = self.lib.shape_batch_tune(self.h, replays, 1 if force else 0, out, C.byref(ms))
        if rc:
            raise RuntimeError("shape_batch_tune failed (%d)" % rc)
        return list(out), ms.value

    def get_tune(self):
        out = (C.c_int * 3)()
        self.lib.shape_batch_get_tune(self.h, out)
        return list(out)

    def set_tune(self, settings):
        arr = (C.c_int * 3)(*settings)
        rc = self.lib.shape_batch_set_tune(self.h, arr)
        if rc:
            raise RuntimeError("shape_batch_set_tune failed (%d)" % rc)

    def launch(self):
        rc = self.lib.shape_batch_launch(self.h)
        if rc:
            raise RuntimeError("shape_batch_launch failed (%d)" % rc)

    def wait(self, want=True):
        dp = C.POINTER(C.c_double)
        if not want:
            rc = self.lib.shape_batch_wait(self.h, None, None)
            if rc:
                raise RuntimeError("shape_batch_wait failed (%d)" % rc)
            return None
        obj, chi = np.zeros(self.nevals), np.zeros(self.nevals)
        rc = self.lib.shape_batch_wait(self.h, obj.ctypes.data_as(dp), chi.ctypes.data_as(dp))
        if rc:
            raise RuntimeError("shape_batch_wait failed (%d)" % rc)
        return obj, chi

    def last_ms(self):
        return float(self.lib.shape_batch_last_ms(self.h))

    def frame_results(self, e):
        nf = self.lib.shape_batch_nframes(self.h, e)
        chi, cal = np.zeros(nf), np.zeros(nf)
        dp = C.POINTER(C.c_double)
        self.lib.shape_batch_frame_results(self.h, e, chi.ctypes.data_as(dp), cal.ctypes.data_as(dp))
        return chi, cal

    def fit(self, e, q, shape):
        out = np.zeros(int(np.prod(shape)))
        rc = self.lib.shape_batch_fit(self.h, e, q, out.ctypes.data_as(C.POINTER(C.c_double)))
        if rc:
            raise RuntimeError("shape_batch_fit failed (%d)" % rc)
        return out.reshape(shape)

    def facet_map(self, e, q, n, v=0):
        """Owner facet of every pixel of view v of frame q of evaluation e
        (POS half-width n; -1: none), from the last replay; needs
        BATCH_FACET_MAP."""
        out = np.zeros((2 * n + 1, 2 * n + 1), dtype=np.int32)
        rc = self.lib.shape_batch_facet_map(self.h, e, q, v, out.ctypes.data_as(C.POINTER(C.c_int)))
        if rc:
            raise RuntimeError("shape_batch_facet_map failed (%d)" % rc)
        return out

    def counters(self):
        out = (C.c_int * 4)()
        self.lib.shape_batch_counters(self.h, out)
        return list(out)

    def sizes(self):
        out = (C.c_longlong * 8)()
        self.lib.shape_batch_sizes(self.h, out)
        return dict(zip(["nevals", "nviews", "nframes", "nslots", "total_pix", "fit_len", "max_nf",
                         "max_pix"], list(out)))

    def frame_terms(self, e):
        """Per frame: o2, m2, om, overflow_o2, overflow_m2, cal, chi2, badradar_logfactor."""
        nf = self.lib.shape_batch_nframes(self.h, e)
        out = np.zeros(8 * max(nf, 1))
        rc = self.lib.shape_batch_frame_terms(self.h, e, out.ctypes.data_as(C.POINTER(C.c_double)))
        if rc:
            raise RuntimeError("shape_batch_frame_terms failed (%d)" % rc)
        return out[:8 * nf].reshape(nf, 8)

    def penalties(self, e):
        """(par->pen.base of evaluation e, their weighted sum) after the last replay."""
        n = self.lib.shape_batch_penalties(self.h, e, None, None)
        if n < 0:
            raise RuntimeError("shape_batch_penalties failed (%d)" % n)
        base = np.zeros(max(n, 1))
        total = C.c_double(0.0)
        self.lib.shape_batch_penalties(self.h, e, base.ctypes.data_as(C.POINTER(C.c_double)), C.byref(total))
        return base[:n], total.value

    def time_kernels(self, on=True):
        """Per-kernel HIP-event timing of eager replays (batches without BATCH_GRAPH)."""
        rc = self.lib.shape_batch_time_kernels(self.h, 1 if on else 0)
        if rc:
            raise RuntimeError("shape_batch_time_kernels failed (%d)" % rc)

    def kernel_stats(self):
        """{kernel name: (total ms, launches)} accumulated since time_kernels(True)."""
        out, idx = {}, 0
        name = C.create_string_buffer(128)
        ms, n = C.c_double(), C.c_longlong()
        while True:
            nk = self.lib.shape_batch_kernel_stat(self.h, idx, name, 128, C.byref(ms), C.byref(n))
            if nk < 0 and idx == 0 and nk != -1:
                raise RuntimeError("shape_batch_kernel_stat failed (%d)" % nk)
            if nk <= idx:
                break
            out[name.value.decode()] = (ms.value, n.value)
            idx += 1
        return out

    def footprint(self):
        out = (C.c_longlong * 7)()
        rc = self.lib.shape_batch_footprint(self.h, out)
        if rc:
            raise RuntimeError("shape_batch_footprint failed (%d)" % rc)
        return dict(zip(["bbox_pix", "grid_pix", "window_cells", "image_cells", "nv", "nf", "claimed"],
                        list(out)))

    def stream(self):
        return self.lib.shape_batch_stream(self.h)

    def close(self):
        if self.h:
            self.lib.shape_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Job(Scene):
    """par/mod/dat read from SHAPE's own input files (parameter, model and obs
    files) by libshape_io.so; the same accessors as a synthetic Scene."""

    def __init__(self, parfile, modfile, obsfile):
        # mod_t and dat_t hold the file names in 80 and 40 characters
        if len(modfile) > 79 or len(obsfile) > 39:
            raise ValueError("model / obs file names longer than 79 / 39 characters: use a shorter "
                             "(relative) path")
        self.lib = load_scene()
        self.io = load_io()
        self.spec = None
        self.h = None
        self._bufs = [C.create_string_buffer(int(self.lib.shp_struct_size(w))) for w in range(3)]
        self.par, self.mod, self.dat = (C.cast(b, C.c_void_p).value for b in self._bufs)
        self._read = []
        self._check(self.io.shape_read_par(parfile.encode(), self.par))
        self._read.append(self.io.shape_io_free_par)
        self.lib.shp_set_name(1, self.mod, modfile.encode())
        self.nfpar_mod = self._check(self.io.shape_read_mod(self.par, self.mod))
        self._read.append(self.io.shape_io_free_mod)
        self.lib.shp_set_name(2, self.dat, obsfile.encode())
        self.nfpar_dat = self._check(self.io.shape_read_dat(self.par, self.mod, self.dat))
        self._read.append(self.io.shape_io_free_dat)

    def _check(self, rc):
        if rc < 0:
            self.close()
            raise ShapeIOError(rc, self.io.shape_io_error().decode())
        return rc

    def close(self):
        roots = [self.par, self.mod, self.dat]
        for free, root in zip(getattr(self, "_read", []), roots):
            free(root)
        self._read = []

    @property
    def nv(self):
        return self.lib.shp_mod_nv(self.mod)

    @property
    def nf(self):
        return self.lib.shp_mod_nf(self.mod)

    def _param(self, i):
        """i: a slot of shp_mod_param_ptr, or ("law", radar, ilaw, which, f) for a
        scattering law's parameter (shp_law_param_ptr)"""
        if isinstance(i, tuple):
            ptr = self.lib.shp_law_param_ptr(self.mod, *[int(x) for x in i[1:]])
        else:
            ptr = self.lib.shp_mod_param_ptr(self.mod, i)
        if not ptr:
            raise IndexError("no parameter slot %d" % i)
        return ptr

    def param_address(self, i):
        return C.cast(self._param(i), C.c_void_p).value

    def get_param(self, i):
        return self._param(i)[0]

    def set_param(self, i, val):
        self._param(i)[0] = val

