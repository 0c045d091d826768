// engine.h -- device-resident model-evaluation engine.
//
// A scene is packed once from the SHAPE ABI structs into flat device arrays:
//   * geometry slots  -- topology (facet vertex indices, vertex->facet CSR),
//                        realization inputs (u, a, r) and realized outputs in
//                        FP64 (x, vertex normals, facet normals) plus FP32
//                        float4 copies for the mixed path;
//   * views           -- one record per (evaluation, frame, view) or
//                        (evaluation, lightcurve point), holding the observer
//                        matrices and, after the setup kernel, the FP64
//                        per-view constants (oa, Doppler gradient, COM shifts);
//   * frames          -- one record per (evaluation, radar frame) or
//                        (evaluation, lightcurve), with offsets into the
//                        obs / weight / fit arenas;
//   * parameter blocks-- one flat FP64 block per evaluation holding every
//                        fittable scalar (spin state, scattering laws, delay
//                        polynomials, scale factors); trial values are written
//                        into these blocks.
// Kernels never call back to the host; one evaluation batch is a fixed
// sequence of launches on one stream, which can be captured into a hipGraph.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "shape_abi.h"

namespace shp {

enum Method { METHOD_FP64 = 0, METHOD_MIXED = 1 };

// ---- flat parameter block layout (doubles) -----------------------------
// spin block
constexpr int PB_ANGLE = 0, PB_OMEGA = 3, PB_OMEGADOT = 6, PB_T0 = 9, PB_LIBAMP = 10,
              PB_LIBFREQ = 11, PB_LIBPHASE = 12, PB_PA = 13, PB_NIMP = 14, PB_TIMP = 15,
              PB_IMP = 25, PB_INERTIA = 55, PB_SPIN_END = 58;
// per radar law (8): type, R, C, cos_cutoff, qsR, qsC, qs_cutoff, spare
constexpr int PB_RAD_STRIDE = 8;
// per optical law (6): type, R, wt, A0, D, k
constexpr int PB_OPT_STRIDE = 6;
// per set (19): angleoff[3], omegaoff[3], dopscale, delcor t0, delcor n, a[0..9]
constexpr int PB_SET_STRIDE = 19, PB_SET_ANGLEOFF = 0, PB_SET_OMEGAOFF = 3, PB_SET_DOPSCALE = 6,
              PB_SET_DCT0 = 7, PB_SET_DCN = 8, PB_SET_DCA = 9;
// per model component: [0..2] scalefactor (VERTEX, HARMONIC) or 2a, a/b, b/c
// (ELLIPSE, OVOID); rot[3]; off[3]; [9] the ovoid's k; from [10] the
// harmonic coefficients, per degree l: a_l0, then (a_lm, b_lm) for m = 1..l
constexpr int PB_COMP_STRIDE = 10, PB_COMP_OVOID_K = 9, PB_COMP_HAR = 10;

struct GeomSlot {
  int nv, nf;
  int model;             // model index (topology / inputs owner)
  int comp_pb;           // parameter-block offset of scalefactor/rot/off
  int eval;              // evaluation whose parameter block realizes this slot (-1: model base)
  const int *fv;         // 3*nf
  const int *af_off;     // nv+1
  const int *af_idx;
  const double *u, *a;   // 3*nv each
  double *r;             // nv (slot-private copy: vertex deviations can be fit parameters)
  double *x, *vn;        // 3*nv
  double *fn, *area, *fx;// 3*nf, nf, 3*nf
  float4 *x32, *vn32;    // nv
  float4 *fn32;          // nf
  float4 *blk;           // mixed method: per 64-facet block, {center, radius} and {cone axis, min cos}
  float4 *sblk;          // mixed method: per 16-facet quarter of each block, {center, radius}
  float4 *fb;            // mixed method: per facet, bounding sphere {center, radius}
  float4 *fvx, *fvn;     // mixed method: per facet, its 3 vertices' x32 / vn32 (one gather, not two)
  unsigned char *seen;   // nf
  const int4 *sides;     // ns: the side (edge) list, {v0, v1, f0, f1} (penalties only)
  int ns;
  int comp_type;         // VERTEX, ELLIPSE, OVOID or HARMONIC
  int nhar;              // HARMONIC: degree L; hbasis holds (L+1)^2 factors per vertex, in the
  const double *hbasis;  //   coefficients' order (afactor[l][0], then afactor[l][m], bfactor[l][m])
  int sf_tie;            // bit j: scale factor j takes factor j-1's value ('=' state)
};

// view types (same numbers as set_t types)
constexpr int VT_DELAY = 0, VT_DOPPLER = 1, VT_LGHTCRV = 3;

struct ViewRec {
  // ---- description (host filled) ----
  int type, set, frm, v, eval, slot, frame_id, is_v0, nviews;
  int n;                 // POS half width
  int smooth, bistatic, host_spin;  // host_spin: ae/spin/deloff/dopoff supplied by host
  int smooth_src;        // smoothing of the source view (calc_fits draws it unsmoothed: 0)
  int exclude_seen;      // 1: do not mark facets seen from this view
  // per-POS entry points: orbit offsets (km, pixels, Doppler bins), the body
  // whose pixels count, rectangles supplied by the caller, and rectangles
  // grown only by pixels this call claims (posvis onto an occupied POS).
  // Zero in the engine's own views.
  double orb[3], orb_x, orb_y, orb_dop;
  int body, keep_rect, rect_by_claim;
  int fp64_skip;         // mixed method: radar view, skipped by the FP64 kernels
  long long pos_off;     // pixel offset of this view's POS buffers
  long long tile_off;    // offset of this view's stale-tile bitmap (mixed method)
  long long bin_off;     // binned mixed raster: first 32x32 tile of this view in the bin arrays
  int bin_ntx, binned;   // tiles per side; 1 when the view takes the binned tile kernel
  int lc_rec_off, lc_rec_cap;  // mixed method, lightcurve views: this view's range of the replay records
  int pb;                // parameter block of the evaluation
  int pb_set, pb_law;    // offsets (within the block) of the set and scattering-law params
  int pb_dc;             // offset of the set block whose delay polynomial applies (= states)
  int law_type;
  // per-evaluation resolved law values (realize_photo): a harmonic or
  // inhomogeneous law's lawk values per facet, or a tabular law's rho table
  const double *lawloc;
  int lawk;
  double kmpp, t;
  double oe[9], se[9], orbspin[3];
  double ae[9], spin[3], intspin[3], deloff, dopoff;
  double delcom_vig, dopcom_vig, dopDC_vig, dopfftlen;
  double Ftx, del_per_pixel, dop_per_pixel;   // dop_per_pixel = dop_per_bin for Doppler frames
  int codemethod, spb, stride, ndel, ndop;
  int nsinc2, sinc2width;
  double frame_dof;
  double phase, intensityfactor;                // lightcurves
  // ---- computed by the setup kernel (FP64) ----
  double oa[9], sa[9], usrc[3];
  double ax, ay, delfact, delshift, dopshift, const1, const2, km2Hz, dop_extra;
  double xincr, yincr, dopdiff_bl, dopdiff_max;
  int badradar_sinc; double badradar_sinc_log;
  int idop0;
  // FP32 copies for the mixed method: rotations, and the delay / Doppler COM
  // bins split into an integer anchor plus an FP32 fraction so that pixel bin
  // coordinates stay small (|rel| < ~1e3) when formed in FP32
  float oa32[9], sa32[9], usrc32[3];
  int D0, P0;
  float dfr, pfr, delfact32, ax32, ay32, const1_32, kmpp32, inv_kmpp32, dcrel32;
  float shade_err32;    // mixed method: FP32 shading error scale (shade_ill)
  float limb32;         // mixed method: cos(scattering) below which the diffuse cosine law
                        // goes to the FP64 replay (C < 1/2: c^(2C-1) grows without bound)
  // ---- accumulated by kernels ----
  int xlim[2], ylim[2], xlim2[2], ylim2[2];
  int outbnd, outbnd2, any_overflow, nstale, nclaimed;  // nclaimed: mixed z-buffer keys claimed
  unsigned long long ov[4], ov2[4];       // ordered doubles: imin, -imax, jmin, -jmax (min-reduced)
  int idellim[2], idoplim[2];
  unsigned long long dellim[2], doplim[2]; // ordered doubles: min delPOS, -max delPOS ...
  double ysum;                             // lightcurve brightness sum
};

struct FrameRec {
  int type, set, frm, eval, first_view, nviews;
  int ndel, ndop;
  long long off;        // offset into the fit arena
  long long obs_off;    // offset into the obs / weight arenas (shared by evaluations)
  int cal_floats;
  double cal_fixed, weight, dof, o2;
  double sdev_sq, lookfact;
  int speckle;
  // lightcurves
  int lc_n, lc_ncalc, lc_nviews;
  long long lc_x_off, lc_t_off; // offsets into lightcurve epoch arenas
  long long lc_y_off;           // offset into the per-evaluation y arena (ncalc+1 doubles)
  long long lc_y2_off;          // offset into the y2 arena (2*(ncalc+1) doubles: y2 and scratch)
  // ---- outputs ----
  double chi2, cal, m2, om;
  double overflow_o2, overflow_m2, overflow_xsec, overflow_dopmean;
  double badradar_logfactor;
  int badradar;
};

struct EvalRec {
  int model, pb, first_frame, nframes;
  int pen_set_off, pen_nsets;  // parameter-block offset and count of the datasets' offset blocks (euleroffs)
  double dof, dof_deldop, dof_doppler, dof_lghtcrv;
  double bad_objfactor;
  // outputs
  double chi2, objective;
  int posbnd, badradar, baddiam;
  double posbnd_logfactor, badradar_logfactor, baddiam_logfactor;
  int photo_job0, photo_njobs;  // this evaluation's realize_photo jobs
  int badphoto;
  double badphoto_logfactor;
};

// realize_photo on the device (k_realize_photo): one job per scattering law
// per evaluation.  Kinds: 0 a homogeneous law (its parameters bounds-checked),
// 1 harmonic (local values from the expansions at the facets' angular
// coordinates), 2 inhomogeneous ('=' locals take the global values),
// 3 tabular ('=' values take the previous one).
constexpr int PHOTO_HOMOG = 0, PHOTO_HARM = 1, PHOTO_INHO = 2, PHOTO_TABLE = 3;
struct PhotoJob {
  int kind, type, k, n;  // k values per facet; n facets (kinds 1, 2) or table length (3)
  int radar;             // a radar law (else optical): the type numbers overlap
  int slot;              // geometry slot whose facet centers kind 1 reads
  long long data;        // absolute parameter offset of the law's data region (kinds 1-3)
  long long loc;         // offset of its resolved values in the local-value arena
  int chk[5], nchk;      // kind 0: absolute parameter offsets checked, in the reference's order
  int coef[5], nhar[5];  // kind 1: each expansion's offset (from data) and degree
  double lo[5], hi[5];   // bounds and checkphotopar mode per value
  int mode[5];
};

}  // namespace shp
