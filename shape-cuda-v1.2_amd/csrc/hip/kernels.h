// kernels.h -- device buffer bundles and kernel launchers shared by the
// engine and the drop-in entry points.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"

namespace shp {

constexpr int STALE_CAND = 64;  // max facets recorded per replayed pixel
constexpr int STALE_TILE = 8;   // stale-pixel bitmap granularity (pixels)
constexpr int SYN_TILE = 16;    // FP64 synthesis: 16 x 16 POS pixels per workgroup
inline unsigned syn_blocks(int n) { unsigned t = (2 * n + 1 + SYN_TILE - 1) / SYN_TILE; return t * t; }

// Binned raster (k_tile32.hip): facets are binned by blocks of BF consecutive
// facets into BT x BT pixel tiles of each view; one workgroup per tile.
constexpr int BT = 32, BT_PIX = BT * BT;
constexpr int BF = 64;           // facets per block: one wave
constexpr int BIN_CAP = 128;     // block ids stored per tile (more: the tile tests every block)
constexpr int BIN_MAX_NTX = 64;  // tiles per side of a binned view (POS width <= 2048)
constexpr int EV_CAP = 2048;     // covering (pixel, facet) events kept per tile
constexpr int TS_CAP = 64;       // replayed pixels per tile whose candidates come from those events
constexpr int TS_CAND = 32;      // candidates kept per such pixel (more: the global raster collects them)
constexpr int WIN_T = 1024;      // FP64 LDS delay-Doppler window of a tile (8 KiB)
constexpr int EV_FACET_BITS = 22;  // event = local pixel << 22 | facet (facets < 2^22)
constexpr int TILE_WG_PER_CU = 6;  // the tile kernel's occupancy (registers and LDS)

// Workgroups are dealt to the chip's 8 XCDs round-robin by linear block id.
constexpr int NXCD = 8;

// A pixel whose final owner has a back-facing smoothed normal; its
// cos(scattering) is whatever an earlier claimant left (posvis.c:243-275).
struct StaleRec {
  int view, pix, oldf, ncand, src;
  double oldz, oldcose, oldcosi;
  int cand[STALE_CAND];
};

// POS buffers of all views, each view at ViewRec::pos_off.  FP64 arrays are
// present for the FP64 method and the drop-ins; the mixed path uses zkey only.
struct PosDev {
  unsigned long long *zkey;  // z-buffer keys (FP64: ordered depth; mixed: packed key)
  int *fbuf;                 // FP64 pass-2 owners
  double *z, *cose, *cosi, *b;
  int *f, *body, *comp;      // body/comp only for the drop-ins (NULL otherwise)
  double *zill, *cosill;
  int *fill, *bodyill, *compill;
  int *stale_idx;            // per pixel: index into stale[], or -1
  unsigned char *stale_tiles;// mixed path: per view (at ViewRec::tile_off), 1 where an 8x8 tile holds a stale pixel
  StaleRec *stale;
  int *nstale;               // replay records in use
  int *ncollect;             // mixed path: records whose candidates the global raster (k_collect32) collects
  int stale_cap;
  int *error;                // bit flags: 1 stale list full, 2 too many candidates
  int *fmap;                 // mixed path, when asked for: each POS pixel's owner facet (-1: none)
  int *lc_count;             // mixed path: replay records in use per lightcurve view (ViewRec::lc_rec_off)
};

// per-view outputs written by k_finish_views (doubles)
constexpr int OUT_OUTBND = 0, OUT_POSBND_LOG = 2, OUT_DELLIM = 4, OUT_DOPLIM = 6, OUT_OVF = 8,
              OUT_BADRADAR = 12, OUT_BADRADAR_LOG = 13, OUT_IDELLIM = 14, OUT_IDOPLIM = 16,
              OUT_KM2HZ = 18, OUT_YSUM = 19, VIEW_OUT = 20;

// ---- moments and penalties (k_pen.hip) ----
// Per geometry slot, MOM_N doubles: the model's moments (realize_mod.c:757-840)
// and what the penalties read from its geometry (penalties.c:597-1208).
constexpr int MOM_AREA = 0, MOM_VOL = 1, MOM_COM = 2, MOM_I = 5, MOM_NONSMOOTH = 14, MOM_CONCAV = 15,
              MOM_LO = 16, MOM_HI = 19, MOM_PM = 22, MOM_AP = 25, MOM_SF = 34, MOM_RDEV = 37, MOM_MAXRDEV = 38,
              MOM_MAXELL = 39, MOM_BIF = 40, MOM_JACOBI_FAIL = 41, MOM_N = 48;
// partial sums per (slot, workgroup): facets and sides / vertices
constexpr int MP1_N = 19, MP2_N = 3 + 2 * 19;
// what the vertex pass computes (flags)
constexpr int PEN_NEED_SIDES = 1, PEN_NEED_VERTS = 2, PEN_NEED_PA = 4;

// The penalty list of an objective (par->pen), shared by every evaluation.
struct PenDev {
  int n;                 // 0: no penalties
  const int *type;       // n
  const double *weight;  // n
  double *base;          // nevals x n: par->pen.base of each evaluation
  double *total;         // nevals: the penalty sum added to chi2/dof
  const double *mom;     // nslots x MOM_N
  const double *law;     // nevals x PEN_LAW_N: the law-reading penalties (k_pen_laws), or NULL
};
// the law-reading penalties' slots in PenDev::law: type - 1 for OPTALBDEL ..
// RAD_C_VAR, then NONCOSINE
constexpr int PEN_LAW_N = 9;
__host__ __device__ inline int pen_law_slot(int type) { return type == NONCOSINE ? 8 : type - 1; }

// which penalty types the kernels evaluate, and what each needs
int pen_supported(int type);
int pen_needs(int type);

__global__ void k_moments_part(const GeomSlot *slots, double *part, int nb);
__global__ void k_moments_fin(const GeomSlot *slots, const double *params, const double *part, int nb, double *mom,
                              const double *given, int need_pa);
__global__ void k_pen_vert_part(const GeomSlot *slots, const double *mom, double *part, int nb, int need);
__global__ void k_pen_vert_fin(const GeomSlot *slots, const double *part, int nb, double *mom, int need);
__global__ void k_pen_eval(const EvalRec *evals, int nevals, const int *eval_slot, const double *params, PenDev pen);
int pen_supported_laws(int type);
__global__ void k_pen_laws(const EvalRec *evals, int nevals, const int *eval_slot, const GeomSlot *slots,
                           const PhotoJob *jobs, const double *lawloc, const double *noncos, int need, double *out);

// ---- FP64 kernels (k_fp64.hip) ----
__global__ void k_setup_views(ViewRec *views, int nviews, const double *params);
__global__ void k_pos_setup(ViewRec *views);
__global__ void k_pos_seed(ViewRec *views, PosDev pd, int src);
__global__ void k_radlaw_one(const double *pl, int type, double c, const double *loc, int f, double *out);
__global__ void k_realize_vertices(GeomSlot *slots, int nslot, const double *params, int *slot_flags,
                                   double *slot_logf);
__global__ void k_realize_harmonic_bad(GeomSlot *slots, const double *params, int *slot_flags, double *slot_logf);
__global__ void k_realize_facets(GeomSlot *slots, int nslot);
__global__ void k_realize_vnormals(GeomSlot *slots, int nslot);
__global__ void k_raster64(ViewRec *views, const GeomSlot *slots, PosDev pd, int src, int pass);
__global__ void k_resolve64(ViewRec *views, const GeomSlot *slots, PosDev pd, int src, int merge, int blank,
                            int body, int comp);
__global__ void k_replay64(ViewRec *views, const GeomSlot *slots, PosDev pd);
__global__ void k_posmask64(ViewRec *views, PosDev pd, double tol);
__global__ void k_photo64(ViewRec *views, PosDev pd, const double *params);
__global__ void k_deldop64(ViewRec *views, FrameRec *frames, PosDev pd, double *fit, const double *params);
__global__ void k_doppler64(ViewRec *views, FrameRec *frames, PosDev pd, double *fit, double *overflow,
                            const double *params);
__global__ void k_finish_views(ViewRec *views, int nviews, double *overflow, const FrameRec *frames,
                               double *view_out);
__global__ void k_spline(FrameRec *frames, int nframes, const double *view_out, const ViewRec *views,
                         const double *lc_x, const double *lc_t, double *lc_y, double *lc_y2, double *fit);
__global__ void k_chi2_frames64(FrameRec *frames, int nframes, double *fit, const double *obs, const double *oov,
                                const double *view_out, int clear_fit, int ovf_from_host, int lc_only);
__global__ void k_objective(EvalRec *evals, int nevals, const FrameRec *frames, const ViewRec *views,
                            const double *view_out, const int *slot_flags, const double *slot_logf,
                            const int *eval_slot, const double *pen_total, const int *job_flag,
                            const double *job_log, int photo_nblk);
__global__ void k_realize_photo(const PhotoJob *jobs, const double *params, const GeomSlot *slots, double *lawloc,
                                int *job_flag, double *job_log);

}  // namespace shp

namespace shp {
__global__ void k_mark_seen(ViewRec *views, GeomSlot *slots, PosDev pd, double mincos);
__global__ void k_fill_i32(int *p, long long n, int v);
__global__ void k_fill_u8(unsigned char *p, long long n);
__global__ void k_fill_f64(double *p, long long n, double v);
__global__ void k_zero_counters(int *p, int n);
}  // namespace shp

namespace shp {
__global__ void k_frame_o2(FrameRec *frames, int nframes, const double *obs, const double *oov);
__global__ void k_chi2_sets(const FrameRec *frames, int nframes, double *set_chi2, int nsets, double *total);
}  // namespace shp
