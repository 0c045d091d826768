// k_mixed.hip -- the FP32/FP64 mixed method: rasterization, chi-square and
// the launch sequence.
//
// FP64 stays where precision is structural: the per-view setup (rotation
// composition at large epoch offsets, COM delay/Doppler offsets; k_fp64.hip)
// and every reduction into chi-square (compensated FP64).  The bulk -- vertex
// projection, rasterization, shading and delay-Doppler binning -- runs in FP32:
//   * raster: one 64-bit atomicMax per covered pixel on a packed (FP32 depth,
//     inverted facet index) key, so the nearest facet wins and equal depths
//     go to the lower facet index, as in the CPU z-test;
//   * binned views (k_tile32.hip): facets are binned into 32x32-pixel tiles
//     and one workgroup per tile rasterizes into an LDS z-buffer, shades the
//     owners and accumulates the tile's delay-Doppler footprint in an LDS
//     window before one flush to HBM;
//   * other views (POS wider than 2048 pixels, or the global path on
//     request): facet per thread into a global z-buffer, then a resolve
//     kernel (k_radar32.hip) per 16x16-pixel tile that reads and clears the
//     keys and synthesizes the same way;
//   * pixels whose owner's smoothed normal faces away, and pixels within the
//     FP32 rounding band of a facet edge, are replayed in FP64 from the facets
//     covering them, reproducing the CPU's owner and inherited
//     cos(scattering);
//   * per-pixel contributions are FP32, every accumulation FP64: the fit
//     image, and chi-square as compensated partial sums over each frame's
//     occupied window, reading and clearing the fit.
// Delay and Doppler bin coordinates are formed relative to an integer anchor
// (ViewRec::D0, P0) so FP32 rounding acts on |x| < ~1e3 bins, not on the
// absolute bin number.
#include "mixed_dev.h"
#include "mixed.h"

namespace shp {

__global__ void k_shade_radar32(ViewRec *__restrict__ views, const GeomSlot *slots, PosDev pd, double *fit, const FrameRec *frames,
                                double *overflow, const double *params, int *counters);
__global__ void k_replay32(ViewRec *__restrict__ views, const GeomSlot *slots, PosDev pd, double *fit, const FrameRec *frames,
                           double *overflow, const double *params);
__global__ void k_block_bounds(GeomSlot *slots, int nslots);
__global__ void k_resolve_lc(ViewRec *views, const GeomSlot *slots, PosDev pd, int src);
__global__ void k_replay_lc(ViewRec *views, const GeomSlot *slots, PosDev pd);
__global__ void k_bin32(const ViewRec *__restrict__ views, const GeomSlot *slots, int *bin_cnt,
                        unsigned int *bin_list, int2 *out_list, int *out_count);
__global__ void k_tile32(const ViewRec *__restrict__ vrd, ViewRec *views, const GeomSlot *__restrict__ slots,
                         PosDev pd, int *bin_cnt, const unsigned int *__restrict__ bin_list, double *fit,
                         const FrameRec *__restrict__ frames, double *overflow, const double *__restrict__ params);
__global__ void k_tile32p(const ViewRec *__restrict__ vrd, ViewRec *views, const GeomSlot *__restrict__ slots,
                          PosDev pd, int *bin_cnt, const unsigned int *__restrict__ bin_list, double *fit,
                          const FrameRec *__restrict__ frames, double *overflow, const double *__restrict__ params,
                          int *work, int ng, int nviews);

// ---------------------------------------------------------------------------
// raster (posvis.c:142-295), one facet per lane:
//   pass 0 (binned views, k_outside32): only the POS extent of facets
//     reaching outside the window (posbnd and the outside extent); the
//     binned tile kernel covers the pixels;
//   pass 1 (views that are not binned, k_raster32): packed-key z-buffer over
//     pixels well inside a facet, marginal pixels recorded for the FP64
//     replay, and the POS rectangles;
//   pass 3 (k_collect32): record the facets covering pixels of the replay
//     list, visiting only facets whose rectangle touches a tile flagged in
//     the stale bitmap.
// Passes 0 and 1 fold the wave's POS rectangle, extent and posbnd flag into
// the view's with one update per wave (raster_limits).
// ---------------------------------------------------------------------------
struct RasterLim {
  int xl0, xl1, yl0, yl1, ob;
  float o0, o1, o2, o3;
};
constexpr int RASTER_CHUNKS = 4;  // chunks of 256 facets per k_raster32 workgroup

__device__ inline void raster_facet(ViewRec &V, const GeomSlot &G, const PosDev &pd, long long sun_off, int src,
                                    int pass, int f, RasterLim &L) {
  const int n = V.n, w = 2 * n + 1;
  L.xl0 = n; L.xl1 = -n; L.yl0 = n; L.yl1 = -n; L.ob = 0;
  L.o0 = 1e30f; L.o1 = -1e30f; L.o2 = 1e30f; L.o3 = -1e30f;
  Proj32 P;
  // pass 3 also admits nearly edge-on facets: the FP64 replay tests facing itself
  if (!(f < G.nf && project32f(V, G, f, src ? V.sa32 : V.oa32, pass == 3 ? -FRONT_TOL : 0.0f, P))) return;
  float imin_d = floorf(fminf(P.x0, fminf(P.x1, P.x2)) + 0.5f);
  float imax_d = floorf(fmaxf(P.x0, fmaxf(P.x1, P.x2)) + 0.5f);
  float jmin_d = floorf(fminf(P.y0, fminf(P.y1, P.y2)) + 0.5f);
  float jmax_d = floorf(fmaxf(P.y0, fmaxf(P.y1, P.y2)) + 0.5f);
  imin_d = fmaxf(imin_d, -1.0e9f); imax_d = fminf(imax_d, 1.0e9f);
  jmin_d = fmaxf(jmin_d, -1.0e9f); jmax_d = fminf(jmax_d, 1.0e9f);
  int imin = (int)imin_d, imax = (int)imax_d, jmin = (int)jmin_d, jmax = (int)jmax_d;
  L.ob = (imin < -n || imax > n || jmin < -n || jmax > n);
  int i1 = max(imin, -n), i2 = min(imax, n), j1 = max(jmin, -n), j2 = min(jmax, n);
  if (i1 > n || i2 < -n || j1 > n || j2 < -n) {
    L.o0 = imin_d; L.o1 = imax_d; L.o2 = jmin_d; L.o3 = jmax_d;
    L.xl0 = min(max(imin, -n), n); L.xl1 = max(min(imax, n), -n);
    L.yl0 = min(max(jmin, -n), n); L.yl1 = max(min(jmax, n), -n);
  } else if (pass == 1) {
    unsigned long long *kb = pd.zkey + V.pos_off + (src ? sun_off : 0);
    // Marginal pixels (near an edge, where FP32 may differ from FP64) are
    // left unclaimed and replayed in FP64: in the observer view, and in a
    // lightcurve point's sun view, whose z-buffer the shadow test reads.
    // (A bistatic radar view's second buffer would claim them instead.)
    const bool mark = !src || V.type == VT_LGHTCRV;
    // one flat loop over the bounding box: a wave runs max(cells) iterations,
    // not max(rows) * max(columns)
    int nc = (i2 - i1 + 1) * (j2 - j1 + 1);
    bool hit = false;
    Edge32 E;
    edge32_init(P, i1, j1, E);
    const float amb = facet_amb(P);
    float di = 0.0f, dj = 0.0f;
    for (int c = 0, i = i1, j = j1; c < nc; c++) {
      float s, t;
      float m = edge32(E, di, dj, s, t);
      long long pix = (long long)(i + n) * w + (j + n);
      if (m >= amb || (!mark && m >= -amb)) {
        hit = true;
        atomicMax(&kb[pix], (unsigned long long)pack_key(depth32(P, s, t), f));
      }
      bool marginal = mark && m >= -amb && m < amb;
      hit = hit || marginal;
      mark_marginal(V, pd, marginal, (int)pix, i, j, src);
      if (++j > j2) { j = j1; i++; dj = 0.0f; di += 1.0f; } else { dj += 1.0f; }
    }
    // The model rectangle grows by the box of every facet that covers a
    // pixel: at most a pixel wider than the covered pixels themselves, which
    // only widens the resolve's tile range.  Pixel positions inside the POS
    // never change the posbnd factor, so the overall extent takes boxes of
    // facets wholly outside only (above).
    if (hit) { L.xl0 = i1; L.xl1 = i2; L.yl0 = j1; L.yl1 = j2; }
  } else if (pass == 3) {
    int tw = (2 * n + STALE_TILE) / STALE_TILE;
    const unsigned char *tb = pd.stale_tiles + V.tile_off;
    bool any = false;
    for (int a = (i1 + n) / STALE_TILE; a <= (i2 + n) / STALE_TILE && !any; a++)
      for (int b = (j1 + n) / STALE_TILE; b <= (j2 + n) / STALE_TILE; b++)
        if (tb[(long long)a * tw + b]) { any = true; break; }
    if (any) {
      int nc = (i2 - i1 + 1) * (j2 - j1 + 1);
      Edge32 E;
      edge32_init(P, i1, j1, E);
      const float amb = facet_amb(P);
      float di = 0.0f, dj = 0.0f;
      for (int c = 0, i = i1, j = j1; c < nc; c++) {
        // candidates: a superset of the facets covering the pixel in FP64
        float s, t;
        if (edge32(E, di, dj, s, t) >= -amb) {
          int si = pd.stale_idx[V.pos_off + (long long)(i + n) * w + (j + n)];
          if (si >= 0) {
            int q = atomicAdd(&pd.stale[si].ncand, 1);
            if (q < STALE_CAND) pd.stale[si].cand[q] = f;
          }
        }
        if (++j > j2) { j = j1; i++; dj = 0.0f; di += 1.0f; } else { dj += 1.0f; }
      }
    }
  }
}

// Fold the lanes' rectangles, extents and posbnd flags into the view's: one
// update per wave, skipped when a read of the current value shows it cannot
// move the limits (they only move outward).  Every lane of the wave calls it.
__device__ inline void raster_limits(ViewRec &V, int src, RasterLim L) {
  const int n = V.n;
  L.xl0 = wmin(L.xl0); L.xl1 = wmax(L.xl1); L.yl0 = wmin(L.yl0); L.yl1 = wmax(L.yl1); L.ob = wmax(L.ob);
  L.o0 = wminf(L.o0); L.o1 = wmaxf(L.o1); L.o2 = wminf(L.o2); L.o3 = wmaxf(L.o3);
  if ((threadIdx.x & 63) == 0) {
    int *xl = src ? V.xlim2 : V.xlim, *yl = src ? V.ylim2 : V.ylim;
    unsigned long long *ov = src ? V.ov2 : V.ov;
    int *obp = src ? &V.outbnd2 : &V.outbnd;
    if (L.xl0 < n && L.xl0 < xl[0]) atomicMin(&xl[0], L.xl0);
    if (L.xl1 > -n && L.xl1 > xl[1]) atomicMax(&xl[1], L.xl1);
    if (L.yl0 < n && L.yl0 < yl[0]) atomicMin(&yl[0], L.yl0);
    if (L.yl1 > -n && L.yl1 > yl[1]) atomicMax(&yl[1], L.yl1);
    if (L.ob && !*obp) atomicOr(obp, 1);
    unsigned long long k;
    if (L.o0 < 1e30f && (k = ord64((double)L.o0)) < ov[0]) atomicMin(&ov[0], k);
    if (L.o1 > -1e30f && (k = ord64(-(double)L.o1)) < ov[1]) atomicMin(&ov[1], k);
    if (L.o2 < 1e30f && (k = ord64((double)L.o2)) < ov[2]) atomicMin(&ov[2], k);
    if (L.o3 > -1e30f && (k = ord64(-(double)L.o3)) < ov[3]) atomicMin(&ov[3], k);
  }
}

// The same for a whole workgroup (every thread calls it): one update of the
// view's limits per workgroup.  Same-address atomics on the view's limits
// serialize at the memory side; with one update per wave they cost more
// than the raster of sub-pixel facets itself.
__device__ inline void raster_limits_wg(ViewRec &V, int src, RasterLim L) {
  __shared__ RasterLim sl[4];
  L.xl0 = wmin(L.xl0); L.xl1 = wmax(L.xl1); L.yl0 = wmin(L.yl0); L.yl1 = wmax(L.yl1); L.ob = wmax(L.ob);
  L.o0 = wminf(L.o0); L.o1 = wmaxf(L.o1); L.o2 = wminf(L.o2); L.o3 = wmaxf(L.o3);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sl[wv] = L;
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
    const RasterLim &M = sl[q];
    L.xl0 = min(L.xl0, M.xl0); L.xl1 = max(L.xl1, M.xl1); L.yl0 = min(L.yl0, M.yl0); L.yl1 = max(L.yl1, M.yl1);
    L.ob |= M.ob;
    L.o0 = fminf(L.o0, M.o0); L.o1 = fmaxf(L.o1, M.o1); L.o2 = fminf(L.o2, M.o2); L.o3 = fmaxf(L.o3, M.o3);
  }
  const int n = V.n;
  int *xl = src ? V.xlim2 : V.xlim, *yl = src ? V.ylim2 : V.ylim;
  unsigned long long *ov = src ? V.ov2 : V.ov;
  int *obp = src ? &V.outbnd2 : &V.outbnd;
  if (L.xl0 < n && L.xl0 < xl[0]) atomicMin(&xl[0], L.xl0);
  if (L.xl1 > -n && L.xl1 > xl[1]) atomicMax(&xl[1], L.xl1);
  if (L.yl0 < n && L.yl0 < yl[0]) atomicMin(&yl[0], L.yl0);
  if (L.yl1 > -n && L.yl1 > yl[1]) atomicMax(&yl[1], L.yl1);
  if (L.ob && !*obp) atomicOr(obp, 1);
  unsigned long long k;
  if (L.o0 < 1e30f && (k = ord64((double)L.o0)) < ov[0]) atomicMin(&ov[0], k);
  if (L.o1 > -1e30f && (k = ord64(-(double)L.o1)) < ov[1]) atomicMin(&ov[1], k);
  if (L.o2 < 1e30f && (k = ord64((double)L.o2)) < ov[2]) atomicMin(&ov[2], k);
  if (L.o3 > -1e30f && (k = ord64(-(double)L.o3)) < ov[3]) atomicMin(&ov[3], k);
}

// pass 1 over every facet of every view that is not binned, RASTER_CHUNKS
// chunks of 256 facets per workgroup: XCD-local eighths of the model (each
// XCD's L2 then keeps one eighth of it).  lc: 0 the radar views, 1 the
// lightcurve points' views (observer, src 0, or sun, src 1).
__global__ __launch_bounds__(256) void k_raster32(ViewRec *__restrict__ views, const GeomSlot *slots, PosDev pd,
                                                  long long sun_off, int src, int lc) {
  ViewRec &V = views[blockIdx.y];
  if (lc ? V.type != VT_LGHTCRV : (V.type != VT_DELAY && V.type != VT_DOPPLER)) return;
  if ((src && !V.bistatic) || V.binned) return;
  const int c0 = xcd_block(blockIdx.x, gridDim.x) * RASTER_CHUNKS;
  const GeomSlot &G = slots[V.slot];
  RasterLim L;
  for (int r = 0; r < RASTER_CHUNKS; r++) {
    RasterLim L1;
    raster_facet(V, G, pd, sun_off, src, 1, (c0 + r) * (int)blockDim.x + (int)threadIdx.x, L1);
    if (r == 0) {
      L = L1;
    } else {
      L.xl0 = min(L.xl0, L1.xl0); L.xl1 = max(L.xl1, L1.xl1); L.yl0 = min(L.yl0, L1.yl0);
      L.yl1 = max(L.yl1, L1.yl1); L.ob |= L1.ob;
      L.o0 = fminf(L.o0, L1.o0); L.o1 = fmaxf(L.o1, L1.o1); L.o2 = fminf(L.o2, L1.o2); L.o3 = fmaxf(L.o3, L1.o3);
    }
  }
  raster_limits_wg(V, src, L);
}

// pass 0 over the bin blocks (view, block) that k_bin32 found reaching
// outside their view's POS window; one wave per block
__global__ __launch_bounds__(256) void k_outside32(ViewRec *__restrict__ views, const GeomSlot *slots, PosDev pd,
                                                   const int2 *list, const int *count) {
  const int nl = *count;
  for (int e = blockIdx.x * 4 + (threadIdx.x >> 6); e < nl; e += gridDim.x * 4) {
    const int2 vb = list[e];
    ViewRec &V = views[vb.x];
    RasterLim L;
    raster_facet(V, slots[V.slot], pd, 0, 0, 0, vb.y * BF + (threadIdx.x & 63), L);
    raster_limits(V, 0, L);
  }
}

// pass 3, grid-stride over (view, 256 facets): a fixed small grid, so an
// empty replay list costs one early exit per workgroup
__global__ __launch_bounds__(256) void k_collect32(ViewRec *__restrict__ views, const GeomSlot *slots, PosDev pd,
                                                   int nviews, int nchunk, int src, int lc) {
  if (*pd.ncollect == 0) return;
  for (long long w = blockIdx.x; w < (long long)nviews * nchunk; w += gridDim.x) {
    ViewRec &V = views[w / nchunk];
    if (lc ? V.type != VT_LGHTCRV : (V.type != VT_DELAY && V.type != VT_DOPPLER)) continue;
    RasterLim L;
    raster_facet(V, slots[V.slot], pd, 0, src, 3, (int)(w % nchunk) * 256 + threadIdx.x, L);
  }
}

// ---------------------------------------------------------------------------
// chi-square of radar frames from the fit image: FP64 compensated partials over
// the frame's occupied window (chi2.c:418-461, :752-785); unless keep is set
// the fit is cleared as it is read, ready for the next evaluation
// ---------------------------------------------------------------------------
__device__ inline void neu(double &s, double &c, double x) {
  double t = s + x;
  if (fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s;
  s = t;
}

__global__ __launch_bounds__(256) void k_chi2_mixed(FrameRec *frames, int nframes, const ViewRec *views, double *fit,
                                                    const float *ow, const float *w32, const double *view_out,
                                                    int keep) {
  FrameRec &F = frames[blockIdx.x];
  if (F.type != VT_DELAY && F.type != VT_DOPPLER) return;
  int r0 = 1, r1 = F.ndel, c0 = 0x7fffffff, c1 = -0x7fffffff;
  int d0 = 0x7fffffff, d1 = -0x7fffffff;
  for (int v = 0; v < F.nviews; v++) {
    const ViewRec &V = views[F.first_view + v];
    c0 = min(c0, V.idoplim[0]); c1 = max(c1, V.idoplim[1]);
    d0 = min(d0, V.idellim[0]); d1 = max(d1, V.idellim[1]);
  }
  if (F.type == VT_DELAY) { r0 = max(1, d0); r1 = min(F.ndel, d1); }
  c0 = max(1, c0); c1 = min(F.ndop, c1);
  int nr = r1 - r0 + 1, nc = c1 - c0 + 1;
  double m2s = 0, m2c = 0, oms = 0, omc = 0;
  double *ft = fit + F.off;
  const float *o = ow + F.obs_off, *wt = w32 + F.obs_off;
  double inv_nv = 1.0 / F.nviews;
  if (nr > 0 && nc > 0) {
    for (int q = threadIdx.x; q < nr * nc; q += blockDim.x) {
      long long idx = (long long)(r0 - 1 + q / nc) * F.ndop + (c0 - 1 + q % nc);
      double fd = ft[idx];
      if (fd != 0.0) {
        if (!keep) ft[idx] = 0.0;
        if (F.nviews > 1) fd *= inv_nv;
        neu(m2s, m2c, fd * fd * (double)wt[idx]);
        neu(oms, omc, fd * (double)o[idx]);
      }
    }
  }
  __shared__ double r[4][256];
  r[0][threadIdx.x] = m2s; r[1][threadIdx.x] = m2c; r[2][threadIdx.x] = oms; r[3][threadIdx.x] = omc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, ac = 0, b = 0, bc = 0;
    for (int q = 0; q < (int)blockDim.x; q++) {
      neu(a, ac, r[0][q]); ac += r[1][q];
      neu(b, bc, r[2][q]); bc += r[3][q];
    }
    double o2 = F.o2, m2 = a + ac, om = b + bc;
    double ovo2 = 0, ovm2 = 0, ovxs = 0, ovdm = 0, blog = 0;
    int bad = 0;
    for (int v = 0; v < F.nviews; v++) {
      const double *vo = view_out + (long long)(F.first_view + v) * VIEW_OUT;
      ovo2 += vo[OUT_OVF]; ovm2 += vo[OUT_OVF + 1]; ovxs += vo[OUT_OVF + 2]; ovdm += vo[OUT_OVF + 3];
      if (vo[OUT_BADRADAR] != 0.0) { bad = 1; blog += vo[OUT_BADRADAR_LOG]; }
    }
    if (F.nviews > 1) { ovo2 /= F.nviews; ovm2 /= F.nviews; ovxs /= F.nviews; ovdm /= F.nviews; }
    F.overflow_o2 = ovo2; F.overflow_m2 = ovm2; F.overflow_xsec = ovxs; F.overflow_dopmean = ovdm;
    o2 += ovo2;
    m2 += ovm2;
    F.badradar = bad;
    F.badradar_logfactor = blog;
    double cal = F.cal_fixed;
    if (F.cal_floats) cal = (om > 0.0) ? om / m2 : 1.0e-40;
    F.cal = cal;
    F.m2 = m2;
    F.om = om;
    F.chi2 = F.weight * (o2 - 2 * cal * om + cal * cal * m2);
  }
}

// ---------------------------------------------------------------------------
// launch sequence of one batch (radar datasets)
// ---------------------------------------------------------------------------
static inline unsigned cdivm(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

int enqueue_mixed(Engine &E) {
  hipStream_t st = E.stream;
  dim3 gf(pad_xcd(cdivm(E.max_nf, 256 * RASTER_CHUNKS)), E.nviews), bt(256);  // (k_raster32)
  int wmax_ = 0;
  for (auto &V : E.h_views) wmax_ = std::max(wmax_, 2 * V.n + 1);
  int tiles = cdivm(wmax_, TILE);
  SHP_LAUNCH(E, "k_fill_u8", k_fill_u8, dim3(256), bt, 0, st, E.pd.stale_tiles, std::max(1LL, E.total_tiles));
  // block and facet bounds, and the facet-major copies every shading kernel reads
  SHP_LAUNCH(E, "k_block_bounds", k_block_bounds, dim3(cdivm(E.max_nblk, 4), E.nslots), bt, 0, st, E.d_slots,
             E.nslots);
  if (E.any_binned) {
    SHP_LAUNCH(E, "k_bin32", k_bin32, dim3(cdivm(E.max_nblk, 256), E.nviews), bt, 0, st, E.d_views, E.d_slots,
               E.d_bin_cnt, E.d_bin_list, E.d_out_list, E.pd.nstale + 4);
    SHP_LAUNCH(E, "k_outside32", k_outside32, dim3(512), bt, 0, st, E.d_views, E.d_slots, E.pd, E.d_out_list,
               E.pd.nstale + 4);
    // one workgroup per tile item, or (tile_grid > 0) the persistent form
    const int ng = (int)pad_xcd((unsigned)(E.max_bin_ntx * E.max_bin_ntx));
    if (E.tune.tile_grid > 0)
      SHP_LAUNCH(E, "k_tile32", k_tile32p, dim3(E.tune.tile_grid), bt, 0, st, E.d_views, E.d_views, E.d_slots, E.pd,
                 E.d_bin_cnt, E.d_bin_list, E.d_fit, E.d_frames, E.d_overflow, E.d_params, E.d_tile_work, ng,
                 E.nviews);
    else
      SHP_LAUNCH(E, "k_tile32", k_tile32, dim3(ng, E.nviews), bt, 0, st, E.d_views, E.d_views, E.d_slots, E.pd,
                 E.d_bin_cnt, E.d_bin_list, E.d_fit, E.d_frames, E.d_overflow, E.d_params);
  }
  if (E.any_global) {
    SHP_LAUNCH(E, "k_raster32", k_raster32, gf, bt, 0, st, E.d_views, E.d_slots, E.pd, E.total_pix, 0, 0);
    SHP_LAUNCH(E, "k_shade_radar32", k_shade_radar32, dim3(tiles * tiles, E.nviews), bt, 0, st, E.d_views, E.d_slots,
               E.pd, E.d_fit, E.d_frames, E.d_overflow, E.d_params, E.pd.nstale);
  }
  SHP_LAUNCH(E, "k_collect32", k_collect32, dim3(std::max(1, E.tune.collect_grid)), bt, 0, st, E.d_views, E.d_slots,
             E.pd, E.nviews, (int)cdivm(E.max_nf, 256), 0, 0);
  SHP_LAUNCH(E, "k_replay32", k_replay32,
             dim3(std::max(1u, std::min(cdivm(E.pd.stale_cap, 256), (unsigned)E.tune.replay_grid))), bt, 0, st,
             E.d_views, E.d_slots, E.pd, E.d_fit, E.d_frames, E.d_overflow, E.d_params);
  if (E.any_lc) {
    // Lightcurve points (k_lc32.hip): the observer view (src 0), then the sun
    // view (src 1), each rasterized in FP32, resolved per pixel in FP64 and
    // its edge-band and faces-away pixels replayed in FP64, through the same
    // key buffer and replay records (the radar replay has finished with
    // them); then the shadow test and the photometry in FP64.
    const int T = 256;
    dim3 gpl(cdivm(E.max_pix, T), E.nviews);
    const dim3 grl(std::max(1u, cdivm(E.max_lc_rec, T)), E.nviews);  // (k_replay_lc: per view)
    for (int src = 0; src < 2; src++) {
      SHP_LAUNCH(E, "k_fill_i32", k_fill_i32, dim3(cdivm(E.nviews, T)), bt, 0, st, E.pd.lc_count, (long long)E.nviews, 0);
      SHP_LAUNCH(E, "k_zero_counters", k_zero_counters, dim3(1), dim3(64), 0, st, E.pd.ncollect, 1);
      SHP_LAUNCH(E, "k_fill_u8", k_fill_u8, dim3(256), bt, 0, st, E.pd.stale_tiles, std::max(1LL, E.total_tiles));
      SHP_LAUNCH(E, src ? "k_raster32/lc/sun" : "k_raster32/lc", k_raster32, gf, bt, 0, st, E.d_views, E.d_slots, E.pd,
                 0LL, src, 1);
      SHP_LAUNCH(E, "k_resolve_lc", k_resolve_lc, gpl, bt, 0, st, E.d_views, E.d_slots, E.pd, src);
      SHP_LAUNCH(E, "k_collect32", k_collect32, dim3(std::max(1, E.tune.collect_grid)), bt, 0, st, E.d_views,
                 E.d_slots, E.pd, E.nviews, (int)cdivm(E.max_nf, 256), src, 1);
      SHP_LAUNCH(E, "k_replay_lc", k_replay_lc, grl, bt, 0, st, E.d_views, E.d_slots, E.pd);
    }
    SHP_LAUNCH(E, "k_posmask64", k_posmask64, gpl, bt, 0, st, E.d_views, E.pd, E.mask_tol);
    SHP_LAUNCH(E, "k_photo64", k_photo64, gpl, bt, 0, st, E.d_views, E.pd, E.d_params);
  }
  SHP_LAUNCH(E, "k_finish_views", k_finish_views, dim3(cdivm(E.nviews, 64)), dim3(64), 0, st, E.d_views, E.nviews, E.d_overflow,
                     E.d_frames, E.d_view_out);
  SHP_LAUNCH(E, "k_chi2_mixed", k_chi2_mixed, dim3(E.nframes), bt, 0, st, E.d_frames, E.nframes, E.d_views, E.d_fit,
                     E.d_obs32, E.d_oov32, E.d_view_out, E.opt.keep_fit);
  if (E.any_lc) {
    SHP_LAUNCH(E, "k_spline", k_spline, dim3(E.nframes), bt, 0, st, E.d_frames, E.nframes, E.d_view_out, E.d_views,
               E.d_lc_x, E.d_lc_t, E.d_lc_y, E.d_lc_y2, E.d_fit);
    SHP_LAUNCH(E, "k_chi2_frames64", k_chi2_frames64, dim3(E.nframes), bt, 0, st, E.d_frames, E.nframes, E.d_fit,
               E.d_obs, E.d_oov, E.d_view_out, 0, 0, 1);
  }
  SHP_LAUNCH(E, "k_objective", k_objective, dim3(std::max(1, E.nevals)), dim3(256), 0, st, E.d_evals, E.nevals, E.d_frames,
                     E.d_views, E.d_view_out, E.d_slot_flags, E.d_slot_logf, E.d_eval_slot, E.pen.n ? E.pen.total : nullptr,
                     E.d_job_flag, E.d_job_log, E.photo_nblk);
  return hipGetLastError() == hipSuccess ? 0 : -42;
}

}  // namespace shp
