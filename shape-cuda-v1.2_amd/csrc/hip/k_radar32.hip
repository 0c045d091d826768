// k_radar32.hip -- mixed method: resolve of the packed-key z-buffer fused with
// delay-Doppler / Doppler synthesis, and the FP64 replay of pixels whose owner
// faces away or that lie on a facet edge.
#include "replay_dev.h"

namespace shp {

// ---------------------------------------------------------------------------
// one workgroup per TILE x TILE tile of a view's POS (TPP pixels per thread):
//   A. read and clear the tile's keys; the claimed pixels' delay rows and
//      Doppler columns bound the tile's footprint window exactly
//   B. shade each owner, compute its delay-Doppler footprint and add it into
//      the FP64 LDS window at once (footprints leaving the frame, or a window
//      too large for LDS, go straight to HBM)
//   C. flush the window to the FP64 fit image; one update of the view's
//      bin limits (the claimed pixels' bins: the chi-square window) and
//      claimed-key count per workgroup
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_shade_radar32(ViewRec *__restrict__ views, const GeomSlot *slots, PosDev pd,
                                                       double *fit, const FrameRec *frames, double *overflow,
                                                       const double *params, int *counters) {
  ViewRec &V = views[blockIdx.y];
  if ((V.type != VT_DELAY && V.type != VT_DOPPLER) || V.binned) return;
  const bool dd = V.type == VT_DELAY;
  const int n = V.n, w = 2 * n + 1;
  const int tx = (w + TILE - 1) / TILE;
  const int ti = blockIdx.x / tx, tj = blockIdx.x % tx;
  if (ti >= tx) return;
  const int i0 = ti * TILE - n, j0 = tj * TILE - n;
  const int xl0 = V.xlim[0], xl1 = V.xlim[1], yl0 = V.ylim[0], yl1 = V.ylim[1];
  if (i0 > xl1 || i0 + TILE - 1 < xl0 || j0 > yl1 || j0 + TILE - 1 < yl0) return;
  __shared__ double win[WIN_MAX];
  __shared__ int shi[2][4];
  __shared__ int shw[4][4];
  const GeomSlot &G = slots[V.slot];
  const double *pl = params + V.pb + V.pb_law;
  const int width = V.sinc2width;

  // Pixel order within the tile: a wave's 64 pixels lie along the axis on
  // which the Doppler coordinate changes faster, so they spread over more
  // Doppler bins and collide less in the LDS window.
  const bool along_j = fabs(V.ay) >= fabs(V.ax);

  // ---- A: keys, claimed pixels and the footprint window ----
  unsigned long long key[TPP];
  int lo_d = 0x7fffffff, hi_d = -0x7fffffff, lo_p = 0x7fffffff, hi_p = -0x7fffffff, ncl = 0;
#pragma unroll
  for (int m = 0; m < TPP; m++) {
    int q = threadIdx.x + m * (int)blockDim.x;
    int i = i0 + (along_j ? q / TILE : q % TILE), j = j0 + (along_j ? q % TILE : q / TILE);
    key[m] = 0ULL;
    if (i <= n && j <= n && i >= xl0 && i <= xl1 && j >= yl0 && j <= yl1) {
      long long p = V.pos_off + (long long)(i + n) * w + (j + n);
      unsigned long long k = pd.zkey[p];
      if (k) {
        pd.zkey[p] = 0ULL;
        ncl++;
        // marginal pixels (near a facet edge) are resolved by the FP64 replay
        if (pd.stale_tiles[stale_tile(V, i, j)] && pd.stale_idx[p] >= 0) k = 0ULL;
      }
      key[m] = k;
      if (k) {
#pragma clang fp contract(off)
        double dop64 = V.ax * (double)i + V.ay * (double)j + V.dopshift;
        lo_p = min(lo_p, (int)floor(dop64 - V.dopdiff_max + 1 - width / 2.0) - V.P0);
        hi_p = max(hi_p, (int)floor(dop64 + V.dopdiff_max + width / 2.0) - V.P0);
        if (dd) {
          double del64 = (double)key_z(k) * V.delfact + V.delshift;
          lo_d = min(lo_d, (int)floor(del64 - V.const1) + 1 - V.D0);
          hi_d = max(hi_d, (int)ceil(del64 + V.const1) - 1 - V.D0);
        }
      }
    }
  }
  {
    int nhp = -hi_p, nhd = -hi_d;
    blk_min4(lo_p, nhp, lo_d, nhd, shw);
    hi_p = -nhp; hi_d = -nhd;
    if (!dd) lo_d = hi_d = 0;
  }
  const bool any = lo_p != 0x7fffffff;
  // the claimed pixels' bins bound the lit pixels' bins: they are the tile's
  // contribution to the frame's delay / Doppler limits (the chi-square window)
  const int lim_d0 = lo_d, lim_d1 = hi_d, lim_p0 = lo_p, lim_p1 = hi_p;
  if (any) {  // one bin of margin: phase B recomputes the same floors
    lo_p--; hi_p++;
    if (dd) { lo_d--; hi_d++; }
  }
  const int wd = hi_d - lo_d + 1, wp = hi_p - lo_p + 1;
  const bool use_lds = any && (long long)wd * wp <= WIN_MAX;
  if (use_lds) {
    for (int q = threadIdx.x; q < wd * wp; q += blockDim.x) win[q] = 0.0;
    __syncthreads();
  }

  // ---- B: shade, footprint, accumulate ----
  double *img = fit + frames[V.frame_id].off;
  double *ov = overflow + (long long)blockIdx.y * 2000;
  int anyo = 0;
  for (int m = 0; m < TPP; m++) {
    unsigned long long k = key[m];
    if (!k) continue;
    int q = threadIdx.x + m * (int)blockDim.x;
    int i = i0 + (along_j ? q / TILE : q % TILE), j = j0 + (along_j ? q % TILE : q / TILE);
    int f = key_facet(k);
    if (pd.fmap) pd.fmap[V.pos_off + (long long)(i + n) * w + (j + n)] = f;  // (inspection)
    Proj32 P;
    corners32(V, G, f, V.oa32, P);
    float s, t;
    bary32(P, i, j, s, t);  // the raster's coordinates, to FP32 rounding
    const float nz = normal32_los(G, f, V.oa32, V.smooth, s, t);
    // cos(scattering) of an owner facing away is inherited from an earlier
    // claimant, an ill-conditioned FP32 cosine needs FP64, and (NaN law) a
    // pixel too close to a quasispecular law's cutoff is beyond FP32: FP64
    // replay (records reserved once per wave)
    const bool ill = !(nz > 0.0f) || shade_ill(V, G, f, P);
    const float cl = ill ? 0.0f : radar_cl(V, G, pl, f, s, t, nz);
    const bool rep = ill || cl != cl;
    int si = stale_slot(pd, rep);
    if (rep) {
      if (si >= 0) stale_init(V, pd, si, blockIdx.y, (i + n) * w + (j + n), i, j);
      continue;
    }
    RadarPix R;
    radar_pixel(V, pl, i, j, (double)key_z(k), cl, R);
    if (!R.inb) anyo = 1;
    if (use_lds && R.inb) {
      radar_scatter_lds(V, R, win + (dd ? (R.id0 - lo_d) * wp : 0) + (R.ip0 - lo_p), wp);
    } else {
      radar_scatter_global(V, R, img, ov);
    }
  }

  // ---- C: flush the window; limits and claimed keys, once per workgroup ----
  if (use_lds) {
    __syncthreads();
    int ndop = V.ndop;
    for (int q = threadIdx.x; q < wd * wp; q += blockDim.x) {
      double v = win[q];
      if (v != 0.0) {
        int r = q / wp, c = q % wp;
        int idel = V.D0 + lo_d + r, idop = V.P0 + lo_p + c;
        if (dd) atomicAdd(&img[(long long)(idel - 1) * ndop + (idop - 1)], v);
        else atomicAdd(&img[idop - 1], v);
      }
    }
  }
  anyo = wmax(anyo);
  ncl = wsum(ncl);
  int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { shi[0][wv] = anyo; shi[1][wv] = ncl; }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int q = 1; q < (int)(blockDim.x >> 6); q++) { anyo |= shi[0][q]; ncl += shi[1][q]; }
  if (ncl) atomicAdd(&V.nclaimed, ncl);
  if (!any) return;
  // the limits only move outward: skip updates a read of the current value
  // shows cannot move them
  if (dd) {
    if (V.D0 + lim_d0 < V.idellim[0]) atomicMin(&V.idellim[0], V.D0 + lim_d0);
    if (V.D0 + lim_d1 > V.idellim[1]) atomicMax(&V.idellim[1], V.D0 + lim_d1);
  }
  if (V.P0 + lim_p0 < V.idoplim[0]) atomicMin(&V.idoplim[0], V.P0 + lim_p0);
  if (V.P0 + lim_p1 > V.idoplim[1]) atomicMax(&V.idoplim[1], V.P0 + lim_p1);
  if (!dd && anyo && !V.any_overflow) atomicOr(&V.any_overflow, 1);
}

__device__ void replay32_one(ViewRec *views, const GeomSlot *slots, const PosDev &pd, double *fit,
                             const FrameRec *frames, double *overflow, const double *params, int si) {
  StaleRec &S = pd.stale[si];
  int vi = S.view;
  ViewRec &V = views[vi];
  const GeomSlot &G = slots[V.slot];
  int n = V.n, w = 2 * n + 1;
  int i = S.pix / w - n, j = S.pix % w - n;
  int nc = min(S.ncand, STALE_CAND);
  if (S.ncand > STALE_CAND) atomicOr(pd.error, 2);
  int c[STALE_CAND];
  for (int q = 0; q < nc; q++) c[q] = S.cand[q];
  sort_candidates(c, nc);
  double zz, cosa;
  const int fac = replay_owner64(V, G, c, nc, i, j, zz, cosa);
  if (pd.fmap) pd.fmap[V.pos_off + S.pix] = fac;  // (inspection)
  pd.stale_idx[V.pos_off + S.pix] = -1;
  if (!(cosa > 0.0)) return;
  replay_emit(V, params + V.pb + V.pb_law, i, j, zz, cosa, fac, fit + frames[V.frame_id].off,
              overflow + (long long)vi * 2000);
}

// ---------------------------------------------------------------------------
// FP64 replay of a pixel whose FP32 owner faces away, or that lies within the
// FP32 rounding band of a facet edge: the candidates claim it in facet order
// with the CPU rule `z > zz || fac < 0`, in the CPU's own FP64 arithmetic, so
// its owner, depth and (inherited) cos(scattering) are the CPU's
// (posvis.c:142-275); then the pixel's power is scattered directly to HBM.
// Grid-stride over the records: the grid is sized for the cap, not the count.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_replay32(ViewRec *__restrict__ views, const GeomSlot *slots, PosDev pd, double *fit,
                                                  const FrameRec *frames, double *overflow, const double *params) {
  int ns = min(*pd.nstale, pd.stale_cap);
  for (int si = blockIdx.x * blockDim.x + threadIdx.x; si < ns; si += gridDim.x * blockDim.x)
    replay32_one(views, slots, pd, fit, frames, overflow, params, si);
}

}  // namespace shp
