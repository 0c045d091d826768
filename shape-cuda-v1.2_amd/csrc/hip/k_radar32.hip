// k_radar32.hip -- mixed method: resolve of the packed-key z-buffer fused with
// delay-Doppler / Doppler synthesis, and the FP64 replay of pixels whose owner
// faces away or that lie on a facet edge.
#include "fp64_dev.h"
#include "mixed_dev.h"

namespace shp {

// Per-view bin limits of lit pixels (pos2deldop.c:337-354) and the count of
// claimed keys: wave reductions, then one update per workgroup.  The limits
// only move outward, so an update that a (possibly stale) read of the current
// value shows cannot move them is skipped: interior tiles issue no atomics.
__device__ inline void view_limits_block(ViewRec &V, bool lit, const RadarPix &R, bool claimed) {
  const bool dd = V.type == VT_DELAY;
  __shared__ int si[4][6];
  __shared__ float sf[4][4];
  int a0 = lit ? R.id0 : 0x7fffffff, a1 = lit ? R.id1 : -0x7fffffff;
  int b0 = lit ? R.ip0 : 0x7fffffff, b1 = lit ? R.ip1 : -0x7fffffff;
  float d0 = lit ? R.delrel : 1e30f, d1 = lit ? R.delrel : -1e30f;
  float p0 = lit ? R.doprel : 1e30f, p1 = lit ? R.doprel : -1e30f;
  a0 = wmin(a0); a1 = wmax(a1); b0 = wmin(b0); b1 = wmax(b1);
  d0 = wminf(d0); d1 = wmaxf(d1); p0 = wminf(p0); p1 = wmaxf(p1);
  int anyo = __ballot(lit && !R.inb) != 0ULL;
  int ncl = __popcll(__ballot(claimed));
  int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    si[wv][0] = a0; si[wv][1] = a1; si[wv][2] = b0; si[wv][3] = b1; si[wv][4] = anyo; si[wv][5] = ncl;
    sf[wv][0] = d0; sf[wv][1] = d1; sf[wv][2] = p0; sf[wv][3] = p1;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
    a0 = min(a0, si[q][0]); a1 = max(a1, si[q][1]); b0 = min(b0, si[q][2]); b1 = max(b1, si[q][3]);
    anyo |= si[q][4]; ncl += si[q][5];
    d0 = fminf(d0, sf[q][0]); d1 = fmaxf(d1, sf[q][1]); p0 = fminf(p0, sf[q][2]); p1 = fmaxf(p1, sf[q][3]);
  }
  if (ncl) atomicAdd(&V.nclaimed, ncl);
  if (a0 == 0x7fffffff) return;
  if (dd) {
    if (V.D0 + a0 < V.idellim[0]) atomicMin(&V.idellim[0], V.D0 + a0);
    if (V.D0 + a1 > V.idellim[1]) atomicMax(&V.idellim[1], V.D0 + a1);
    unsigned long long k0 = ord64((double)V.D0 + (double)d0), k1 = ord64(-((double)V.D0 + (double)d1));
    if (k0 < V.dellim[0]) atomicMin(&V.dellim[0], k0);
    if (k1 < V.dellim[1]) atomicMin(&V.dellim[1], k1);
  }
  if (V.P0 + b0 < V.idoplim[0]) atomicMin(&V.idoplim[0], V.P0 + b0);
  if (V.P0 + b1 > V.idoplim[1]) atomicMax(&V.idoplim[1], V.P0 + b1);
  unsigned long long k2 = ord64((double)V.P0 + (double)p0), k3 = ord64(-((double)V.P0 + (double)p1));
  if (k2 < V.doplim[0]) atomicMin(&V.doplim[0], k2);
  if (k3 < V.doplim[1]) atomicMin(&V.doplim[1], k3);
  if (!dd && anyo && !V.any_overflow) atomicOr(&V.any_overflow, 1);
}

// ---------------------------------------------------------------------------
// one workgroup per 16x16 tile of a view's POS: read and clear the keys,
// shade the owners, accumulate the tile's delay-Doppler footprint in LDS,
// flush once to the FP64 fit image
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_shade_radar32(ViewRec *views, const GeomSlot *slots, PosDev pd,
                                                       double *fit, const FrameRec *frames, double *overflow,
                                                       const double *params, int *counters) {
  ViewRec &V = views[blockIdx.y];
  if (V.type != VT_DELAY && V.type != VT_DOPPLER) return;
  const bool dd = V.type == VT_DELAY;
  int n = V.n, w = 2 * n + 1;
  int tx = (w + TILE - 1) / TILE;
  int ti = blockIdx.x / tx, tj = blockIdx.x % tx;
  if (ti >= tx) return;
  int i0 = ti * TILE - n, j0 = tj * TILE - n;
  if (i0 > V.xlim[1] || i0 + TILE - 1 < V.xlim[0] || j0 > V.ylim[1] || j0 + TILE - 1 < V.ylim[0]) return;
  int i = i0 + (threadIdx.x >> 4), j = j0 + (threadIdx.x & 15);
  const GeomSlot &G = slots[V.slot];
  const FrameRec &F = frames[V.frame_id];
  const double *pl = params + V.pb + V.pb_law;
  bool lit = false;
  RadarPix R;
  R.inb = false; R.wide = false; R.amp = 0.0f; R.delrel = R.doprel = 0.0f;
  R.id0 = R.id1 = R.ip0 = R.ip1 = 0;
  bool claimed = false;
  if (i <= n && j <= n && i >= V.xlim[0] && i <= V.xlim[1] && j >= V.ylim[0] && j <= V.ylim[1]) {
    long long pix = (long long)(i + n) * w + (j + n);
    long long p = V.pos_off + pix;
    unsigned long long k = pd.zkey[p];
    if (k) pd.zkey[p] = 0ULL;
    claimed = k != 0ULL;
    // marginal pixels (near a facet edge) are resolved by the FP64 replay
    if (k && pd.stale_tiles[stale_tile(V, i, j)] && pd.stale_idx[p] >= 0) k = 0ULL;
    if (k) {
      int f = key_facet(k);
      float z = key_z(k);
      Proj32 P;
      corners32(V, G, f, V.oa32, P);
      float s, t;
      bary32(P, i, j, s, t);  // same arithmetic as the raster
      float nx, ny, nz;
      normal32(G, f, V.oa32, V.smooth, s, t, nx, ny, nz);
      if (nz > 0.0f) {
        lit = true;
        radar_pixel(V, pl, i, j, z, nz, R);
        if (R.wide) atomicAdd(&counters[3], 1);
      } else {
        // cos(scattering) is inherited from an earlier claimant: replay later
        int si = atomicAdd(pd.nstale, 1);
        if (si < pd.stale_cap) {
          StaleRec &S = pd.stale[si];
          S.view = blockIdx.y; S.pix = (int)pix; S.oldf = -1; S.oldz = -1.0e20; S.oldcose = 0.0;
          S.oldcosi = 0.0; S.ncand = 0; S.src = 0;
          pd.stale_idx[p] = si;
          pd.stale_tiles[stale_tile(V, i, j)] = 1;
        } else {
          atomicOr(pd.error, 1);
        }
      }
    }
  }
  view_limits_block(V, lit, R, claimed);
  // ---- LDS window over the tile's in-bounds footprints ----
  // FP32 contributions, FP64 accumulation: a bin gathers thousands of pixels
  // and chi-square later cancels o2 - om^2/m2, so the sums must not round at
  // FP32 precision
  __shared__ double win[WIN_MAX];
  __shared__ int wb[4];
  if (threadIdx.x == 0) { wb[0] = 0x7fffffff; wb[1] = -0x7fffffff; wb[2] = 0x7fffffff; wb[3] = -0x7fffffff; }
  __syncthreads();
  bool scat = lit && R.inb && !R.wide;
  if (scat) {
    if (dd) { atomicMin(&wb[0], R.id0); atomicMax(&wb[1], R.id1); }
    atomicMin(&wb[2], R.ip0); atomicMax(&wb[3], R.ip1);
  }
  __syncthreads();
  int lo_d = dd ? wb[0] : 0, hi_d = dd ? wb[1] : 0, lo_p = wb[2], hi_p = wb[3];
  int wd = hi_d - lo_d + 1, wp = hi_p - lo_p + 1;
  bool use_lds = (lo_p != 0x7fffffff) && (long long)wd * wp <= WIN_MAX;
  double *img = fit + F.off;
  int ndop = V.ndop;
  double *ov = overflow + (long long)blockIdx.y * 2000;
  if (use_lds) {
    for (int q = threadIdx.x; q < wd * wp; q += blockDim.x) win[q] = 0.0;
    __syncthreads();
    if (scat) {
      int np = R.ip1 - R.ip0 + 1, nd = dd ? R.id1 - R.id0 + 1 : 1;
#pragma unroll
      for (int a = 0; a < DW; a++) {
        if (a < nd) {
          float ad = dd ? R.amp * R.dw[a] : R.amp;
          double *row = win + (dd ? (R.id0 + a - lo_d) * wp : 0) + (R.ip0 - lo_p);
#pragma unroll
          for (int b = 0; b < PW; b++)
            if (b < np) atomicAdd(&row[b], (double)(ad * R.pw[b]));
        }
      }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < wd * wp; q += blockDim.x) {
      double v = win[q];
      if (v != 0.0) {
        int r = q / wp, c = q % wp;
        int idel = V.D0 + lo_d + r, idop = V.P0 + lo_p + c;
        if (dd) atomicAdd(&img[(long long)(idel - 1) * ndop + (idop - 1)], v);
        else atomicAdd(&img[idop - 1], v);
      }
    }
    if (lit && !R.inb && !dd) radar_scatter_global(V, R, img, ov);
  } else if (lit) {
    radar_scatter_global(V, R, img, ov);
  }
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
  for (int a = 1; a < nc; a++) {
    int v = c[a], b = a - 1;
    while (b >= 0 && c[b] > v) { c[b + 1] = c[b]; b--; }
    c[b + 1] = v;
  }
  double zz = -1.0e20, cosa = 0.0;
  int fac = -1;
  for (int q = 0; q < nc; q++) {
    if (q > 0 && c[q] == c[q - 1]) continue;
    FacetProj P;
    project_facet(V, G, c[q], 0, P);
    if (!P.front) continue;
    double s, t, z;
    if (!pixel_hit(P, V.kmpp, i, j, s, t, z)) continue;
    if (!((z > zz) || (fac < 0))) continue;
    zz = z;
    double nz, cb;
    shade64(V, G, c[q], 0, V.smooth, s, t, nz, cb);
    if (nz > 0.0) cosa = nz;
    fac = c[q];
  }
  pd.stale_idx[V.pos_off + S.pix] = -1;
  if (!(cosa > 0.0)) return;
  RadarPix R;
  const double *pl = params + V.pb + V.pb_law;
  radar_pixel(V, pl, i, j, zz, (float)cosa, R);
  radar_scatter_global(V, R, fit + frames[V.frame_id].off, overflow + (long long)vi * 2000);
  // limits: a replayed pixel is lit like any other
  const bool dd = V.type == VT_DELAY;
  if (dd) {
    if (V.D0 + R.id0 < V.idellim[0]) atomicMin(&V.idellim[0], V.D0 + R.id0);
    if (V.D0 + R.id1 > V.idellim[1]) atomicMax(&V.idellim[1], V.D0 + R.id1);
    atomic_min_d(&V.dellim[0], (double)V.D0 + (double)R.delrel);
    atomic_min_d(&V.dellim[1], -((double)V.D0 + (double)R.delrel));
  }
  if (V.P0 + R.ip0 < V.idoplim[0]) atomicMin(&V.idoplim[0], V.P0 + R.ip0);
  if (V.P0 + R.ip1 > V.idoplim[1]) atomicMax(&V.idoplim[1], V.P0 + R.ip1);
  atomic_min_d(&V.doplim[0], (double)V.P0 + (double)R.doprel);
  atomic_min_d(&V.doplim[1], -((double)V.P0 + (double)R.doprel));
  if (!dd && !R.inb) atomicOr(&V.any_overflow, 1);
}

// ---------------------------------------------------------------------------
// FP64 replay of a pixel whose FP32 owner faces away, or that lies within the
// FP32 rounding band of a facet edge: the candidates claim it in facet order
// with the CPU rule `z > zz || fac < 0`, in the CPU's own FP64 arithmetic, so
// its owner, depth and (inherited) cos(scattering) are the CPU's
// (posvis.c:142-275); then the pixel's power is scattered directly to HBM.
// Grid-stride over the records: the grid is sized for the cap, not the count.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_replay32(ViewRec *views, const GeomSlot *slots, PosDev pd, double *fit,
                                                  const FrameRec *frames, double *overflow, const double *params) {
  int ns = min(*pd.nstale, pd.stale_cap);
  for (int si = blockIdx.x * blockDim.x + threadIdx.x; si < ns; si += gridDim.x * blockDim.x)
    replay32_one(views, slots, pd, fit, frames, overflow, params, si);
}

}  // namespace shp
