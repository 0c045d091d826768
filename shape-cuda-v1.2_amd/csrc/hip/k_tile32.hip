// k_tile32.hip -- mixed method, binned raster: no z-buffer key reaches HBM.
//
//   k_block_bounds  per slot: bounding sphere and normal cone of every block of
//                   BF consecutive facets and a sphere per facet (FP32, from
//                   the realized model), and facet-major copies of each
//                   facet's vertices and vertex normals
//   k_bin32         per view: each block's conservative pixel rectangle, entered
//                   in the lists of the BT x BT tiles it touches
//   k_tile32        per tile: rasterize the listed blocks into an LDS z-buffer
//                   (the packed keys of the global path), shade the owners
//                   straight from it into an FP64 LDS delay-Doppler window, and
//                   hand the tile's edge and faces-away pixels, with the facets
//                   covering them (an LDS list of covering events), to the
//                   FP64 replay (k_replay32)
//
// Facets wholly or partly outside the POS window (posbnd, posvis.c:172-196)
// are handled by k_outside32, over the blocks k_bin32 lists as reaching
// outside the window.  A tile whose lists overflow falls back to testing
// every block, or to the global replay records of k_collect32 /
// k_replay32; results do not depend on which route a pixel takes.
#include "mixed_dev.h"

namespace shp {

// ---------------------------------------------------------------------------
// block bounds: one wave per block, lane = facet
// ---------------------------------------------------------------------------
__device__ inline float wsumf(float v) { for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64); return v; }
// minimum / maximum over each group of 16 lanes
__device__ inline float seg16_minf(float v) { for (int o = 8; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64)); return v; }
__device__ inline float seg16_maxf(float v) { for (int o = 8; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64)); return v; }

__global__ __launch_bounds__(256) void k_block_bounds(GeomSlot *slots, int nslots) {
  const GeomSlot &G = slots[blockIdx.y];
  const int nb = (G.nf + BF - 1) / BF;
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= nb) return;  // whole wave
  const int f = b * BF + (threadIdx.x & 63);
  const bool ok = f < G.nf;
  float4 va = {0, 0, 0, 0}, vb = va, vc = va, fn = va;
  if (ok) {
    const int *fv = G.fv + 3 * f;
    va = G.x32[fv[0]]; vb = G.x32[fv[1]]; vc = G.x32[fv[2]];
    fn = G.fn32[f];
    // facet-major copies for the tile kernel: one gather per facet, not two
    // (.w: the facet's longest edge in km and the larger vertex-normal
    // difference along its barycentric axes, b - a and c - b; shade_ill)
    const float4 na = G.vn32[fv[0]], nb = G.vn32[fv[1]], nc = G.vn32[fv[2]];
    const float e0 = sqrtf((vb.x - va.x) * (vb.x - va.x) + (vb.y - va.y) * (vb.y - va.y) + (vb.z - va.z) * (vb.z - va.z));
    const float e1 = sqrtf((vc.x - vb.x) * (vc.x - vb.x) + (vc.y - vb.y) * (vc.y - vb.y) + (vc.z - vb.z) * (vc.z - vb.z));
    const float e2 = sqrtf((va.x - vc.x) * (va.x - vc.x) + (va.y - vc.y) * (va.y - vc.y) + (va.z - vc.z) * (va.z - vc.z));
    const float d0 = sqrtf((nb.x - na.x) * (nb.x - na.x) + (nb.y - na.y) * (nb.y - na.y) + (nb.z - na.z) * (nb.z - na.z));
    const float d1 = sqrtf((nc.x - nb.x) * (nc.x - nb.x) + (nc.y - nb.y) * (nc.y - nb.y) + (nc.z - nb.z) * (nc.z - nb.z));
    G.fvx[3 * f] = make_float4(va.x, va.y, va.z, fmaxf(e0, fmaxf(e1, e2)));
    G.fvx[3 * f + 1] = vb; G.fvx[3 * f + 2] = vc;
    G.fvn[3 * f] = make_float4(na.x, na.y, na.z, fmaxf(d0, d1));
    G.fvn[3 * f + 1] = nb; G.fvn[3 * f + 2] = nc;
  }
  float lo[3], hi[3];
  {
    float a[3] = {va.x, va.y, va.z}, bb[3] = {vb.x, vb.y, vb.z}, c[3] = {vc.x, vc.y, vc.z};
    for (int k = 0; k < 3; k++) {
      lo[k] = ok ? fminf(a[k], fminf(bb[k], c[k])) : 1.0e30f;
      hi[k] = ok ? fmaxf(a[k], fmaxf(bb[k], c[k])) : -1.0e30f;
      lo[k] = wminf(lo[k]);
      hi[k] = wmaxf(hi[k]);
    }
  }
  float cx = 0.5f * (lo[0] + hi[0]), cy = 0.5f * (lo[1] + hi[1]), cz = 0.5f * (lo[2] + hi[2]);
  float r2 = 0.0f;
  if (ok) {
    float d0x = va.x - cx, d0y = va.y - cy, d0z = va.z - cz;
    float d1x = vb.x - cx, d1y = vb.y - cy, d1z = vb.z - cz;
    float d2x = vc.x - cx, d2y = vc.y - cy, d2z = vc.z - cz;
    r2 = fmaxf(d0x * d0x + d0y * d0y + d0z * d0z,
               fmaxf(d1x * d1x + d1y * d1y + d1z * d1z, d2x * d2x + d2y * d2y + d2z * d2z));
  }
  if (ok) {  // the facet's own sphere: centroid and farthest corner
    float fx = (va.x + vb.x + vc.x) * (1.0f / 3.0f), fy = (va.y + vb.y + vc.y) * (1.0f / 3.0f),
          fz = (va.z + vb.z + vc.z) * (1.0f / 3.0f);
    float e0 = (va.x - fx) * (va.x - fx) + (va.y - fy) * (va.y - fy) + (va.z - fz) * (va.z - fz);
    float e1 = (vb.x - fx) * (vb.x - fx) + (vb.y - fy) * (vb.y - fy) + (vb.z - fz) * (vb.z - fz);
    float e2 = (vc.x - fx) * (vc.x - fx) + (vc.y - fy) * (vc.y - fy) + (vc.z - fz) * (vc.z - fz);
    G.fb[f] = make_float4(fx, fy, fz, sqrtf(fmaxf(e0, fmaxf(e1, e2))) * 1.00001f + 1.0e-9f);
  }
  {
    // the quarter's sphere (16 consecutive facets: lanes 16q .. 16q + 15)
    float qlo[3], qhi[3];
    float a[3] = {va.x, va.y, va.z}, bb[3] = {vb.x, vb.y, vb.z}, c[3] = {vc.x, vc.y, vc.z};
    for (int k = 0; k < 3; k++) {
      qlo[k] = seg16_minf(ok ? fminf(a[k], fminf(bb[k], c[k])) : 1.0e30f);
      qhi[k] = seg16_maxf(ok ? fmaxf(a[k], fmaxf(bb[k], c[k])) : -1.0e30f);
    }
    float qx = 0.5f * (qlo[0] + qhi[0]), qy = 0.5f * (qlo[1] + qhi[1]), qz = 0.5f * (qlo[2] + qhi[2]);
    float q2 = 0.0f;
    if (ok) {
      float d0x = va.x - qx, d0y = va.y - qy, d0z = va.z - qz;
      float d1x = vb.x - qx, d1y = vb.y - qy, d1z = vb.z - qz;
      float d2x = vc.x - qx, d2y = vc.y - qy, d2z = vc.z - qz;
      q2 = fmaxf(d0x * d0x + d0y * d0y + d0z * d0z,
                 fmaxf(d1x * d1x + d1y * d1y + d1z * d1z, d2x * d2x + d2y * d2y + d2z * d2z));
    }
    q2 = seg16_maxf(q2);
    if ((threadIdx.x & 15) == 0)
      G.sblk[4 * b + ((threadIdx.x & 63) >> 4)] = make_float4(qx, qy, qz, sqrtf(q2) * 1.00001f + 1.0e-9f);
  }
  r2 = wmaxf(r2);
  float r = sqrtf(r2) * 1.00001f + 1.0e-9f;
  float ax = wsumf(fn.x), ay = wsumf(fn.y), az = wsumf(fn.z);
  float al = sqrtf(ax * ax + ay * ay + az * az);
  float cmin = -1.0f;  // no cone: the block may face every way
  if (al > 1.0e-6f) {
    ax /= al; ay /= al; az /= al;
    float c = ok ? ax * fn.x + ay * fn.y + az * fn.z : 1.0f;
    cmin = wminf(c);
    cmin = fmaxf(cmin - 1.0e-5f, -1.0f);  // rounding of the FP32 normals
  }
  if ((threadIdx.x & 63) == 0) {
    G.blk[2 * b] = make_float4(cx, cy, cz, r);
    G.blk[2 * b + 1] = make_float4(ax, ay, az, cmin);
  }
}

// ---------------------------------------------------------------------------
// binning: thread per block.  A workgroup's blocks are neighbours on the model,
// so entries are counted in LDS over the workgroup's tile rectangle and each
// touched tile's list is reserved with one global atomic.
// ---------------------------------------------------------------------------
constexpr int BIN_LDS = 1024;  // tile counters of a workgroup rectangle held in LDS

// a, c: workgroup minima; b, d: workgroup maxima
__device__ inline void blk_minmax4(int &a, int &b, int &c, int &d, int (*sh)[4]) {
  int nb = -b, nd = -d;
  blk_min4(a, nb, c, nd, sh);
  b = -nb; d = -nd;
}

static_assert(BF == 64, "a block is one wave of 64 facets, in four quarters of 16 lanes");
// Which quarters of a block touch tile (x, y): bit q for quarter q, from the
// quarters' conservative rectangles qr[q] = {i0, i1, j0, j1}
__device__ inline unsigned quarter_mask(const int (&qr)[4][4], int x, int y, int n) {
  const int ti0 = x * BT - n, tj0 = y * BT - n;
  unsigned mk = 0;
#pragma unroll
  for (int q = 0; q < 4; q++)
    if (!(qr[q][0] > ti0 + BT - 1 || qr[q][1] < ti0 || qr[q][2] > tj0 + BT - 1 || qr[q][3] < tj0)) mk |= 1u << q;
  return mk;
}

__global__ __launch_bounds__(256) void k_bin32(const ViewRec *__restrict__ views, const GeomSlot *slots,
                                               int *bin_cnt, unsigned int *bin_list, int2 *out_list,
                                               int *out_count) {
  const ViewRec &V = views[blockIdx.y];
  if (!V.binned) return;
  const GeomSlot &G = slots[V.slot];
  const int nb = (G.nf + BF - 1) / BF;
  if ((int)blockIdx.x * 256 >= nb) return;
  const int b = blockIdx.x * 256 + threadIdx.x;
  const int n = V.n, ntx = V.bin_ntx;
  int tx0 = 0x7fffffff, tx1 = -1, ty0 = 0x7fffffff, ty1 = -1;
  int i0, i1, j0, j1;
  int qr[4][4];
  if (b < nb && block_rect(V, G, b, i0, i1, j0, j1)) {
    // a block reaching outside the POS window: its facets' posbnd and
    // outside extent are taken by k_outside32 (rare; one slot per block
    // of every view exists)
    if (i0 < -n || i1 > n || j0 < -n || j1 > n) out_list[atomicAdd(out_count, 1)] = make_int2(blockIdx.y, b);
    i0 = max(i0, -n); i1 = min(i1, n); j0 = max(j0, -n); j1 = min(j1, n);
    if (i0 <= i1 && j0 <= j1) {
      tx0 = (i0 + n) / BT; tx1 = (i1 + n) / BT;
      ty0 = (j0 + n) / BT; ty1 = (j1 + n) / BT;
#pragma unroll
      for (int q = 0; q < 4; q++) sphere_rect(V, G.sblk[4 * b + q], qr[q][0], qr[q][1], qr[q][2], qr[q][3]);
    }
  }
  __shared__ int shr[4][4];
  __shared__ int lc[BIN_LDS];
  int rx0 = tx0, rx1 = tx1, ry0 = ty0, ry1 = ty1;
  blk_minmax4(rx0, rx1, ry0, ry1, shr);
  if (rx1 < rx0) return;  // nothing of this workgroup lands in the window
  const int rw = ry1 - ry0 + 1, rn = (rx1 - rx0 + 1) * rw;
  const long long base = V.bin_off;
  // an entry is the block id and the mask of its quarters that touch the
  // tile; tiles that no quarter touches get no entry
  if (rn > BIN_LDS) {
    // a scattered workgroup (unordered facets): one global atomic per entry
    for (int x = tx0; x <= tx1; x++)
      for (int y = ty0; y <= ty1; y++) {
        const unsigned mk = quarter_mask(qr, x, y, n);
        if (!mk) continue;
        long long t = base + (long long)x * ntx + y;
        int q = atomicAdd(&bin_cnt[t], 1);
        if (q < BIN_CAP) bin_list[t * BIN_CAP + q] = ((unsigned)b << 4) | mk;
      }
    return;
  }
  for (int q = threadIdx.x; q < rn; q += blockDim.x) lc[q] = 0;
  __syncthreads();
  for (int x = tx0; x <= tx1; x++)
    for (int y = ty0; y <= ty1; y++)
      if (quarter_mask(qr, x, y, n)) atomicAdd(&lc[(x - rx0) * rw + (y - ry0)], 1);
  __syncthreads();
  for (int q = threadIdx.x; q < rn; q += blockDim.x) {
    int c = lc[q];
    if (c) lc[q] = atomicAdd(&bin_cnt[base + (long long)(rx0 + q / rw) * ntx + (ry0 + q % rw)], c);
  }
  __syncthreads();
  for (int x = tx0; x <= tx1; x++)
    for (int y = ty0; y <= ty1; y++) {
      const unsigned mk = quarter_mask(qr, x, y, n);
      if (!mk) continue;
      long long t = base + (long long)x * ntx + y;
      int q = atomicAdd(&lc[(x - rx0) * rw + (y - ry0)], 1);
      if (q < BIN_CAP) bin_list[t * BIN_CAP + q] = ((unsigned)b << 4) | mk;
    }
}

// ---------------------------------------------------------------------------
// the tile kernel
// ---------------------------------------------------------------------------
// Append value e to list[] (capacity cap; the count may run past it) for the
// lanes of the calling wave that need it: one LDS atomic per wave.  Every
// active lane of the wave must call it.
__device__ inline void lds_append(bool need, unsigned int e, void *list, int *cnt, int cap) {
  unsigned long long mask = __ballot(need);
  if (!mask) return;
  int lane = threadIdx.x & 63, leader = __ffsll((long long)mask) - 1, base = 0;
  if (lane == leader) base = atomicAdd(cnt, __popcll(mask));
  base = __builtin_amdgcn_readlane(base, leader);
  if (need) {
    int q = base + __popcll(mask & ((1ULL << lane) - 1ULL));
    if (q < cap) static_cast<unsigned int *>(list)[q] = e;
  }
}

// Record pixel (i, j) (local index lp) for the FP64 replay: in the tile's LDS
// list while its events are complete and the list has room, else in the
// global replay records (k_collect32 collects its candidates).
__device__ inline void tile_stale(const ViewRec &V, int vi, const PosDev &pd, bool ev_ok, int lp, int i, int j,
                                  unsigned char *flag, short *spix, int *nts) {
  if (ev_ok) {
    int s = atomicAdd(nts, 1);
    if (s < TS_CAP) {
      spix[s] = (short)lp;
      flag[lp] = (unsigned char)(0x80 | (s << 1) | (flag[lp] & 1));
      return;
    }
  }
  int si = atomicAdd(pd.nstale, 1);
  atomicAdd(pd.ncollect, 1);
  if (si >= pd.stale_cap) {
    atomicOr(pd.error, 1);
    return;
  }
  stale_init(V, pd, si, vi, (i + V.n) * (2 * V.n + 1) + (j + V.n), i, j);
}

// Pixel (a, c) of the tile taken by this thread in its m-th turn: each wave
// takes an 8 x 8 grid of pixels at stride 4 across the whole tile, so its
// lanes spread over the tile's range of delay rows and Doppler bins and
// their LDS window atomics collide far less than along one pixel row.
static_assert(BT == 32 && BT_PIX == 4 * 256, "tile_pixel assumes 32x32 tiles, 4 pixels per thread");
__device__ inline void tile_pixel(int m, int &a, int &c) {
  const int g = (threadIdx.x >> 6) * (BT_PIX / 256) + m, l = threadIdx.x & 63;
  a = (l >> 3) * 4 + (g >> 2);
  c = (l & 7) * 4 + (g & 3);
}

// One tile item of view vi: item b of the view's ng items (ng a multiple of
// NXCD; the workgroup runs on XCD b % NXCD).  vrd and views are the same
// view records: every read goes through the read-only, non-aliased vrd (the
// fields it reads stay fixed during the launch), so the compiler keeps them
// in scalar loads instead of reloading them per pixel after the kernel's
// global stores; the limits that workgroups update are read and written
// through views only.  Every return is taken by the whole workgroup, except
// the last, after which no thread reaches a barrier of this item.
__device__ __forceinline__ void tile32_item(int vi, int b, int ng, const ViewRec *__restrict__ vrd, ViewRec *views,
                                            const GeomSlot *__restrict__ slots, const PosDev &pd, int *bin_cnt,
                                            const unsigned int *__restrict__ bin_list, double *fit,
                                            const FrameRec *__restrict__ frames, double *overflow,
                                            const double *__restrict__ params) {
  const ViewRec &V = vrd[vi];
  ViewRec &Vw = views[vi];
  if (!V.binned) return;
  const int n = V.n, ntx = V.bin_ntx;
  // each XCD takes a contiguous eighth of the view's tiles (neighbouring
  // tiles share blocks, whose vertices then stay in one L2); which eighth
  // rotates with the view, so that the edge bands, mostly empty sky, are
  // spread over all XCDs instead of idling two of them
  const int per = ng / NXCD;
  const int bx = (int)((b % NXCD + vi) % NXCD) * per + (int)(b / NXCD);
  if (bx >= ntx * ntx) return;
  const long long bin = V.bin_off + bx;
  const int cnt = bin_cnt[bin];
  if (cnt == 0) return;
  const int i0 = (bx / ntx) * BT - n, j0 = (bx % ntx) * BT - n;
  const GeomSlot &G = slots[V.slot];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;

  __shared__ unsigned long long key[BT_PIX];
  __shared__ unsigned char flag[BT_PIX];  // bit 0: marginal; bit 7: LDS replay slot in bits 1-6
  __shared__ unsigned int ev[EV_CAP + 1];  // (pixel << 22 | facet); the last slot takes overflowing writes
  __shared__ unsigned int blist[BIN_CAP];
  __shared__ double wbuf[WIN_T];          // delay-Doppler window, then replay candidates
  static_assert(TS_CAP * TS_CAND * sizeof(int) <= sizeof(wbuf), "candidates must fit the window buffer");
  static_assert(TS_CAP <= 64, "replay slot must fit flag bits 1-6");
  __shared__ short spix[TS_CAP];
  __shared__ int scnt[TS_CAP];
  __shared__ int nev, nts, sbase, nfl;
  __shared__ int shw[4][4];
  __shared__ int shr[4][8];

  const bool full = cnt > BIN_CAP;
  if (!full)
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) blist[q] = bin_list[bin * BIN_CAP + q];
  for (int q = threadIdx.x; q < BT_PIX; q += blockDim.x) { key[q] = 0ULL; flag[q] = 0; }
  if (threadIdx.x == 0) { nev = 0; nts = 0; }
  __syncthreads();

  // ---- raster, in rounds: (1) waves test the facets of up to FL_CAP / BF
  // listed blocks against the tile and compact the survivors into an LDS
  // list (the window buffer, free until phase A); (2) every thread
  // rasterizes listed facets, so lanes stay busy in the pixel loops ----
  int xl0 = n, xl1 = -n, yl0 = n, yl1 = -n;
  {
    constexpr int FL_CAP = WIN_T * 2;  // ints in the window buffer
    int *flist = reinterpret_cast<int *>(wbuf);
    const int nb = (G.nf + BF - 1) / BF;
    const int nbl = full ? nb : cnt;
    const float *m = V.oa32;
    const float sc = V.inv_kmpp32;
    for (int kb = 0; kb < nbl; kb += FL_CAP / BF) {
      const int kend = min(nbl, kb + FL_CAP / BF);
      if (threadIdx.x == 0) nfl = 0;
      __syncthreads();
      for (int k = kb + wv; k < kend; k += 4) {
        int b = k;
        unsigned mk = 15u;  // quarters of the block touching the tile
        if (full) {  // lists overflowed: test every block (wave-uniform)
          int r0, r1, c0, c1;
          if (!block_rect(V, G, b, r0, r1, c0, c1) || r0 > i0 + BT - 1 || r1 < i0 || c0 > j0 + BT - 1 || c1 < j0)
            continue;
        } else {
          b = (int)(blist[k] >> 4);
          mk = blist[k] & 15u;
        }
        const int f = b * BF + lane;
        bool keep = false;
        if (f < G.nf && ((mk >> (lane >> 4)) & 1u)) {
          // the facet's sphere against the tile, before any dependent gather
          // (its bounding box, posvis.c:172-180, lies within the projected
          // sphere), and its facing: nearly edge-on facets are replay
          // candidates only, the FP64 replay tests facing itself
          // (posvis.c:152).  Both loads are coalesced and issued together.
          const float4 fb = gload4(G.fb + f), fn = gload4(G.fn32 + f);
          const float px = (m[0] * fb.x + m[1] * fb.y + m[2] * fb.z) * sc;
          const float py = (m[3] * fb.x + m[4] * fb.y + m[5] * fb.z) * sc;
          const float R = fb.w * sc * 1.0001f + 0.01f + 1.0e-6f * (fabsf(px) + fabsf(py));
          keep = !(floorf(px + R + 0.5f) < (float)i0 || floorf(px - R + 0.5f) > (float)(i0 + BT - 1) ||
                   floorf(py + R + 0.5f) < (float)j0 || floorf(py - R + 0.5f) > (float)(j0 + BT - 1)) &&
                 facing32(m, fn) > -FRONT_TOL;  // as project32nf
        }
        lds_append(keep, f, flist, &nfl, FL_CAP);
      }
      __syncthreads();
      const int nl = min(nfl, FL_CAP);
      // every lane of a wave takes each trip (q >= nl: idle lane), so the
      // wave-wide event reservation below sees all 64 lanes
      for (int q0 = threadIdx.x & ~63; q0 < nl; q0 += blockDim.x) {
        const int q = q0 + lane;
        Proj32 P;
        Edge32 E;
        float nz = 0.0f, amb = AMB_TOL;
        int f = 0, ti1 = 0, ti2 = -1, tj1 = 0, tj2 = -1, wi1 = 0, wi2 = 0, wj1 = 0, wj2 = 0;
        if (q < nl) {
          f = flist[q];
          if (project32nf(V, G, f, V.oa32, -FRONT_TOL, P, nz)) {
            float imin_d = floorf(fminf(P.x0, fminf(P.x1, P.x2)) + 0.5f);
            float imax_d = floorf(fmaxf(P.x0, fmaxf(P.x1, P.x2)) + 0.5f);
            float jmin_d = floorf(fminf(P.y0, fminf(P.y1, P.y2)) + 0.5f);
            float jmax_d = floorf(fmaxf(P.y0, fmaxf(P.y1, P.y2)) + 0.5f);
            int imin = (int)fmaxf(imin_d, -1.0e9f), imax = (int)fminf(imax_d, 1.0e9f);
            int jmin = (int)fmaxf(jmin_d, -1.0e9f), jmax = (int)fminf(jmax_d, 1.0e9f);
            // window-clipped box (the POS rectangle grows by it), then tile-clipped
            wi1 = max(imin, -n); wi2 = min(imax, n); wj1 = max(jmin, -n); wj2 = min(jmax, n);
            ti1 = max(wi1, i0); ti2 = min(wi2, i0 + BT - 1); tj1 = max(wj1, j0); tj2 = min(wj2, j0 + BT - 1);
            edge32_init(P, wi1, wj1, E);  // anchored as in the global raster
            amb = facet_amb(P);
          }
        }
        const bool front = nz > 0.0f;
        const int nc = (ti1 <= ti2 && tj1 <= tj2) ? (ti2 - ti1 + 1) * (tj2 - tj1 + 1) : 0;
        // covered cells (cell c of the box: bit c).  A box over 64 cells
        // (a facet wider than 8 pixels, none at config 4) leaves the tile's
        // event list incomplete: its replayed pixels then take the global
        // candidate collection (k_collect32)
        if (nc > 64) atomicMax(&nev, EV_CAP + 1);  // (other waves add to nev concurrently)
        // covered cells, shifted in: cell c of the box (row-major) is bit
        // nc - 1 - c, for the box's last 64 cells
        unsigned long long covm = 0ULL;
        bool hit = false;
        const int wj = tj2 - tj1 + 1;
        const int lp0 = (ti1 - i0) * BT + (tj1 - j0);
        {
          const float dj0 = (float)(tj1 - wj1), dj2 = (float)(tj2 - wj1);
          float di = (float)(ti1 - wi1), dj = dj0;
          // a row advances lp by BT, so the box ends at lp0 + rows * BT
          const int lp_end = nc > 0 ? lp0 + (ti2 - ti1 + 1) * BT : lp0;
          for (int lp = lp0; lp < lp_end;) {
            float s, t;
            const float mg = edge32(E, di, dj, s, t);
            const bool cov = mg >= -amb;
            if (front && mg >= amb) {
              atomicMax(&key[lp], (unsigned long long)pack_key(depth32(P, s, t), f));
              hit = true;
            } else if (front && cov) {
              flag[lp] = 1;  // marginal: claimed by no facet, replayed in FP64
              hit = true;
            }
            covm = covm + covm + (cov ? 1ULL : 0ULL);
            const bool wrap = dj >= dj2;
            dj = wrap ? dj0 : dj + 1.0f;
            di = wrap ? di + 1.0f : di;
            lp += wrap ? BT - wj + 1 : 1;
          }
        }
        if (hit) {
          xl0 = min(xl0, wi1); xl1 = max(xl1, wi2);
          yl0 = min(yl0, wj1); yl1 = max(yl1, wj2);
        }
        // one reservation for the whole wave's events: exclusive prefix sum
        // of the lanes' counts, one LDS atomic by the last lane
        const int ne = __popcll(covm);
        const int incl = wave_incl_scan(ne);
        int base = 0;
        if (lane == 63) base = atomicAdd(&nev, incl);
        base = __builtin_amdgcn_readlane(base, 63) + incl - ne;
        // row of cell c: (c + 0.5) / wj lies at least 0.5 / wj from an
        // integer, far beyond the FP32 error, so the float quotient is exact
        // for the boxes whose events count (nc <= 64)
        const float inv_wj = __builtin_amdgcn_rcpf((float)wj);
        while (covm) {
          const int c = nc - 1 - __builtin_ctzll(covm);
          covm &= covm - 1ULL;
          const int r = (int)(((float)c + 0.5f) * inv_wj);
          const int lp = lp0 + c + __mul24(r, BT - wj);
          ev[min(base, EV_CAP)] = ((unsigned int)lp << EV_FACET_BITS) | (unsigned int)f;  // slot EV_CAP: spill
          base++;
        }
      }
      __syncthreads();
    }
  }

  // ---- A: keys, marginal pixels and the footprint window ----
  const bool dd = V.type == VT_DELAY;
  const int width = V.sinc2width;
  const bool ev_ok = nev <= EV_CAP;
  constexpr int TP = BT_PIX / 256;
  int lo_d = 0x7fffffff, hi_d = -0x7fffffff, lo_p = 0x7fffffff, hi_p = -0x7fffffff, ncl = 0;
  unsigned long long kk[TP];
#pragma unroll
  for (int m = 0; m < TP; m++) {
    int a, c;
    tile_pixel(m, a, c);
    const int i = i0 + a, j = j0 + c, lp = a * BT + c;
    unsigned long long k = 0ULL;
    if (i <= n && j <= n) {
      k = key[lp];
      if (k) ncl++;
      if (flag[lp] & 1) {
        k = 0ULL;  // not shaded: replayed in FP64
        tile_stale(V, vi, pd, ev_ok, lp, i, j, flag, spix, &nts);
      }
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
    kk[m] = k;
  }
  {
    int nhp = -hi_p, nhd = -hi_d;
    blk_min4(lo_p, nhp, lo_d, nhd, shw);  // (its barriers also end every read of the keys)
    hi_p = -nhp; hi_d = -nhd;
    if (!dd) lo_d = hi_d = 0;
  }
  // Each wave lists its pixels to shade contiguously in its own quarter of
  // the key array, so that the shading loop runs no lane idle for an
  // unclaimed pixel.  A facet below 2^22 leaves the top ten bits of ~facet
  // set: they carry the pixel's tile index meanwhile.
  static_assert(BT_PIX == 4 * 256 && EV_FACET_BITS == 22 && BT_PIX <= 1024, "a wave's list and packed pixel index");
  constexpr unsigned long long PIX_BITS = 0x3FFULL << EV_FACET_BITS;
  unsigned long long *shade = key + wv * 256;
  int nsh = 0;
#pragma unroll
  for (int m = 0; m < TP; m++) {
    const unsigned long long mask = __ballot(kk[m] != 0ULL);
    if (kk[m]) {
      int a, c;
      tile_pixel(m, a, c);
      shade[nsh + __popcll(mask & ((1ULL << lane) - 1ULL))] =
          (kk[m] & ~PIX_BITS) | ((unsigned long long)(a * BT + c) << EV_FACET_BITS);
    }
    nsh += __popcll(mask);
  }
  const bool any = lo_p != 0x7fffffff;
  const int lim_d0 = lo_d, lim_d1 = hi_d, lim_p0 = lo_p, lim_p1 = hi_p;
  if (any) {  // one bin of margin: phase B recomputes the same floors
    lo_p--; hi_p++;
    if (dd) { lo_d--; hi_d++; }
  }
  const int wd = hi_d - lo_d + 1, wp = hi_p - lo_p + 1;
  const bool use_lds = any && (long long)wd * wp <= WIN_T;
  double *win = wbuf;
  if (use_lds) {
    for (int q = threadIdx.x; q < wd * wp; q += blockDim.x) win[q] = 0.0;
    __syncthreads();
  }

  // ---- B: shade, footprint, accumulate ----
  const double *pl = params + V.pb + V.pb_law;
  double *img = fit + frames[V.frame_id].off;
  double *ov = overflow + (long long)vi * 2000;
  int anyo = 0;
  for (int q = lane; q < nsh; q += 64) {
    const unsigned long long e = shade[q];
    const int lp = (int)((e & PIX_BITS) >> EV_FACET_BITS), a = lp / BT, c = lp % BT;
    const unsigned long long k = e | PIX_BITS;
    const int i = i0 + a, j = j0 + c;
    const int f = key_facet(k);
    if (pd.fmap) pd.fmap[V.pos_off + (long long)(i + n) * (2 * n + 1) + (j + n)] = f;  // (inspection)
    Proj32 P;
    const float4 xa = G.fvx[3 * f];
    corners32p(V, xa, G.fvx[3 * f + 1], G.fvx[3 * f + 2], V.oa32, P);
    float s, t;
    bary32(P, i, j, s, t);  // the raster's coordinates, to FP32 rounding
    float nz, spread = 0.0f;
    if (V.smooth) {
      const float4 na = G.fvn[3 * f];
      nz = normal_los(V.oa32, na, G.fvn[3 * f + 1], G.fvn[3 * f + 2], s, t);
      spread = na.w;
    } else {
      nz = facing32(V.oa32, G.fn32[f]);
    }
    if (!(nz > V.limb32) || shade_ill(V, P, xa.w, spread)) {
      // cos(scattering) of an owner facing away is inherited from an earlier
      // claimant; an ill-conditioned FP32 cosine needs FP64 (and so does a
      // limb pixel under a cosine law with C < 1/2): FP64 replay
      tile_stale(V, vi, pd, ev_ok, a * BT + c, i, j, flag, spix, &nts);
      continue;
    }
    RadarPix R;
    radar_pixel(V, pl, i, j, (double)key_z(k), nz, R);  // (binned views: diffuse cosine law)
    if (!R.inb) anyo = 1;
    if (use_lds && R.inb) {
      radar_scatter_lds(V, R, win + (dd ? (R.id0 - lo_d) * wp : 0) + (R.ip0 - lo_p), wp);
    } else {
      radar_scatter_global(V, R, img, ov);
    }
  }

  // ---- C: flush the window ----
  __syncthreads();
  if (use_lds) {
    const int ndop = V.ndop;
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
  __syncthreads();

  // ---- D: the tile's recorded pixels join the global replay records, with
  // the candidates gathered from the tile's events; k_replay32 replays them
  // in FP64 (its own launch keeps this kernel's registers for the bulk) ----
  const int nt = min(nts, TS_CAP);
  if (nt > 0) {
    int (*cand)[TS_CAND] = reinterpret_cast<int (*)[TS_CAND]>(wbuf);
    for (int q = threadIdx.x; q < nt; q += blockDim.x) scnt[q] = 0;
    __syncthreads();
    const int ne = min(nev, EV_CAP);
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      const unsigned int x = ev[e];
      const unsigned char fl = flag[x >> EV_FACET_BITS];
      if (fl & 0x80) {
        const int s = (fl >> 1) & 63;
        const int q = atomicAdd(&scnt[s], 1);
        if (q < TS_CAND) cand[s][q] = (int)(x & ((1u << EV_FACET_BITS) - 1u));
      }
    }
    if (threadIdx.x == 0) {
      int b0 = atomicAdd(pd.nstale, nt);
      if (b0 + nt > pd.stale_cap) {
        atomicOr(pd.error, 1);
        b0 = -1;
      }
      sbase = b0;
    }
    __syncthreads();
    const int base = sbase;
    if (base >= 0) {
      const int w = 2 * n + 1;
      for (int s = wv; s < nt; s += 4) {
        StaleRec &S = pd.stale[base + s];
        const int nc = scnt[s];
        const bool own = nc <= TS_CAND;  // else k_collect32 collects them
        if (own)
          for (int q = lane; q < nc; q += 64) S.cand[q] = cand[s][q];
        if (lane == 0) {
          const int lp = spix[s], i = i0 + lp / BT, j = j0 + lp % BT;
          S.view = vi;
          S.pix = (i + n) * w + (j + n);
          S.oldf = -1; S.oldz = -1.0e20; S.oldcose = 0.0; S.oldcosi = 0.0;
          S.src = 0;
          S.ncand = own ? nc : 0;
          if (!own) {
            pd.stale_idx[V.pos_off + S.pix] = base + s;
            pd.stale_tiles[stale_tile(V, i, j)] = 1;
            atomicAdd(pd.ncollect, 1);
          }
        }
      }
    }
  }

  // ---- E: rectangle, limits and claimed keys, once per workgroup ----
  anyo = wmax(anyo);
  ncl = wsum(ncl);
  xl0 = wmin(xl0); xl1 = wmax(xl1); yl0 = wmin(yl0); yl1 = wmax(yl1);
  if (lane == 0) {
    shr[wv][0] = anyo; shr[wv][1] = ncl; shr[wv][2] = xl0; shr[wv][3] = xl1;
    shr[wv][4] = yl0; shr[wv][5] = yl1;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  bin_cnt[bin] = 0;  // ready for the next replay's binning
  for (int q = 1; q < 4; q++) {
    anyo |= shr[q][0]; ncl += shr[q][1];
    xl0 = min(xl0, shr[q][2]); xl1 = max(xl1, shr[q][3]);
    yl0 = min(yl0, shr[q][4]); yl1 = max(yl1, shr[q][5]);
  }
  if (ncl) atomicAdd(&Vw.nclaimed, ncl);
  if (xl0 < n && xl0 < Vw.xlim[0]) atomicMin(&Vw.xlim[0], xl0);
  if (xl1 > -n && xl1 > Vw.xlim[1]) atomicMax(&Vw.xlim[1], xl1);
  if (yl0 < n && yl0 < Vw.ylim[0]) atomicMin(&Vw.ylim[0], yl0);
  if (yl1 > -n && yl1 > Vw.ylim[1]) atomicMax(&Vw.ylim[1], yl1);
  if (!dd && anyo && !Vw.any_overflow) atomicOr(&Vw.any_overflow, 1);
  if (!any) return;
  if (dd) {
    if (V.D0 + lim_d0 < Vw.idellim[0]) atomicMin(&Vw.idellim[0], V.D0 + lim_d0);
    if (V.D0 + lim_d1 > Vw.idellim[1]) atomicMax(&Vw.idellim[1], V.D0 + lim_d1);
  }
  if (V.P0 + lim_p0 < Vw.idoplim[0]) atomicMin(&Vw.idoplim[0], V.P0 + lim_p0);
  if (V.P0 + lim_p1 > Vw.idoplim[1]) atomicMax(&Vw.idoplim[1], V.P0 + lim_p1);
}

// One workgroup per tile item: grid (ng, nviews).  Six workgroups per CU (LDS
// 26 KiB each): the register cap of six waves per SIMD hides more of the
// gathers' latency than a larger register budget would (measured).
__global__ __launch_bounds__(256, TILE_WG_PER_CU) void k_tile32(const ViewRec *__restrict__ vrd, ViewRec *views,
                                                             const GeomSlot *__restrict__ slots, PosDev pd,
                                                             int *bin_cnt, const unsigned int *__restrict__ bin_list,
                                                             double *fit, const FrameRec *__restrict__ frames,
                                                             double *overflow, const double *__restrict__ params) {
  tile32_item(blockIdx.y, blockIdx.x, gridDim.x, vrd, views, slots, pd, bin_cnt, bin_list, fit, frames, overflow,
              params);
}

// Persistent form: gridDim.x workgroups (a multiple of NXCD; the hardware
// deals workgroup g to XCD g % NXCD) take the nviews x ng tile items from one
// work counter per XCD, so a workgroup on XCD x takes only items b with
// b % NXCD == x and the tile-to-XCD mapping of the per-tile grid holds.  Each
// workgroup fetches its next item while it works on the current one.
// work[0..NXCD): the counters; work[NXCD..2 NXCD): workgroups done per XCD.
// The last workgroup of an XCD to finish resets both, so they start every
// launch at zero (they are zeroed once when allocated).
__global__ __launch_bounds__(256, TILE_WG_PER_CU) void k_tile32p(const ViewRec *__restrict__ vrd, ViewRec *views,
                                                              const GeomSlot *__restrict__ slots, PosDev pd,
                                                              int *bin_cnt, const unsigned int *__restrict__ bin_list,
                                                              double *fit, const FrameRec *__restrict__ frames,
                                                              double *overflow, const double *__restrict__ params,
                                                              int *work, int ng, int nviews) {
  const int x = blockIdx.x % NXCD;
  const long long nitems = (long long)(ng / NXCD) * nviews;  // items of this XCD
  __shared__ int s_next;
  if (threadIdx.x == 0) s_next = atomicAdd(&work[x], 1);
  __syncthreads();
  int k = s_next;
  while (k < nitems) {
    int nxt = 0;
    if (threadIdx.x == 0) nxt = atomicAdd(&work[x], 1);
    const long long it = (long long)k * NXCD + x;
    tile32_item((int)(it / ng), (int)(it % ng), ng, vrd, views, slots, pd, bin_cnt, bin_list, fit, frames, overflow,
                params);
    __syncthreads();  // every thread has left the item and read s_next
    if (threadIdx.x == 0) s_next = nxt;
    __syncthreads();
    k = s_next;
  }
  if (threadIdx.x == 0 && atomicAdd(&work[NXCD + x], 1) == (int)(gridDim.x / NXCD) - 1) {
    // every workgroup of this XCD has taken its last item index
    atomicExch(&work[x], 0);
    atomicExch(&work[NXCD + x], 0);
  }
}

}  // namespace shp
