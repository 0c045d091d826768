// mixed_dev.h -- FP32 device helpers of the mixed method (projection, barycentric
// test, owner shading, delay/Doppler responses, scattering laws).
#pragma once

#include "common.h"
#include "kernels.h"

namespace shp {

constexpr int TILE = 16;         // resolve tiles: 16x16 pixels per workgroup of 256 threads
constexpr int TPP = TILE * TILE / 256;  // pixels per thread
constexpr int WIN_MAX = 2048;    // LDS delay-Doppler window (FP64 bins, 16 KiB)
// The LDS scatter adds a footprint's delay rows DW at a time, unrolled (a
// row past the footprint is a masked lane op, not a skipped one).  One
// sample per baud puts a pixel's power into two delay rows (the triangle's
// half-width is one row), so DW = 2 wastes no adds there; more samples per
// baud take more groups.
constexpr int DW = 2;

// With the grid's x size nb a multiple of NXCD (pad_xcd), remap so that XCD k
// runs the k-th contiguous eighth of [0, nb): in the raster, each XCD's
// private L2 then keeps one eighth of the model's facets and vertices across
// all views instead of streaming the whole model for every view (measured:
// half the raster time on a 327,680-facet model).
__device__ inline int xcd_block(int b, int nb) { return (b % NXCD) * (nb / NXCD) + b / NXCD; }
inline unsigned pad_xcd(unsigned n) { return (n + NXCD - 1) / NXCD * NXCD; }

// Wave-wide reductions (every lane of the wave active), returned to every
// lane: DPP lane shifts within each 16-lane row leave the row's result in
// its lane 15 (a lane whose DPP source falls outside its row takes the
// identity); the four rows' results are read out and combined.  No LDS
// round trips.
template <int CTRL>
__device__ inline int dpp_i(int v, int id) { return __builtin_amdgcn_update_dpp(id, v, CTRL, 0xf, 0xf, false); }
// Five wait states between a write of EXEC (where diverged lanes rejoin)
// and a DPP op that reads v: without them the DPP op may still run under the
// old EXEC, leaving rejoining lanes out of the lane shifts.  The compiler
// inserts them after vector writes of EXEC but not after scalar ones.
__device__ inline int dpp_exec_guard(int v) {
  asm volatile("s_nop 4" : "+v"(v));
  return v;
}
template <typename Op>
__device__ inline int wave_reduce_i(int v, int id, Op op) {
  v = dpp_exec_guard(v);
  v = op(v, dpp_i<0x111>(v, id));  // row_shr:1
  v = op(v, dpp_i<0x112>(v, id));  // row_shr:2
  v = op(v, dpp_i<0x114>(v, id));  // row_shr:4
  v = op(v, dpp_i<0x118>(v, id));  // row_shr:8
  return op(op(__builtin_amdgcn_readlane(v, 15), __builtin_amdgcn_readlane(v, 31)),
            op(__builtin_amdgcn_readlane(v, 47), __builtin_amdgcn_readlane(v, 63)));
}
__device__ inline int wmin(int v) { return wave_reduce_i(v, 0x7fffffff, [](int a, int b) { return min(a, b); }); }
__device__ inline int wmax(int v) { return wave_reduce_i(v, -0x7fffffff - 1, [](int a, int b) { return max(a, b); }); }
__device__ inline int wsum(int v) { return wave_reduce_i(v, 0, [](int a, int b) { return a + b; }); }
__device__ inline float wminf(float v) {
  return __int_as_float(wave_reduce_i(__float_as_int(v), __float_as_int(HUGE_VALF),
                                      [](int a, int b) { return __float_as_int(fminf(__int_as_float(a), __int_as_float(b))); }));
}
__device__ inline float wmaxf(float v) {
  return __int_as_float(wave_reduce_i(__float_as_int(v), __float_as_int(-HUGE_VALF),
                                      [](int a, int b) { return __float_as_int(fmaxf(__int_as_float(a), __int_as_float(b))); }));
}

// Inclusive prefix sum over the 64 lanes of a wave (every lane active): DPP
// lane shifts within each 16-lane row, then the totals of the rows before a
// lane's own, read out of lanes 15, 31 and 47.
__device__ inline int wave_incl_scan(int v0) {
  v0 = dpp_exec_guard(v0);
  int v = v0 + __builtin_amdgcn_update_dpp(0, v0, 0x111, 0xf, 0xf, true);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v0, 0x112, 0xf, 0xf, true);          // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v0, 0x113, 0xf, 0xf, true);          // row_shr:3
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xe, true);           // row_shr:4, lanes 4-15 of a row
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xc, true);           // row_shr:8, lanes 8-15
  const int t0 = __builtin_amdgcn_readlane(v, 15), t1 = __builtin_amdgcn_readlane(v, 31),
            t2 = __builtin_amdgcn_readlane(v, 47);
  const int row = (threadIdx.x & 63) >> 4;
  return v + (row > 0 ? t0 : 0) + (row > 1 ? t1 : 0) + (row > 2 ? t2 : 0);
}

// block-wide minimum of four ints at once (one barrier pair)
__device__ inline void blk_min4(int &a, int &b, int &c, int &d, int (*sh)[4]) {
  a = wmin(a); b = wmin(b); c = wmin(c); d = wmin(d);
  int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sh[wv][0] = a; sh[wv][1] = b; sh[wv][2] = c; sh[wv][3] = d; }
  __syncthreads();
  a = sh[0][0]; b = sh[0][1]; c = sh[0][2]; d = sh[0][3];
  for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
    a = min(a, sh[q][0]); b = min(b, sh[q][1]); c = min(c, sh[q][2]); d = min(d, sh[q][3]);
  }
  __syncthreads();
}

// The FP32 geometry below (projection, barycentric test, depth, normals)
// spells out its fused multiply-adds and turns contraction off otherwise:
// left to the compiler, contraction differs between kernels and inlining
// sites, and a pixel at a claim threshold (AMB_TOL, facing) would then take
// different routes in the binned and global-z-buffer rasters.  Written out,
// every kernel computes the same bits.
__device__ inline void rot3f(const float m[9], float x, float y, float z, float &a, float &b, float &c) {
#pragma clang fp contract(off)
  a = fmaf(m[2], z, fmaf(m[1], y, m[0] * x));
  b = fmaf(m[5], z, fmaf(m[4], y, m[3] * x));
  c = fmaf(m[8], z, fmaf(m[7], y, m[6] * x));
}

struct Proj32 {
  float x0, y0, z0, x1, y1, z1, x2, y2, z2;  // vertices: x, y in pixel units, z in km
  float den;
};

// Loads of geometry through the pointers a GeomSlot holds (device
// allocations).  The compiler cannot see their address space and emits flat
// loads, which count against the LDS counter too: a wait for an LDS result
// then also waits for them.  The tile raster's facet loads use these global
// loads (measured faster); in its shading phase they cost register spills,
// so the other helpers keep plain loads.  HIP's float4 copies through a
// generic reference, so the load goes through the native vector type.
typedef float gvec4 __attribute__((ext_vector_type(4)));
__device__ inline float4 gload4(const float4 *p) {
  const gvec4 v = *(const __attribute__((address_space(1))) gvec4 *)p;
  return make_float4(v.x, v.y, v.z, v.w);
}

// project a facet's corners a, b, c (posvis.c:159-166) in pixel units
__device__ inline void corners32p(const ViewRec &V, float4 a, float4 b, float4 c, const float *m, Proj32 &P) {
#pragma clang fp contract(off)
  float s = V.inv_kmpp32;
  rot3f(m, a.x, a.y, a.z, P.x0, P.y0, P.z0);
  rot3f(m, b.x, b.y, b.z, P.x1, P.y1, P.z1);
  rot3f(m, c.x, c.y, c.z, P.x2, P.y2, P.z2);
  P.x0 *= s; P.y0 *= s; P.x1 *= s; P.y1 *= s; P.x2 *= s; P.y2 *= s;
  // hardware reciprocal (1 ulp): every raster and shader takes the same value
  P.den = __builtin_amdgcn_rcpf(fmaf(P.x1 - P.x0, P.y2 - P.y1, -((P.x2 - P.x1) * (P.y1 - P.y0))));
}
__device__ inline void corners32(const ViewRec &V, const GeomSlot &G, int f, const float *m, Proj32 &P) {
  const int *fv = G.fv + 3 * f;
  corners32p(V, G.x32[fv[0]], G.x32[fv[1]], G.x32[fv[2]], m, P);
}
// the same from the facet-major vertex copies (one gather level)
__device__ inline void corners32f(const ViewRec &V, const GeomSlot &G, int f, const float *m, Proj32 &P) {
  corners32p(V, G.fvx[3 * f], G.fvx[3 * f + 1], G.fvx[3 * f + 2], m, P);
}

constexpr float FRONT_TOL = 1.0e-4f;
// line-of-sight component of facet normal fn (the facing test of the raster)
__device__ inline float facing32(const float *m, float4 fn) {
#pragma clang fp contract(off)
  return fmaf(m[8], fn.z, fmaf(m[7], fn.y, m[6] * fn.x));
}
// as project32 from the facet-major vertex copies, also returning the facet
// normal's line-of-sight component
__device__ inline bool project32nf(const ViewRec &V, const GeomSlot &G, int f, const float *m, float nz_min,
                                   Proj32 &P, float &nz) {
#pragma clang fp contract(off)
  float4 fn = gload4(G.fn32 + f);
  nz = facing32(m, fn);
  if (!(nz > nz_min)) return false;
  corners32p(V, gload4(G.fvx + 3 * f), gload4(G.fvx + 3 * f + 1), gload4(G.fvx + 3 * f + 2), m, P);
  return true;
}
// as project32 from the facet-major copies (one coalesced gather level; the
// same values, so the same bits)
__device__ inline bool project32f(const ViewRec &V, const GeomSlot &G, int f, const float *m, float nz_min,
                                  Proj32 &P) {
#pragma clang fp contract(off)
  float4 fn = G.fn32[f];
  float nz = facing32(m, fn);
  if (!(nz > nz_min)) return false;  // back-facing (posvis.c:152)
  corners32f(V, G, f, m, P);
  return true;
}
__device__ inline bool project32(const ViewRec &V, const GeomSlot &G, int f, const float *m, float nz_min, Proj32 &P) {
#pragma clang fp contract(off)
  float4 fn = G.fn32[f];
  float nz = facing32(m, fn);
  if (!(nz > nz_min)) return false;  // back-facing (posvis.c:152)
  corners32(V, G, f, m, P);
  return true;
}

// Conservative pixel rectangle of a body-frame sphere cr = {center, radius}
// in view V: it contains the bounding box (posvis.c:172-180) of every facet
// whose corners lie in the sphere, widened for FP32 rounding.
__device__ inline void sphere_rect(const ViewRec &V, float4 cr, int &i0, int &i1, int &j0, int &j1) {
  const float *m = V.oa32;
  float cx, cy, cz;
  rot3f(m, cr.x, cr.y, cr.z, cx, cy, cz);
  float s = V.inv_kmpp32, px = cx * s, py = cy * s;
  float R = cr.w * s * 1.0001f + 0.01f + 1.0e-6f * (fabsf(px) + fabsf(py));
  i0 = (int)fminf(fmaxf(floorf(px - R + 0.5f), -1.0e9f), 1.0e9f);
  i1 = (int)fminf(fmaxf(floorf(px + R + 0.5f), -1.0e9f), 1.0e9f);
  j0 = (int)fminf(fmaxf(floorf(py - R + 0.5f), -1.0e9f), 1.0e9f);
  j1 = (int)fminf(fmaxf(floorf(py + R + 0.5f), -1.0e9f), 1.0e9f);
}

// Conservative pixel rectangle [i0, i1] x [j0, j1] of bin block b in view V:
// it contains the bounding box (posvis.c:172-180) of every facet of the
// block, whose corners lie within the block's radius of its center, widened
// for FP32 rounding.  Not clipped to the POS window.  False when every facet
// of the block faces away by more than FRONT_TOL (normal cone).
__device__ inline bool block_rect(const ViewRec &V, const GeomSlot &G, int b, int &i0, int &i1, int &j0, int &j1) {
  float4 cr = G.blk[2 * b], ax = G.blk[2 * b + 1];
  const float *m = V.oa32;
  if (ax.w > -1.0f) {
    // every normal is within acos(ax.w) of the axis, so its line-of-sight
    // component is at most cos(angle(axis, sight) - acos(ax.w))
    float cphi = m[6] * ax.x + m[7] * ax.y + m[8] * ax.z;
    float d = acosf(fminf(fmaxf(cphi, -1.0f), 1.0f)) - acosf(fminf(fmaxf(ax.w, -1.0f), 1.0f));
    if (d > 0.0f && cosf(d) < -2.0f * FRONT_TOL) return false;
  }
  sphere_rect(V, cr, i0, i1, j0, j1);
  return true;
}

// Barycentric coordinates of pixel (i, j) (posvis.c:217-228): the facet
// covers the pixel center when t >= 0, t <= s and s <= 1, i.e. when
// margin = min(t, s - t, 1 - s) >= 0.
//
// FP32 vertex positions carry ~1e-4 px of rounding (|x| up to 512 px), so a
// center within that distance of an edge may be covered by a different facet
// than in FP64, or by none.  Such a center is "marginal": no facet claims it
// in the FP32 z-buffer; it is recorded and resolved in FP64, in facet order,
// exactly as the CPU does.  The band is facet_amb(P) in barycentric units:
// AMB_TOL, >= 5x the FP32 error for well-shaped facets of >= 2 px, widened
// for a facet whose barycentric coordinates amplify that error (a sliver, or
// a facet seen nearly edge-on: its projected area is small).
constexpr float AMB_TOL = 2.0e-4f;
// The FP32 error of a barycentric coordinate is about the vertex error
// (a few ulps of the largest projected coordinate M) times |grad s| <=
// den x longest edge; the band takes 4x that, and never less than AMB_TOL.
__device__ inline float facet_amb(const Proj32 &P) {
  const float M = fmaxf(fmaxf(fmaxf(fabsf(P.x0), fabsf(P.y0)), fmaxf(fabsf(P.x1), fabsf(P.y1))),
                        fmaxf(fabsf(P.x2), fabsf(P.y2)));
  const float e01 = fabsf(P.x1 - P.x0) + fabsf(P.y1 - P.y0), e12 = fabsf(P.x2 - P.x1) + fabsf(P.y2 - P.y1),
              e20 = fabsf(P.x0 - P.x2) + fabsf(P.y0 - P.y2);
  const float dx = 3.0f * 5.9604645e-8f * (M + 1.0f);  // 3 ulps of the largest coordinate (px)
  return fmaxf(AMB_TOL, 4.0f * dx * fabsf(P.den) * fmaxf(e01, fmaxf(e12, e20)));
}
__device__ inline float bary32(const Proj32 &P, int i, int j, float &s, float &t) {
#pragma clang fp contract(off)
  float xi = (float)i, yj = (float)j;
  s = fmaf(xi - P.x0, P.y2 - P.y1, -((P.x2 - P.x1) * (yj - P.y0))) * P.den;
  t = fmaf(P.x1 - P.x0, yj - P.y0, -((xi - P.x0) * (P.y1 - P.y0))) * P.den;
  return fminf(t, fminf(s - t, 1.0f - s));
}
// The rasters' barycentric coordinates over a facet's bounding box. s and t
// are linear in the pixel indices, so each pixel takes two FMAs per
// coordinate from the values at an anchor pixel: (di, dj) = pixel - anchor,
// both small integers held exactly in float.  Every raster of a view anchors
// a facet at its window-clipped box corner, so the tile and global-z-buffer
// rasters and the candidate collection make identical decisions.  Rounding
// stays ~1e-7 barycentric, far inside AMB_TOL.
struct Edge32 {
  float s0, t0, dsi, dsj, dti, dtj;
};
__device__ inline void edge32_init(const Proj32 &P, int ia, int ja, Edge32 &E) {
#pragma clang fp contract(off)
  bary32(P, ia, ja, E.s0, E.t0);
  E.dsi = (P.y2 - P.y1) * P.den;
  E.dsj = -((P.x2 - P.x1) * P.den);
  E.dti = -((P.y1 - P.y0) * P.den);
  E.dtj = (P.x1 - P.x0) * P.den;
}
__device__ inline float edge32(const Edge32 &E, float di, float dj, float &s, float &t) {
#pragma clang fp contract(off)
  s = fmaf(di, E.dsi, fmaf(dj, E.dsj, E.s0));
  t = fmaf(di, E.dti, fmaf(dj, E.dtj, E.t0));
  return fminf(t, fminf(s - t, 1.0f - s));
}
__device__ inline float depth32(const Proj32 &P, float s, float t) {
#pragma clang fp contract(off)
  return fmaf(t, P.z2 - P.z1, fmaf(s, P.z1 - P.z0, P.z0));
}

// tile flag of pixel (i, j) in the view's 8x8-tile bitmap
__device__ inline long long stale_tile(const ViewRec &V, int i, int j) {
  int tw = (2 * V.n + STALE_TILE) / STALE_TILE;
  return V.tile_off + (long long)((i + V.n) / STALE_TILE) * tw + (j + V.n) / STALE_TILE;
}

// Reserve replay records for the lanes of the calling wave that need one:
// one atomic on the shared counter per wave (a per-lane atomic on that single
// address serializes).  Every active lane of the wave must call it.  Returns
// the lane's record index, or -1 when it needs none or the list is full.
__device__ inline int stale_slot(const PosDev &pd, bool need) {
  unsigned long long mask = __ballot(need);
  if (!mask) return -1;
  int lane = threadIdx.x & 63, leader = __ffsll((long long)mask) - 1, base = 0;
  if (lane == leader) {
    base = atomicAdd(pd.nstale, __popcll(mask));
    atomicAdd(pd.ncollect, __popcll(mask));  // candidates come from the global raster
  }
  base = __builtin_amdgcn_readlane(base, leader);
  if (!need) return -1;
  int si = base + __popcll(mask & ((1ULL << lane) - 1ULL));
  if (si >= pd.stale_cap) {
    atomicOr(pd.error, 1);
    return -1;
  }
  return si;
}

// As stale_slot for a lightcurve view (index vi): its records come from the
// view's own range of the list (ViewRec::lc_rec_off, lc_rec_cap), reserved on
// the view's own counter; ncollect is only set nonzero (k_collect32 skips an
// empty pass).  Every active lane of the wave must call it.
__device__ inline int stale_slot_lc(const ViewRec &V, const PosDev &pd, int vi, bool need) {
  unsigned long long mask = __ballot(need);
  if (!mask) return -1;
  int lane = threadIdx.x & 63, leader = __ffsll((long long)mask) - 1, base = 0;
  if (lane == leader) {
    base = atomicAdd(&pd.lc_count[vi], __popcll(mask));
    *pd.ncollect = 1;
  }
  base = __builtin_amdgcn_readlane(base, leader);
  if (!need) return -1;
  int off = base + __popcll(mask & ((1ULL << lane) - 1ULL));
  if (off >= V.lc_rec_cap) {
    atomicOr(pd.error, 1);
    return -1;
  }
  return V.lc_rec_off + off;
}

// an empty replay record: the pixel's claims start from blank sky and are
// replayed in facet order
// (src: 0 observer view, 1 a lightcurve point's sun view)
__device__ inline void stale_init(const ViewRec &V, const PosDev &pd, int si, int view, int pix, int i, int j,
                                  int src = 0) {
  StaleRec &S = pd.stale[si];
  S.view = view; S.pix = pix; S.oldf = -1; S.oldz = -1.0e20; S.oldcose = 0.0;
  S.oldcosi = 0.0; S.ncand = 0; S.src = src;
  pd.stale_idx[V.pos_off + pix] = si;
  pd.stale_tiles[stale_tile(V, i, j)] = 1;
}

// Record a marginal pixel for the FP64 replay, once: the first facet to find
// it swaps its stale index from -1.  Every active lane of the wave calls this
// (marginal or not) so that records are reserved once per wave.
__device__ inline void mark_marginal(const ViewRec &V, const PosDev &pd, bool marginal, int pix, int i, int j,
                                     int src = 0) {
  long long p = V.pos_off + pix;
  bool first = marginal && atomicCAS(&pd.stale_idx[p], -1, -2) == -1;
  int si = V.type == VT_LGHTCRV ? stale_slot_lc(V, pd, blockIdx.y, first) : stale_slot(pd, first);
  if (si >= 0) stale_init(V, pd, si, blockIdx.y, pix, i, j, src);
  else if (first) pd.stale_idx[p] = -1;  // list full: the error flag is set
}

// Whether a shaded pixel's FP32 cos(scattering) may be far off.  The
// smoothed normal interpolates the facet's vertex normals at the pixel's
// FP32 barycentric coordinates, whose error is the FP32 rounding of the
// projected corners (V.shade_err32: 3 ulps of the window's largest
// coordinate, twice, per km of edge) times the facet's longest edge over
// twice its projected area (|P.den|); times the larger vertex-normal
// difference along the barycentric axes, that bounds the cosine's error.
// A facet seen nearly edge-on on a rough surface (vertex normals far apart)
// makes it large: on the 3%-noise model one pixel's cosine was off by 7.6e-4
// and the model image pixel it fed by 1.7e-4 of itself.  Pixels whose
// bound exceeds SHADE_TOL are shaded by the FP64 replay instead.  The bound
// runs ~250x above the actual errors, which it ranks well: emulating the
// FP32 shading of every lit pixel of one frame, SHADE_TOL = 3e-4 sends
// 0.09% (large-model workload, 0.2% noise) and 1.05% (3% noise) of them
// to the replay and leaves the rest within 4e-5 and 1.3e-5 of the FP64
// cosine.  The per-facet edge and normal spread sit in the .w lanes of the
// facet-major copies (k_block_bounds).
constexpr float SHADE_TOL = 3.0e-4f;
// edge: the facet's longest edge (km); spread: its vertex-normal spread, or
// 0 for unsmoothed shading (the facet normal does not depend on s and t)
__device__ inline bool shade_ill(const ViewRec &V, const Proj32 &P, float edge, float spread) {
  return V.shade_err32 * fabsf(P.den) * edge * spread > SHADE_TOL;
}
__device__ inline bool shade_ill(const ViewRec &V, const GeomSlot &G, int f, const Proj32 &P) {
  return shade_ill(V, P, G.fvx[3 * f].w, V.smooth ? G.fvn[3 * f].w : 0.0f);
}

// Line-of-sight component of the normal at (s, t) of facet f in observer
// coordinates (posvis.c:243-258): the radar's cos(scattering).  A smoothed
// normal interpolates the facet's vertex normals na, nb, nc and is
// normalized by its body-frame length (a rotation keeps lengths), with the
// hardware reciprocal square root (1 ulp; the length is near 1).
__device__ inline float normal_los(const float *m, float4 na, float4 nb, float4 nc, float s, float t) {
#pragma clang fp contract(off)
  const float bx = fmaf(t, nc.x - nb.x, fmaf(s, nb.x - na.x, na.x));
  const float by = fmaf(t, nc.y - nb.y, fmaf(s, nb.y - na.y, na.y));
  const float bz = fmaf(t, nc.z - nb.z, fmaf(s, nb.z - na.z, na.z));
  const float nz = fmaf(m[8], bz, fmaf(m[7], by, m[6] * bx));
  return nz * __builtin_amdgcn_rsqf(fmaf(bz, bz, fmaf(by, by, bx * bx)));
}
// from the facet-major vertex-normal copies
__device__ inline float normal32f_los(const GeomSlot &G, int f, const float *m, int smooth, float s, float t) {
  if (smooth) return normal_los(m, G.fvn[3 * f], G.fvn[3 * f + 1], G.fvn[3 * f + 2], s, t);
  return facing32(m, G.fn32[f]);
}
// from the vertex normals through the facet's vertex indices
__device__ inline float normal32_los(const GeomSlot &G, int f, const float *m, int smooth, float s, float t) {
  if (smooth) {
    const int *fv = G.fv + 3 * f;
    return normal_los(m, G.vn32[fv[0]], G.vn32[fv[1]], G.vn32[fv[2]], s, t);
  }
  return facing32(m, G.fn32[f]);
}

__device__ inline float sinc2f(float x) {
  float t = fabsf(3.14159265358979f * x);
  if (t > 1.0e-4f) { t = __sinf(t) / t; return t * t; }
  return 1.0f;
}
__device__ inline float trif(float x) { float t = 1.0f - fabsf(x); return t > 0.0f ? t : 0.0f; }

// hw - |x| for x = hi + lo + c (hi + lo: an FP64 fraction in [0, 1) held as
// an FP32 pair; c and hw small multiples of 1/4), clamped at zero: the
// unit triangle of half-width hw is this over hw.  Near the triangle's feet,
// where |x| ~ hw and the difference cancels, (hw -+ c) -+ hi is exact in
// FP32, so the small weights of a footprint's edge rows keep FP32 relative
// precision; formed from one rounded FP32 argument they would carry its
// absolute error (~3e-8) instead.
__device__ inline float tri_rem(float hi, float lo, float c, float hw) {
#pragma clang fp contract(off)
  const float r = (hi + c) + lo >= 0.0f ? ((hw - c) - hi) - lo : ((hw + c) + hi) + lo;
  return r > 0.0f ? r : 0.0f;
}

// delay response of image row idel = delPOS - (hi + lo) - k
// (pos2deldop.c:371-424; delPOS - idel = hi + lo + k).  One sample per baud
// (the common case) inline, the rest out of line.
__device__ __attribute__((noinline)) float delay_w32_other(const ViewRec &V, float hi, float lo, int k) {
  const int sos = V.spb / V.stride;
  const float hw = (float)sos, fk = (float)k;
  if (V.codemethod != 1) {
    const float c2 = (float)V.const2;
    if (V.spb == 2) {  // TRI(arg) + TRI(arg - 1/2), arg = (delPOS - idel + const2) / sos
      const float d = (tri_rem(hi, lo, fk + c2, hw) + tri_rem(hi, lo, fk + c2 - 0.5f * hw, hw)) / hw;
      return d * (d / 4.0f);
    }
    // sample m's triangle is centred m/stride - const2 rows after the row centre
    const float drel = (hi + fk) + lo;
    const int m_min = max((int)floorf((drel - c2) * V.stride), 0);
    const int m_max = min((int)ceilf((drel + (float)V.const1) * V.stride), V.spb) - 1;
    const float step = hw / V.spb;
    float d = 0.0f;
    for (int m = m_min; m <= m_max; m++) d += tri_rem(hi, lo, fk + c2 - m * step, hw);
    d /= hw;
    return d * (d / (V.spb * V.spb));
  }
  const float t = tri_rem(hi, lo, fk, hw) / hw;  // long code, original reduction
  return t * t;
}
__device__ inline float delay_w32(const ViewRec &V, float hi, float lo, int k) {
  if (V.codemethod != 1 && V.spb == 1) { const float t = tri_rem(hi, lo, (float)k, 1.0f); return t * t; }
  return delay_w32_other(V, hi, lo, k);
}

// c^e for 0 < c <= 1 by the hardware base-2 log and exponential (~1 ulp
// each; the per-pixel FP32 contributions carry ~1e-7 relative anyway)
__device__ inline float powc(float c, float e) { return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(c)); }

// The diffuse cosine law R(C+1) cos^(2C-1) (radlaw.c:51-54); pl = [type, R, C, ...].
__device__ inline float cosine_law32(const double *pl, float c) {
  return (float)pl[1] * ((float)pl[2] + 1.0f) * powc(c, 2.0f * (float)pl[2] - 1.0f);
}

// Half-width, in cos(scattering), of the band around a quasispecular law's
// cutoff inside which a pixel's FP32 cosine cannot settle the cutoff test:
// well above the FP32 error of the LOS normal.
constexpr float CUT_BAND = 2e-5f;

// sin^2 of the scattering angle of facet f at barycentric (s, t) from the
// LOS-frame normal's x and y components (accurate where 1 - c^2 cancels):
// the body-frame normal is the facet's, or its vertex normals' at (s, t)
__device__ inline float los_sin2(const ViewRec &V, const GeomSlot &G, int f, float s, float t) {
  float bx, by, bz;
  if (V.smooth) {
    const int *fv = G.fv + 3 * f;
    const float4 na = G.vn32[fv[0]], nb = G.vn32[fv[1]], nc = G.vn32[fv[2]];
    bx = fmaf(t, nc.x - nb.x, fmaf(s, nb.x - na.x, na.x));
    by = fmaf(t, nc.y - nb.y, fmaf(s, nb.y - na.y, na.y));
    bz = fmaf(t, nc.z - nb.z, fmaf(s, nb.z - na.z, na.z));
  } else {
    const float4 n = G.fn32[f];
    bx = n.x; by = n.y; bz = n.z;
  }
  const float *m = V.oa32;
  const float x = fmaf(m[2], bz, fmaf(m[1], by, m[0] * bx));
  const float y = fmaf(m[5], bz, fmaf(m[4], by, m[3] * bx));
  return fmaf(x, x, y * y) / fmaf(bz, bz, fmaf(by, by, bx * bx));
}

// Below this cos(scattering) the tabular and per-facet laws go to the FP64
// replay: rho / c (and c^(2C-1) for C < 1/2) grows without bound at the limb,
// where the FP32 cosine's ~1e-7 absolute error becomes a large relative one.
// A few pixels in a thousand on a convex body.
constexpr float LIMB_REPLAY = 0.05f;

// The tabular law and the per-facet cosine laws (radlaw.c:56-66, :118-135)
// in FP32 from the evaluation's resolved values (V.lawloc): the tabular
// law's incidence as atan2(sin, cos) with sin^2 from the normal (acos of an
// FP32 cosine near 1 loses half its digits), its index clamped to n - 2.
// NaN (FP64 replay) for a limb pixel.
__device__ __attribute__((noinline)) float radlaw32_local(const ViewRec &V, const GeomSlot &G, const double *pl,
                                                          int f, float s, float t, float c) {
  if (c < LIMB_REPLAY) return __builtin_nanf("");
  if (V.law_type == TABULARLAW) {
    const int n = (int)pl[2];
    const float inc = atan2f(sqrtf(los_sin2(V, G, f, s, t)), c);
    const float ratio = inc * ((float)(n - 1) / 1.57079632679489662f);
    const int i = min((int)floorf(ratio), n - 2);
    const float r1 = (float)V.lawloc[i], r2 = (float)V.lawloc[i + 1];
    return (r1 + (ratio - (float)i) * (r2 - r1)) / c;
  }
  const float R = (float)V.lawloc[2 * f], Cc = (float)V.lawloc[2 * f + 1];
  return R * (Cc + 1.0f) * powc(c, 2.0f * Cc - 1.0f);
}

// The other homogeneous laws (radlaw.c:64-117) for a lit pixel of facet f
// with FP32 cos(scattering) c at barycentric (s, t); pl = [type, R, C,
// cutoff, qsR, qsC, qs_cutoff].  Their quasispecular terms peak steeply at
// c = 1, where 1 - c^2 cancels: the Hagfors argument c^4 + C sin^2, the
// Gaussian tan^2 = sin^2 / c^2 and cos^(2C-1) = exp((C - 1/2) log1p(-sin^2))
// take sin^2 from the LOS-frame normal's own x and y components instead.
// Returns NaN for a pixel within CUT_BAND of the cutoff: the caller hands it
// to the FP64 replay, which evaluates the law from the CPU's cosine.
// Out of line: no configuration of the hot path uses these laws.
__device__ __attribute__((noinline)) float radlaw32_steep(const ViewRec &V, const GeomSlot &G, const double *pl, int f,
                                                           float s, float t, float c) {
#pragma clang fp contract(off)
  const int type = V.law_type;
  const bool hybrid = type >= GAUSSIAN_COSINE;
  const float cut = (float)pl[hybrid ? 6 : 3];
  if (fabsf(c - cut) <= CUT_BAND) return __builtin_nanf("");
  float d = 0.0f, R = (float)pl[1], Cc = (float)pl[2];
  if (hybrid) {
    d = R * (Cc + 1.0f) * powc(c, 2.0f * Cc - 1.0f);  // the diffuse component
    R = (float)pl[4];
    Cc = (float)pl[5];
  }
  if (!(c >= cut)) return d;
  const float s2 = los_sin2(V, G, f, s, t);
  const float c2 = c * c;
  switch (type) {
  case GAUSSIANLAW: case GAUSSIAN_COSINE:
    d += R * Cc * expf(-Cc * (s2 / c2)) / (c2 * c2 * c);
    break;
  case HAGFORSLAW: case HAGFORS_COSINE: {
    const float q = c2 * c2 + Cc * s2;
    d += 0.5f * R * Cc / (q * sqrtf(q) * c);
    break;
  }
  default:  // COSINELAW_QS, COSINE_COSINE
    d += R * (Cc + 1.0f) * expf((Cc - 0.5f) * log1pf(-s2));
  }
  return d;
}

// radar_pixel's cl for a lit pixel with FP32 cos(scattering) c: c itself
// under the diffuse cosine law, else the law's value (NaN: FP64 replay).
__device__ inline float radar_cl(const ViewRec &V, const GeomSlot &G, const double *pl, int f, float s, float t,
                                 float c) {
  if (V.law_type == COSINELAW_DIFF) return c < V.limb32 ? __builtin_nanf("") : c;
  if (V.lawloc) return radlaw32_local(V, G, pl, f, s, t, c);
  return radlaw32_steep(V, G, pl, f, s, t, c);
}

// One lit POS pixel's delay-Doppler footprint (pos2deldop.c:320-482 /
// pos2doppler.c:259-362): bins relative to the anchors D0 / P0, the fractional
// bin positions its response weights are computed from (dop_w / del_w, at
// the point of use, so no weight array stays live) and the amplitude of
// cross section per unit weight.
struct RadarPix {
  bool inb;
  bool sinc1;        // single-sample sinc^2 Doppler response
  float amp, delrel, doprel;
  int id0, id1, ip0, ip1;
  float pfr, s2;     // fractional Doppler bin; sin^2(pi pfr) / pi^2 (sinc1)
  int pm0;           // Doppler coordinate - first bin = pfr + pm0
  float dfr, dlo;    // fractional delay bin (FP64) as an FP32 pair dfr + dlo
  int dm0;           // delay coordinate - first row = dfr + dlo + dm0
};

// Doppler weight of bin b of a single-sample sinc^2 footprint
__device__ inline float dop_weight(float pfr, int pm0, float s2, int b) {
  float a = pfr + (float)(pm0 - b);
  // hardware reciprocal (1 ulp): the weights are FP32 and normalized by their
  // sum.  Within 1e-4 / pi of the center sinc^2 is 1 to 3e-9.
  return (fabsf(a) > 1.0e-4f / 3.14159265358979f) ? s2 * __builtin_amdgcn_rcpf(a * a) : 1.0f;
}

// Doppler response of bin ip0 + b averaged over nsinc2 x nsinc2 subpixels
// (pos2deldop.c:437-455); out of line: only nsinc2 > 1 takes it
__device__ __attribute__((noinline)) float dop_w_sub(const ViewRec &V, float pfr, int pm0, int b) {
  const int ns = V.nsinc2;
  const float xi = (float)V.xincr, yi = (float)V.yincr;
  float al = pfr + (float)(pm0 - b) + (float)V.dopdiff_bl, acc = 0.0f;
  for (int q = 0; q < ns; q++) {
    float g = al;
    for (int r = 0; r < ns; r++) { acc += sinc2f(g); g += xi; }
    al += yi;
  }
  return acc / (ns * ns);
}

// Doppler response of bin ip0 + b: sinc^2 at the pixel center, or its
// subpixel mean
__device__ inline float dop_w(const ViewRec &V, const RadarPix &R, int b) {
  return R.sinc1 ? dop_weight(R.pfr, R.pm0, R.s2, b) : dop_w_sub(V, R.pfr, R.pm0, b);
}

// delay response of row id0 + a
__device__ inline float del_w(const ViewRec &V, const RadarPix &R, int a) {
  return delay_w32(V, R.dfr, R.dlo, R.dm0 - a);
}

// cl: the pixel's FP32 cos(scattering) in a view with the diffuse cosine law
// (the law is evaluated here, last), or the law's value itself otherwise.
__device__ inline void radar_pixel(const ViewRec &V, const double *pl, int i, int j, double z, float cl,
                                   RadarPix &R) {
  // (FP64 bin coordinates in the CPU's order of operations: no contraction)
#pragma clang fp contract(off)
  const bool dd = V.type == VT_DELAY;
  int width = V.sinc2width, ns = V.nsinc2;
  // Bin coordinates and the integer footprint decisions in FP64, as the CPU
  // forms them (pos2deldop.c:320-329): a footprint that straddles a frame edge
  // moves a pixel's whole power between the image and the overflow, so these
  // floors must not flip under FP32 rounding.  The weights take FP32 offsets
  // built from the FP64 fractional bin coordinate plus a small integer, so an
  // offset carries one FP32 rounding of itself, not of the absolute bin number.
  double dop64 = V.ax * (double)i + V.ay * (double)j + V.dopshift;
  R.ip0 = (int)floor(dop64 - V.dopdiff_max + 1 - width / 2.0) - V.P0;
  R.ip1 = (int)floor(dop64 + V.dopdiff_max + width / 2.0) - V.P0;
  R.doprel = (float)(dop64 - (double)V.P0);
  double pfl = floor(dop64);
  R.pfr = (float)(dop64 - pfl);
  R.pm0 = (int)pfl - (V.P0 + R.ip0);
  const int np = R.ip1 - R.ip0 + 1;
  float codef = 1.0f;
  if (dd && V.codemethod == 0) codef = sinc2f((R.doprel + V.dcrel32) / (float)V.dopfftlen);
  R.s2 = 0.0f;
  R.sinc1 = ns == 1;
  if (R.sinc1) {
    // sin(pi (x - k)) = +-sin(pi x): one sine serves every Doppler bin.  The
    // hardware sine (argument in revolutions) suffices: s2 scales every
    // weight alike and cancels in their normalization, except for a bin
    // within 1e-4 of the center, whose neighbours then weigh < 1e-8.
    float sp = __builtin_amdgcn_sinf(0.5f * R.pfr);
    R.s2 = sp * sp * (1.0f / (3.14159265358979f * 3.14159265358979f));
  }
  float sump = 0.0f;
  if (R.sinc1)
    for (int b = 0; b < np; b++) sump += dop_weight(R.pfr, R.pm0, R.s2, b);
  else
    for (int b = 0; b < np; b++) sump += dop_w_sub(V, R.pfr, R.pm0, b);
  float sumd = 1.0f;
  R.delrel = 0.0f;
  R.id0 = 0;
  R.id1 = 0;
  R.dfr = 0.0f;
  R.dlo = 0.0f;
  R.dm0 = 0;
  if (dd) {
    double del64 = (double)z * V.delfact + V.delshift;
    R.id0 = (int)floor(del64 - V.const1) + 1 - V.D0;
    R.id1 = (int)ceil(del64 + V.const1) - 1 - V.D0;
    R.delrel = (float)(del64 - (double)V.D0);
    double dfl = floor(del64);
    const double dfrac = del64 - dfl;
    R.dfr = (float)dfrac;
    R.dlo = (float)(dfrac - (double)R.dfr);
    R.dm0 = (int)dfl - (V.D0 + R.id0);
    const int nd = R.id1 - R.id0 + 1;
    sumd = 0.0f;
    for (int a = 0; a < nd; a++) sumd += del_w(V, R, a);
    R.inb = (V.D0 + R.id0 >= 1) && (V.D0 + R.id1 <= V.ndel) && (V.P0 + R.ip0 >= 1) && (V.P0 + R.ip1 <= V.ndop);
  } else {
    R.inb = (V.P0 + R.ip0 >= 1) && (V.P0 + R.ip1 <= V.ndop);
  }
  // (an empty Doppler range, sinc2width 0, adds nothing, as on the CPU)
  R.amp = np < 1 ? 0.0f
                 : (V.law_type == COSINELAW_DIFF ? cosine_law32(pl, cl) : cl) * V.kmpp32 * V.kmpp32 * codef *
                       __builtin_amdgcn_rcpf(sumd * sump);
}

// scatter one pixel straight into the HBM fit image (FP64 atomics) or, for
// Doppler pixels beyond the spectrum, into the FP64 overflow spectrum
__device__ inline void radar_scatter_global(const ViewRec &V, const RadarPix &R, double *img, double *ov) {
  const bool dd = V.type == VT_DELAY;
  int np = R.ip1 - R.ip0 + 1, nd = dd ? R.id1 - R.id0 + 1 : 1, ndop = V.ndop;
  if (R.inb) {
    for (int a = 0; a < nd; a++) {
      const float ad = dd ? R.amp * del_w(V, R, a) : R.amp;
      const long long row = (long long)((dd ? V.D0 + R.id0 + a : 1) - 1) * ndop + (V.P0 + R.ip0 - 1);
      for (int b = 0; b < np; b++) atomicAdd(&img[row + b], (double)(ad * dop_w(V, R, b)));
    }
  } else if (!dd) {
    int idop0 = V.idop0;
    int lo = max(V.P0 + R.ip0, -idop0), hi = min(V.P0 + R.ip1, -idop0 + 2000 - 1);
    for (int b = 0; b < np; b++) {
      int idop = V.P0 + R.ip0 + b;
      if (idop >= lo && idop <= hi) atomicAdd(&ov[idop + idop0], (double)(R.amp * dop_w(V, R, b)));
    }
  }
}

// Add one pixel's footprint into an FP64 LDS window (rows of wp cells; origin
// = window cell of the footprint's first row and bin).  Lanes whose
// footprints start at the same bin would add into the same cells in lockstep,
// and same-address FP64 LDS atomics serialize, so each lane walks its Doppler
// bins from its own starting offset.
__device__ inline void radar_scatter_lds(const ViewRec &V, const RadarPix &R, double *origin, int wp) {
  const bool dd = V.type == VT_DELAY;
  const int np = R.ip1 - R.ip0 + 1, nd = dd ? R.id1 - R.id0 + 1 : 1;
  int b0 = threadIdx.x & 15;
  while (b0 >= np) b0 -= np;
  // rows in groups of DW
  for (int a0 = 0; a0 < nd; a0 += DW) {
    float ad[DW];
#pragma unroll
    for (int a = 0; a < DW; a++) ad[a] = (a0 + a < nd) ? (dd ? R.amp * del_w(V, R, a0 + a) : R.amp) : 0.0f;
    double *org = origin + a0 * wp;
    int b = b0;
    if (R.sinc1) {
      for (int st = 0; st < np; st++) {
        const float pwb = dop_weight(R.pfr, R.pm0, R.s2, b);
#pragma unroll
        for (int a = 0; a < DW; a++)
          if (a0 + a < nd) atomicAdd(&org[a * wp + b], (double)(ad[a] * pwb));
        if (++b == np) b = 0;
      }
    } else {
      for (int st = 0; st < np; st++) {
        const float pwb = dop_w_sub(V, R.pfr, R.pm0, b);
#pragma unroll
        for (int a = 0; a < DW; a++)
          if (a0 + a < nd) atomicAdd(&org[a * wp + b], (double)(ad[a] * pwb));
        if (++b == np) b = 0;
      }
    }
  }
}

}  // namespace shp
