// replay_dev.h -- FP64 replay of one POS pixel of the mixed method, shared by
// the global replay list (k_replay32) and the binned tile kernel (k_tile32).
#pragma once

#include "fp64_dev.h"
#include "mixed_dev.h"

namespace shp {

// Sort a pixel's candidate facets ascending (insertion sort; at most
// STALE_CAND entries).  c may live in LDS.
__device__ inline void sort_candidates(int *c, int nc) {
  for (int a = 1; a < nc; a++) {
    int v = c[a], b = a - 1;
    while (b >= 0 && c[b] > v) { c[b + 1] = c[b]; b--; }
    c[b + 1] = v;
  }
}

// The candidates (sorted ascending) claim pixel (i, j) in facet order with the
// CPU rule `z > zz || fac < 0`, in the CPU's own FP64 arithmetic, so the
// owner, depth and (inherited) cos(scattering) are the CPU's
// (posvis.c:142-275).  Returns the owner (-1: none); cosa > 0 when lit.
__device__ inline int replay_owner64(const ViewRec &V, const GeomSlot &G, const int *c, int nc, int i, int j,
                                     double &zz, double &cosa) {
  zz = -1.0e20;
  cosa = 0.0;
  int fac = -1;
  for (int q = 0; q < nc; q++) {
    if (q > 0 && c[q] == c[q - 1]) continue;
    FacetProj P;
    project_facet_px(V, G, c[q], 0, P);
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
  return fac;
}

// Lower the ordered FP64 limit at a to v, reading it first: the limits only
// move outward, so a value read stale can only cost an atomic, never skip
// one that is needed, and most replayed pixels leave the limits as they are
// (same-address atomics serialize at the memory side).
__device__ inline void lower_limit_d(unsigned long long *a, double v) {
  const unsigned long long o = (unsigned long long)ord64(v);
  if (o < *(volatile unsigned long long *)a) atomicMin(a, o);
}

// radlaw64 for the laws other than the diffuse cosine law, out of line so
// that it costs the replay kernel no registers where those laws are not used
__device__ __attribute__((noinline)) float radlaw64_other(const double *pl, int type, double c, const double *loc,
                                                          int f) {
  return (float)radlaw64x(pl, type, c, loc, f);
}

// Scatter a replayed lit pixel straight into the HBM fit image (or the
// overflow spectrum) and grow the view's bin limits by it: a replayed pixel
// is lit like any other.
__device__ inline void replay_emit(ViewRec &V, const double *pl, int i, int j, double zz, double cosa, int fac,
                                   double *img, double *ov) {
  RadarPix R;
  // the other laws in FP64 from the CPU's cosine: their cutoff test and
  // steep quasispecular terms need it
  const float cl = V.law_type == COSINELAW_DIFF ? (float)cosa : radlaw64_other(pl, V.law_type, cosa, V.lawloc, fac);
  radar_pixel(V, pl, i, j, zz, cl, R);
  radar_scatter_global(V, R, img, ov);
  const bool dd = V.type == VT_DELAY;
  if (dd) {
    if (V.D0 + R.id0 < V.idellim[0]) atomicMin(&V.idellim[0], V.D0 + R.id0);
    if (V.D0 + R.id1 > V.idellim[1]) atomicMax(&V.idellim[1], V.D0 + R.id1);
    lower_limit_d(&V.dellim[0], (double)V.D0 + (double)R.delrel);
    lower_limit_d(&V.dellim[1], -((double)V.D0 + (double)R.delrel));
  }
  if (V.P0 + R.ip0 < V.idoplim[0]) atomicMin(&V.idoplim[0], V.P0 + R.ip0);
  if (V.P0 + R.ip1 > V.idoplim[1]) atomicMax(&V.idoplim[1], V.P0 + R.ip1);
  lower_limit_d(&V.doplim[0], (double)V.P0 + (double)R.doprel);
  lower_limit_d(&V.doplim[1], -((double)V.P0 + (double)R.doprel));
  if (!dd && !R.inb) atomicOr(&V.any_overflow, 1);
}

}  // namespace shp
