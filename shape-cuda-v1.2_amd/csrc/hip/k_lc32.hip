// k_lc32.hip -- the mixed method for lightcurve points (calc_fits.c:1509-1959:
// posvis.c for the observer, src 0, and the sun, src 1; posmask.c;
// apply_photo.c).
//
// The facet-by-pixel work -- projecting and covering every facet in both
// views -- is the FP32 raster of the radar views (k_raster32 with lc = 1):
// packed (depth, facet) keys, and pixels within the FP32 band of a facet
// edge recorded for the FP64 replay instead of claimed.  What the shadow test
// and the photometry read is then formed in FP64, per pixel:
//   k_resolve_lc  the owner a pixel's FP32 key names is exact off the edge
//                 band, so its depth and cos(emission) / cos(incidence) come
//                 from that facet in FP64 (posvis.c:217-275); a pixel whose
//                 owner's smoothed normal faces away inherits its cosine from
//                 an earlier claimant and joins the replay records;
//   k_replay_lc   every recorded pixel replays the facets covering it (found
//                 by k_collect32) in facet order with the CPU's claim rule, in
//                 FP64: owner, depth and cosines as the CPU leaves them.
// The observer and sun views take these steps one after the other through
// the same key buffer and replay records; k_posmask64 and k_photo64 then run
// on the FP64 arrays exactly as in the FP64 method.
#include "replay_dev.h"

namespace shp {

__global__ __launch_bounds__(256) void k_resolve_lc(ViewRec *views, const GeomSlot *slots, PosDev pd, int src) {
  ViewRec &V = views[blockIdx.y];
  if (V.type != VT_LGHTCRV) return;  // whole workgroup
  const int n = V.n, w = 2 * n + 1;
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = pix / w - n, j = pix % w - n;
  const int *xl = src ? V.xlim2 : V.xlim, *yl = src ? V.ylim2 : V.ylim;
  // pixels outside the view's rectangle are never read (k_posmask64,
  // k_photo64 and the sun-depth lookups treat them as blank sky)
  const bool in = pix < w * w && i >= xl[0] && i <= xl[1] && j >= yl[0] && j <= yl[1];
  bool rec = false;
  if (in) {
    const long long p = V.pos_off + pix;
    const unsigned long long k = pd.zkey[p];
    pd.zkey[p] = 0ULL;  // ready for the next view's raster
    double *zz = src ? pd.zill : pd.z, *cosa = src ? pd.cosill : pd.cose;
    int *fac = src ? pd.fill : pd.f;
    if (pd.stale_idx[p] < 0) {  // (a recorded pixel is written by the replay)
      zz[p] = -1.0e20;
      cosa[p] = 0.0;
      fac[p] = -1;
      if (!src) pd.cosi[p] = 0.0;
      if (k) {
        const GeomSlot &G = slots[V.slot];
        const int fw = key_facet(k);
        FacetProj P;
        project_facet_px(V, G, fw, src, P);
        double s, t, z;
        if (P.front && pixel_hit(P, V.kmpp, i, j, s, t, z)) {
          double nz, cosb;
          shade64(V, G, fw, src, src ? V.smooth_src : V.smooth, s, t, nz, cosb);
          zz[p] = z;
          fac[p] = fw;
          if (nz > 0.0) {
            cosa[p] = nz;
            if (!src && V.bistatic) {
              pd.cosi[p] = cosb;
              if (cosb <= 0.0) cosa[p] = 0.0;
            }
          } else {
            rec = true;  // cos inherited from an earlier claimant
          }
        } else {
          // the FP32 claim does not hold in FP64 (its band is 5x the FP32
          // error, so this is not expected): the replay decides
          rec = true;
        }
      }
    }
  }
  const int si = stale_slot_lc(V, pd, blockIdx.y, rec);  // (every lane of the wave takes this call)
  if (si >= 0) stale_init(V, pd, si, blockIdx.y, pix, i, j, src);
}

// grid (x, view): the view's records, from its own range of the list
__global__ __launch_bounds__(256) void k_replay_lc(ViewRec *views, const GeomSlot *slots, PosDev pd) {
  const ViewRec &Vb = views[blockIdx.y];
  if (Vb.type != VT_LGHTCRV) return;
  const int nrec = min(pd.lc_count[blockIdx.y], Vb.lc_rec_cap);
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrec; r += gridDim.x * blockDim.x) {
    StaleRec &S = pd.stale[Vb.lc_rec_off + r];
    const ViewRec &V = views[S.view];
    const GeomSlot &G = slots[V.slot];
    const int n = V.n, w = 2 * n + 1;
    const int i = S.pix / w - n, j = S.pix % w - n;
    const long long p = V.pos_off + S.pix;
    const int nc = min(S.ncand, STALE_CAND);
    if (S.ncand > STALE_CAND) atomicOr(pd.error, 2);
    int c[STALE_CAND];
    for (int q = 0; q < nc; q++) c[q] = S.cand[q];
    sort_candidates(c, nc);
    const int src = S.src, smooth = src ? V.smooth_src : V.smooth;
    double zz = S.oldz, cosa = S.oldcose, cosb = S.oldcosi;
    int fac = S.oldf;
    for (int q = 0; q < nc; q++) {
      if (q > 0 && c[q] == c[q - 1]) continue;
      FacetProj P;
      project_facet_px(V, G, c[q], src, P);
      if (!P.front) continue;
      double s, t, z;
      if (!pixel_hit(P, V.kmpp, i, j, s, t, z)) continue;
      if (!((z > zz) || (fac < 0))) continue;
      zz = z;
      double nz, cb;
      shade64(V, G, c[q], src, smooth, s, t, nz, cb);
      if (nz > 0.0) {
        cosa = nz;
        if (!src && V.bistatic) {
          cosb = cb;
          if (cb <= 0.0) cosa = 0.0;
        }
      }
      fac = c[q];
    }
    (src ? pd.zill : pd.z)[p] = zz;
    (src ? pd.fill : pd.f)[p] = fac;
    (src ? pd.cosill : pd.cose)[p] = cosa;
    if (!src) pd.cosi[p] = V.bistatic ? cosb : 0.0;
    pd.stale_idx[p] = -1;
  }
}

}  // namespace shp
