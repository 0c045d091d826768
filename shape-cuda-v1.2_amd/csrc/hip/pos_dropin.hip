// pos_dropin.hip -- the reference's per-POS entry points on the GPU: posclr,
// posvis, posmask, radlaw, pos2deldop, pos2doppler and apply_photo.
//
// Each call moves the caller's Numerical-Recipes POS arrays ([-n..n]^2, one
// allocation per row) into one device block per array, runs the same kernels
// the engine uses (k_fp64.hip) on a single view, and writes back exactly the
// fields the reference writes.  There is no CPU fallback: a failed call
// records its code for shape_hip_last_error() and leaves the caller's
// outputs untouched.
#include <math.h>
#include <string.h>

#include <vector>

#include "common.h"
#include "engine_host.h"
#include "shape_hip.h"

using namespace shp;

namespace {

// device allocations of one call, released on scope exit
struct Scratch {
  std::vector<void *> ptrs;
  ~Scratch() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  T *get(size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(T) * (n ? n : 1)) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return (T *)p;
  }
};

inline long long pos_pixels(const struct pos_t *pos) { return (2LL * pos->n + 1) * (2LL * pos->n + 1); }

// host NR matrix <-> contiguous device block, row by row through one staging copy
template <class T>
int upload_mat(T *dst, T **m, int n) {
  int w = 2 * n + 1;
  std::vector<T> h((size_t)w * w);
  for (int i = -n; i <= n; i++) memcpy(&h[(size_t)(i + n) * w], &m[i][-n], sizeof(T) * w);
  return hipMemcpy(dst, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice) == hipSuccess ? 0 : -3;
}
template <class T>
int download_mat(T **m, const T *src, int n) {
  int w = 2 * n + 1;
  std::vector<T> h((size_t)w * w);
  if (hipMemcpy(h.data(), src, sizeof(T) * h.size(), hipMemcpyDeviceToHost) != hipSuccess) return -4;
  for (int i = -n; i <= n; i++) memcpy(&m[i][-n], &h[(size_t)(i + n) * w], sizeof(T) * w);
  return 0;
}

void copy9(double *d, double m[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) d[3 * i + j] = m[i][j];
}

// a view describing one caller POS; ordered-double accumulators start empty
void pos_view(ViewRec &V, struct pos_t *pos) {
  memset(&V, 0, sizeof(V));
  V.type = -1;
  V.n = pos->n;
  V.kmpp = pos->km_per_pixel;
  V.bistatic = pos->bistatic;
  V.host_spin = 1;
  V.keep_rect = 1;
  copy9(V.ae, pos->ae);
  copy9(V.oe, pos->oe);
  copy9(V.se, pos->se);
  V.xlim[0] = pos->xlim[0]; V.xlim[1] = pos->xlim[1];
  V.ylim[0] = pos->ylim[0]; V.ylim[1] = pos->ylim[1];
  V.xlim2[0] = pos->xlim2[0]; V.xlim2[1] = pos->xlim2[1];
  V.ylim2[0] = pos->ylim2[0]; V.ylim2[1] = pos->ylim2[1];
  for (int q = 0; q < 4; q++) V.ov[q] = V.ov2[q] = ord64(1.0e20);
  V.dellim[0] = V.dellim[1] = V.doplim[0] = V.doplim[1] = ord64(1.0e20);
}

// radar law parameters in the engine's block layout: type, R, C, cutoff, qsR, qsC, qs_cutoff
int pack_radlaw(double pl[8], struct photo_t *photo, int ilaw) {
  memset(pl, 0, sizeof(double) * 8);
  if (ilaw < 0 || ilaw >= photo->nradlaws) return -13;
  union radscat_t *r = &photo->radar[ilaw];
  int t = photo->radtype[ilaw];
  pl[0] = t;
  switch (t) {
  case COSINELAW_DIFF: pl[1] = r->RC.R.val; pl[2] = r->RC.C.val; return 0;
  case GAUSSIANLAW: case HAGFORSLAW: case COSINELAW_QS:
    pl[1] = r->quasispec.R.val; pl[2] = r->quasispec.C.val; pl[3] = r->quasispec.cos_cutoff;
    return 0;
  case GAUSSIAN_COSINE: case HAGFORS_COSINE: case COSINE_COSINE:
    pl[1] = r->hybrid.diff.R.val; pl[2] = r->hybrid.diff.C.val;
    pl[4] = r->hybrid.qs.R.val; pl[5] = r->hybrid.qs.C.val; pl[6] = r->hybrid.qs.cos_cutoff;
    return 0;
  case TABULARLAW: pl[2] = r->tabular.n; return 0;
  case HARMCOSINE_DIFF: case INHOCOSINE_DIFF: return 0;
  default:
    return -13;
  }
}

// The caller's resolved values of a tabular or per-facet law (its
// realize_photo has run): a tabular law's rho table (k = 1), or k values per
// facet of component 0 for facets 0 .. nf-1.  Empty for the other laws.
void law_values(struct photo_t *photo, int radar, int ilaw, int nf, std::vector<double> &loc, int &k) {
  loc.clear();
  k = 0;
  const int t = radar ? photo->radtype[ilaw] : photo->opttype[ilaw];
  char *base = nullptr;
  size_t elem = 0;
  if (radar) {
    union radscat_t *r = &photo->radar[ilaw];
    if (t == TABULARLAW) {
      k = 1;
      for (int i = 0; i < r->tabular.n; i++) loc.push_back(r->tabular.rho[i].val);
      return;
    }
    if (t == HARMCOSINE_DIFF) base = (char *)r->harmcosine.local[0];
    else if (t == INHOCOSINE_DIFF) base = (char *)r->inhocosine.local[0];
    else return;
    k = 2;
    elem = sizeof(struct RC_t);
  } else {
    union optscat_t *o = &photo->optical[ilaw];
    switch (t) {
    case HARMLAMBERT: case HARMLOMMEL: base = (char *)o->harmR.local[0]; k = 1; elem = sizeof(struct R_t); break;
    case INHOLAMBERT: case INHOLOMMEL: base = (char *)o->inhoR.local[0]; k = 1; elem = sizeof(struct R_t); break;
    case HARMHAPKE: base = (char *)o->harmhapke.local[0]; k = 5; elem = sizeof(struct hapke_t); break;
    case INHOHAPKE: base = (char *)o->inhohapke.local[0]; k = 5; elem = sizeof(struct hapke_t); break;
    case HARMKAAS: base = (char *)o->harmkaas.local[0]; k = 5; elem = sizeof(struct kaas_t); break;
    case INHOKAAS: base = (char *)o->inhokaas.local[0]; k = 5; elem = sizeof(struct kaas_t); break;
    default: return;
    }
  }
  loc.resize((size_t)k * std::max(nf, 1), 0.0);
  for (int f = 0; f < nf; f++) {
    const struct param_t *p = (const struct param_t *)(base + elem * f);
    for (int j = 0; j < k; j++) loc[(size_t)k * f + j] = p[j].val;
  }
}

// facets a POS names: one more than the largest pos->f in its rectangle
int pos_nf(struct pos_t *pos) {
  int m = -1;
  for (int i = pos->xlim[0]; i <= pos->xlim[1]; i++)
    for (int j = pos->ylim[0]; j <= pos->ylim[1]; j++) m = std::max(m, pos->f[i][j]);
  return m + 1;
}

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------------------
// posclr (posclr.c:24): blank sky over the whole POS
// ---------------------------------------------------------------------------
extern "C" void posclr(struct pos_t *pos) {
  int n = pos->n;
  long long P = pos_pixels(pos);
  Scratch S;
  double *dz = S.get<double>(P), *d0 = S.get<double>(P);
  int *dm = S.get<int>(P);
  if (!dz || !d0 || !dm) { record_error(-2, "posclr allocation"); return; }
  hipLaunchKernelGGL(k_fill_f64, dim3(cdiv(P, 256)), dim3(256), 0, 0, dz, P, -HUGENUMBER);
  hipLaunchKernelGGL(k_fill_f64, dim3(cdiv(P, 256)), dim3(256), 0, 0, d0, P, 0.0);
  hipLaunchKernelGGL(k_fill_i32, dim3(cdiv(P, 256)), dim3(256), 0, 0, dm, P, -1);
  if (hipDeviceSynchronize() != hipSuccess) { record_error(-4, "posclr"); return; }
  int rc = download_mat(pos->b, d0, n) | download_mat(pos->cose, d0, n) | download_mat(pos->cosi, d0, n) |
           download_mat(pos->z, dz, n) | download_mat(pos->body, dm, n) | download_mat(pos->comp, dm, n) |
           download_mat(pos->f, dm, n);
  if (!rc && pos->bistatic)
    rc = download_mat(pos->cosill, d0, n) | download_mat(pos->zill, dz, n) | download_mat(pos->bodyill, dm, n) |
         download_mat(pos->compill, dm, n) | download_mat(pos->fill, dm, n);
  if (rc) { record_error(rc, "posclr download"); return; }
  pos->xlim[0] = pos->ylim[0] = n;
  pos->xlim[1] = pos->ylim[1] = -n;
  if (pos->bistatic) {
    pos->xlim2[0] = pos->ylim2[0] = n;
    pos->xlim2[1] = pos->ylim2[1] = -n;
  }
}

// ---------------------------------------------------------------------------
// posvis (posvis.c:84): rasterize a realized vertex set into the POS.  Pixels
// the POS already owns are claimed only by nearer facets (the `z > zz ||
// fac < 0` rule), so several bodies or components can be drawn in sequence.
// Returns outbnd; sets pos->posbnd_logfactor.
// ---------------------------------------------------------------------------
extern "C" int posvis(struct vertices_t *verts, double orbit_offset[3], struct pos_t *pos, int smooth, int src,
                      int body, int comp) {
  int n = pos->n, nv = verts->nv, nf = verts->nf;
  long long P = pos_pixels(pos);
  if (src && !pos->bistatic) { record_error(-16, "posvis (src = 1 needs a bistatic POS)"); return 0; }
  // geometry
  std::vector<double> x(3 * (size_t)nv), vn(3 * (size_t)nv), fn(3 * (size_t)nf);
  std::vector<int> fv(3 * (size_t)nf);
  for (int i = 0; i < nv; i++)
    for (int k = 0; k < 3; k++) { x[3 * i + k] = verts->v[i].x[k]; vn[3 * i + k] = verts->v[i].n[k]; }
  for (int f = 0; f < nf; f++)
    for (int k = 0; k < 3; k++) { fn[3 * f + k] = verts->f[f].n[k]; fv[3 * f + k] = verts->f[f].v[k]; }
  Scratch S;
  GeomSlot G;
  memset(&G, 0, sizeof(G));
  G.nv = nv; G.nf = nf;
  G.x = S.get<double>(x.size()); G.vn = S.get<double>(vn.size()); G.fn = S.get<double>(fn.size());
  int *dfv = S.get<int>(fv.size());
  G.fv = dfv;
  GeomSlot *dG = S.get<GeomSlot>(1);
  // POS arrays of the side being drawn (observer or source)
  PosDev pd;
  memset(&pd, 0, sizeof(pd));
  pd.zkey = S.get<unsigned long long>(P);
  pd.fbuf = S.get<int>(P);
  pd.stale_idx = S.get<int>(P);
  pd.stale_cap = (int)std::max<long long>(P / 8, 4096);
  pd.stale = S.get<StaleRec>(pd.stale_cap);
  pd.nstale = S.get<int>(4);
  pd.error = pd.nstale + 1;
  pd.ncollect = pd.nstale + 2;
  double *zz = S.get<double>(P), *cosa = S.get<double>(P);
  int *fac = S.get<int>(P), *bod = S.get<int>(P), *cmp = S.get<int>(P);
  double *cosb = (!src && pos->bistatic) ? S.get<double>(P) : nullptr;
  ViewRec V;
  pos_view(V, pos);
  V.smooth = smooth;
  V.smooth_src = smooth;
  V.rect_by_claim = 1;
  for (int k = 0; k < 3; k++) V.orb[k] = orbit_offset[k];
  ViewRec *dV = S.get<ViewRec>(1);
  double *vout = S.get<double>(VIEW_OUT);
  if (!G.x || !G.vn || !G.fn || !dfv || !dG || !pd.zkey || !pd.fbuf || !pd.stale_idx || !pd.stale || !pd.nstale ||
      !zz || !cosa || !fac || !bod || !cmp || (!src && pos->bistatic && !cosb) || !dV || !vout) {
    record_error(-2, "posvis allocation");
    return 0;
  }
  // body/comp are always set: the resolve kernel writes body/comp (or their
  // source-view arrays) when they are present
  pd.body = bod; pd.comp = cmp;
  if (src) { pd.zill = zz; pd.cosill = cosa; pd.fill = fac; pd.bodyill = bod; pd.compill = cmp; }
  else { pd.z = zz; pd.cose = cosa; pd.f = fac; pd.cosi = cosb; }
  int rc = 0;
  rc |= hipMemcpy(G.x, x.data(), sizeof(double) * x.size(), hipMemcpyHostToDevice) != hipSuccess;
  rc |= hipMemcpy(G.vn, vn.data(), sizeof(double) * vn.size(), hipMemcpyHostToDevice) != hipSuccess;
  rc |= hipMemcpy(G.fn, fn.data(), sizeof(double) * fn.size(), hipMemcpyHostToDevice) != hipSuccess;
  rc |= hipMemcpy(dfv, fv.data(), sizeof(int) * fv.size(), hipMemcpyHostToDevice) != hipSuccess;
  rc |= hipMemcpy(dG, &G, sizeof(G), hipMemcpyHostToDevice) != hipSuccess;
  rc |= hipMemcpy(dV, &V, sizeof(V), hipMemcpyHostToDevice) != hipSuccess;
  if (rc) { record_error(-3, "posvis upload"); return 0; }
  rc = src ? (upload_mat(zz, pos->zill, n) | upload_mat(cosa, pos->cosill, n) | upload_mat(fac, pos->fill, n) |
              upload_mat(bod, pos->bodyill, n) | upload_mat(cmp, pos->compill, n))
           : (upload_mat(zz, pos->z, n) | upload_mat(cosa, pos->cose, n) | upload_mat(fac, pos->f, n) |
              upload_mat(bod, pos->body, n) | upload_mat(cmp, pos->comp, n) |
              (cosb ? upload_mat(cosb, pos->cosi, n) : 0));
  if (rc) { record_error(rc, "posvis upload"); return 0; }
  const int T = 256;
  dim3 gf(cdiv(std::max(nf, 1), T), 1), gp(cdiv(P, T), 1);
  hipLaunchKernelGGL(k_pos_setup, dim3(1), dim3(64), 0, 0, dV);
  hipLaunchKernelGGL(k_pos_seed, gp, dim3(T), 0, 0, dV, pd, src);
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, 0, pd.nstale, 4);
  hipLaunchKernelGGL(k_raster64, gf, dim3(T), 0, 0, dV, (const GeomSlot *)dG, pd, src, 1);
  hipLaunchKernelGGL(k_raster64, gf, dim3(T), 0, 0, dV, (const GeomSlot *)dG, pd, src, 2);
  hipLaunchKernelGGL(k_resolve64, gp, dim3(T), 0, 0, dV, (const GeomSlot *)dG, pd, src, 1, 0, body, comp);
  hipLaunchKernelGGL(k_raster64, gf, dim3(T), 0, 0, dV, (const GeomSlot *)dG, pd, src, 3);
  hipLaunchKernelGGL(k_replay64, dim3(std::min(cdiv(pd.stale_cap, T), 2048u)), dim3(T), 0, 0, dV,
                     (const GeomSlot *)dG, pd);
  hipLaunchKernelGGL(k_finish_views, dim3(1), dim3(64), 0, 0, dV, 1, (double *)nullptr, (const FrameRec *)nullptr,
                     vout);
  int cnt[4];
  double out[VIEW_OUT];
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(&V, dV, sizeof(V), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(cnt, pd.nstale, sizeof(cnt), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(out, vout, sizeof(out), hipMemcpyDeviceToHost) != hipSuccess) {
    record_error(-4, "posvis kernels");
    return 0;
  }
  if (cnt[1]) { record_error(-30, "posvis (replay capacity)"); return 0; }
  rc = src ? (download_mat(pos->zill, zz, n) | download_mat(pos->cosill, cosa, n) | download_mat(pos->fill, fac, n) |
              download_mat(pos->bodyill, bod, n) | download_mat(pos->compill, cmp, n))
           : (download_mat(pos->z, zz, n) | download_mat(pos->cose, cosa, n) | download_mat(pos->f, fac, n) |
              download_mat(pos->body, bod, n) | download_mat(pos->comp, cmp, n) |
              (cosb ? download_mat(pos->cosi, cosb, n) : 0));
  if (rc) { record_error(rc, "posvis download"); return 0; }
  int *xl = src ? pos->xlim2 : pos->xlim, *yl = src ? pos->ylim2 : pos->ylim;
  const int *vx = src ? V.xlim2 : V.xlim, *vy = src ? V.ylim2 : V.ylim;
  xl[0] = vx[0]; xl[1] = vx[1]; yl[0] = vy[0]; yl[1] = vy[1];
  pos->posbnd_logfactor = out[OUT_POSBND_LOG + src];
  return (int)out[OUT_OUTBND + src];
}

// ---------------------------------------------------------------------------
// posmask (posmask.c:58): zero cos(scattering) where the source view shows a
// different surface element nearer than tol
// ---------------------------------------------------------------------------
extern "C" void posmask(struct pos_t *pos, double tol) {
  int n = pos->n;
  long long P = pos_pixels(pos);
  Scratch S;
  PosDev pd;
  memset(&pd, 0, sizeof(pd));
  pd.z = S.get<double>(P); pd.cose = S.get<double>(P); pd.zill = S.get<double>(P);
  pd.f = S.get<int>(P); pd.comp = S.get<int>(P); pd.body = S.get<int>(P);
  pd.fill = S.get<int>(P); pd.compill = S.get<int>(P); pd.bodyill = S.get<int>(P);
  ViewRec V;
  pos_view(V, pos);
  // every pixel of both arrays is read as it stands (posmask.c:88-166)
  V.xlim[0] = V.ylim[0] = V.xlim2[0] = V.ylim2[0] = -n;
  V.xlim[1] = V.ylim[1] = V.xlim2[1] = V.ylim2[1] = n;
  ViewRec *dV = S.get<ViewRec>(1);
  if (!pd.z || !pd.cose || !pd.zill || !pd.f || !pd.comp || !pd.body || !pd.fill || !pd.compill || !pd.bodyill ||
      !dV) {
    record_error(-2, "posmask allocation");
    return;
  }
  int rc = upload_mat(pd.z, pos->z, n) | upload_mat(pd.cose, pos->cose, n) | upload_mat(pd.zill, pos->zill, n) |
           upload_mat(pd.f, pos->f, n) | upload_mat(pd.comp, pos->comp, n) | upload_mat(pd.body, pos->body, n) |
           upload_mat(pd.fill, pos->fill, n) | upload_mat(pd.compill, pos->compill, n) |
           upload_mat(pd.bodyill, pos->bodyill, n);
  if (rc || hipMemcpy(dV, &V, sizeof(V), hipMemcpyHostToDevice) != hipSuccess) {
    record_error(-3, "posmask upload");
    return;
  }
  hipLaunchKernelGGL(k_posmask64, dim3(cdiv(P, 256), 1), dim3(256), 0, 0, dV, pd, tol);
  if (hipDeviceSynchronize() != hipSuccess) { record_error(-4, "posmask kernel"); return; }
  rc = download_mat(pos->cose, pd.cose, n);
  if (rc) record_error(rc, "posmask download");
}

// ---------------------------------------------------------------------------
// radlaw (radlaw.c:44): one scattering-law value for cos(incidence) cosinc
// ---------------------------------------------------------------------------
extern "C" double radlaw(struct photo_t *photo, int ilaw, double cosinc, int c, int f) {
  (void)c;  // component 0: the library's models have one
  double pl[8];
  int rc = pack_radlaw(pl, photo, ilaw);
  if (rc) { record_error(rc, "radlaw (law type)"); return 0.0; }
  // a per-facet law's values of facet f alone (passed as facet 0)
  std::vector<double> loc;
  int k = 0, fl = f;
  law_values(photo, 1, ilaw, f + 1, loc, k);
  if (k == 2 && f >= 0) {
    loc.erase(loc.begin(), loc.begin() + 2 * f);
    fl = 0;
  }
  Scratch S;
  double *dpl = S.get<double>(8), *dout = S.get<double>(1), *dloc = loc.empty() ? nullptr : S.get<double>(loc.size());
  double v = 0.0;
  if (!dpl || !dout || (!loc.empty() && !dloc) || hipMemcpy(dpl, pl, sizeof(pl), hipMemcpyHostToDevice) != hipSuccess ||
      (dloc && hipMemcpy(dloc, loc.data(), sizeof(double) * loc.size(), hipMemcpyHostToDevice) != hipSuccess)) {
    record_error(-3, "radlaw upload");
    return 0.0;
  }
  hipLaunchKernelGGL(k_radlaw_one, dim3(1), dim3(64), 0, 0, (const double *)dpl, (int)pl[0], cosinc,
                     (const double *)dloc, fl, dout);
  if (hipMemcpy(&v, dout, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
    record_error(-4, "radlaw kernel");
    return 0.0;
  }
  return v;
}

// ---------------------------------------------------------------------------
// pos2deldop (pos2deldop.c:176) and pos2doppler (pos2doppler.c:149): add one
// view's echo power, from the frame's POS, into the caller's fit (the fit is
// accumulated, as the reference does); set the frame's limits, km2Hz, the
// overflow terms and the bad-radar log factor; return badradar
// ---------------------------------------------------------------------------
namespace {

int radar_view(struct par_t *par, struct photo_t *photo, double orbit_xoff, double orbit_yoff, double orbit_dopoff,
               int is_dd, struct deldop_t *D, struct doppler_t *Q, int body, int frm, int v) {
  struct pos_t *pos = is_dd ? &D->frame[frm].pos : &Q->frame[frm].pos;
  int n = pos->n;
  long long P = pos_pixels(pos);
  // parameter block: set block (dopscale) at 0, radar law after it
  std::vector<double> pb(PB_SET_STRIDE + PB_RAD_STRIDE, 0.0);
  pb[PB_SET_DOPSCALE] = is_dd ? D->dopscale.val : Q->dopscale.val;
  int ilaw = is_dd ? D->iradlaw : Q->iradlaw;
  int rc = pack_radlaw(&pb[PB_SET_STRIDE], photo, ilaw);
  if (rc) return rc;
  ViewRec V;
  pos_view(V, pos);
  V.type = is_dd ? VT_DELAY : VT_DOPPLER;
  V.pb = 0; V.pb_set = 0; V.pb_dc = 0; V.pb_law = PB_SET_STRIDE;
  V.law_type = (int)pb[PB_SET_STRIDE];
  V.nsinc2 = par->nsinc2; V.sinc2width = par->sinc2width;
  V.orb_x = orbit_xoff; V.orb_y = orbit_yoff; V.orb_dop = orbit_dopoff;
  V.body = body;
  FrameRec F;
  memset(&F, 0, sizeof(F));
  F.type = V.type; F.nviews = 1; F.first_view = 0;
  long long N;
  if (is_dd) {
    struct deldopfrm_t *fr = &D->frame[frm];
    struct deldopview_t *w = &fr->view[v];
    copy9(V.oe, w->oe);
    for (int j = 0; j < 3; j++) V.spin[j] = w->spin[j];
    V.deloff = w->deloff; V.dopoff = w->dopoff;
    V.delcom_vig = fr->delcom_vig; V.dopcom_vig = fr->dopcom_vig; V.dopDC_vig = fr->dopDC_vig;
    V.dopfftlen = D->dopfftlen; V.Ftx = D->Ftx; V.del_per_pixel = D->del_per_pixel;
    V.dop_per_pixel = D->dop_per_pixel; V.codemethod = D->codemethod; V.spb = D->spb; V.stride = D->stride;
    V.ndel = fr->ndel; V.ndop = fr->ndop;
    F.ndel = fr->ndel; F.ndop = fr->ndop;
    N = (long long)fr->ndel * fr->ndop;
  } else {
    struct dopfrm_t *fr = &Q->frame[frm];
    struct dopview_t *w = &fr->view[v];
    copy9(V.oe, w->oe);
    for (int j = 0; j < 3; j++) V.spin[j] = w->spin[j];
    V.deloff = 0.0; V.dopoff = w->dopoff;
    V.dopcom_vig = fr->dopcom_vig;
    V.Ftx = Q->Ftx; V.dop_per_pixel = Q->dop_per_bin; V.ndel = 1; V.ndop = fr->ndop;
    V.spb = 1; V.stride = 1; V.codemethod = LONG_ORIG;
    F.ndel = 1; F.ndop = fr->ndop;
    F.sdev_sq = fr->sdev * fr->sdev; F.lookfact = fr->nlooks > 0.0 ? 1.0 / fr->nlooks : 0.0;
    F.speckle = par->speckle;
    N = fr->ndop;
  }
  // the caller's fit, accumulated into
  std::vector<double> fit(N);
  if (is_dd) {
    struct deldopfrm_t *fr = &D->frame[frm];
    for (int i = 1; i <= fr->ndel; i++) memcpy(&fit[(size_t)(i - 1) * fr->ndop], &fr->fit[i][1], sizeof(double) * fr->ndop);
  } else {
    memcpy(fit.data(), &Q->frame[frm].fit[1], sizeof(double) * N);
  }
  std::vector<double> loc;
  int lawk = 0;
  law_values(photo, 1, ilaw, pos_nf(pos), loc, lawk);
  Scratch S;
  PosDev pd;
  memset(&pd, 0, sizeof(pd));
  pd.z = S.get<double>(P); pd.cose = S.get<double>(P); pd.body = S.get<int>(P);
  double *dloc = nullptr;
  if (lawk) {
    pd.f = S.get<int>(P);
    dloc = S.get<double>(loc.size());
    if (!pd.f || !dloc || upload_mat(pd.f, pos->f, n) ||
        hipMemcpy(dloc, loc.data(), sizeof(double) * loc.size(), hipMemcpyHostToDevice) != hipSuccess)
      return -3;
    V.lawloc = dloc;
    V.lawk = lawk;
  }
  ViewRec *dV = S.get<ViewRec>(1);
  FrameRec *dF = S.get<FrameRec>(1);
  double *dpb = S.get<double>(pb.size()), *dfit = S.get<double>(N), *dov = S.get<double>(2000),
         *vout = S.get<double>(VIEW_OUT);
  if (!pd.z || !pd.cose || !pd.body || !dV || !dF || !dpb || !dfit || !dov || !vout) return -2;
  rc = upload_mat(pd.z, pos->z, n) | upload_mat(pd.cose, pos->cose, n) | upload_mat(pd.body, pos->body, n);
  if (rc) return rc;
  if (hipMemcpy(dV, &V, sizeof(V), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dF, &F, sizeof(F), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dpb, pb.data(), sizeof(double) * pb.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dfit, fit.data(), sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess)
    return -3;
  hipLaunchKernelGGL(k_fill_f64, dim3(8), dim3(256), 0, 0, dov, 2000LL, 0.0);
  hipLaunchKernelGGL(k_setup_views, dim3(1), dim3(64), 0, 0, dV, 1, (const double *)dpb);
  if (is_dd)
    hipLaunchKernelGGL(k_deldop64, dim3(syn_blocks(n), 1), dim3(256), 0, 0, dV, dF, pd, dfit, (const double *)dpb);
  else
    hipLaunchKernelGGL(k_doppler64, dim3(syn_blocks(n), 1), dim3(256), 0, 0, dV, dF, pd, dfit, dov,
                       (const double *)dpb);
  hipLaunchKernelGGL(k_finish_views, dim3(1), dim3(64), 0, 0, dV, 1, dov, (const FrameRec *)dF, vout);
  double out[VIEW_OUT];
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(out, vout, sizeof(out), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(fit.data(), dfit, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess)
    return -4;
  if (is_dd) {
    struct deldopfrm_t *fr = &D->frame[frm];
    for (int i = 1; i <= fr->ndel; i++) memcpy(&fr->fit[i][1], &fit[(size_t)(i - 1) * fr->ndop], sizeof(double) * fr->ndop);
    fr->idellim[0] = (int)out[OUT_IDELLIM]; fr->idellim[1] = (int)out[OUT_IDELLIM + 1];
    fr->idoplim[0] = (int)out[OUT_IDOPLIM]; fr->idoplim[1] = (int)out[OUT_IDOPLIM + 1];
    fr->dellim[0] = out[OUT_DELLIM]; fr->dellim[1] = out[OUT_DELLIM + 1];
    fr->doplim[0] = out[OUT_DOPLIM]; fr->doplim[1] = out[OUT_DOPLIM + 1];
    fr->view[v].km2Hz = out[OUT_KM2HZ];
    fr->badradar_logfactor = out[OUT_BADRADAR_LOG];
    fr->overflow_o2 = fr->overflow_m2 = fr->overflow_xsec = fr->overflow_delmean = fr->overflow_dopmean = 0.0;
  } else {
    struct dopfrm_t *fr = &Q->frame[frm];
    memcpy(&fr->fit[1], fit.data(), sizeof(double) * N);
    fr->idoplim[0] = (int)out[OUT_IDOPLIM]; fr->idoplim[1] = (int)out[OUT_IDOPLIM + 1];
    fr->doplim[0] = out[OUT_DOPLIM]; fr->doplim[1] = out[OUT_DOPLIM + 1];
    fr->view[v].km2Hz = out[OUT_KM2HZ];
    fr->badradar_logfactor = out[OUT_BADRADAR_LOG];
    fr->overflow_o2 = out[OUT_OVF]; fr->overflow_m2 = out[OUT_OVF + 1];
    fr->overflow_xsec = out[OUT_OVF + 2]; fr->overflow_dopmean = out[OUT_OVF + 3];
  }
  return out[OUT_BADRADAR] != 0.0 ? 1 : 0;
}

}  // namespace

extern "C" int pos2deldop(struct par_t *par, struct photo_t *photo, double orbit_xoff, double orbit_yoff,
                          double orbit_dopoff, struct deldop_t *deldop, int body, int set, int frm, int v) {
  (void)set;
  int r = radar_view(par, photo, orbit_xoff, orbit_yoff, orbit_dopoff, 1, deldop, nullptr, body, frm, v);
  if (r < 0) { record_error(r, "pos2deldop"); return 0; }
  return r;
}

extern "C" int pos2doppler(struct par_t *par, struct photo_t *photo, double orbit_xoff, double orbit_yoff,
                           double orbit_dopoff, struct doppler_t *doppler, int body, int set, int frm, int v) {
  (void)set;
  int r = radar_view(par, photo, orbit_xoff, orbit_yoff, orbit_dopoff, 0, nullptr, doppler, body, frm, v);
  if (r < 0) { record_error(r, "pos2doppler"); return 0; }
  return r;
}

// ---------------------------------------------------------------------------
// apply_photo (apply_photo.c:78): brightness of every lit, illuminated pixel
// of the given body into pos->b; returns their sum (1e-40 when zero)
// ---------------------------------------------------------------------------
extern "C" double apply_photo(struct mod_t *mod, int ilaw, double phase, double intensityfactor, struct pos_t *pos,
                              int body, int s, int frm) {
  (void)s;
  (void)frm;
  if (ilaw < 0 || ilaw >= mod->photo.noptlaws) { record_error(-14, "apply_photo (law index)"); return 1.0e-40; }
  int t = mod->photo.opttype[ilaw];
  union optscat_t *o = &mod->photo.optical[ilaw];
  double pl[PB_OPT_STRIDE] = {(double)t, 0, 0, 0, 0, 0};
  if (t == LAMBERTLAW || t == LOMMEL || t == GEOMETRICAL) pl[1] = o->R.R.val;
  else if (t == KAASALAINEN) { pl[1] = o->kaas.R.val; pl[2] = o->kaas.wt.val; pl[3] = o->kaas.A0.val; pl[4] = o->kaas.D.val; pl[5] = o->kaas.k.val; }
  else if (t == HAPKE) { const struct param_t *q = &o->hapke.w; for (int j = 0; j < 5; j++) pl[1 + j] = q[j].val; }
  else if (t < HARMLAMBERT || t > INHOKAAS) { record_error(-14, "apply_photo (law type)"); return 1.0e-40; }
  std::vector<double> loc;
  int lawk = 0;
  law_values(&mod->photo, 0, ilaw, pos_nf(pos), loc, lawk);
  int n = pos->n;
  long long P = pos_pixels(pos);
  Scratch S;
  PosDev pd;
  memset(&pd, 0, sizeof(pd));
  pd.cose = S.get<double>(P); pd.cosi = S.get<double>(P); pd.b = S.get<double>(P); pd.body = S.get<int>(P);
  ViewRec V;
  pos_view(V, pos);
  V.type = VT_LGHTCRV;
  V.pb = 0; V.pb_law = 0; V.law_type = t;
  V.phase = phase; V.intensityfactor = intensityfactor; V.body = body;
  V.ysum = 0.0;
  if (lawk) {
    pd.f = S.get<int>(P);
    double *dloc = S.get<double>(loc.size());
    if (!pd.f || !dloc || upload_mat(pd.f, pos->f, n) ||
        hipMemcpy(dloc, loc.data(), sizeof(double) * loc.size(), hipMemcpyHostToDevice) != hipSuccess) {
      record_error(-3, "apply_photo upload");
      return 1.0e-40;
    }
    V.lawloc = dloc;
    V.lawk = lawk;
  }
  ViewRec *dV = S.get<ViewRec>(1);
  double *dpl = S.get<double>(PB_OPT_STRIDE);
  if (!pd.cose || !pd.cosi || !pd.b || !pd.body || !dV || !dpl) { record_error(-2, "apply_photo allocation"); return 1.0e-40; }
  int rc = upload_mat(pd.cose, pos->cose, n) | upload_mat(pd.cosi, pos->cosi, n) | upload_mat(pd.b, pos->b, n) |
           upload_mat(pd.body, pos->body, n);
  if (rc || hipMemcpy(dV, &V, sizeof(V), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dpl, pl, sizeof(pl), hipMemcpyHostToDevice) != hipSuccess) {
    record_error(-3, "apply_photo upload");
    return 1.0e-40;
  }
  hipLaunchKernelGGL(k_photo64, dim3(cdiv(P, 256), 1), dim3(256), 0, 0, dV, pd, (const double *)dpl);
  if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(&V, dV, sizeof(V), hipMemcpyDeviceToHost) != hipSuccess) {
    record_error(-4, "apply_photo kernel");
    return 1.0e-40;
  }
  rc = download_mat(pos->b, pd.b, n);
  if (rc) { record_error(rc, "apply_photo download"); return 1.0e-40; }
  return V.ysum == 0.0 ? 1.0e-40 : V.ysum;
}
