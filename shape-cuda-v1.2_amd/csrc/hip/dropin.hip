// dropin.hip -- the reference's C entry points, computed on the GPU.
//
// Each entry keeps the reference's signature and side effects on the caller's
// structs (include/shape_hip.h cites the routine each one replaces).  The
// computation runs in HIP kernels; the host code here only moves the caller's
// Numerical-Recipes arrays to and from device memory.  There is no CPU
// fallback: when the GPU path fails the entry reports the failure through
// shape_hip_last_error() (and, where the reference returns a flag, returns -1)
// and leaves outputs untouched.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <memory>
#include <vector>

#include "common.h"
#include "engine_host.h"
#include "shape_hip.h"

using namespace shp;

static int g_err = 0;
static int g_strict_pos = 0;
static int g_last_nstale = 0;

static void set_err(int rc, const char *where) {
  if (rc) {
    g_err = rc;
    fprintf(stderr, "shape_hip: %s failed (code %d)\n", where, rc);
  }
}

namespace shp {
void record_error(int rc, const char *where) { set_err(rc, where); }
}  // namespace shp

extern "C" int shape_hip_last_error(void) { return g_err; }
extern "C" void shape_hip_clear_error(void) { g_err = 0; }
extern "C" void shape_hip_set_strict(int write_pos) { g_strict_pos = write_pos; }

extern "C" int shape_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// ---------------------------------------------------------------------------
// calc_fits (calc_fits.c:340): delay-Doppler, Doppler and lightcurve fits for
// every dataset from the already-realized model and spin state
// ---------------------------------------------------------------------------
extern "C" void calc_fits(struct par_t *par, struct mod_t *mod, struct dat_t *dat) {
  Engine E;
  RunOptions o;
  o.method = METHOD_FP64;
  o.host_spin = 1;
  o.host_geometry = 1;
  o.keep_fit = 1;
  o.with_body = 0;
  std::vector<EvalSpec> ev(1);
  ev[0].model = 0;
  int rc = E.build(par, &mod, &dat, 1, ev, o);
  if (!rc) rc = E.upload_params();
  if (!rc) rc = E.enqueue();
  if (!rc) rc = E.sync();
  if (!rc && E.error_flags()) rc = -30;
  g_last_nstale = E.last_nstale;
  if (!rc) rc = E.writeback(0, par, mod, dat, g_strict_pos);
  set_err(rc, "calc_fits");
}

// number of POS pixels whose cos(scattering) was inherited from an earlier
// claimant in the last calc_fits (the replayed pixels); diagnostic only
extern "C" int shape_hip_last_replayed(void) { return g_last_nstale; }

// ---------------------------------------------------------------------------
// chi2 (chi2.c:240): chi-square of the fits currently held by the caller
// ---------------------------------------------------------------------------
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  int get(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 8) == hipSuccess ? 0 : -2; }
};

extern "C" double chi2(struct par_t *par, struct dat_t *dat, int list_breakdown) {
  (void)par;
  (void)list_breakdown;
  // frame records first (offsets into flat arrays), then the fit,
  // observation and weight arrays copied from the frames' row arrays in
  // parallel
  std::vector<FrameRec> fr;
  std::vector<const void *> src;  // the frame (deldopfrm_t, dopfrm_t) or lightcurve (lghtcrv_t)
  long long ncell = 0;
  auto add = [&](FrameRec &F, int s, long long N, const void *h) {
    F.set = s;
    F.nviews = 1;
    F.off = ncell;
    F.obs_off = ncell;
    ncell += N;
    fr.push_back(F);
    src.push_back(h);
  };
  for (int s = 0; s < dat->nsets; s++) {
    struct set_t *set = &dat->set[s];
    if (set->type == DELAY) {
      for (int f = 0; f < set->desc.deldop.nframes; f++) {
        struct deldopfrm_t *h = &set->desc.deldop.frame[f];
        FrameRec F;
        memset(&F, 0, sizeof(F));
        F.type = VT_DELAY; F.frm = f; F.ndel = h->ndel; F.ndop = h->ndop;
        F.cal_floats = h->cal.state == 'f'; F.cal_fixed = h->cal.val; F.weight = h->weight;
        F.overflow_o2 = h->overflow_o2; F.overflow_m2 = h->overflow_m2;
        add(F, s, (long long)h->ndel * h->ndop, h);
      }
    } else if (set->type == DOPPLER) {
      for (int f = 0; f < set->desc.doppler.nframes; f++) {
        struct dopfrm_t *h = &set->desc.doppler.frame[f];
        FrameRec F;
        memset(&F, 0, sizeof(F));
        F.type = VT_DOPPLER; F.frm = f; F.ndel = 1; F.ndop = h->ndop;
        F.cal_floats = h->cal.state == 'f'; F.cal_fixed = h->cal.val; F.weight = h->weight;
        F.overflow_o2 = h->overflow_o2; F.overflow_m2 = h->overflow_m2;
        add(F, s, h->ndop, h);
      }
    } else if (set->type == LGHTCRV) {
      struct lghtcrv_t *L = &set->desc.lghtcrv;
      FrameRec F;
      memset(&F, 0, sizeof(F));
      F.type = VT_LGHTCRV; F.lc_n = L->n;
      F.cal_floats = L->cal.state == 'f'; F.cal_fixed = L->cal.val; F.weight = L->weight;
      add(F, s, L->n, L);
    } else {
      set_err(-15, "chi2 (plane-of-sky sets are outside this library)");
      return dat->chi2;
    }
  }
  // fit, obs and 1/var back to back in the thread's pinned stage (plain
  // host memory if it cannot be pinned)
  std::vector<double> plain;
  double *fit = host_stage(3 * (size_t)ncell);
  if (!fit) {
    plain.resize(3 * (size_t)ncell);
    fit = plain.data();
  }
  double *obs = fit + ncell, *oov = obs + ncell;
  host_parallel_for((int)fr.size(), [&](int k) {
    const FrameRec &F = fr[k];
    double *a = fit + F.off, *o = obs + F.off, *w = oov + F.off;
    if (F.type == VT_DELAY) {
      const struct deldopfrm_t *h = (const struct deldopfrm_t *)src[k];
      for (int i = 1; i <= h->ndel; i++)
        for (int j = 1; j <= h->ndop; j++) { *a++ = h->fit[i][j]; *o++ = h->obs[i][j]; *w++ = h->oneovervar[i][j]; }
    } else if (F.type == VT_DOPPLER) {
      const struct dopfrm_t *h = (const struct dopfrm_t *)src[k];
      for (int j = 1; j <= h->ndop; j++) { *a++ = h->fit[j]; *o++ = h->obs[j]; *w++ = h->oneovervar[j]; }
    } else {
      const struct lghtcrv_t *L = (const struct lghtcrv_t *)src[k];
      for (int i = 1; i <= L->n; i++) { *a++ = L->fit[i]; *o++ = L->obs[i]; *w++ = L->oneovervar[i]; }
    }
  });
  int nf = (int)fr.size(), nsets = dat->nsets;
  DevBuf dfr, dfit, dobs, doov, dset, dtot;
  const size_t nb = sizeof(double) * (size_t)ncell;
  int rc = dfr.get(sizeof(FrameRec) * nf) | dfit.get(nb) | dobs.get(nb) | doov.get(nb) | dset.get(sizeof(double) * nsets) |
           dtot.get(sizeof(double));
  if (rc) { set_err(rc, "chi2 allocation"); return dat->chi2; }
  bool ok = hipMemcpy(dfr.p, fr.data(), sizeof(FrameRec) * nf, hipMemcpyHostToDevice) == hipSuccess &&
            (ncell == 0 || (hipMemcpy(dfit.p, fit, nb, hipMemcpyHostToDevice) == hipSuccess &&
                            hipMemcpy(dobs.p, obs, nb, hipMemcpyHostToDevice) == hipSuccess &&
                            hipMemcpy(doov.p, oov, nb, hipMemcpyHostToDevice) == hipSuccess));
  if (!ok) { set_err(-3, "chi2 upload"); return dat->chi2; }
  FrameRec *df = (FrameRec *)dfr.p;
  hipLaunchKernelGGL(k_frame_o2, dim3(std::max(1, nf)), dim3(256), 0, 0, df, nf, (double *)dobs.p, (double *)doov.p);
  hipLaunchKernelGGL(k_chi2_frames64, dim3(std::max(1, nf)), dim3(256), 0, 0, df, nf, (double *)dfit.p,
                     (double *)dobs.p, (double *)doov.p, (const double *)nullptr, 0, 1, 0);
  hipLaunchKernelGGL(k_chi2_sets, dim3(1), dim3(64), 0, 0, (const FrameRec *)df, nf, (double *)dset.p, nsets,
                     (double *)dtot.p);
  std::vector<double> setc(nsets);
  double total = 0.0;
  ok = hipDeviceSynchronize() == hipSuccess &&
       hipMemcpy(fr.data(), dfr.p, sizeof(FrameRec) * nf, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(setc.data(), dset.p, sizeof(double) * nsets, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(&total, dtot.p, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) { set_err(-4, "chi2 kernels"); return dat->chi2; }
  for (auto &F : fr) {
    struct set_t *set = &dat->set[F.set];
    if (F.type == VT_DELAY) { set->desc.deldop.frame[F.frm].chi2 = F.chi2; set->desc.deldop.frame[F.frm].cal.val = F.cal; }
    else if (F.type == VT_DOPPLER) { set->desc.doppler.frame[F.frm].chi2 = F.chi2; set->desc.doppler.frame[F.frm].cal.val = F.cal; }
    else set->desc.lghtcrv.cal.val = F.cal;
  }
  for (int s = 0; s < nsets; s++) dat->set[s].chi2 = setc[s];
  dat->chi2 = total;
  return total;
}

// ---------------------------------------------------------------------------
// realize_mod (realize_mod.c:153): vertex coordinates, facet normals, areas,
// centroids and vertex normals of a single-component VERTEX model
// ---------------------------------------------------------------------------
extern "C" void realize_mod(struct par_t *par, struct mod_t *mod) {
  if (mod->shape.ncomp != 1 || mod->shape.comp[0].type != VERTEX) { set_err(-10, "realize_mod"); return; }
  struct comp_t *comp = &mod->shape.comp[0];
  struct vertices_t *rv = &comp->real;
  int nv = rv->nv, nf = rv->nf;
  std::vector<int> fv(3 * (size_t)nf), af_off(nv + 1, 0), af_idx;
  std::vector<double> u(3 * (size_t)nv), a(3 * (size_t)nv), r(nv), pc(9);
  for (int f = 0; f < nf; f++)
    for (int k = 0; k < 3; k++) fv[3 * f + k] = rv->f[f].v[k];
  for (int i = 0; i < nv; i++) {
    struct vertex_t *v = &rv->v[i];
    if (v->r.state == '=') v->r.val = rv->v[v->v_mirror].r.val;  // realize_mod.c:450-454
    af_off[i + 1] = af_off[i] + v->naf;
    for (int q = 0; q < v->naf; q++) af_idx.push_back(v->af[q]);
    for (int k = 0; k < 3; k++) { u[3 * i + k] = v->u[k]; a[3 * i + k] = v->a[k]; }
    r[i] = v->r.val;
  }
  for (int j = 0; j < 3; j++) {
    if (j > 0 && comp->desc.ver.scalefactor[j].state == '=')
      comp->desc.ver.scalefactor[j].val = comp->desc.ver.scalefactor[j - 1].val;
    pc[j] = comp->desc.ver.scalefactor[j].val;
    pc[3 + j] = comp->rot[j].val;
    pc[6 + j] = comp->off[j].val;
  }
  DevBuf dfv, doff, didx, du, da, dr, dpc, dx, dvn, dfn, dfx, dar, dx32, dvn32, dfn32, dslot, dflag, dlogf;
  int rc = dfv.get(sizeof(int) * fv.size()) | doff.get(sizeof(int) * af_off.size()) |
           didx.get(sizeof(int) * af_idx.size()) | du.get(sizeof(double) * u.size()) | da.get(sizeof(double) * a.size()) |
           dr.get(sizeof(double) * nv) | dpc.get(sizeof(double) * 9) | dx.get(sizeof(double) * 3 * nv) |
           dvn.get(sizeof(double) * 3 * nv) | dfn.get(sizeof(double) * 3 * nf) | dfx.get(sizeof(double) * 3 * nf) |
           dar.get(sizeof(double) * nf) | dx32.get(sizeof(float4) * nv) | dvn32.get(sizeof(float4) * nv) |
           dfn32.get(sizeof(float4) * nf) | dslot.get(sizeof(GeomSlot)) | dflag.get(sizeof(int)) | dlogf.get(sizeof(double));
  if (rc) { set_err(rc, "realize_mod allocation"); return; }
  GeomSlot G;
  memset(&G, 0, sizeof(G));
  G.nv = nv; G.nf = nf; G.comp_pb = 0; G.eval = -1;
  G.fv = (int *)dfv.p; G.af_off = (int *)doff.p; G.af_idx = (int *)didx.p; G.u = (double *)du.p; G.a = (double *)da.p;
  G.r = (double *)dr.p; G.x = (double *)dx.p; G.vn = (double *)dvn.p; G.fn = (double *)dfn.p; G.fx = (double *)dfx.p;
  G.area = (double *)dar.p; G.x32 = (float4 *)dx32.p; G.vn32 = (float4 *)dvn32.p; G.fn32 = (float4 *)dfn32.p;
  bool ok = hipMemcpy(dfv.p, fv.data(), sizeof(int) * fv.size(), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(doff.p, af_off.data(), sizeof(int) * af_off.size(), hipMemcpyHostToDevice) == hipSuccess &&
            (af_idx.empty() || hipMemcpy(didx.p, af_idx.data(), sizeof(int) * af_idx.size(), hipMemcpyHostToDevice) == hipSuccess) &&
            hipMemcpy(du.p, u.data(), sizeof(double) * u.size(), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(da.p, a.data(), sizeof(double) * a.size(), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dr.p, r.data(), sizeof(double) * nv, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dpc.p, pc.data(), sizeof(double) * 9, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dslot.p, &G, sizeof(G), hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) { set_err(-3, "realize_mod upload"); return; }
  GeomSlot *ds = (GeomSlot *)dslot.p;
  hipLaunchKernelGGL(k_realize_vertices, dim3((nv + 255) / 256, 1), dim3(256), 0, 0, ds, 1, (const double *)dpc.p,
                     (int *)dflag.p, (double *)dlogf.p);
  hipLaunchKernelGGL(k_realize_facets, dim3((nf + 255) / 256, 1), dim3(256), 0, 0, ds, 1);
  hipLaunchKernelGGL(k_realize_vnormals, dim3((nv + 255) / 256, 1), dim3(256), 0, 0, ds, 1);
  std::vector<double> x(3 * (size_t)nv), vn(3 * (size_t)nv), fn(3 * (size_t)nf), fx(3 * (size_t)nf), ar(nf);
  int flag = 0;
  double logf = 0.0;
  ok = hipDeviceSynchronize() == hipSuccess &&
       hipMemcpy(x.data(), dx.p, sizeof(double) * x.size(), hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(vn.data(), dvn.p, sizeof(double) * vn.size(), hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(fn.data(), dfn.p, sizeof(double) * fn.size(), hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(fx.data(), dfx.p, sizeof(double) * fx.size(), hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(ar.data(), dar.p, sizeof(double) * nf, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(&flag, dflag.p, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(&logf, dlogf.p, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) { set_err(-4, "realize_mod kernels"); return; }
  for (int i = 0; i < nv; i++)
    for (int k = 0; k < 3; k++) { rv->v[i].x[k] = x[3 * i + k]; rv->v[i].n[k] = vn[3 * i + k]; }
  for (int f = 0; f < nf; f++) {
    for (int k = 0; k < 3; k++) { rv->f[f].n[k] = fn[3 * f + k]; rv->f[f].x[k] = fx[3 * f + k]; }
    rv->f[f].area = ar[f];
  }
  for (int j = 0; j < 3; j++) {
    double s = pc[j];
    if (s <= 0.01) s = 0.01 / (1 + 0.01 - s);
    rv->scalefactor[j].val = s;
  }
  par->baddiam = (unsigned char)flag;
  par->baddiam_logfactor = logf;
}

// ---------------------------------------------------------------------------
// realize_spin (realize_spin.c:41) and realize_delcor (realize_delcor.c:50):
// per-view spin state and COM offsets, from the setup kernel
// ---------------------------------------------------------------------------
static int views_pass(struct par_t *par, struct mod_t *mod, struct dat_t *dat, int want_spin, int want_off) {
  Engine E;
  RunOptions o;
  o.method = METHOD_FP64;
  o.host_spin = 0;
  o.views_only = 1;
  std::vector<EvalSpec> ev(1);
  ev[0].model = 0;
  int rc = E.build(par, &mod, &dat, 1, ev, o);
  if (rc) return rc;
  hipLaunchKernelGGL(k_setup_views, dim3((E.nviews + 63) / 64), dim3(64), 0, E.stream, E.d_views, E.nviews, E.d_params);
  std::vector<ViewRec> vw;
  if ((rc = E.download_views(vw))) return rc;
  for (const ViewRec &V : vw) {
    struct set_t *set = &dat->set[V.set];
    double (*ae)[3] = nullptr;
    double *spin = nullptr, *intspin = nullptr, *deloff = nullptr, *dopoff = nullptr;
    if (V.type == VT_DELAY) {
      struct deldopview_t *w = &set->desc.deldop.frame[V.frm].view[V.v];
      ae = w->ae; spin = w->spin; intspin = w->intspin; deloff = &w->deloff; dopoff = &w->dopoff;
    } else if (V.type == VT_DOPPLER) {
      struct dopview_t *w = &set->desc.doppler.frame[V.frm].view[V.v];
      ae = w->ae; spin = w->spin; intspin = w->intspin; dopoff = &w->dopoff;
    } else {
      struct crvrend_t *w = &set->desc.lghtcrv.rend[V.frm + 1];
      ae = w->ae; spin = w->spin; intspin = w->intspin;
    }
    if (want_spin) {
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) ae[i][j] = V.ae[3 * i + j];
      for (int j = 0; j < 3; j++) { spin[j] = V.spin[j]; intspin[j] = V.intspin[j]; }
    }
    if (want_off) {
      if (deloff) *deloff = V.deloff;
      if (dopoff) *dopoff = V.dopoff;
    }
  }
  return 0;
}

extern "C" void realize_spin(struct par_t *par, struct mod_t *mod, struct dat_t *dat) {
  // '=' states of the per-set angle and spin offsets (realize_spin.c:283-342)
  for (int j = 0; j <= 2; j++) {
    int sa = -1, so = -1;
    for (int s = 0; s < dat->nsets; s++) {
      if (dat->set[s].angleoff[j].state != '=') sa = s;
      else if (sa >= 0) dat->set[s].angleoff[j].val = dat->set[sa].angleoff[j].val;
      if (dat->set[s].omegaoff[j].state != '=') so = s;
      else if (so >= 0) dat->set[s].omegaoff[j].val = dat->set[so].omegaoff[j].val;
    }
  }
  set_err(views_pass(par, mod, dat, 1, 0), "realize_spin");
}

// realize_delcor (realize_delcor.c:50) takes only the data; the offsets do
// not depend on the model, so the views are packed against a placeholder
// model (no facets, principal-axis spin, cosine/Lambert laws) whose values
// the offset computation never reads.
extern "C" void realize_delcor(struct dat_t *dat, double delta_delcor0, int delcor0_mode) {
  int nrad = 1, nopt = 1;
  for (int s = 0; s < dat->nsets; s++) {
    struct set_t *set = &dat->set[s];
    if (set->type == DELAY) nrad = std::max(nrad, set->desc.deldop.iradlaw + 1);
    if (set->type == DOPPLER) nrad = std::max(nrad, set->desc.doppler.iradlaw + 1);
    if (set->type == LGHTCRV) nopt = std::max(nopt, set->desc.lghtcrv.ioptlaw + 1);
  }
  struct mod_t m;
  memset(&m, 0, sizeof(m));
  struct comp_t c;
  memset(&c, 0, sizeof(c));
  c.type = VERTEX;
  m.shape.ncomp = 1;
  m.shape.comp = &c;
  m.spin.pa = 1;
  std::vector<unsigned char> rt(nrad, COSINELAW_DIFF), ot(nopt, LAMBERTLAW);
  std::vector<union radscat_t> rl(nrad);
  std::vector<union optscat_t> ol(nopt);
  memset(rl.data(), 0, sizeof(union radscat_t) * nrad);
  memset(ol.data(), 0, sizeof(union optscat_t) * nopt);
  m.photo.nradlaws = nrad; m.photo.radtype = rt.data(); m.photo.radar = rl.data();
  m.photo.noptlaws = nopt; m.photo.opttype = ot.data(); m.photo.optical = ol.data();
  struct par_t p;
  memset(&p, 0, sizeof(p));
  p.nsinc2 = 1;
  p.sinc2width = 6;
  p.exclude_seen = -1;
  shape_hip_realize_delcor(&p, &m, dat, delta_delcor0, delcor0_mode);
}

extern "C" void shape_hip_realize_delcor(struct par_t *par, struct mod_t *mod, struct dat_t *dat,
                                         double delta_delcor0, int delcor0_mode) {
  // polynomial bookkeeping (realize_delcor.c:67-140): '=' states and delcor0 modes
  int s_dc = -1, type_dc = -1;
  for (int s = 0; s < dat->nsets; s++) {
    struct set_t *set = &dat->set[s];
    if (set->type != DELAY && set->type != DOPPLER) continue;
    struct delcor_t *dc = set->type == DELAY ? &set->desc.deldop.delcor : &set->desc.doppler.delcor;
    if (dc->a[0].state != '=') {
      s_dc = s;
      type_dc = set->type;
      if (set->type == DELAY && dc->a[0].state == 'f') {
        if (delcor0_mode != 0) dc->a[0].val = dc->delcor0_save + delta_delcor0;
        if (delcor0_mode != 1) dc->delcor0_save = dc->a[0].val;
      }
    } else if (s_dc >= 0) {
      struct set_t *src = &dat->set[s_dc];
      struct delcor_t *sdc = type_dc == DELAY ? &src->desc.deldop.delcor : &src->desc.doppler.delcor;
      dc->t0 = sdc->t0;
      for (int i = 0; i <= dc->n; i++) dc->a[i].val = sdc->a[i].val;
    }
  }
  set_err(views_pass(par, mod, dat, 0, 1), "realize_delcor");
}
