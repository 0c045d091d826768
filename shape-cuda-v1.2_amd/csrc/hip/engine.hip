// engine.hip -- packing of SHAPE ABI structs into device arenas and the
// launch sequence of one evaluation batch.
#include <math.h>
#include <string.h>

#include <algorithm>

#include "common.h"
#include "engine_host.h"
#include "mixed.h"

extern "C" int shp_fail(hipError_t e) { return (int)e == 0 ? -1 : (int)e; }

namespace shp {

Engine::~Engine() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (hipEvent_t e : t_events) (void)hipEventDestroy(e);
  for (void *p : allocs) (void)hipFree(p);
  if (h_pin) (void)hipHostFree(h_pin);
  if (stream) (void)hipStreamDestroy(stream);
}

int Engine::alloc(void **p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  SHP_CHECK(hipMalloc(p, bytes));
  allocs.push_back(*p);
  dev_bytes += bytes;
  return 0;
}

template <class T>
static int upload(Engine *E, T **dptr, const std::vector<T> &h) {
  void *p = nullptr;
  int rc = 0;
  size_t bytes = h.size() * sizeof(T);
  if (bytes == 0) bytes = sizeof(T);
  if (hipMalloc(&p, bytes) != hipSuccess) return -2;
  E->allocs.push_back(p);
  E->dev_bytes += bytes;
  if (!h.empty() && hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) rc = -3;
  *dptr = (T *)p;
  return rc;
}

// ---------------------------------------------------------------------------
// per-model packing state
// ---------------------------------------------------------------------------
struct ModelPack {
  std::vector<double> base;  // parameter block
  int comp_off = 0;
  // topology / inputs (host)
  std::vector<int> fv, af_off, af_idx;
  std::vector<double> u, a, r;
  std::vector<int4> sides;  // v0, v1, f0, f1 (penalties)
  std::vector<double> hbasis;  // harmonic factors per vertex, in the coefficients' order
  int comp_type = VERTEX, nhar = 0, sf_tie = 0;
  int nv = 0, nf = 0;
  int set_off = 0, nsets = 0;  // the datasets' blocks in the parameter block
  // realized geometry from the host (drop-in mode)
  std::vector<double> hx, hvn, hfn, hfx;
  // scattering laws as realize_photo sees them: the job template of each
  // law, its parameter-block data and its share of the evaluation's
  // resolved values (loc_len doubles per evaluation in all)
  struct LawPack {
    int pb_law;          // the law's header slots in the parameter block
    int has_loc;         // the views of this law read resolved values
    PhotoJob job;        // data / loc / chk relative to the block / the evaluation
    int data_len, loc_len;
  };
  std::vector<LawPack> laws;
  long long loc_len = 0;
  int photo_off = 0;  // start of the laws' data regions (after the component block)
  // device topology
  int *d_fv = nullptr, *d_af_off = nullptr, *d_af_idx = nullptr;
  double *d_u = nullptr, *d_a = nullptr;
  int4 *d_sides = nullptr;
  double *d_hbasis = nullptr;
  // view / frame templates
  std::vector<ViewRec> views;
  std::vector<FrameRec> frames;
  std::vector<int> frame_first_view;  // template view index of each frame's first view
  long long obs_off = 0;              // into the shared obs arena
  // observations and weights to copy (after the pack, in parallel): the
  // source frame and its offset in the model's observations
  struct ObsJob {
    int kind;         // VT_DELAY, VT_DOPPLER, VT_LGHTCRV
    const void *src;  // deldopfrm_t, dopfrm_t or lghtcrv_t
    long long off, n;
  };
  std::vector<ObsJob> obs_jobs;
  long long nobs = 0;
  const struct vertex_t *verts = nullptr;  // the model's vertices: deviation targets by address
  std::vector<double> lc_x, lc_t;
  long long lc_x_off = 0, lc_t_off = 0;
  double dof = 0, dof_dd = 0, dof_dop = 0, dof_lc = 0;
};

// Packing scratch, per host thread: the one-dataset drop-ins (calc_deldop
// and co.) are called from several host threads at once.
static thread_local std::vector<ModelPack> g_packs;

namespace {
struct HostStage {
  double *p = nullptr;
  size_t n = 0;
  ~HostStage() {
    if (p) (void)hipHostFree(p);
  }
};
thread_local HostStage t_stage;
}  // namespace

double *host_stage(size_t n) {
  if (n <= t_stage.n) return t_stage.p;
  if (t_stage.p) (void)hipHostFree(t_stage.p);
  t_stage.p = nullptr;
  t_stage.n = 0;
  size_t want = std::max(n, (size_t)1 << 20);
  if (hipHostMalloc((void **)&t_stage.p, sizeof(double) * want, hipHostMallocDefault) != hipSuccess) {
    t_stage.p = nullptr;
    return nullptr;
  }
  t_stage.n = want;
  return t_stage.p;
}

static void set_param_target(std::map<const double *, ParamTarget> &m, const double *addr, int model, int off) {
  m[addr] = ParamTarget{0, model, off};
}

// repack: only what changes between calls of a drop-in (parameters, views,
// frame scalars, realized geometry); the topology stays as built.
int Engine::pack_model(int mi, struct par_t *par, struct mod_t *mod, struct dat_t *dat, bool repack) {
  ModelPack &P = g_packs[mi];
  auto &tmap = model_params[mi];
  if (mod->shape.ncomp != 1) return -10;  // the engine handles single-component models
  struct comp_t *comp = &mod->shape.comp[0];
  struct vertices_t *rv = &comp->real;
  // ---- parameter block layout ----
  int nrad = mod->photo.nradlaws, nopt = mod->photo.noptlaws, nsets = dat->nsets;
  int off_rad = PB_SPIN_END, off_opt = off_rad + PB_RAD_STRIDE * nrad;
  int off_set = off_opt + PB_OPT_STRIDE * nopt, off_comp = off_set + PB_SET_STRIDE * nsets;
  const int ctype = comp->type;
  if (ctype != VERTEX && ctype != ELLIPSE && ctype != OVOID && ctype != HARMONIC) return -10;
  const int nhar = ctype == HARMONIC ? comp->desc.har.nhar : 0;
  const int nbasis = ctype == HARMONIC ? (nhar + 1) * (nhar + 1) : 0;
  // ---- scattering laws (realize_photo): job templates and data regions ----
  const int nf_law = rv->nf;
  int off_photo = off_comp + PB_COMP_STRIDE + nbasis, photo_len = 0;
  P.laws.clear();
  P.loc_len = 0;
  auto add_law = [&](int radar, int l) {
    ModelPack::LawPack L;
    memset(&L, 0, sizeof(L));
    PhotoJob &J = L.job;
    const int t = radar ? mod->photo.radtype[l] : mod->photo.opttype[l];
    L.pb_law = radar ? off_rad + PB_RAD_STRIDE * l : off_opt + PB_OPT_STRIDE * l;
    J.type = t;
    J.radar = radar;
    auto bounds = [&](int j, double lo, double hi, int mode) { J.lo[j] = lo; J.hi[j] = hi; J.mode[j] = mode; };
    auto chk = [&](int slot, double lo, double hi, int mode) {
      bounds(J.nchk, lo, hi, mode);
      J.chk[J.nchk++] = L.pb_law + slot;
    };
    if (t == NOLAW) return;
    if (radar) {
      J.k = 2;
      bounds(0, par->rad_R_min, par->rad_R_max, 3);
      bounds(1, par->rad_C_min, par->rad_C_max, 3);
      switch (t) {
      case COSINELAW_DIFF: case GAUSSIANLAW: case HAGFORSLAW: case COSINELAW_QS:
        J.kind = PHOTO_HOMOG; J.nchk = 0;
        chk(1, par->rad_R_min, par->rad_R_max, 3);
        chk(2, par->rad_C_min, par->rad_C_max, 3);
        break;
      case GAUSSIAN_COSINE: case HAGFORS_COSINE: case COSINE_COSINE:  // qs R, qs C, diffuse R, C
        J.kind = PHOTO_HOMOG; J.nchk = 0;
        chk(4, par->rad_R_min, par->rad_R_max, 3);
        chk(5, par->rad_C_min, par->rad_C_max, 3);
        chk(1, par->rad_R_min, par->rad_R_max, 3);
        chk(2, par->rad_C_min, par->rad_C_max, 3);
        break;
      case TABULARLAW:
        J.kind = PHOTO_TABLE; J.k = 1; J.n = mod->photo.radar[l].tabular.n;
        bounds(0, par->rad_rho_min, par->rad_rho_max, 3);
        L.data_len = 2 * J.n; L.loc_len = J.n;
        break;
      case HARMCOSINE_DIFF: {
        J.kind = PHOTO_HARM; J.n = nf_law;
        struct harmonic_t *h = &mod->photo.radar[l].harmcosine.R;
        for (int j = 0; j < 2; j++) {
          J.nhar[j] = h[j].nhar; J.coef[j] = L.data_len;
          L.data_len += (h[j].nhar + 1) * (h[j].nhar + 1);
        }
        L.loc_len = 2 * nf_law;
        break;
      }
      case INHOCOSINE_DIFF:
        J.kind = PHOTO_INHO; J.n = nf_law;
        L.data_len = 2 + 4 * nf_law; L.loc_len = 2 * nf_law;
        break;
      default:
        return;  // not evaluated by the engine; refused if a set uses it
      }
    } else {
      const int hk = (t == HAPKE || t == HARMHAPKE || t == INHOHAPKE);
      const int kk = (t == KAASALAINEN || t == HARMKAAS || t == INHOKAAS);
      J.k = (hk || kk) ? 5 : 1;
      if (hk) {
        bounds(0, par->opt_w_min, par->opt_w_max, 3);
        bounds(1, par->opt_h_min, par->opt_h_max, 3);
        bounds(2, par->opt_B0_min, par->opt_B0_max, 2);
        bounds(3, par->opt_g_min, par->opt_g_max, 3);
        bounds(4, par->opt_theta_min, par->opt_theta_max, 2);
      } else if (kk) {
        bounds(0, par->opt_R_min, par->opt_R_max, 3);
        bounds(1, par->opt_wt_min, par->opt_wt_max, 0);
        bounds(2, par->opt_A0_min, par->opt_A0_max, 2);
        bounds(3, par->opt_D_min, par->opt_D_max, 3);
        bounds(4, par->opt_k_min, par->opt_k_max, 1);
      } else {
        bounds(0, par->opt_R_min, par->opt_R_max, 3);
      }
      if (t == GEOMETRICAL || t == LAMBERTLAW || t == LOMMEL || t == HAPKE || t == KAASALAINEN) {
        J.kind = PHOTO_HOMOG;
        const PhotoJob b = J;
        J.nchk = 0;
        for (int j = 0; j < b.k; j++) chk(1 + j, b.lo[j], b.hi[j], b.mode[j]);
      } else if (t == HARMLAMBERT || t == HARMLOMMEL || t == HARMHAPKE || t == HARMKAAS) {
        J.kind = PHOTO_HARM; J.n = nf_law;
        union optscat_t *o = &mod->photo.optical[l];
        struct harmonic_t *h = hk ? &o->harmhapke.w : kk ? &o->harmkaas.R : &o->harmR.R;
        for (int j = 0; j < J.k; j++) {
          J.nhar[j] = h[j].nhar; J.coef[j] = L.data_len;
          L.data_len += (h[j].nhar + 1) * (h[j].nhar + 1);
        }
        L.loc_len = J.k * nf_law;
      } else if (t == INHOLAMBERT || t == INHOLOMMEL || t == INHOHAPKE || t == INHOKAAS) {
        J.kind = PHOTO_INHO; J.n = nf_law;
        L.data_len = J.k + 2 * J.k * nf_law; L.loc_len = J.k * nf_law;
      } else {
        return;
      }
    }
    L.has_loc = L.loc_len > 0;
    J.data = off_photo + photo_len;  // relative to the block until the evaluation's is known
    J.loc = P.loc_len;               // relative to the evaluation's resolved values
    photo_len += L.data_len;
    P.loc_len += L.loc_len;
    P.laws.push_back(L);
  };
  for (int l = 0; l < nopt; l++) add_law(0, l);  // realize_photo's order: optical, then radar
  for (int l = 0; l < nrad; l++) add_law(1, l);
  int size = off_photo + photo_len;
  P.photo_off = off_photo;
  P.base.assign(size, 0.0);
  P.comp_off = off_comp;
  P.set_off = off_set;
  P.nsets = nsets;
  double *b = P.base.data();
  struct spin_t *sp = &mod->spin;
  for (int j = 0; j < 3; j++) {
    b[PB_ANGLE + j] = sp->angle[j].val; set_param_target(tmap, &sp->angle[j].val, mi, PB_ANGLE + j);
    b[PB_OMEGA + j] = sp->omega[j].val; set_param_target(tmap, &sp->omega[j].val, mi, PB_OMEGA + j);
    b[PB_OMEGADOT + j] = sp->omegadot[j].val; set_param_target(tmap, &sp->omegadot[j].val, mi, PB_OMEGADOT + j);
  }
  b[PB_T0] = sp->t0;
  b[PB_LIBAMP] = sp->lib_amp.val; set_param_target(tmap, &sp->lib_amp.val, mi, PB_LIBAMP);
  b[PB_LIBFREQ] = sp->lib_freq.val; set_param_target(tmap, &sp->lib_freq.val, mi, PB_LIBFREQ);
  b[PB_LIBPHASE] = sp->lib_phase.val; set_param_target(tmap, &sp->lib_phase.val, mi, PB_LIBPHASE);
  // 1: principal-axis kinematics; NPA by Euler's equations, 0 periodic / -1 brute
  b[PB_PA] = sp->pa ? 1.0 : (par->int_method == BRUTE ? -1.0 : 0.0);
  for (int j = 0; j < 3; j++) {
    b[PB_INERTIA + j] = sp->inertia[j].val;
    set_param_target(tmap, &sp->inertia[j].val, mi, PB_INERTIA + j);
  }
  b[PB_NIMP] = sp->n_impulse;
  for (int q = 0; q < sp->n_impulse && q < MAXIMP; q++) {
    b[PB_TIMP + q] = sp->t_impulse[q];
    for (int j = 0; j < 3; j++) {
      b[PB_IMP + 3 * q + j] = sp->impulse[q][j].val;
      set_param_target(tmap, &sp->impulse[q][j].val, mi, PB_IMP + 3 * q + j);
    }
  }
  for (int l = 0; l < nrad; l++) {
    double *pl = b + off_rad + PB_RAD_STRIDE * l;
    union radscat_t *r = &mod->photo.radar[l];
    int t = mod->photo.radtype[l];
    pl[0] = t;
    switch (t) {
    case COSINELAW_DIFF:
      pl[1] = r->RC.R.val; pl[2] = r->RC.C.val;
      set_param_target(tmap, &r->RC.R.val, mi, off_rad + PB_RAD_STRIDE * l + 1);
      set_param_target(tmap, &r->RC.C.val, mi, off_rad + PB_RAD_STRIDE * l + 2);
      break;
    case GAUSSIANLAW: case HAGFORSLAW: case COSINELAW_QS:
      pl[1] = r->quasispec.R.val; pl[2] = r->quasispec.C.val; pl[3] = r->quasispec.cos_cutoff;
      set_param_target(tmap, &r->quasispec.R.val, mi, off_rad + PB_RAD_STRIDE * l + 1);
      set_param_target(tmap, &r->quasispec.C.val, mi, off_rad + PB_RAD_STRIDE * l + 2);
      break;
    case GAUSSIAN_COSINE: case HAGFORS_COSINE: case COSINE_COSINE:
      pl[1] = r->hybrid.diff.R.val; pl[2] = r->hybrid.diff.C.val;
      pl[4] = r->hybrid.qs.R.val; pl[5] = r->hybrid.qs.C.val; pl[6] = r->hybrid.qs.cos_cutoff;
      set_param_target(tmap, &r->hybrid.diff.R.val, mi, off_rad + PB_RAD_STRIDE * l + 1);
      set_param_target(tmap, &r->hybrid.diff.C.val, mi, off_rad + PB_RAD_STRIDE * l + 2);
      set_param_target(tmap, &r->hybrid.qs.R.val, mi, off_rad + PB_RAD_STRIDE * l + 4);
      set_param_target(tmap, &r->hybrid.qs.C.val, mi, off_rad + PB_RAD_STRIDE * l + 5);
      break;
    case TABULARLAW:
      pl[2] = r->tabular.n;
      break;
    case HARMCOSINE_DIFF: case INHOCOSINE_DIFF:
      break;
    default:
      pl[0] = -1;  // not evaluated by the engine; rejected below if a set uses it
    }
  }
  for (int l = 0; l < nopt; l++) {
    double *po = b + off_opt + PB_OPT_STRIDE * l;
    union optscat_t *o = &mod->photo.optical[l];
    int t = mod->photo.opttype[l];
    po[0] = t;
    int base = off_opt + PB_OPT_STRIDE * l;
    if (t == LAMBERTLAW || t == LOMMEL || t == GEOMETRICAL) {
      po[1] = o->R.R.val;
      set_param_target(tmap, &o->R.R.val, mi, base + 1);
    } else if (t == KAASALAINEN) {
      po[1] = o->kaas.R.val; po[2] = o->kaas.wt.val; po[3] = o->kaas.A0.val; po[4] = o->kaas.D.val; po[5] = o->kaas.k.val;
      set_param_target(tmap, &o->kaas.R.val, mi, base + 1);
      set_param_target(tmap, &o->kaas.wt.val, mi, base + 2);
      set_param_target(tmap, &o->kaas.A0.val, mi, base + 3);
      set_param_target(tmap, &o->kaas.D.val, mi, base + 4);
      set_param_target(tmap, &o->kaas.k.val, mi, base + 5);
    } else if (t == HAPKE) {
      struct param_t *q = &o->hapke.w;
      for (int j = 0; j < 5; j++) {
        po[1 + j] = q[j].val;
        set_param_target(tmap, &q[j].val, mi, base + 1 + j);
      }
    } else if (t < HARMLAMBERT || t > INHOKAAS) {
      po[0] = -1;
    }
  }
  // the laws' data regions: expansion coefficients in the model file's order;
  // an inhomogeneous law's k global values, then k per facet, then k tie flags
  // per facet ('=' state); a tabular law's values, then its tie flags
  for (auto &L : P.laws) {
    const PhotoJob &J = L.job;
    const int d0 = (int)J.data;
    double *d = b + d0;
    const bool radar = L.pb_law < off_opt;
    const int l = radar ? (L.pb_law - off_rad) / PB_RAD_STRIDE : (L.pb_law - off_opt) / PB_OPT_STRIDE;
    auto put = [&](int q, struct param_t *p) {
      d[q] = p->val;
      set_param_target(tmap, &p->val, mi, d0 + q);
    };
    if (J.kind == PHOTO_TABLE) {
      struct tabular_t *tb = &mod->photo.radar[l].tabular;
      for (int i = 0; i < J.n; i++) {
        put(i, &tb->rho[i]);
        d[J.n + i] = tb->rho[i].state == '=' ? 1.0 : 0.0;
      }
    } else if (J.kind == PHOTO_HARM) {
      struct harmonic_t *h;
      if (radar) h = &mod->photo.radar[l].harmcosine.R;
      else {
        union optscat_t *o = &mod->photo.optical[l];
        const int t = J.type;
        h = t == HARMHAPKE ? &o->harmhapke.w : t == HARMKAAS ? &o->harmkaas.R : &o->harmR.R;
      }
      for (int j = 0; j < J.k; j++) {
        int q = J.coef[j];
        for (int ll = 0; ll <= h[j].nhar; ll++) {
          put(q++, &h[j].a[ll][0]);
          for (int m = 1; m <= ll; m++) {
            put(q++, &h[j].a[ll][m]);
            put(q++, &h[j].b[ll][m]);
          }
        }
      }
    } else if (J.kind == PHOTO_INHO) {
      struct param_t *g;
      char *loc;
      size_t elem;
      if (radar) {
        g = &mod->photo.radar[l].inhocosine.global.R;
        loc = (char *)mod->photo.radar[l].inhocosine.local[0];
        elem = sizeof(struct RC_t);
      } else {
        union optscat_t *o = &mod->photo.optical[l];
        const int t = J.type;
        g = t == INHOHAPKE ? &o->inhohapke.global.w : t == INHOKAAS ? &o->inhokaas.global.R : &o->inhoR.global.R;
        loc = t == INHOHAPKE ? (char *)o->inhohapke.local[0] : t == INHOKAAS ? (char *)o->inhokaas.local[0]
                                                                               : (char *)o->inhoR.local[0];
        elem = t == INHOHAPKE ? sizeof(struct hapke_t) : t == INHOKAAS ? sizeof(struct kaas_t) : sizeof(struct R_t);
      }
      const int k = J.k;
      for (int j = 0; j < k; j++) put(j, &g[j]);
      for (int f = 0; f < J.n; f++) {
        struct param_t *p = (struct param_t *)(loc + elem * f);
        for (int j = 0; j < k; j++) {
          put(k + k * f + j, &p[j]);
          d[k + k * J.n + k * f + j] = p[j].state == '=' ? 1.0 : 0.0;
        }
      }
    }
  }
  // sets: offsets ('=' states of the delay polynomial resolved to the source set)
  std::vector<int> dc_src(nsets, -1);
  int last_dc = -1;
  for (int s = 0; s < nsets; s++) {
    struct set_t *set = &dat->set[s];
    double *ps = b + off_set + PB_SET_STRIDE * s;
    int base = off_set + PB_SET_STRIDE * s;
    for (int j = 0; j < 3; j++) {
      ps[PB_SET_ANGLEOFF + j] = set->angleoff[j].val;
      ps[PB_SET_OMEGAOFF + j] = set->omegaoff[j].val;
      set_param_target(tmap, &set->angleoff[j].val, mi, base + PB_SET_ANGLEOFF + j);
      set_param_target(tmap, &set->omegaoff[j].val, mi, base + PB_SET_OMEGAOFF + j);
    }
    struct delcor_t *dc = nullptr;
    struct param_t *dsc = nullptr;
    if (set->type == DELAY) { dc = &set->desc.deldop.delcor; dsc = &set->desc.deldop.dopscale; }
    else if (set->type == DOPPLER) { dc = &set->desc.doppler.delcor; dsc = &set->desc.doppler.dopscale; }
    if (dsc) { ps[PB_SET_DOPSCALE] = dsc->val; set_param_target(tmap, &dsc->val, mi, base + PB_SET_DOPSCALE); }
    if (dc) {
      if (dc->n > 9) return -12;
      ps[PB_SET_DCT0] = dc->t0;
      ps[PB_SET_DCN] = dc->n;
      for (int q = 0; q <= dc->n; q++) {
        ps[PB_SET_DCA + q] = dc->a[q].val;
        set_param_target(tmap, &dc->a[q].val, mi, base + PB_SET_DCA + q);
      }
      if (dc->a[0].state != '=') { last_dc = s; dc_src[s] = s; }
      else dc_src[s] = last_dc >= 0 ? last_dc : s;
    }
  }
  // component: its shape parameters (scale factors, whose '=' ties the
  // realization applies, or the ellipsoid / ovoid dimensions), rotation,
  // offset, the ovoid's k, the harmonic coefficients
  double *pc = b + off_comp;
  P.comp_type = ctype;
  P.nhar = nhar;
  P.sf_tie = 0;
  struct param_t *shp3 = ctype == ELLIPSE  ? &comp->desc.ell.two_a
                         : ctype == OVOID  ? &comp->desc.ovoid.two_a
                         : ctype == HARMONIC ? comp->desc.har.scalefactor
                                             : comp->desc.ver.scalefactor;
  for (int j = 0; j < 3; j++) {
    // (ELLIPSE / OVOID: two_a, a_over_b, b_over_c are consecutive param_t members)
    struct param_t *q = shp3 + j;
    pc[j] = q->val;
    set_param_target(tmap, &q->val, mi, off_comp + j);
    if ((ctype == VERTEX || ctype == HARMONIC) && j > 0 && q->state == '=') P.sf_tie |= 1 << j;
    pc[3 + j] = comp->rot[j].val; set_param_target(tmap, &comp->rot[j].val, mi, off_comp + 3 + j);
    pc[6 + j] = comp->off[j].val; set_param_target(tmap, &comp->off[j].val, mi, off_comp + 6 + j);
  }
  if (ctype == OVOID) {
    pc[PB_COMP_OVOID_K] = comp->desc.ovoid.k.val;
    set_param_target(tmap, &comp->desc.ovoid.k.val, mi, off_comp + PB_COMP_OVOID_K);
  }
  for (int l = 0, q = PB_COMP_HAR; l <= nhar && ctype == HARMONIC; l++) {
    pc[q] = comp->desc.har.a[l][0].val;
    set_param_target(tmap, &comp->desc.har.a[l][0].val, mi, off_comp + q);
    q++;
    for (int m = 1; m <= l; m++, q += 2) {
      pc[q] = comp->desc.har.a[l][m].val;
      set_param_target(tmap, &comp->desc.har.a[l][m].val, mi, off_comp + q);
      pc[q + 1] = comp->desc.har.b[l][m].val;
      set_param_target(tmap, &comp->desc.har.b[l][m].val, mi, off_comp + q + 1);
    }
  }
  // ---- topology and inputs (none for the views-only drop-ins) ----
  P.nv = rv->nv;
  P.nf = rv->nf;
  P.verts = rv->v;
  if (opt.views_only) P.nv = P.nf = 0;
  if (!repack) {
  P.fv.resize(3 * (size_t)P.nf);
  for (int f = 0; f < P.nf; f++)
    for (int k = 0; k < 3; k++) P.fv[3 * f + k] = rv->f[f].v[k];
  P.af_off.resize(P.nv + 1);
  P.u.resize(3 * (size_t)P.nv);
  P.a.resize(3 * (size_t)P.nv);
  P.r.resize(P.nv);
  P.af_off[0] = 0;
  for (int i = 0; i < P.nv; i++) {
    struct vertex_t *v = &rv->v[i];
    P.af_off[i + 1] = P.af_off[i] + v->naf;
    for (int q = 0; q < v->naf; q++) P.af_idx.push_back(v->af[q]);
    for (int k = 0; k < 3; k++) { P.u[3 * i + k] = v->u[k]; P.a[3 * i + k] = v->a[k]; }
    int src = v->r.state == '=' ? v->v_mirror : i;
    P.r[i] = rv->v[src].r.val;
  }
  P.hbasis.clear();
  if (ctype == HARMONIC) {
    P.hbasis.resize((size_t)P.nv * nbasis);
    for (int i = 0; i < P.nv; i++) {
      double *h = &P.hbasis[(size_t)i * nbasis];
      for (int l = 0, q = 0; l <= nhar; l++) {
        h[q++] = rv->v[i].afactor[l][0];
        for (int m = 1; m <= l; m++) { h[q++] = rv->v[i].afactor[l][m]; h[q++] = rv->v[i].bfactor[l][m]; }
      }
    }
  }
  P.sides.clear();
  if (opt.penalties && rv->s)
    for (int q = 0; q < rv->ns; q++) P.sides.push_back(make_int4(rv->s[q].v[0], rv->s[q].v[1], rv->s[q].f[0], rv->s[q].f[1]));
  }
  if (opt.host_geometry) {
    P.hx.resize(3 * (size_t)P.nv); P.hvn.resize(3 * (size_t)P.nv); P.hfn.resize(3 * (size_t)P.nf);
    for (int i = 0; i < P.nv; i++)
      for (int k = 0; k < 3; k++) { P.hx[3 * i + k] = rv->v[i].x[k]; P.hvn[3 * i + k] = rv->v[i].n[k]; }
    P.hfx.resize(3 * (size_t)P.nf);
    for (int f = 0; f < P.nf; f++)
      for (int k = 0; k < 3; k++) { P.hfn[3 * f + k] = rv->f[f].n[k]; P.hfx[3 * f + k] = rv->f[f].x[k]; }
  }
  // ---- views and frames ----
  auto fill_view_common = [&](ViewRec &V, struct pos_t *pos) {
    memset(&V, 0, sizeof(V));
    V.n = pos->n;
    V.kmpp = pos->km_per_pixel;
    V.smooth = par->pos_smooth;
    V.host_spin = opt.host_spin;
    V.nsinc2 = par->nsinc2;
    V.sinc2width = par->sinc2width;
  };
  auto copy9 = [](double *d, double m[3][3]) { for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) d[3 * i + j] = m[i][j]; };
  for (int s = 0; s < nsets; s++) {
    if (opt.only_set >= 0 && s != opt.only_set) continue;
    struct set_t *set = &dat->set[s];
    int pb_set = off_set + PB_SET_STRIDE * s;
    int pb_dc = dc_src[s] >= 0 ? off_set + PB_SET_STRIDE * dc_src[s] : pb_set;
    if (set->type == DELAY || set->type == DOPPLER) {
      bool dd = set->type == DELAY;
      struct deldop_t *D = &set->desc.deldop;
      struct doppler_t *Q = &set->desc.doppler;
      int nfr = dd ? D->nframes : Q->nframes, nv = dd ? D->nviews : Q->nviews, v0 = dd ? D->v0 : Q->v0;
      int ilaw = dd ? D->iradlaw : Q->iradlaw;
      if (mod->photo.radtype[ilaw] < COSINELAW_DIFF || mod->photo.radtype[ilaw] > TABULARLAW) return -13;
      for (int f = 0; f < nfr; f++) {
        FrameRec F;
        memset(&F, 0, sizeof(F));
        F.type = set->type;
        F.set = s;
        F.frm = f;
        F.nviews = nv;
        F.first_view = (int)P.views.size();
        struct pos_t *pos = dd ? &D->frame[f].pos : &Q->frame[f].pos;
        for (int k = 0; k < nv; k++) {
          ViewRec V;
          fill_view_common(V, pos);
          V.type = set->type;
          V.set = s; V.frm = f; V.v = k; V.is_v0 = (k == v0); V.nviews = nv;
          V.exclude_seen = (s == par->exclude_seen);
          V.pb_set = pb_set; V.pb_dc = pb_dc;
          V.pb_law = off_rad + PB_RAD_STRIDE * ilaw;
          V.law_type = mod->photo.radtype[ilaw];
          V.frame_id = (int)P.frames.size();
          if (dd) {
            struct deldopfrm_t *fr = &D->frame[f];
            struct deldopview_t *w = &fr->view[k];
            V.t = w->t; copy9(V.oe, w->oe); copy9(V.ae, w->ae);
            for (int j = 0; j < 3; j++) { V.orbspin[j] = w->orbspin[j]; V.spin[j] = w->spin[j]; V.intspin[j] = w->intspin[j]; }
            V.deloff = w->deloff; V.dopoff = w->dopoff;
            V.delcom_vig = fr->delcom_vig; V.dopcom_vig = fr->dopcom_vig; V.dopDC_vig = fr->dopDC_vig;
            V.dopfftlen = D->dopfftlen; V.Ftx = D->Ftx; V.del_per_pixel = D->del_per_pixel;
            V.dop_per_pixel = D->dop_per_pixel; V.codemethod = D->codemethod; V.spb = D->spb;
            V.stride = D->stride; V.ndel = fr->ndel; V.ndop = fr->ndop; V.frame_dof = fr->dof;
          } else {
            struct dopfrm_t *fr = &Q->frame[f];
            struct dopview_t *w = &fr->view[k];
            V.t = w->t; copy9(V.oe, w->oe); copy9(V.ae, w->ae);
            for (int j = 0; j < 3; j++) { V.orbspin[j] = w->orbspin[j]; V.spin[j] = w->spin[j]; V.intspin[j] = w->intspin[j]; }
            V.deloff = 0.0; V.dopoff = w->dopoff;
            V.dopcom_vig = fr->dopcom_vig;
            V.Ftx = Q->Ftx; V.dop_per_pixel = Q->dop_per_bin; V.ndel = 1; V.ndop = fr->ndop;
            V.frame_dof = fr->dof; V.spb = 1; V.stride = 1; V.codemethod = LONG_ORIG;
          }
          P.views.push_back(V);
        }
        long long N;
        if (dd) {
          struct deldopfrm_t *fr = &D->frame[f];
          F.ndel = fr->ndel; F.ndop = fr->ndop;
          N = (long long)fr->ndel * fr->ndop;
          F.obs_off = P.nobs;
          if (!opt.views_only) {
            P.obs_jobs.push_back({VT_DELAY, fr, P.nobs, N});
            P.nobs += N;
          }
          F.cal_floats = fr->cal.state == 'f'; F.cal_fixed = fr->cal.val; F.weight = fr->weight; F.dof = fr->dof;
          F.sdev_sq = fr->sdev * fr->sdev; F.lookfact = fr->nlooks > 0.0 ? 1.0 / fr->nlooks : 0.0;
          F.overflow_o2 = fr->overflow_o2; F.overflow_m2 = fr->overflow_m2;
        } else {
          struct dopfrm_t *fr = &Q->frame[f];
          F.ndel = 1; F.ndop = fr->ndop;
          N = fr->ndop;
          F.obs_off = P.nobs;
          if (!opt.views_only) {
            P.obs_jobs.push_back({VT_DOPPLER, fr, P.nobs, N});
            P.nobs += N;
          }
          F.cal_floats = fr->cal.state == 'f'; F.cal_fixed = fr->cal.val; F.weight = fr->weight; F.dof = fr->dof;
          F.sdev_sq = fr->sdev * fr->sdev; F.lookfact = fr->nlooks > 0.0 ? 1.0 / fr->nlooks : 0.0;
          F.overflow_o2 = fr->overflow_o2; F.overflow_m2 = fr->overflow_m2;
        }
        F.speckle = par->speckle;
        F.off = N;  // temporarily: element count
        P.frames.push_back(F);
      }
    } else if (set->type == LGHTCRV) {
      struct lghtcrv_t *L = &set->desc.lghtcrv;
      int t = mod->photo.opttype[L->ioptlaw];
      if (t < GEOMETRICAL || t > INHOKAAS) return -14;
      FrameRec F;
      memset(&F, 0, sizeof(F));
      F.type = VT_LGHTCRV;
      F.set = s;
      F.nviews = 1;
      F.first_view = (int)P.views.size();
      F.lc_n = L->n; F.lc_ncalc = L->ncalc; F.lc_nviews = L->nviews;
      for (int i = 1; i <= L->ncalc; i++) {
        struct crvrend_t *rd = &L->rend[i];
        ViewRec V;
        fill_view_common(V, &rd->pos);
        V.type = VT_LGHTCRV;
        V.set = s; V.frm = i - 1; V.v = 0; V.is_v0 = 1; V.nviews = 1;
        V.exclude_seen = (s == par->exclude_seen);
        V.bistatic = 1;
        V.pb_set = pb_set; V.pb_dc = pb_dc;
        V.pb_law = off_opt + PB_OPT_STRIDE * L->ioptlaw;
        V.law_type = t;
        V.frame_id = (int)P.frames.size();
        V.t = L->x[i];
        copy9(V.oe, rd->oe); copy9(V.se, rd->se); copy9(V.ae, rd->ae);
        for (int j = 0; j < 3; j++) { V.orbspin[j] = rd->orbspin[j]; V.spin[j] = rd->spin[j]; V.intspin[j] = rd->intspin[j]; }
        V.phase = L->solar_phase[i];
        V.intensityfactor = pow(rd->pos.km_per_pixel / AU, 2.0);
        P.views.push_back(V);
      }
      F.obs_off = P.nobs;
      if (!opt.views_only) {
        P.obs_jobs.push_back({VT_LGHTCRV, L, P.nobs, (long long)L->n});
        P.nobs += L->n;
      }
      F.cal_floats = L->cal.state == 'f'; F.cal_fixed = L->cal.val; F.weight = L->weight; F.dof = L->dof;
      F.lc_x_off = (long long)P.lc_x.size();
      P.lc_x.push_back(0.0);
      for (int i = 1; i <= L->ncalc; i++) P.lc_x.push_back(L->x[i]);
      F.lc_t_off = (long long)P.lc_t.size();
      for (int i = 1; i <= L->n; i++)
        for (int v = 0; v < L->nviews; v++) P.lc_t.push_back(L->t[i][v]);
      F.off = L->n;
      P.frames.push_back(F);
    } else {
      return -15;  // plane-of-sky sets are outside this path
    }
  }
  P.dof = dat->dof; P.dof_dd = dat->dof_deldop; P.dof_dop = dat->dof_doppler; P.dof_lc = dat->dof_lghtcrv;
  return 0;
}

int Engine::build(struct par_t *par, struct mod_t **mods, struct dat_t **dats, int nmodels,
                  const std::vector<EvalSpec> &evals, const RunOptions &o) {
  opt = o;
  nsinc2 = par->nsinc2;
  sinc2width = par->sinc2width;
  mask_tol = par->mask_tol;
  mincosine_seen = par->mincosine_seen;
  exclude_seen = par->exclude_seen;
  if (!stream) SHP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  g_packs.assign(nmodels, ModelPack());
  model_params.assign(nmodels, {});
  int rc;
  for (int m = 0; m < nmodels; m++)
    if ((rc = pack_model(m, par, mods[m], dats[m])) != 0) return rc;
  nevals = (int)evals.size();
  any_harmonic = false;
  for (auto &P : g_packs) any_harmonic = any_harmonic || P.comp_type == HARMONIC;
  pb_stride = 0;
  for (auto &P : g_packs) pb_stride = std::max(pb_stride, (int)P.base.size());
  pb_stride = (pb_stride + 7) & ~7;
  // parameter blocks: evaluations first, then one base block per model
  h_params.assign((size_t)pb_stride * (nevals + nmodels), 0.0);
  for (int m = 0; m < nmodels; m++)
    std::copy(g_packs[m].base.begin(), g_packs[m].base.end(), h_params.begin() + (size_t)pb_stride * (nevals + m));
  // per-model device topology
  for (int m = 0; m < nmodels; m++) {
    ModelPack &P = g_packs[m];
    if ((rc = upload(this, &P.d_fv, P.fv))) return rc;
    if ((rc = upload(this, &P.d_af_off, P.af_off))) return rc;
    if ((rc = upload(this, &P.d_af_idx, P.af_idx))) return rc;
    if ((rc = upload(this, &P.d_u, P.u))) return rc;
    if ((rc = upload(this, &P.d_a, P.a))) return rc;
    if (!P.sides.empty() && (rc = upload(this, &P.d_sides, P.sides))) return rc;
    if (!P.hbasis.empty() && (rc = upload(this, &P.d_hbasis, P.hbasis))) return rc;
  }
  // evaluations, their slots, views and frames
  eval_targets.assign(nevals, {});
  h_slots.clear();
  std::vector<std::vector<double>> slot_r;
  for (int m = 0; m < nmodels; m++) {  // base slots
    GeomSlot G;
    memset(&G, 0, sizeof(G));
    G.model = m; G.eval = -1; G.nv = g_packs[m].nv; G.nf = g_packs[m].nf;
    G.comp_pb = pb_stride * (nevals + m) + g_packs[m].comp_off;
    h_slots.push_back(G);
    slot_r.push_back(g_packs[m].r);
  }
  h_eval_slot.assign(nevals, 0);
  h_views.clear();
  h_frames.clear();
  h_evals.clear();
  long long fit_cursor = 0, pos_cursor = 0, y_cursor = 0, y2_cursor = 0, tile_cursor = 0, bin_cursor = 0;
  h_photo_jobs.clear();
  lawloc_len = 0;
  std::vector<long long> view_loc;  // per view: offset of its law's resolved values, or -1
  long long lc_rec_cursor = 0;
  std::vector<long long> model_obs_base(nmodels, 0), model_x_base(nmodels, 0), model_t_base(nmodels, 0);
  long long ob = 0, xb = 0, tb = 0;
  for (int m = 0; m < nmodels; m++) {
    model_obs_base[m] = ob; ob += g_packs[m].nobs;
    model_x_base[m] = xb; xb += (long long)g_packs[m].lc_x.size();
    model_t_base[m] = tb; tb += (long long)g_packs[m].lc_t.size();
  }
  for (int e = 0; e < nevals; e++) {
    const EvalSpec &S = evals[e];
    int m = S.model;
    ModelPack &P = g_packs[m];
    int pb = pb_stride * e;
    std::copy(P.base.begin(), P.base.end(), h_params.begin() + pb);
    bool own_slot = false;
    std::vector<double> r;
    for (size_t q = 0; q < S.targets.size(); q++) {
      auto it = model_params[m].find(S.targets[q]);
      ParamTarget T;
      if (it != model_params[m].end()) {
        T = it->second;
      } else {
        // a vertex deviation: the r.val field of one of the model's vertices
        const ptrdiff_t d = (const char *)S.targets[q] - (const char *)&P.verts[0].r.val;
        if (P.nv == 0 || d < 0 || d % (ptrdiff_t)sizeof(struct vertex_t) != 0 ||
            d / (ptrdiff_t)sizeof(struct vertex_t) >= P.nv)
          return -16;  // not a parameter of this model
        T = ParamTarget{1, m, (int)(d / (ptrdiff_t)sizeof(struct vertex_t))};
      }
      eval_targets[e].push_back(T);
      if (T.kind == 0) {
        h_params[pb + T.offset] = S.values[q];
        if (T.offset >= P.comp_off && T.offset < P.photo_off) own_slot = true;
      } else {
        if (!own_slot || r.empty()) r = P.r;
        own_slot = true;
      }
    }
    if (own_slot) {
      if (r.empty()) r = P.r;
      for (size_t q = 0; q < S.targets.size(); q++)
        if (eval_targets[e][q].kind == 1) r[eval_targets[e][q].offset] = S.values[q];
      GeomSlot G;
      memset(&G, 0, sizeof(G));
      G.model = m; G.eval = e; G.nv = P.nv; G.nf = P.nf;
      G.comp_pb = pb + P.comp_off;
      h_eval_slot[e] = (int)h_slots.size();
      h_slots.push_back(G);
      slot_r.push_back(r);
    } else {
      h_eval_slot[e] = m;
    }
    EvalRec E;
    memset(&E, 0, sizeof(E));
    E.model = m; E.pb = pb;
    E.first_frame = (int)h_frames.size();
    E.nframes = (int)P.frames.size();
    E.dof = P.dof; E.dof_deldop = P.dof_dd; E.dof_doppler = P.dof_dop; E.dof_lghtcrv = P.dof_lc;
    E.bad_objfactor = par->bad_objfactor;
    E.pen_set_off = P.set_off;
    E.pen_nsets = P.nsets;
    const long long loc_base = lawloc_len;
    E.photo_job0 = (int)h_photo_jobs.size();
    E.photo_njobs = (int)P.laws.size();
    for (const auto &L : P.laws) {
      PhotoJob J = L.job;
      J.data += pb;
      J.loc += loc_base;
      J.slot = h_eval_slot[e];
      for (int q = 0; q < J.nchk; q++) J.chk[q] += pb;
      h_photo_jobs.push_back(J);
    }
    lawloc_len += P.loc_len;
    h_evals.push_back(E);
    int view_base = (int)h_views.size(), frame_base = (int)h_frames.size();
    for (size_t q = 0; q < P.views.size(); q++) {
      ViewRec V = P.views[q];
      V.fp64_skip = (opt.method == METHOD_MIXED && V.type != VT_LGHTCRV) ? 1 : 0;
      V.eval = e;
      V.slot = h_eval_slot[e];
      V.pb = pb;
      V.frame_id += frame_base;
      V.pos_off = pos_cursor;
      pos_cursor += (long long)(2 * V.n + 1) * (2 * V.n + 1);
      V.tile_off = tile_cursor;
      tile_cursor += (long long)((2 * V.n + STALE_TILE) / STALE_TILE) * ((2 * V.n + STALE_TILE) / STALE_TILE);
      V.binned = 0; V.bin_ntx = 0; V.bin_off = 0;
      V.lc_rec_off = 0; V.lc_rec_cap = 0;
      V.lawloc = nullptr; V.lawk = 0;
      long long vloc = -1;
      for (const auto &L : P.laws)
        if (L.pb_law == V.pb_law && L.has_loc) { vloc = loc_base + L.job.loc; V.lawk = L.job.k; }
      view_loc.push_back(vloc);
      if (opt.method == METHOD_MIXED && V.type == VT_LGHTCRV) {
        // a range of the replay records per lightcurve view: its records are
        // reserved on a counter of its own (one shared counter serializes)
        V.lc_rec_off = (int)lc_rec_cursor;
        V.lc_rec_cap = (2 * V.n + 1) * (2 * V.n + 1) / 16 + 64;
        lc_rec_cursor += V.lc_rec_cap;
        max_lc_rec = std::max(max_lc_rec, V.lc_rec_cap);
      }
      if (opt.method == METHOD_MIXED && V.type != VT_LGHTCRV) {
        // binned tile kernel unless the POS is wider than its tile grid, the
        // facet index does not fit an event, the global z-buffer is asked for,
        // or the view's radar law is not the diffuse cosine law (the tile
        // kernel evaluates that law only; the global path has the others)
        int ntx = (2 * V.n + 1 + BT - 1) / BT;
        if (!opt.global_zbuf && ntx <= BIN_MAX_NTX && P.nf <= (1 << EV_FACET_BITS) &&
            V.law_type == COSINELAW_DIFF) {
          V.binned = 1; V.bin_ntx = ntx; V.bin_off = bin_cursor;
          bin_cursor += (long long)ntx * ntx;
          max_bin_ntx = std::max(max_bin_ntx, ntx);
          any_binned = true;
        } else {
          any_global = true;
        }
      }
      max_pix = std::max(max_pix, (2 * V.n + 1) * (2 * V.n + 1));
      if (V.type == VT_LGHTCRV) any_lc = true;
      if (V.type == VT_DELAY) any_dd = true;
      if (V.type == VT_DOPPLER) any_dop = true;
      h_views.push_back(V);
    }
    for (size_t q = 0; q < P.frames.size(); q++) {
      FrameRec F = P.frames[q];
      long long N = F.off;
      F.eval = e;
      F.first_view += view_base;
      F.off = fit_cursor;
      fit_cursor += N;
      F.obs_off += model_obs_base[m];
      if (F.type == VT_LGHTCRV) {
        F.lc_x_off += model_x_base[m];
        F.lc_t_off += model_t_base[m];
        F.lc_y_off = y_cursor; y_cursor += F.lc_ncalc + 1;
        F.lc_y2_off = y2_cursor; y2_cursor += 2 * (F.lc_ncalc + 1);
      }
      h_frames.push_back(F);
    }
  }
  nviews = (int)h_views.size();
  nframes = (int)h_frames.size();
  nslots = (int)h_slots.size();
  total_pix = pos_cursor;
  total_tiles = tile_cursor;
  total_bins = bin_cursor;
  lc_rec_total = lc_rec_cursor;
  fit_len = fit_cursor;
  // ---- device arenas ----
  max_nf = 0;
  for (int s = 0; s < nslots; s++) {
    GeomSlot &G = h_slots[s];
    ModelPack &P = g_packs[G.model];
    max_nf = std::max(max_nf, G.nf);
    G.fv = P.d_fv; G.af_off = P.d_af_off; G.af_idx = P.d_af_idx; G.u = P.d_u; G.a = P.d_a;
    G.sides = P.d_sides; G.ns = (int)P.sides.size();
    G.comp_type = P.comp_type; G.nhar = P.nhar; G.hbasis = P.d_hbasis; G.sf_tie = P.sf_tie;
    if ((rc = upload(this, &G.r, slot_r[s]))) return rc;
    if ((rc = alloc((void **)&G.x, sizeof(double) * 3 * G.nv))) return rc;
    if ((rc = alloc((void **)&G.vn, sizeof(double) * 3 * G.nv))) return rc;
    if ((rc = alloc((void **)&G.fn, sizeof(double) * 3 * G.nf))) return rc;
    if ((rc = alloc((void **)&G.fx, sizeof(double) * 3 * G.nf))) return rc;
    if ((rc = alloc((void **)&G.area, sizeof(double) * G.nf))) return rc;
    if ((rc = alloc((void **)&G.x32, sizeof(float4) * G.nv))) return rc;
    if ((rc = alloc((void **)&G.vn32, sizeof(float4) * G.nv))) return rc;
    if ((rc = alloc((void **)&G.fn32, sizeof(float4) * G.nf))) return rc;
    if ((rc = alloc((void **)&G.seen, G.nf))) return rc;
    if (opt.method == METHOD_MIXED) {
      int nblk = (G.nf + BF - 1) / BF;
      max_nblk = std::max(max_nblk, nblk);
      if ((rc = alloc((void **)&G.blk, sizeof(float4) * 2 * (size_t)nblk))) return rc;
      if ((rc = alloc((void **)&G.sblk, sizeof(float4) * 4 * (size_t)nblk))) return rc;
      if ((rc = alloc((void **)&G.fb, sizeof(float4) * (size_t)G.nf))) return rc;
      if ((rc = alloc((void **)&G.fvx, sizeof(float4) * 3 * (size_t)G.nf))) return rc;
      if ((rc = alloc((void **)&G.fvn, sizeof(float4) * 3 * (size_t)G.nf))) return rc;
    }
    SHP_CHECK(hipMemset(G.seen, 0, G.nf));
    if (opt.host_geometry && (rc = upload_host_geometry(s))) return rc;
  }
  if ((rc = setup_penalties(par))) return rc;
  // realize_photo: resolved values, jobs and their outputs
  if ((rc = alloc((void **)&d_lawloc, sizeof(double) * (size_t)std::max(1LL, lawloc_len)))) return rc;
  SHP_CHECK(hipMemset(d_lawloc, 0, sizeof(double) * (size_t)std::max(1LL, lawloc_len)));
  for (int q = 0; q < nviews; q++)
    if (view_loc[q] >= 0) h_views[q].lawloc = d_lawloc + view_loc[q];
  photo_nblk = 1;
  for (const auto &J : h_photo_jobs)
    if (J.kind == PHOTO_HARM || J.kind == PHOTO_INHO) photo_nblk = std::max(photo_nblk, (J.n + 255) / 256);
  if (!h_photo_jobs.empty()) {
    if ((rc = upload(this, &d_photo_jobs, h_photo_jobs))) return rc;
    if ((rc = alloc((void **)&d_job_flag, sizeof(int) * h_photo_jobs.size() * photo_nblk))) return rc;
    if ((rc = alloc((void **)&d_job_log, sizeof(double) * h_photo_jobs.size() * photo_nblk))) return rc;
  }
  if ((rc = upload(this, &d_slots, h_slots))) return rc;
  if ((rc = upload(this, &d_views, h_views))) return rc;
  if ((rc = upload(this, &d_frames, h_frames))) return rc;
  if ((rc = upload(this, &d_evals, h_evals))) return rc;
  if ((rc = upload(this, &d_eval_slot, h_eval_slot))) return rc;
  if ((rc = upload(this, &d_params, h_params))) return rc;
  SHP_CHECK(hipHostMalloc((void **)&h_pin, sizeof(double) * h_params.size(), hipHostMallocDefault));
  memcpy(h_pin, h_params.data(), sizeof(double) * h_params.size());
  if (opt.views_only) { g_packs.clear(); return 0; }
  if ((rc = alloc((void **)&d_slot_flags, sizeof(int) * nslots))) return rc;
  if ((rc = alloc((void **)&d_slot_logf, sizeof(double) * nslots))) return rc;
  SHP_CHECK(hipMemset(d_slot_flags, 0, sizeof(int) * nslots));
  SHP_CHECK(hipMemset(d_slot_logf, 0, sizeof(double) * nslots));
  std::vector<double> x_all, t_all;
  std::vector<std::pair<const ModelPack::ObsJob *, long long>> jobs;  // (job, offset in obs_all)
  long long nobs_all = 0;
  for (auto &P : g_packs) {
    for (auto &J : P.obs_jobs) jobs.push_back({&J, nobs_all + J.off});
    nobs_all += P.nobs;
    x_all.insert(x_all.end(), P.lc_x.begin(), P.lc_x.end());
    t_all.insert(t_all.end(), P.lc_t.begin(), P.lc_t.end());
  }
  // observations and weights back to back in the thread's pinned stage
  // (plain host memory if it cannot be pinned)
  std::vector<double> plain;
  double *obs_all = host_stage(2 * (size_t)nobs_all);
  if (!obs_all) {
    plain.resize(2 * (size_t)nobs_all);
    obs_all = plain.data();
  }
  double *oov_all = obs_all + nobs_all;
  host_parallel_for((int)jobs.size(), [&](int k) {
    const ModelPack::ObsJob &J = *jobs[k].first;
    double *o = obs_all + jobs[k].second, *w = oov_all + jobs[k].second;
    if (J.kind == VT_DELAY) {
      const struct deldopfrm_t *fr = (const struct deldopfrm_t *)J.src;
      for (int i = 1; i <= fr->ndel; i++)
        for (int j = 1; j <= fr->ndop; j++) { *o++ = fr->obs[i][j]; *w++ = fr->oneovervar[i][j]; }
    } else if (J.kind == VT_DOPPLER) {
      const struct dopfrm_t *fr = (const struct dopfrm_t *)J.src;
      for (int j = 1; j <= fr->ndop; j++) { *o++ = fr->obs[j]; *w++ = fr->oneovervar[j]; }
    } else {
      const struct lghtcrv_t *L = (const struct lghtcrv_t *)J.src;
      for (int i = 1; i <= L->n; i++) { *o++ = L->obs[i]; *w++ = L->oneovervar[i]; }
    }
  });
  obs_len = nobs_all;
  if ((rc = alloc((void **)&d_obs, sizeof(double) * std::max(1LL, obs_len)))) return rc;
  if ((rc = alloc((void **)&d_oov, sizeof(double) * std::max(1LL, obs_len)))) return rc;
  if (obs_len) {
    SHP_CHECK(hipMemcpy(d_obs, obs_all, sizeof(double) * obs_len, hipMemcpyHostToDevice));
    SHP_CHECK(hipMemcpy(d_oov, oov_all, sizeof(double) * obs_len, hipMemcpyHostToDevice));
  }
  if ((rc = upload(this, &d_lc_x, x_all))) return rc;
  if ((rc = upload(this, &d_lc_t, t_all))) return rc;
  lcx_len = (long long)x_all.size();
  lct_len = (long long)t_all.size();
  if ((rc = alloc((void **)&d_lc_y, sizeof(double) * std::max(1LL, y_cursor)))) return rc;
  if ((rc = alloc((void **)&d_lc_y2, sizeof(double) * std::max(1LL, y2_cursor)))) return rc;
  if ((rc = alloc((void **)&d_fit, sizeof(double) * std::max(1LL, fit_len)))) return rc;
  SHP_CHECK(hipMemset(d_fit, 0, sizeof(double) * std::max(1LL, fit_len)));
  if ((rc = alloc((void **)&d_overflow, sizeof(double) * 2000 * (size_t)std::max(1, nviews)))) return rc;
  SHP_CHECK(hipMemset(d_overflow, 0, sizeof(double) * 2000 * (size_t)std::max(1, nviews)));
  if ((rc = alloc((void **)&d_view_out, sizeof(double) * VIEW_OUT * (size_t)std::max(1, nviews)))) return rc;
  // mixed-path frame data (FP32 obs*w and w)
  if (opt.method == METHOD_MIXED) {
    std::vector<float> ow(obs_len), w32(obs_len);
    for (long long q = 0; q < obs_len; q++) { ow[q] = (float)(obs_all[q] * oov_all[q]); w32[q] = (float)oov_all[q]; }
    if ((rc = upload(this, &d_obs32, ow))) return rc;
    if ((rc = upload(this, &d_oov32, w32))) return rc;
  }
  // POS buffers
  size_t P = (size_t)std::max(1LL, total_pix);
  memset(&pd, 0, sizeof(pd));
  if (opt.facet_map && opt.method == METHOD_MIXED)
    if ((rc = alloc((void **)&pd.fmap, sizeof(int) * P))) return rc;
  if ((rc = alloc((void **)&pd.zkey, sizeof(unsigned long long) * P))) return rc;
  SHP_CHECK(hipMemset(pd.zkey, 0, sizeof(unsigned long long) * P));
  if (opt.method == METHOD_FP64 || any_lc) {
    if ((rc = alloc((void **)&pd.fbuf, sizeof(int) * P))) return rc;
    if ((rc = alloc((void **)&pd.z, sizeof(double) * P))) return rc;
    if ((rc = alloc((void **)&pd.cose, sizeof(double) * P))) return rc;
    if ((rc = alloc((void **)&pd.f, sizeof(int) * P))) return rc;
    if ((rc = alloc((void **)&pd.stale_idx, sizeof(int) * P))) return rc;
    hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(256), 0, stream, pd.fbuf, (long long)P, 0x7fffffff);
    hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(256), 0, stream, pd.stale_idx, (long long)P, -1);
    hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(256), 0, stream, pd.f, (long long)P, -1);
    SHP_CHECK(hipMemsetAsync(pd.cose, 0, sizeof(double) * P, stream));
    hipLaunchKernelGGL(k_fill_f64, dim3(1024), dim3(256), 0, stream, pd.z, (long long)P, -1.0e20);
    if (any_lc) {
      if ((rc = alloc((void **)&pd.cosi, sizeof(double) * P))) return rc;
      if ((rc = alloc((void **)&pd.b, sizeof(double) * P))) return rc;
      if ((rc = alloc((void **)&pd.zill, sizeof(double) * P))) return rc;
      if ((rc = alloc((void **)&pd.cosill, sizeof(double) * P))) return rc;
      if ((rc = alloc((void **)&pd.fill, sizeof(int) * P))) return rc;
      SHP_CHECK(hipMemsetAsync(pd.cosi, 0, sizeof(double) * P, stream));
      SHP_CHECK(hipMemsetAsync(pd.b, 0, sizeof(double) * P, stream));
      SHP_CHECK(hipMemsetAsync(pd.cosill, 0, sizeof(double) * P, stream));
      hipLaunchKernelGGL(k_fill_f64, dim3(1024), dim3(256), 0, stream, pd.zill, (long long)P, -1.0e20);
      hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(256), 0, stream, pd.fill, (long long)P, -1);
    }
    if (opt.with_body) {
      if ((rc = alloc((void **)&pd.body, sizeof(int) * P))) return rc;
      if ((rc = alloc((void **)&pd.comp, sizeof(int) * P))) return rc;
      hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(256), 0, stream, pd.body, (long long)P, -1);
      hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(256), 0, stream, pd.comp, (long long)P, -1);
      if (any_lc) {
        if ((rc = alloc((void **)&pd.bodyill, sizeof(int) * P))) return rc;
        if ((rc = alloc((void **)&pd.compill, sizeof(int) * P))) return rc;
        hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(256), 0, stream, pd.bodyill, (long long)P, -1);
        hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(256), 0, stream, pd.compill, (long long)P, -1);
      }
    }
  }
  if (opt.method == METHOD_MIXED && !pd.stale_idx) {
    if ((rc = alloc((void **)&pd.stale_idx, sizeof(int) * P))) return rc;
    hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(256), 0, stream, pd.stale_idx, (long long)P, -1);
  }
  if (any_binned) {
    // tile lists: counts start (and are left by every tile) at zero
    if ((rc = alloc((void **)&d_bin_cnt, sizeof(int) * (size_t)total_bins))) return rc;
    SHP_CHECK(hipMemsetAsync(d_bin_cnt, 0, sizeof(int) * (size_t)total_bins, stream));
    if ((rc = alloc((void **)&d_bin_list, sizeof(unsigned int) * BIN_CAP * (size_t)total_bins))) return rc;
    // every block of every binned view may reach outside its window
    long long out_cap = 0;
    for (auto &V : h_views)
      if (V.binned) out_cap += (h_slots[V.slot].nf + BF - 1) / BF;
    if ((rc = alloc((void **)&d_out_list, sizeof(int2) * (size_t)std::max(1LL, out_cap)))) return rc;
    // the persistent tile kernel's work counters start at zero and every launch leaves them there
    if ((rc = alloc((void **)&d_tile_work, sizeof(int) * 2 * NXCD))) return rc;
    SHP_CHECK(hipMemsetAsync(d_tile_work, 0, sizeof(int) * 2 * NXCD, stream));
  }
  if (opt.method == METHOD_MIXED) {
    if ((rc = alloc((void **)&pd.stale_tiles, (size_t)std::max(1LL, total_tiles)))) return rc;
    SHP_CHECK(hipMemsetAsync(pd.stale_tiles, 0, (size_t)std::max(1LL, total_tiles), stream));
  }
  // replay records: 1 per 32 POS pixels (a limb band and the facet-edge band
  // of every view take well under that), at least 4096, at most 64M
  pd.stale_cap = (int)std::min<long long>(std::max<long long>({(long long)P / 32, 4096LL, lc_rec_total}), 64LL << 20);
  if (lc_rec_total > pd.stale_cap) return -31;  // lightcurve record ranges beyond the record list's bound
  if (any_lc && opt.method == METHOD_MIXED) {
    if ((rc = alloc((void **)&pd.lc_count, sizeof(int) * (size_t)std::max(1, nviews)))) return rc;
    SHP_CHECK(hipMemsetAsync(pd.lc_count, 0, sizeof(int) * (size_t)std::max(1, nviews), stream));
  }
  if ((rc = alloc((void **)&pd.stale, sizeof(StaleRec) * (size_t)pd.stale_cap))) return rc;
  // counters: [0] replay records, [1] error flags, [2] records to collect,
  // [3] unused (0), [4] blocks reaching outside their POS window
  if ((rc = alloc((void **)&pd.nstale, sizeof(int) * 8))) return rc;
  pd.error = pd.nstale + 1;
  pd.ncollect = pd.nstale + 2;
  SHP_CHECK(hipMemset(pd.nstale, 0, sizeof(int) * 8));
  hipLaunchKernelGGL(k_frame_o2, dim3(std::max(1, nframes)), dim3(256), 0, stream, d_frames, nframes, d_obs, d_oov);
  SHP_CHECK(hipStreamSynchronize(stream));
  g_packs.clear();
  return 0;
}

// the realized geometry the caller's structs hold (drop-in mode), FP64 and
// the FP32 copies, into slot s (g_packs holds the packed model)
int Engine::upload_host_geometry(int s) {
  GeomSlot &G = h_slots[s];
  const ModelPack &P = g_packs[G.model];
  SHP_CHECK(hipMemcpy(G.x, P.hx.data(), sizeof(double) * 3 * G.nv, hipMemcpyHostToDevice));
  SHP_CHECK(hipMemcpy(G.vn, P.hvn.data(), sizeof(double) * 3 * G.nv, hipMemcpyHostToDevice));
  SHP_CHECK(hipMemcpy(G.fn, P.hfn.data(), sizeof(double) * 3 * G.nf, hipMemcpyHostToDevice));
  SHP_CHECK(hipMemcpy(G.fx, P.hfx.data(), sizeof(double) * 3 * G.nf, hipMemcpyHostToDevice));
  std::vector<float4> t4(std::max(G.nv, G.nf));
  for (int i = 0; i < G.nv; i++) t4[i] = make_float4((float)P.hx[3 * i], (float)P.hx[3 * i + 1], (float)P.hx[3 * i + 2], 0.f);
  SHP_CHECK(hipMemcpy(G.x32, t4.data(), sizeof(float4) * G.nv, hipMemcpyHostToDevice));
  for (int i = 0; i < G.nv; i++) t4[i] = make_float4((float)P.hvn[3 * i], (float)P.hvn[3 * i + 1], (float)P.hvn[3 * i + 2], 0.f);
  SHP_CHECK(hipMemcpy(G.vn32, t4.data(), sizeof(float4) * G.nv, hipMemcpyHostToDevice));
  for (int i = 0; i < G.nf; i++) t4[i] = make_float4((float)P.hfn[3 * i], (float)P.hfn[3 * i + 1], (float)P.hfn[3 * i + 2], 0.f);
  SHP_CHECK(hipMemcpy(G.fn32, t4.data(), sizeof(float4) * G.nf, hipMemcpyHostToDevice));
  return 0;
}

int Engine::refresh(struct par_t *par, struct mod_t *mod, struct dat_t *dat) {
  if (nevals != 1 || h_evals.size() != 1 || opt.views_only) return 1;
  g_packs.assign(1, ModelPack());
  if (model_params.empty()) model_params.assign(1, {});
  int rc = pack_model(0, par, mod, dat, true);
  if (rc) {
    g_packs.clear();
    return rc;
  }
  ModelPack &P = g_packs[0];
  const GeomSlot &G0 = h_slots[h_eval_slot[0]];
  bool same = (int)P.views.size() == nviews && (int)P.frames.size() == nframes && (int)P.base.size() <= pb_stride &&
              P.loc_len == lawloc_len && P.laws.size() == h_photo_jobs.size() &&
              P.nv == G0.nv && P.nf == G0.nf && (long long)P.lc_x.size() == lcx_len &&
              (long long)P.lc_t.size() == lct_len;
  // the laws' jobs as a fresh build would make them (same kinds, offsets, bounds)
  for (size_t q = 0; same && q < P.laws.size(); q++) {
    PhotoJob J = P.laws[q].job;
    const PhotoJob &O = h_photo_jobs[q];
    J.data += h_evals[0].pb;
    J.slot = O.slot;
    for (int c = 0; c < J.nchk; c++) J.chk[c] += h_evals[0].pb;
    same = J.kind == O.kind && J.type == O.type && J.k == O.k && J.n == O.n && J.data == O.data && J.loc == O.loc &&
           J.nchk == O.nchk;
    for (int c = 0; same && c < 5; c++)
      same = J.chk[c] == O.chk[c] && J.coef[c] == O.coef[c] && J.nhar[c] == O.nhar[c] && J.lo[c] == O.lo[c] &&
             J.hi[c] == O.hi[c] && J.mode[c] == O.mode[c];
  }
  for (int q = 0; same && q < nviews; q++) {
    const ViewRec &A = h_views[q], &B = P.views[q];
    same = A.type == B.type && A.n == B.n && A.ndel == B.ndel && A.ndop == B.ndop && A.nviews == B.nviews &&
           A.law_type == B.law_type && A.set == B.set && A.frm == B.frm && A.v == B.v;
  }
  for (int q = 0; same && q < nframes; q++) {
    const FrameRec &A = h_frames[q], &B = P.frames[q];
    same = A.type == B.type && A.ndel == B.ndel && A.ndop == B.ndop && A.nviews == B.nviews && A.lc_n == B.lc_n &&
           A.lc_ncalc == B.lc_ncalc && A.lc_nviews == B.lc_nviews;
  }
  if (!same) {
    g_packs.clear();
    return 1;
  }
  // views: the caller's description, the engine's offsets
  for (int q = 0; q < nviews; q++) {
    ViewRec V = P.views[q];
    const ViewRec &O = h_views[q];
    V.eval = O.eval; V.slot = O.slot; V.pb = O.pb; V.frame_id = O.frame_id; V.pos_off = O.pos_off;
    V.tile_off = O.tile_off; V.bin_off = O.bin_off; V.bin_ntx = O.bin_ntx; V.binned = O.binned;
    V.fp64_skip = O.fp64_skip;
    V.lawloc = O.lawloc; V.lawk = O.lawk;
    h_views[q] = V;
  }
  for (int q = 0; q < nframes; q++) {
    FrameRec F = P.frames[q];
    const FrameRec &O = h_frames[q];
    F.eval = O.eval; F.first_view = O.first_view; F.off = O.off; F.obs_off = O.obs_off;
    F.lc_x_off = O.lc_x_off; F.lc_t_off = O.lc_t_off; F.lc_y_off = O.lc_y_off; F.lc_y2_off = O.lc_y2_off;
    h_frames[q] = F;
  }
  EvalRec &E = h_evals[0];
  E.dof = P.dof; E.dof_deldop = P.dof_dd; E.dof_doppler = P.dof_dop; E.dof_lghtcrv = P.dof_lc;
  E.bad_objfactor = par->bad_objfactor;
  std::copy(P.base.begin(), P.base.end(), h_params.begin());
  std::copy(P.base.begin(), P.base.end(), h_params.begin() + (size_t)pb_stride);
  if (h_pin) memcpy(h_pin, h_params.data(), sizeof(double) * h_params.size());
  nsinc2 = par->nsinc2;
  sinc2width = par->sinc2width;
  mask_tol = par->mask_tol;
  mincosine_seen = par->mincosine_seen;
  exclude_seen = par->exclude_seen;
  SHP_CHECK(hipMemcpy(d_views, h_views.data(), sizeof(ViewRec) * nviews, hipMemcpyHostToDevice));
  SHP_CHECK(hipMemcpy(d_frames, h_frames.data(), sizeof(FrameRec) * nframes, hipMemcpyHostToDevice));
  SHP_CHECK(hipMemcpy(d_evals, h_evals.data(), sizeof(EvalRec), hipMemcpyHostToDevice));
  if (!P.lc_x.empty()) SHP_CHECK(hipMemcpy(d_lc_x, P.lc_x.data(), sizeof(double) * P.lc_x.size(), hipMemcpyHostToDevice));
  if (!P.lc_t.empty()) SHP_CHECK(hipMemcpy(d_lc_t, P.lc_t.data(), sizeof(double) * P.lc_t.size(), hipMemcpyHostToDevice));
  if (opt.host_geometry && (rc = upload_host_geometry(h_eval_slot[0]))) {
    g_packs.clear();
    return rc;
  }
  // o2 of every frame from the observations uploaded at build
  hipLaunchKernelGGL(k_frame_o2, dim3(std::max(1, nframes)), dim3(256), 0, stream, d_frames, nframes, d_obs, d_oov);
  SHP_CHECK(hipGetLastError());
  g_packs.clear();
  return 0;
}

int Engine::set_values(int e, const std::vector<double> &values) {
  if (e < 0 || e >= nevals || values.size() != eval_targets[e].size()) return -20;
  // a replay in flight reads the pinned copy in its first step
  SHP_CHECK(hipStreamSynchronize(stream));
  for (size_t q = 0; q < values.size(); q++) {
    const ParamTarget &T = eval_targets[e][q];
    if (T.kind == 0) {
      h_params[(size_t)pb_stride * e + T.offset] = values[q];
      if (h_pin) h_pin[(size_t)pb_stride * e + T.offset] = values[q];
    } else {
      int sl = h_eval_slot[e];
      SHP_CHECK(hipMemcpyAsync(h_slots[sl].r + T.offset, &values[q], sizeof(double), hipMemcpyHostToDevice, stream));
      SHP_CHECK(hipStreamSynchronize(stream));
    }
  }
  return 0;
}

int Engine::upload_params() {
  const double *src = h_pin ? h_pin : h_params.data();
  SHP_CHECK(hipMemcpyAsync(d_params, src, sizeof(double) * h_params.size(), hipMemcpyHostToDevice, stream));
  return 0;
}

static inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

// The penalty list of par (penalties.c): every type must be one the kernels
// evaluate (-17 otherwise) and the side list must be present when a penalty
// reads sides (-18).  Workgroups per slot for the moment reductions grow with
// the facet count up to 64; the count is fixed for the engine's life, so the
// sums repeat from replay to replay.
int Engine::setup_penalties(struct par_t *par) {
  pen = PenDev{};
  pen_need = 0;
  pen_law_need = 0;
  // the joint adjustments of vary_params are not part of the batched objective
  if (opt.penalties && (par->vary_delcor0 || par->vary_radalb || par->vary_optalb || par->vary_dopscale)) return -19;
  if (!opt.penalties || par->pen.n <= 0 || opt.views_only) return 0;
  const int n = par->pen.n;
  std::vector<int> type(n);
  std::vector<double> weight(n);
  for (int i = 0; i < n; i++) {
    type[i] = par->pen.type[i + 1];
    weight[i] = par->pen.weight[i + 1];
    if (!pen_supported(type[i]) && !pen_supported_laws(type[i])) return -17;
    if (pen_supported_laws(type[i])) pen_law_need |= 1 << pen_law_slot(type[i]);
    // rdev needs a VERTEX component (penalties.c:725-726 bails out otherwise)
    if (type[i] == RDEV)
      for (auto &P : g_packs)
        if (P.comp_type != VERTEX) return -17;
    pen_need |= pen_needs(type[i]);
  }
  if (pen_need & PEN_NEED_SIDES)
    for (auto &G : h_slots)
      if (G.ns <= 0 || !G.sides) return -18;
  pen_nb = std::max(1, std::min(64, (int)cdiv(max_nf, 2048)));
  int rc;
  int *dt;
  double *dw;
  if ((rc = upload(this, &dt, type))) return rc;
  if ((rc = upload(this, &dw, weight))) return rc;
  pen.n = n;
  pen.type = dt;
  pen.weight = dw;
  if ((rc = alloc((void **)&pen.base, sizeof(double) * (size_t)n * std::max(1, nevals)))) return rc;
  if ((rc = alloc((void **)&pen.total, sizeof(double) * std::max(1, nevals)))) return rc;
  if ((rc = alloc((void **)&d_mom, sizeof(double) * MOM_N * (size_t)nslots))) return rc;
  if ((rc = alloc((void **)&d_mom_part, sizeof(double) * std::max(MP1_N, MP2_N) * (size_t)nslots * pen_nb))) return rc;
  pen.mom = d_mom;
  if (pen_law_need) {
    // every model must carry a law the requested penalties read (the
    // reference bails out otherwise); noncosine's abscissae from par, set up
    // from the first tabular law (every tabular law must have its length)
    noncos_n = 0;
    for (const auto &J : h_photo_jobs)
      if (J.radar && J.kind == PHOTO_TABLE && !noncos_n) noncos_n = J.n - 1;
    for (int e = 0; e < nevals; e++) {
      int have = 0;
      for (int q = 0; q < h_evals[e].photo_njobs; q++) {
        const PhotoJob &J = h_photo_jobs[h_evals[e].photo_job0 + q];
        const bool loc = J.kind == PHOTO_HARM || J.kind == PHOTO_INHO;
        if (loc && !J.radar) have |= 1 << 0;
        if (loc && !J.radar && J.type != HARMHAPKE) have |= 1 << 1;  // optalbvar counts no harmonic-Hapke facet
        if (loc && !J.radar && (J.type == HARMHAPKE || J.type == INHOHAPKE)) have |= (1 << 2) | (1 << 3);
        if (loc && J.radar) have |= (1 << 4) | (1 << 5) | (1 << 6) | (1 << 7);
        if (J.radar && J.kind == PHOTO_TABLE) {
          have |= 1 << 8;
          if (!par->pen.x_noncosine || J.n != noncos_n + 1) return -17;
        }
      }
      if ((pen_law_need & have) != pen_law_need) return -17;
    }
    if (pen_law_need >> 8 & 1) {
      std::vector<double> xc(noncos_n + 1);
      for (int i = 0; i < noncos_n; i++) xc[i] = par->pen.x_noncosine[i];
      xc[noncos_n] = par->pen.xmean_noncosine;
      if ((rc = upload(this, &d_noncos, xc))) return rc;
    }
    if ((rc = alloc((void **)&d_pen_law, sizeof(double) * PEN_LAW_N * (size_t)std::max(1, nevals)))) return rc;
    pen.law = d_pen_law;
  }
  return 0;
}

void Engine::enqueue_penalties() {
  if (!pen.n) return;
  const bool pa = (pen_need & PEN_NEED_PA) != 0;
  SHP_LAUNCH(*this, "k_moments_part", k_moments_part, dim3(pen_nb, nslots), dim3(256), 0, stream, d_slots, d_mom_part, pen_nb);
  SHP_LAUNCH(*this, "k_moments_fin", k_moments_fin, dim3(nslots), dim3(64), 0, stream, d_slots, d_params, d_mom_part, pen_nb,
             d_mom, nullptr, pa ? 1 : 0);
  if (pen_need & PEN_NEED_VERTS) {
    SHP_LAUNCH(*this, "k_pen_vert_part", k_pen_vert_part, dim3(pen_nb, nslots), dim3(256), 0, stream, d_slots, d_mom, d_mom_part,
               pen_nb, pen_need);
    SHP_LAUNCH(*this, "k_pen_vert_fin", k_pen_vert_fin, dim3(nslots), dim3(64), 0, stream, d_slots, d_mom_part, pen_nb, d_mom,
               pen_need);
  }
  if (pen_law_need)
    SHP_LAUNCH(*this, "k_pen_laws", k_pen_laws, dim3(std::max(1, nevals)), dim3(256), 0, stream, d_evals, nevals,
               d_eval_slot, d_slots, d_photo_jobs, d_lawloc, d_noncos, pen_law_need, d_pen_law);
  SHP_LAUNCH(*this, "k_pen_eval", k_pen_eval, dim3(std::max(1, nevals)), dim3(64), 0, stream, d_evals, nevals, d_eval_slot,
             d_params, pen);
}

int Engine::enqueue() {
  const int T = 256;
  // an error left by an earlier, already reported call must not fail this one
  (void)hipGetLastError();
  dim3 gv(cdiv(nviews, 64)), bv(64);
  SHP_LAUNCH(*this, "k_zero_counters", k_zero_counters, dim3(1), dim3(64), 0, stream, pd.nstale, 8);
  if (opt.facet_map) {  // owners of this replay only
    int *fm = pd.fmap ? pd.fmap : pd.f;
    if (fm) SHP_LAUNCH(*this, "k_fill_i32", k_fill_i32, dim3(1024), dim3(256), 0, stream, fm, std::max(1LL, total_pix), -1);
  }
  SHP_LAUNCH(*this, "k_setup_views", k_setup_views, gv, bv, 0, stream, d_views, nviews, d_params);
  const bool photo = !h_photo_jobs.empty() && !opt.views_only;
  if (!opt.host_geometry) {
    int max_nv = 0;
    for (auto &G : h_slots) max_nv = std::max(max_nv, G.nv);
    SHP_LAUNCH(*this, "k_realize_vertices", k_realize_vertices, dim3(cdiv(max_nv, T), nslots), dim3(T), 0, stream, d_slots, nslots,
                       d_params, d_slot_flags, d_slot_logf);
    if (any_harmonic)
      SHP_LAUNCH(*this, "k_realize_harmonic_bad", k_realize_harmonic_bad, dim3(nslots), dim3(256), 0, stream, d_slots, d_params,
                 d_slot_flags, d_slot_logf);
    SHP_LAUNCH(*this, "k_realize_facets", k_realize_facets, dim3(cdiv(max_nf, T), nslots), dim3(T), 0, stream, d_slots, nslots);
    SHP_LAUNCH(*this, "k_realize_vnormals", k_realize_vnormals, dim3(cdiv(max_nv, T), nslots), dim3(T), 0, stream, d_slots, nslots);
  }
  if (photo)
    SHP_LAUNCH(*this, "k_realize_photo", k_realize_photo, dim3(photo_nblk, (unsigned)h_photo_jobs.size()), dim3(T), 0, stream,
               d_photo_jobs, d_params, d_slots, d_lawloc, d_job_flag, d_job_log);
  enqueue_penalties();
  if (opt.method == METHOD_MIXED) return enqueue_mixed(*this);
  dim3 gf(cdiv(max_nf, T), nviews), gp(cdiv(max_pix, T), nviews), bt(T);
  // observer view
  SHP_LAUNCH(*this, "k_raster64/1", k_raster64, gf, bt, 0, stream, d_views, d_slots, pd, 0, 1);
  SHP_LAUNCH(*this, "k_raster64/2", k_raster64, gf, bt, 0, stream, d_views, d_slots, pd, 0, 2);
  SHP_LAUNCH(*this, "k_resolve64", k_resolve64, gp, bt, 0, stream, d_views, d_slots, pd, 0, 0, 1, 0, 0);
  SHP_LAUNCH(*this, "k_raster64/3", k_raster64, gf, bt, 0, stream, d_views, d_slots, pd, 0, 3);
  SHP_LAUNCH(*this, "k_replay64", k_replay64, dim3(std::min(cdiv(pd.stale_cap, T), 2048u)), bt, 0, stream, d_views, d_slots, pd);
  if (any_lc) {
    SHP_LAUNCH(*this, "k_zero_counters", k_zero_counters, dim3(1), dim3(64), 0, stream, pd.nstale, 1);
    SHP_LAUNCH(*this, "k_raster64/1/sun", k_raster64, gf, bt, 0, stream, d_views, d_slots, pd, 1, 1);
    SHP_LAUNCH(*this, "k_raster64/2/sun", k_raster64, gf, bt, 0, stream, d_views, d_slots, pd, 1, 2);
    SHP_LAUNCH(*this, "k_resolve64", k_resolve64, gp, bt, 0, stream, d_views, d_slots, pd, 1, 0, 1, 0, 0);
    SHP_LAUNCH(*this, "k_posmask64", k_posmask64, gp, bt, 0, stream, d_views, pd, mask_tol);
    SHP_LAUNCH(*this, "k_photo64", k_photo64, gp, bt, 0, stream, d_views, pd, d_params);
  }
  SHP_LAUNCH(*this, "k_mark_seen", k_mark_seen, gp, bt, 0, stream, d_views, d_slots, pd, mincosine_seen);
  int max_n = 0;
  for (auto &V : h_views) max_n = std::max(max_n, V.n);
  dim3 gs(syn_blocks(max_n), nviews);
  if (any_dd) SHP_LAUNCH(*this, "k_deldop64", k_deldop64, gs, bt, 0, stream, d_views, d_frames, pd, d_fit, d_params);
  if (any_dop) SHP_LAUNCH(*this, "k_doppler64", k_doppler64, gs, bt, 0, stream, d_views, d_frames, pd, d_fit, d_overflow, d_params);
  SHP_LAUNCH(*this, "k_finish_views", k_finish_views, gv, bv, 0, stream, d_views, nviews, d_overflow, d_frames, d_view_out);
  if (any_lc)
    SHP_LAUNCH(*this, "k_spline", k_spline, dim3(nframes), bt, 0, stream, d_frames, nframes, d_view_out, d_views, d_lc_x,
                       d_lc_t, d_lc_y, d_lc_y2, d_fit);
  SHP_LAUNCH(*this, "k_chi2_frames64", k_chi2_frames64, dim3(nframes), bt, 0, stream, d_frames, nframes, d_fit, d_obs, d_oov,
                     d_view_out, opt.keep_fit ? 0 : 1, 0, 0);
  SHP_LAUNCH(*this, "k_objective", k_objective, dim3(std::max(1, nevals)), dim3(256), 0, stream, d_evals, nevals, d_frames, d_views,
                     d_view_out, d_slot_flags, d_slot_logf, d_eval_slot, pen.n ? pen.total : nullptr, d_job_flag,
                     d_job_log, photo_nblk);
  SHP_CHECK(hipGetLastError());
  return 0;
}

int Engine::capture_graph() {
  if (graph_exec) { (void)hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
  if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
  SHP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  int rc = upload_params();
  if (!rc) rc = enqueue();
  hipError_t e = hipStreamEndCapture(stream, &graph);
  if (rc) return rc;
  SHP_CHECK(e);
  SHP_CHECK(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
  return 0;
}

int Engine::launch_graph() {
  SHP_CHECK(hipGraphLaunch(graph_exec, stream));
  return 0;
}

int Engine::sync() {
  SHP_CHECK(hipStreamSynchronize(stream));
  return 0;
}

void Engine::tbegin(const char *name) {
  if (!time_kernels) return;
  if ((size_t)(2 * t_used + 2) > t_events.size()) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    t_events.push_back(a);
    t_events.push_back(b);
  }
  if ((int)t_names.size() <= t_used) t_names.resize(t_used + 1);
  t_names[t_used] = name;
  (void)hipEventRecord(t_events[2 * t_used], stream);
}

void Engine::tend() {
  if (!time_kernels || (size_t)(2 * t_used + 2) > t_events.size()) return;
  (void)hipEventRecord(t_events[2 * t_used + 1], stream);
  t_used++;
}

int Engine::collect_kernel_times() {
  SHP_CHECK(hipStreamSynchronize(stream));
  for (int q = 0; q < t_used; q++) {
    float ms = 0.0f;
    SHP_CHECK(hipEventElapsedTime(&ms, t_events[2 * q], t_events[2 * q + 1]));
    size_t k = 0;
    while (k < kt_names.size() && kt_names[k] != t_names[q]) k++;
    if (k == kt_names.size()) {
      kt_names.push_back(t_names[q]);
      kt_ms.push_back(0.0);
      kt_count.push_back(0);
    }
    kt_ms[k] += ms;
    kt_count[k]++;
  }
  t_used = 0;
  return 0;
}

int Engine::download_evals(std::vector<EvalRec> &out) {
  out.resize(nevals);
  SHP_CHECK(hipMemcpyAsync(out.data(), d_evals, sizeof(EvalRec) * nevals, hipMemcpyDeviceToHost, stream));
  SHP_CHECK(hipStreamSynchronize(stream));
  return 0;
}
int Engine::download_penalties(int e, std::vector<double> &base, double *total) {
  base.assign(pen.n, 0.0);
  if (!pen.n) { if (total) *total = 0.0; return 0; }
  SHP_CHECK(hipMemcpyAsync(base.data(), pen.base + (size_t)e * pen.n, sizeof(double) * pen.n, hipMemcpyDeviceToHost, stream));
  if (total) SHP_CHECK(hipMemcpyAsync(total, pen.total + e, sizeof(double), hipMemcpyDeviceToHost, stream));
  SHP_CHECK(hipStreamSynchronize(stream));
  return 0;
}
int Engine::download_frames(std::vector<FrameRec> &out) {
  out.resize(nframes);
  SHP_CHECK(hipMemcpyAsync(out.data(), d_frames, sizeof(FrameRec) * nframes, hipMemcpyDeviceToHost, stream));
  SHP_CHECK(hipStreamSynchronize(stream));
  return 0;
}
int Engine::download_view_out(std::vector<double> &out) {
  out.resize((size_t)VIEW_OUT * nviews);
  SHP_CHECK(hipMemcpyAsync(out.data(), d_view_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, stream));
  SHP_CHECK(hipStreamSynchronize(stream));
  return 0;
}
int Engine::download_views(std::vector<ViewRec> &out) {
  out.resize(nviews);
  SHP_CHECK(hipMemcpyAsync(out.data(), d_views, sizeof(ViewRec) * nviews, hipMemcpyDeviceToHost, stream));
  SHP_CHECK(hipStreamSynchronize(stream));
  return 0;
}
int Engine::error_flags() {
  int h[4] = {0, 0, 0, 0};
  if (hipMemcpy(h, pd.nstale, sizeof(int) * 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
  last_nstale = h[0];
  return h[1];
}

// copy a device POS array of one view into an NR [-n..n][-n..n] host matrix
template <class T>
static int pos_to_host(hipStream_t st, T **dst, const T *src, long long off, int n) {
  int w = 2 * n + 1;
  std::vector<T> tmp((size_t)w * w);
  SHP_CHECK(hipMemcpyAsync(tmp.data(), src + off, sizeof(T) * tmp.size(), hipMemcpyDeviceToHost, st));
  SHP_CHECK(hipStreamSynchronize(st));
  for (int i = -n; i <= n; i++) memcpy(&dst[i][-n], &tmp[(size_t)(i + n) * w], sizeof(T) * w);
  return 0;
}

int Engine::writeback(int e, struct par_t *par, struct mod_t *mod, struct dat_t *dat, int write_pos, int accumulate) {
  std::vector<FrameRec> fr;
  std::vector<EvalRec> ev;
  std::vector<double> vo;
  std::vector<ViewRec> vw;
  int rc;
  if ((rc = download_frames(fr)) || (rc = download_evals(ev)) || (rc = download_view_out(vo)) ||
      (rc = download_views(vw)))
    return rc;
  // the fit images through the thread's pinned stage (plain host memory if
  // it cannot be pinned)
  std::vector<double> plain;
  double *fitp = host_stage((size_t)std::max(1LL, fit_len));
  if (!fitp) {
    plain.resize((size_t)std::max(1LL, fit_len));
    fitp = plain.data();
  }
  SHP_CHECK(hipMemcpy(fitp, d_fit, sizeof(double) * fit_len, hipMemcpyDeviceToHost));
  const double *fit = fitp;
  const EvalRec &E = ev[e];
  // the radar fit images into the frames' row arrays, frames in parallel
  host_parallel_for(E.nframes, [&](int q) {
    const FrameRec &F = fr[E.first_frame + q];
    struct set_t *set = &dat->set[F.set];
    const double *img = fit + F.off;
    const double inv = 1.0 / F.nviews;
    if (F.type == VT_DELAY) {
      struct deldopfrm_t *h = &set->desc.deldop.frame[F.frm];
      for (int i = 1; i <= F.ndel; i++)
        for (int j = 1; j <= F.ndop; j++) {
          double v = img[(size_t)(i - 1) * F.ndop + (j - 1)];
          h->fit[i][j] = F.nviews > 1 ? v * inv : v;
        }
    } else if (F.type == VT_DOPPLER) {
      struct dopfrm_t *h = &set->desc.doppler.frame[F.frm];
      for (int j = 1; j <= F.ndop; j++) h->fit[j] = F.nviews > 1 ? img[j - 1] * inv : img[j - 1];
    }
  });
  for (int q = 0; q < E.nframes; q++) {
    const FrameRec &F = fr[E.first_frame + q];
    struct set_t *set = &dat->set[F.set];
    const double *img = fit + F.off;
    if (F.type == VT_DELAY) {
      struct deldopfrm_t *h = &set->desc.deldop.frame[F.frm];
      h->overflow_o2 = h->overflow_m2 = h->overflow_xsec = h->overflow_delmean = h->overflow_dopmean = 0.0;
      h->badradar_logfactor = 0.0;
    } else if (F.type == VT_DOPPLER) {
      struct dopfrm_t *h = &set->desc.doppler.frame[F.frm];
      h->overflow_o2 = F.overflow_o2; h->overflow_m2 = F.overflow_m2;
      h->overflow_xsec = F.overflow_xsec; h->overflow_dopmean = F.overflow_dopmean;
    } else if (F.type == VT_LGHTCRV) {
      struct lghtcrv_t *L = &set->desc.lghtcrv;
      for (int i = 1; i <= F.lc_n; i++) L->fit[i] = img[i - 1];
      std::vector<double> y(F.lc_ncalc + 1), y2(2 * (F.lc_ncalc + 1));
      SHP_CHECK(hipMemcpy(y.data(), d_lc_y + F.lc_y_off, sizeof(double) * y.size(), hipMemcpyDeviceToHost));
      SHP_CHECK(hipMemcpy(y2.data(), d_lc_y2 + F.lc_y2_off, sizeof(double) * y2.size(), hipMemcpyDeviceToHost));
      for (int i = 1; i <= F.lc_ncalc; i++) { L->y[i] = y[i]; L->y2[i] = y2[i]; }
    }
    // view-level side effects; the frame keeps the values of view v0
    for (int v = 0; v < F.nviews && F.type != VT_LGHTCRV; v++) {
      const ViewRec &V = vw[F.first_view + v];
      const double *o = vo.data() + (size_t)(F.first_view + v) * VIEW_OUT;
      if (F.type == VT_DELAY) {
        struct deldopfrm_t *h = &set->desc.deldop.frame[F.frm];
        struct deldopview_t *w = &h->view[V.v];
        w->km2Hz = o[OUT_KM2HZ];
        if (!opt.host_spin) {
          for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) w->ae[i][j] = V.ae[3 * i + j];
          for (int j = 0; j < 3; j++) { w->spin[j] = V.spin[j]; w->intspin[j] = V.intspin[j]; }
          w->deloff = V.deloff; w->dopoff = V.dopoff;
        }
        if (V.is_v0) {
          h->idellim[0] = (int)o[OUT_IDELLIM]; h->idellim[1] = (int)o[OUT_IDELLIM + 1];
          h->idoplim[0] = (int)o[OUT_IDOPLIM]; h->idoplim[1] = (int)o[OUT_IDOPLIM + 1];
          h->dellim[0] = o[OUT_DELLIM]; h->dellim[1] = o[OUT_DELLIM + 1];
          h->doplim[0] = o[OUT_DOPLIM]; h->doplim[1] = o[OUT_DOPLIM + 1];
          h->badradar_logfactor = o[OUT_BADRADAR_LOG];
          h->pos.posbnd_logfactor = o[OUT_POSBND_LOG];
        }
      } else {
        struct dopfrm_t *h = &set->desc.doppler.frame[F.frm];
        struct dopview_t *w = &h->view[V.v];
        w->km2Hz = o[OUT_KM2HZ];
        if (!opt.host_spin) {
          for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) w->ae[i][j] = V.ae[3 * i + j];
          for (int j = 0; j < 3; j++) { w->spin[j] = V.spin[j]; w->intspin[j] = V.intspin[j]; }
          w->dopoff = V.dopoff;
        }
        if (V.is_v0) {
          h->idoplim[0] = (int)o[OUT_IDOPLIM]; h->idoplim[1] = (int)o[OUT_IDOPLIM + 1];
          h->doplim[0] = o[OUT_DOPLIM]; h->doplim[1] = o[OUT_DOPLIM + 1];
          h->badradar_logfactor = o[OUT_BADRADAR_LOG];
          h->pos.posbnd_logfactor = o[OUT_POSBND_LOG];
        }
      }
    }
  }
  if (!accumulate) {
    // par flags (calc_fits.c:349-354, 407-409)
    par->posbnd = (unsigned char)E.posbnd;
    par->badradar = (unsigned char)E.badradar;
    par->badposet = 0;
    par->posbnd_logfactor = E.posbnd_logfactor;
    par->badradar_logfactor = E.badradar_logfactor;
    par->badposet_logfactor = 0.0 / dat->dof_poset;
  } else {
    // one dataset: flags raised and un-normalized log factors added
    // (calc_fits.c:509-513, :539-550, :1650-1660, :1947)
    for (int q = 0; q < E.nframes; q++) {
      const FrameRec &F = fr[E.first_frame + q];
      if (F.type == VT_LGHTCRV) {
        double lf = 0.0;
        for (int i = 0; i < F.lc_ncalc; i++) {
          const double *o = vo.data() + (size_t)(F.first_view + i) * VIEW_OUT;
          if (o[OUT_OUTBND] != 0.0) { par->posbnd = 1; lf += 0.5 * o[OUT_POSBND_LOG]; }
          if (o[OUT_OUTBND + 1] != 0.0) { par->posbnd = 1; lf += 0.5 * o[OUT_POSBND_LOG + 1]; }
        }
        par->posbnd_logfactor += F.dof * (lf / F.lc_ncalc);
        continue;
      }
      for (int v = 0; v < F.nviews; v++) {
        const ViewRec &V = vw[F.first_view + v];
        const double *o = vo.data() + (size_t)(F.first_view + v) * VIEW_OUT;
        if (V.is_v0 && o[OUT_OUTBND] != 0.0) {
          par->posbnd = 1;
          par->posbnd_logfactor += F.dof * o[OUT_POSBND_LOG];
        }
        if (o[OUT_BADRADAR] != 0.0) {
          par->badradar = 1;
          par->badradar_logfactor += F.dof * o[OUT_BADRADAR_LOG] / F.nviews;
        }
      }
    }
  }
  // seen flags (calc_fits resets them first; one dataset only raises them)
  const GeomSlot &G = h_slots[h_eval_slot[e]];
  std::vector<unsigned char> seen(G.nf);
  SHP_CHECK(hipMemcpy(seen.data(), G.seen, G.nf, hipMemcpyDeviceToHost));
  struct vertices_t *rv = &mod->shape.comp[0].real;
  for (int f = 0; f < G.nf && f < rv->nf; f++) rv->f[f].seen = accumulate ? (rv->f[f].seen | seen[f]) : seen[f];
  // POS arrays of every view (with pos_scope GLOBAL the last view's survive)
  if (write_pos && pd.z) {
    for (int v = 0; v < nviews; v++) {
      const ViewRec &V = vw[v];
      // a smeared frame's POS holds its view v0, which the CPU's view loop
      // (v0 + 1, ..., v0 + nviews mod nviews) renders last
      if (V.eval != e || !V.is_v0) continue;
      struct set_t *set = &dat->set[V.set];
      struct pos_t *pos = V.type == VT_DELAY ? &set->desc.deldop.frame[V.frm].pos
                          : V.type == VT_DOPPLER ? &set->desc.doppler.frame[V.frm].pos
                                                 : &set->desc.lghtcrv.rend[V.frm + 1].pos;
      if ((rc = pos_to_host(stream, pos->z, pd.z, V.pos_off, V.n))) return rc;
      if ((rc = pos_to_host(stream, pos->cose, pd.cose, V.pos_off, V.n))) return rc;
      if ((rc = pos_to_host(stream, pos->f, pd.f, V.pos_off, V.n))) return rc;
      if (V.bistatic) {
        if ((rc = pos_to_host(stream, pos->cosi, pd.cosi, V.pos_off, V.n))) return rc;
        if ((rc = pos_to_host(stream, pos->b, pd.b, V.pos_off, V.n))) return rc;
        if ((rc = pos_to_host(stream, pos->zill, pd.zill, V.pos_off, V.n))) return rc;
        if ((rc = pos_to_host(stream, pos->cosill, pd.cosill, V.pos_off, V.n))) return rc;
        if ((rc = pos_to_host(stream, pos->fill, pd.fill, V.pos_off, V.n))) return rc;
      }
      // Pixels outside the rectangles are blank sky, as posclr leaves them
      // (the device arrays may hold an earlier call's pixels there); body /
      // component maps of a single-body, single-component model
      // (posvis.c:284-286).
      for (int i = -V.n; i <= V.n; i++)
        for (int j = -V.n; j <= V.n; j++) {
          if (i < V.xlim[0] || i > V.xlim[1] || j < V.ylim[0] || j > V.ylim[1]) {
            pos->z[i][j] = -1.0e20;
            pos->cose[i][j] = 0.0;
            pos->f[i][j] = -1;
            if (V.bistatic) { pos->cosi[i][j] = 0.0; pos->b[i][j] = 0.0; }
          }
          // brightness only where lit and illuminated (apply_photo.c:86-93)
          if (V.bistatic && !(pos->cose[i][j] > 0.0 && pos->cosi[i][j] > 0.0)) pos->b[i][j] = 0.0;
          if (V.bistatic && (i < V.xlim2[0] || i > V.xlim2[1] || j < V.ylim2[0] || j > V.ylim2[1])) {
            pos->zill[i][j] = -1.0e20;
            pos->cosill[i][j] = 0.0;
            pos->fill[i][j] = -1;
          }
          int own = pos->f[i][j] >= 0 ? 0 : -1;
          pos->body[i][j] = pos->comp[i][j] = own;
          if (V.bistatic) pos->bodyill[i][j] = pos->compill[i][j] = pos->fill[i][j] >= 0 ? 0 : -1;
        }
      pos->xlim[0] = V.xlim[0]; pos->xlim[1] = V.xlim[1]; pos->ylim[0] = V.ylim[0]; pos->ylim[1] = V.ylim[1];
      if (V.bistatic) {
        pos->xlim2[0] = V.xlim2[0]; pos->xlim2[1] = V.xlim2[1]; pos->ylim2[0] = V.ylim2[0]; pos->ylim2[1] = V.ylim2[1];
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) { pos->ae[i][j] = V.ae[3 * i + j]; pos->oe[i][j] = V.oe[3 * i + j]; }
    }
  }
  return 0;
}

}  // namespace shp
