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

#include <atomic>
#include <list>
#include <memory>
#include <mutex>
#include <vector>

#include "common.h"
#include "engine_host.h"
#include "shape_hip.h"

using namespace shp;

static std::atomic<int> g_err{0};
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
// Device arenas kept across calls.  A fit calls calc_fits (or, under the
// HostMultiThreading driver, calc_deldop / calc_doppler / calc_lghtcrv per
// dataset) on the same structs again and again with new parameter values.
// Each call site -- (entry, model, dataset) -- keeps its engine: built on
// the first call, refreshed from the structs on later ones (views, spin
// state, frame scalars, parameters, realized geometry), rebuilt when the
// structs' shape or the model's topology changed.  Engines not in use are
// freed, least recently used first, when the cache holds more than its
// device-memory budget (a quarter of the device); shape_hip_release_cache
// frees them all.
// ---------------------------------------------------------------------------
namespace {
enum { DK_FITS = 0, DK_DELDOP = 1, DK_DOPPLER = 2, DK_LGHTCRV = 3 };

struct CacheSlot {
  int kind = 0;
  const void *mod = nullptr, *key = nullptr;
  unsigned long long topo = 0;
  std::unique_ptr<Engine> E;
  unsigned long long tick = 0;
  bool busy = false;
};
std::mutex g_cache_mu;
std::list<CacheSlot> g_cache;
unsigned long long g_tick = 0;
size_t g_cache_budget = 0;

size_t cache_budget() {
  if (!g_cache_budget) {
    size_t fr = 0, tot = 0;
    g_cache_budget = hipMemGetInfo(&fr, &tot) == hipSuccess && tot ? tot / 4 : (size_t)32 << 30;
  }
  return g_cache_budget;
}

// FNV-1a over the model's facet corners and vertex -> facet lists
unsigned long long topo_hash(const struct mod_t *mod) {
  unsigned long long h = 1469598103934665603ULL;
  auto mix = [&h](long long v) { h = (h ^ (unsigned long long)v) * 1099511628211ULL; };
  mix(mod->shape.ncomp);
  for (int c = 0; c < mod->shape.ncomp; c++) {
    const struct vertices_t *rv = &mod->shape.comp[c].real;
    mix(rv->nv);
    mix(rv->nf);
    for (int f = 0; f < rv->nf; f++) { mix(rv->f[f].v[0]); mix(rv->f[f].v[1]); mix(rv->f[f].v[2]); }
    for (int i = 0; i < rv->nv; i++) {
      mix(rv->v[i].naf);
      for (int q = 0; q < rv->v[i].naf; q++) mix(rv->v[i].af[q]);
    }
  }
  return h;
}

// free idle engines, least recently used first, down to `keep` bytes
// (called with g_cache_mu held)
void cache_trim(size_t keep) {
  for (;;) {
    size_t held = 0;
    for (auto &c : g_cache)
      if (c.E) held += c.E->dev_bytes;
    if (held <= keep) return;
    auto victim = g_cache.end();
    for (auto it = g_cache.begin(); it != g_cache.end(); ++it)
      if (!it->busy && it->E && (victim == g_cache.end() || it->tick < victim->tick)) victim = it;
    if (victim == g_cache.end()) return;
    g_cache.erase(victim);
  }
}

// The engine of this call site, built or refreshed; nullptr and *rc on failure.
CacheSlot *acquire(int kind, struct par_t *par, struct mod_t *mod, struct dat_t *dat, const void *key, int only_set,
                   int *rc) {
  const unsigned long long topo = topo_hash(mod);
  CacheSlot *slot = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (auto &c : g_cache)
      if (c.kind == kind && c.mod == mod && c.key == key && !c.busy) { slot = &c; break; }
    if (!slot) {
      g_cache.emplace_back();
      slot = &g_cache.back();
      slot->kind = kind; slot->mod = mod; slot->key = key;
    }
    slot->busy = true;
  }
  *rc = 0;
  bool built = false;
  if (slot->E && slot->topo == topo) {
    const int r = slot->E->refresh(par, mod, dat);
    if (r < 0) *rc = r;
    built = r == 0;
  }
  if (!built && !*rc) {
    slot->E.reset();  // its device memory first
    {
      std::lock_guard<std::mutex> lk(g_cache_mu);
      cache_trim(cache_budget());
    }
    slot->E.reset(new Engine());
    RunOptions o;
    o.method = METHOD_FP64;
    o.host_spin = 1;
    o.host_geometry = 1;
    o.keep_fit = 1;
    o.only_set = only_set;
    std::vector<EvalSpec> ev(1);
    ev[0].model = 0;
    *rc = slot->E->build(par, &mod, &dat, 1, ev, o);
    slot->topo = topo;
  }
  if (*rc) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (auto it = g_cache.begin(); it != g_cache.end(); ++it)
      if (&*it == slot) { g_cache.erase(it); break; }
    return nullptr;
  }
  return slot;
}

void release(CacheSlot *slot) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  slot->busy = false;
  slot->tick = ++g_tick;
  cache_trim(cache_budget());
}

// One evaluation on a drop-in engine: fits and seen flags start from zero
// (calc_fits.c:358-361, :538, :919), as the reference's clrmat and reset do.
int run_fits(Engine &E, struct par_t *par, struct mod_t *mod, struct dat_t *dat, int accumulate) {
  SHP_CHECK(hipMemsetAsync(E.d_fit, 0, sizeof(double) * std::max(1LL, E.fit_len), E.stream));
  const GeomSlot &G = E.h_slots[E.h_eval_slot[0]];
  SHP_CHECK(hipMemsetAsync(G.seen, 0, std::max(1, G.nf), E.stream));
  int rc = E.upload_params();
  if (!rc) rc = E.enqueue();
  if (!rc) rc = E.sync();
  if (!rc && E.error_flags()) rc = -30;
  g_last_nstale = E.last_nstale;
  if (!rc) rc = E.writeback(0, par, mod, dat, g_strict_pos, accumulate);
  return rc;
}

// A dat_t holding one caller dataset at index s (earlier entries are
// placeholders the engine skips), for the one-dataset entries: the engine
// packs it as dataset s, so exclude_seen and the parameter layout keep the
// caller's numbering.  The dataset's description is copied shallowly:
// every array the fits are written to stays the caller's.
struct OneSet {
  std::vector<struct set_t> sets;
  struct dat_t dat;
  OneSet(int s, int type, const void *desc, size_t desc_size) : sets((size_t)s + 1) {
    memset((void *)sets.data(), 0, sizeof(struct set_t) * sets.size());
    for (int k = 0; k < s; k++) sets[k].type = 255;
    sets[s].type = (unsigned char)type;
    memcpy((void *)&sets[s].desc, desc, desc_size);
    memset((void *)&dat, 0, sizeof(dat));
    dat.nsets = s + 1;
    dat.set = sets.data();
    dat.dof = dat.dof_deldop = dat.dof_doppler = dat.dof_poset = dat.dof_lghtcrv = 1.0;
  }
};

void one_set_fits(int kind, int type, const char *name, struct par_t *par, struct mod_t *mod, const void *desc,
                  size_t desc_size, int s) {
  if (s < 0) { set_err(-15, name); return; }
  OneSet one(s, type, desc, desc_size);
  int rc = 0;
  CacheSlot *slot = acquire(kind, par, mod, &one.dat, desc, s, &rc);
  if (slot) {
    rc = run_fits(*slot->E, par, mod, &one.dat, 1);
    release(slot);
  }
  set_err(rc, name);
}
}  // namespace

// ---------------------------------------------------------------------------
// calc_fits (calc_fits.c:340): delay-Doppler, Doppler and lightcurve fits for
// every dataset from the already-realized model and spin state
// ---------------------------------------------------------------------------
extern "C" void calc_fits(struct par_t *par, struct mod_t *mod, struct dat_t *dat) {
  int rc = 0;
  CacheSlot *slot = acquire(DK_FITS, par, mod, dat, dat, -1, &rc);
  if (slot) {
    rc = run_fits(*slot->E, par, mod, dat, 0);
    release(slot);
  }
  set_err(rc, "calc_fits");
}

// calc_fits.c:440 / :830 / :1509: the fits of dataset s alone; the flags in
// par are raised and the un-normalized log factors added to, the facets'
// seen flags or-ed in (the caller resets them, as calc_fits does).
// Disjoint datasets may be computed from several host threads at once.
extern "C" void calc_deldop(struct par_t *par, struct mod_t *mod, struct deldop_t *deldop, int s) {
  one_set_fits(DK_DELDOP, DELAY, "calc_deldop", par, mod, deldop, sizeof(*deldop), s);
}
extern "C" void calc_doppler(struct par_t *par, struct mod_t *mod, struct doppler_t *doppler, int s) {
  one_set_fits(DK_DOPPLER, DOPPLER, "calc_doppler", par, mod, doppler, sizeof(*doppler), s);
}
extern "C" void calc_lghtcrv(struct par_t *par, struct mod_t *mod, struct lghtcrv_t *lghtcrv, int s) {
  one_set_fits(DK_LGHTCRV, LGHTCRV, "calc_lghtcrv", par, mod, lghtcrv, sizeof(*lghtcrv), s);
}

// ---------------------------------------------------------------------------
// vary_params (vary_params.c:87): the means behind bestfit's joint
// adjustments, from the first dataset (the reference's loop runs s < 1): the
// sub-radar point's distance toward Earth and the radar cross section of each
// frame whose delay constant floats or whose calibration is absolute, the
// absolutely calibrated lightcurve's brightness at the observation epochs,
// and cos(sub-radar latitude) where the Doppler scale floats.  The rendering
// and synthesis run on the cached one-dataset engine; one workgroup per frame
// then reduces its POS and fit.  The reference also leaves these frames' fit
// and POS arrays behind as scratch; they are not written back here.
// ---------------------------------------------------------------------------
namespace {
// per frame: [0] max z over the POS rectangle's pixels with cos(scattering) > 0
// (-1e20 if none), [1] the fit's sum plus the overflow cross section
__global__ __launch_bounds__(256) void k_vary_frames(const FrameRec *frames, const ViewRec *views, PosDev pd,
                                                     const double *fit, double *out) {
  const FrameRec &F = frames[blockIdx.x];
  __shared__ double sz[256], sf[256];
  double zmax = -1.0e20, sum = 0.0;
  const long long ncell = F.type == VT_DELAY ? (long long)F.ndel * F.ndop : F.type == VT_DOPPLER ? F.ndop : F.lc_n;
  for (long long q = threadIdx.x; q < ncell; q += 256) sum += fit[F.off + q];
  if (F.type != VT_LGHTCRV) {
    const ViewRec &V = views[F.first_view];
    const int w = 2 * V.n + 1, nx = V.xlim[1] - V.xlim[0] + 1, ny = V.ylim[1] - V.ylim[0] + 1;
    if (nx > 0 && ny > 0)
      for (int q = threadIdx.x; q < nx * ny; q += 256) {
        const int i = V.xlim[0] + q / ny, j = V.ylim[0] + q % ny;
        const long long p = V.pos_off + (long long)(i + V.n) * w + (j + V.n);
        if (pd.cose[p] > 0.0 && pd.z[p] > zmax) zmax = pd.z[p];
      }
  }
  sz[threadIdx.x] = zmax;
  sf[threadIdx.x] = sum;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      sz[threadIdx.x] = sz[threadIdx.x] > sz[threadIdx.x + k] ? sz[threadIdx.x] : sz[threadIdx.x + k];
      sf[threadIdx.x] += sf[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = sz[0];
    out[2 * blockIdx.x + 1] = sf[0] + (F.type == VT_LGHTCRV ? 0.0 : F.overflow_xsec);
  }
}

// cos(sub-radar latitude) of a view: the target-to-Earth vector in body
// coordinates is the third row of oe ae^T (vary_params.c:177-184)
double cos_subradar(double ae[3][3], double oe[3][3]) {
  double t[3];
  for (int j = 0; j < 3; j++) {
    t[j] = 0.0;
    for (int k = 0; k < 3; k++) t[j] += oe[2][k] * ae[j][k];
  }
  return sqrt(t[0] * t[0] + t[1] * t[1]);
}

// render and synthesize dataset 0 on its one-dataset engine; per frame
// (zmax, fit sum + overflow) into out.  Returns 0 or an error code.
int vary_set0(int kind, int type, struct par_t *par, struct mod_t *mod, const void *desc, size_t desc_size,
              std::vector<double> &out) {
  OneSet one(0, type, desc, desc_size);
  int rc = 0;
  CacheSlot *slot = acquire(kind, par, mod, &one.dat, desc, 0, &rc);
  if (!slot) return rc ? rc : -2;
  Engine &E = *slot->E;
  double *d_out = nullptr;
  rc = hipMemsetAsync(E.d_fit, 0, sizeof(double) * std::max(1LL, E.fit_len), E.stream) == hipSuccess ? 0 : -4;
  if (!rc) rc = E.upload_params();
  if (!rc) rc = E.enqueue();
  if (!rc && hipMalloc(&d_out, sizeof(double) * 2 * std::max(1, E.nframes)) != hipSuccess) rc = -2;
  if (!rc) {
    hipLaunchKernelGGL(k_vary_frames, dim3(E.nframes), dim3(256), 0, E.stream, E.d_frames, E.d_views, E.pd, E.d_fit,
                       d_out);
    rc = E.sync();
  }
  if (!rc && E.error_flags()) rc = -30;
  out.assign(2 * (size_t)E.nframes, 0.0);
  if (!rc && hipMemcpy(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost) != hipSuccess) rc = -4;
  if (d_out) (void)hipFree(d_out);
  release(slot);
  return rc;
}
}  // namespace

extern "C" void vary_params(struct par_t *par, struct mod_t *mod, struct dat_t *dat, int action, double *deldop_zmax,
                            double *rad_xsec, double *opt_brightness, double *cos_subradarlat) {
  (void)action;
  double s_zmax = 0.0, s_xsec = 0.0, s_bright = 0.0, s_cos = 0.0;
  int rc = 0;
  if (dat->nsets > 0) {
    struct set_t *set = &dat->set[0];
    std::vector<double> fr;
    if (set->type == DELAY) {
      struct deldop_t *dd = &set->desc.deldop;
      bool any = false;
      for (int f = 0; f < dd->nframes; f++)
        any = any || (par->vary_delcor0 != 0 && dd->delcor.a[0].state != 'c') ||
              (par->vary_radalb != 0 && dd->frame[f].cal.state == 'c');
      if (any && dd->nviews != 1) rc = -19;  // one view per frame (view v0 = view 0)
      if (any && !rc) rc = vary_set0(DK_DELDOP, DELAY, par, mod, dd, sizeof(*dd), fr);
      for (int f = 0; f < dd->nframes && !rc; f++) {
        if (par->vary_delcor0 != 0 && dd->delcor.a[0].state != 'c') s_zmax += fr[2 * f] * dd->frame[f].weight;
        if (par->vary_radalb != 0 && dd->frame[f].cal.state == 'c')
          s_xsec += fr[2 * f + 1] * dd->frame[f].cal.val * dd->frame[f].weight;
        if (par->vary_dopscale != 0 && dd->dopscale.state != 'c')
          s_cos += cos_subradar(dd->frame[f].view[dd->v0].ae, dd->frame[f].view[dd->v0].oe) * dd->frame[f].weight;
      }
    } else if (set->type == DOPPLER) {
      struct doppler_t *dp = &set->desc.doppler;
      bool any = false;
      for (int f = 0; f < dp->nframes; f++) any = any || (par->vary_radalb != 0 && dp->frame[f].cal.state == 'c');
      if (any && dp->nviews != 1) rc = -19;
      if (any && !rc) rc = vary_set0(DK_DOPPLER, DOPPLER, par, mod, dp, sizeof(*dp), fr);
      for (int f = 0; f < dp->nframes && !rc; f++) {
        if (par->vary_radalb != 0 && dp->frame[f].cal.state == 'c')
          s_xsec += fr[2 * f + 1] * dp->frame[f].cal.val * dp->frame[f].weight;
        if (par->vary_dopscale != 0 && dp->dopscale.state != 'c')
          s_cos += cos_subradar(dp->frame[f].view[dp->v0].ae, dp->frame[f].view[dp->v0].oe) * dp->frame[f].weight;
      }
    } else if (set->type == LGHTCRV) {
      struct lghtcrv_t *lc = &set->desc.lghtcrv;
      if (par->vary_optalb != 0 && lc->cal.state == 'c') {
        if (lc->nviews != 1) rc = -19;
        if (!rc) rc = vary_set0(DK_LGHTCRV, LGHTCRV, par, mod, lc, sizeof(*lc), fr);
        if (!rc) s_bright += fr[1] * lc->weight;
      }
    }
  }
  if (rc) {
    set_err(rc, "vary_params");
    return;
  }
  *deldop_zmax = dat->sum_deldop_zmax_weights > 0.0 ? s_zmax / dat->sum_deldop_zmax_weights : 0.0;
  *rad_xsec = dat->sum_rad_xsec_weights > 0.0 ? s_xsec / dat->sum_rad_xsec_weights : 0.0;
  *opt_brightness = dat->sum_opt_brightness_weights > 0.0 ? s_bright / dat->sum_opt_brightness_weights : 0.0;
  *cos_subradarlat = dat->sum_cos_subradarlat_weights > 0.0 ? s_cos / dat->sum_cos_subradarlat_weights : 0.0;
}

extern "C" void shape_hip_release_cache(void) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  cache_trim(0);
}

// number of POS pixels whose cos(scattering) was inherited from an earlier
// claimant in the last calc_fits (the replayed pixels); diagnostic only
extern "C" int shape_hip_last_replayed(void) { return g_last_nstale; }

// ---------------------------------------------------------------------------
// realize_angleoff / realize_omegaoff (realize_spin.c:286, :317): the '='
// state of each component of the datasets' angle and spin offsets takes the
// value of the last earlier dataset whose state is not '='.  Bookkeeping of
// the caller's structs, no arithmetic; the reference's bailout for an '='
// on the first dataset becomes an error code.
// ---------------------------------------------------------------------------
static int resolve_offsets(struct dat_t *dat, int which) {
  for (int j = 0; j <= 2; j++) {
    int src = -1;
    for (int s = 0; s < dat->nsets; s++) {
      struct param_t *p = which ? &dat->set[s].omegaoff[j] : &dat->set[s].angleoff[j];
      if (p->state != '=') src = s;
      else if (src < 0) return -17;
      else p->val = (which ? dat->set[src].omegaoff[j] : dat->set[src].angleoff[j]).val;
    }
  }
  return 0;
}
extern "C" void realize_angleoff(struct dat_t *dat) { set_err(resolve_offsets(dat, 0), "realize_angleoff"); }
extern "C" void realize_omegaoff(struct dat_t *dat) { set_err(resolve_offsets(dat, 1), "realize_omegaoff"); }

// Device buffers of the chi2 and realize_mod drop-ins, grown on demand and
// kept across calls.
namespace {
struct GrowBuf {
  void *p = nullptr;
  size_t cap = 0;
  ~GrowBuf() { if (p) (void)hipFree(p); }
  int get(size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (bytes <= cap) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, bytes) != hipSuccess) return -2;
    cap = bytes;
    return 0;
  }
};
std::mutex g_chi2_mu;
GrowBuf g_c_fr, g_c_fit, g_c_obs, g_c_oov, g_c_set, g_c_tot;
}  // namespace

// ---------------------------------------------------------------------------
// chi2 (chi2.c:240): chi-square of the fits currently held by the caller
// ---------------------------------------------------------------------------

// The chi-square frame records of every dataset in set order, with offsets
// into flat fit / observation / weight arrays (each frame's cells row-major,
// as the drop-ins flatten them); src gets each record's host frame
// (deldopfrm_t, dopfrm_t) or lightcurve (lghtcrv_t).
static int chi2_frames(struct dat_t *dat, std::vector<FrameRec> &fr, std::vector<const void *> &src,
                       long long &ncell) {
  ncell = 0;
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
      return -15;  // plane-of-sky sets are outside this library
    }
  }
  return 0;
}

// o2 per frame, the frames' chi-square terms (overflow terms from the host
// frames, chi2.c:430-431) and the set and total sums, on the null stream
static void chi2_kernels(FrameRec *df, int nf, const double *fit, const double *obs, const double *oov,
                         double *dset, int nsets, double *dtot) {
  hipLaunchKernelGGL(k_frame_o2, dim3(std::max(1, nf)), dim3(256), 0, 0, df, nf, obs, oov);
  hipLaunchKernelGGL(k_chi2_frames64, dim3(std::max(1, nf)), dim3(256), 0, 0, df, nf, (double *)fit, obs, oov,
                     (const double *)nullptr, 0, 1, 0);
  hipLaunchKernelGGL(k_chi2_sets, dim3(1), dim3(64), 0, 0, (const FrameRec *)df, nf, dset, nsets, dtot);
}

// results into the caller's structs (chi2.c:270-300, :430-460)
static double chi2_store(struct dat_t *dat, const std::vector<FrameRec> &fr, const std::vector<double> &setc,
                         double total) {
  for (auto &F : fr) {
    struct set_t *set = &dat->set[F.set];
    if (F.type == VT_DELAY) { set->desc.deldop.frame[F.frm].chi2 = F.chi2; set->desc.deldop.frame[F.frm].cal.val = F.cal; }
    else if (F.type == VT_DOPPLER) { set->desc.doppler.frame[F.frm].chi2 = F.chi2; set->desc.doppler.frame[F.frm].cal.val = F.cal; }
    else set->desc.lghtcrv.cal.val = F.cal;
  }
  for (int s = 0; s < dat->nsets; s++) dat->set[s].chi2 = setc[s];
  dat->chi2 = total;
  return total;
}

extern "C" double chi2(struct par_t *par, struct dat_t *dat, int list_breakdown) {
  (void)par;
  (void)list_breakdown;
  // frame records first (offsets into flat arrays), then the fit,
  // observation and weight arrays copied from the frames' row arrays in
  // parallel
  std::vector<FrameRec> fr;
  std::vector<const void *> src;
  long long ncell = 0;
  if (chi2_frames(dat, fr, src, ncell)) {
    set_err(-15, "chi2 (plane-of-sky sets are outside this library)");
    return dat->chi2;
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
  std::lock_guard<std::mutex> lk(g_chi2_mu);
  GrowBuf &dfr = g_c_fr, &dfit = g_c_fit, &dobs = g_c_obs, &doov = g_c_oov, &dset = g_c_set, &dtot = g_c_tot;
  const size_t nb = sizeof(double) * (size_t)ncell;
  int rc = dfr.get(sizeof(FrameRec) * nf) | dfit.get(nb) | dobs.get(nb) | doov.get(nb) | dset.get(sizeof(double) * nsets) |
           dtot.get(sizeof(double));
  if (rc) { set_err(rc, "chi2 allocation"); return dat->chi2; }
  bool ok = hipMemcpy(dfr.p, fr.data(), sizeof(FrameRec) * nf, hipMemcpyHostToDevice) == hipSuccess &&
            (ncell == 0 || (hipMemcpy(dfit.p, fit, nb, hipMemcpyHostToDevice) == hipSuccess &&
                            hipMemcpy(dobs.p, obs, nb, hipMemcpyHostToDevice) == hipSuccess &&
                            hipMemcpy(doov.p, oov, nb, hipMemcpyHostToDevice) == hipSuccess));
  if (!ok) { set_err(-3, "chi2 upload"); return dat->chi2; }
  chi2_kernels((FrameRec *)dfr.p, nf, (const double *)dfit.p, (const double *)dobs.p, (const double *)doov.p,
               (double *)dset.p, nsets, (double *)dtot.p);
  std::vector<double> setc(nsets);
  double total = 0.0;
  ok = hipDeviceSynchronize() == hipSuccess &&
       hipMemcpy(fr.data(), dfr.p, sizeof(FrameRec) * nf, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(setc.data(), dset.p, sizeof(double) * nsets, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(&total, dtot.p, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) { set_err(-4, "chi2 kernels"); return dat->chi2; }
  return chi2_store(dat, fr, setc, total);
}

// ---------------------------------------------------------------------------
// shape_chi2_noalloc: chi2 on fits, observations and weights the caller
// already holds in device memory, for callers that cannot allocate device
// memory (code under graph capture, serving loops).  The frame records and
// the set and total sums live in `scratch`, one file-scope device array of
// fixed size per GPU; the address of each GPU's copy is taken once, with
// that GPU current, on the first call there.  Every call on a GPU shares its
// array: overlapping calls on one GPU race.  Nothing in it needs clearing
// between calls -- each call writes the frame records it reads, and
// k_chi2_sets zeroes the set sums before adding.
// ---------------------------------------------------------------------------
namespace {
constexpr int NOALLOC_MAX_FRAMES = 8192, NOALLOC_MAX_SETS = 8192;
constexpr size_t NOALLOC_SETS_OFF = sizeof(FrameRec) * (size_t)NOALLOC_MAX_FRAMES;
constexpr size_t NOALLOC_BYTES = NOALLOC_SETS_OFF + sizeof(double) * ((size_t)NOALLOC_MAX_SETS + 1);
static_assert(NOALLOC_SETS_OFF % 256 == 0, "set sums aligned");
__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[NOALLOC_MAX_DEVICES];

// this GPU's copy of scratch (its device address, not the host-side name)
int scratch_ptr(unsigned char **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return -7;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "shape_chi2_noalloc: hipGetSymbolAddress failed: %s\n", hipGetErrorString(e));
      return -7;
    }
    g_scratch_dev[dev] = p;
  }
  *out = (unsigned char *)g_scratch_dev[dev];
  return 0;
}
}  // namespace

extern "C" double shape_chi2_noalloc(struct par_t *par, struct dat_t *dat, const double *d_fit, const double *d_obs,
                                     const double *d_oov) {
  (void)par;
  std::vector<FrameRec> fr;
  std::vector<const void *> src;
  long long ncell = 0;
  if (chi2_frames(dat, fr, src, ncell)) {
    set_err(-15, "shape_chi2_noalloc (plane-of-sky sets are outside this library)");
    return dat->chi2;
  }
  const int nf = (int)fr.size(), nsets = dat->nsets;
  if (nf > NOALLOC_MAX_FRAMES || nsets > NOALLOC_MAX_SETS) {
    fprintf(stderr, "shape_chi2_noalloc: %d frames in %d datasets exceed its scratch (at most %d frames, %d datasets)\n",
            nf, nsets, NOALLOC_MAX_FRAMES, NOALLOC_MAX_SETS);
    set_err(-18, "shape_chi2_noalloc");
    return dat->chi2;
  }
  if (ncell > 0 && (!d_fit || !d_obs || !d_oov)) {
    set_err(-19, "shape_chi2_noalloc (null device array)");
    return dat->chi2;
  }
  unsigned char *sp = nullptr;
  int rc = scratch_ptr(&sp);
  if (rc) { set_err(rc, "shape_chi2_noalloc"); return dat->chi2; }
  FrameRec *df = (FrameRec *)sp;
  double *dset = (double *)(sp + NOALLOC_SETS_OFF), *dtot = dset + NOALLOC_MAX_SETS;
  if (nf > 0 && hipMemcpy(df, fr.data(), sizeof(FrameRec) * nf, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    set_err(-3, "shape_chi2_noalloc upload");
    return dat->chi2;
  }
  chi2_kernels(df, nf, d_fit, d_obs, d_oov, dset, nsets, dtot);
  std::vector<double> setc(nsets);
  double total = 0.0;
  const bool ok = hipStreamSynchronize(0) == hipSuccess &&
                  (nf == 0 || hipMemcpy(fr.data(), df, sizeof(FrameRec) * nf, hipMemcpyDeviceToHost) == hipSuccess) &&
                  (nsets == 0 || hipMemcpy(setc.data(), dset, sizeof(double) * nsets, hipMemcpyDeviceToHost) == hipSuccess) &&
                  hipMemcpy(&total, dtot, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    set_err(-4, "shape_chi2_noalloc kernels");
    return dat->chi2;
  }
  return chi2_store(dat, fr, setc, total);
}

// ---------------------------------------------------------------------------
// realize_mod (realize_mod.c:153): vertex coordinates, facet normals, areas,
// centroids and vertex normals of a single-component VERTEX model
// ---------------------------------------------------------------------------
namespace {
// realize_mod's device arrays, kept across calls; the topology (facet
// corners, vertex -> facet lists) is uploaded again only when the model or
// its topology changes
struct RealizeBufs {
  const void *mod = nullptr;
  unsigned long long topo = 0;
  GrowBuf fv, off, idx, u, a, r, pc, x, vn, fn, fx, ar, x32, vn32, fn32, slot, flag, mom, part, hb;
};
// workgroups per model for the moment reductions (as the engine's)
int moment_blocks(int nf) { return std::max(1, std::min(64, (nf + 2047) / 2048)); }
std::mutex g_realize_mu;
RealizeBufs g_rb;
}  // namespace

extern "C" void realize_mod(struct par_t *par, struct mod_t *mod) {
  if (mod->shape.ncomp != 1) { set_err(-10, "realize_mod"); return; }
  struct comp_t *comp = &mod->shape.comp[0];
  const int ctype = comp->type;
  if (ctype != VERTEX && ctype != ELLIPSE && ctype != OVOID && ctype != HARMONIC) { set_err(-10, "realize_mod"); return; }
  struct vertices_t *rv = &comp->real;
  const int nv = rv->nv, nf = rv->nf;
  const int nhar = ctype == HARMONIC ? comp->desc.har.nhar : 0;
  const int nbasis = ctype == HARMONIC ? (nhar + 1) * (nhar + 1) : 0;
  const int npc = PB_COMP_STRIDE + nbasis;
  std::vector<double> in(7 * (size_t)nv + npc);  // u, a (3 nv each), r (nv), the component block
  double *u = in.data(), *a = u + 3 * (size_t)nv, *r = a + 3 * (size_t)nv, *pc = r + nv;
  for (int i = 0; i < nv; i++) {
    struct vertex_t *v = &rv->v[i];
    if (ctype == VERTEX && v->r.state == '=') v->r.val = rv->v[v->v_mirror].r.val;  // realize_mod.c:450-454
    for (int k = 0; k < 3; k++) { u[3 * i + k] = v->u[k]; a[3 * i + k] = v->a[k]; }
    r[i] = v->r.val;
  }
  // the component block as the engine lays it out; '=' ties of scale factors
  // are applied to the description here as realize_mod.c:455-457 / :425-428 do
  int sf_tie = 0;
  struct param_t *shp3 = ctype == ELLIPSE    ? &comp->desc.ell.two_a
                         : ctype == OVOID    ? &comp->desc.ovoid.two_a
                         : ctype == HARMONIC ? comp->desc.har.scalefactor
                                             : comp->desc.ver.scalefactor;
  for (int j = 0; j < 3; j++) {
    if ((ctype == VERTEX || ctype == HARMONIC) && j > 0 && shp3[j].state == '=') {
      shp3[j].val = shp3[j - 1].val;
      sf_tie |= 1 << j;
    }
    pc[j] = shp3[j].val;
    pc[3 + j] = comp->rot[j].val;
    pc[6 + j] = comp->off[j].val;
  }
  pc[PB_COMP_OVOID_K] = ctype == OVOID ? comp->desc.ovoid.k.val : 0.0;
  for (int l = 0, q = PB_COMP_HAR; l <= nhar && ctype == HARMONIC; l++) {
    pc[q++] = comp->desc.har.a[l][0].val;
    for (int m = 1; m <= l; m++) { pc[q++] = comp->desc.har.a[l][m].val; pc[q++] = comp->desc.har.b[l][m].val; }
  }
  const unsigned long long topo = topo_hash(mod);
  std::lock_guard<std::mutex> lk(g_realize_mu);
  RealizeBufs &B = g_rb;
  const bool new_topo = B.mod != mod || B.topo != topo;
  int rc = B.u.get(sizeof(double) * 3 * nv) | B.a.get(sizeof(double) * 3 * nv) | B.r.get(sizeof(double) * nv) |
           B.pc.get(sizeof(double) * npc) | B.x.get(sizeof(double) * 3 * nv) | B.vn.get(sizeof(double) * 3 * nv) |
           B.fn.get(sizeof(double) * 3 * nf) | B.fx.get(sizeof(double) * 3 * nf) | B.ar.get(sizeof(double) * nf) |
           B.x32.get(sizeof(float4) * nv) | B.vn32.get(sizeof(float4) * nv) | B.fn32.get(sizeof(float4) * nf) |
           B.slot.get(sizeof(GeomSlot)) | B.flag.get(2 * sizeof(double)) |  // baddiam flag, log factor at +8
           B.mom.get(sizeof(double) * MOM_N) | B.part.get(sizeof(double) * MP1_N * moment_blocks(nf)) |
           B.hb.get(sizeof(double) * std::max((size_t)1, (size_t)nv * nbasis));
  if (rc) { B.mod = nullptr; set_err(rc, "realize_mod allocation"); return; }
  bool ok = true;
  if (new_topo) {
    std::vector<int> fv(3 * (size_t)nf), af_off(nv + 1, 0), af_idx;
    for (int f = 0; f < nf; f++)
      for (int k = 0; k < 3; k++) fv[3 * f + k] = rv->f[f].v[k];
    for (int i = 0; i < nv; i++) {
      af_off[i + 1] = af_off[i] + rv->v[i].naf;
      for (int q = 0; q < rv->v[i].naf; q++) af_idx.push_back(rv->v[i].af[q]);
    }
    B.mod = nullptr;
    ok = B.fv.get(sizeof(int) * fv.size()) == 0 && B.off.get(sizeof(int) * af_off.size()) == 0 &&
         B.idx.get(sizeof(int) * af_idx.size()) == 0 &&
         hipMemcpy(B.fv.p, fv.data(), sizeof(int) * fv.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B.off.p, af_off.data(), sizeof(int) * af_off.size(), hipMemcpyHostToDevice) == hipSuccess &&
         (af_idx.empty() ||
          hipMemcpy(B.idx.p, af_idx.data(), sizeof(int) * af_idx.size(), hipMemcpyHostToDevice) == hipSuccess);
    if (ok) { B.mod = mod; B.topo = topo; }
  }
  if (nbasis) {
    // the harmonic factors every call: the topology key does not cover the
    // degree or the vertex directions they come from, and the buffer may
    // have grown (moved) since the last upload
    std::vector<double> hb((size_t)nv * nbasis);
    for (int i = 0; i < nv; i++) {
      double *h = &hb[(size_t)i * nbasis];
      for (int l = 0, q = 0; l <= nhar; l++) {
        h[q++] = rv->v[i].afactor[l][0];
        for (int m = 1; m <= l; m++) { h[q++] = rv->v[i].afactor[l][m]; h[q++] = rv->v[i].bfactor[l][m]; }
      }
    }
    ok = ok && hipMemcpy(B.hb.p, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice) == hipSuccess;
  }
  // the slot names every buffer (they may have moved when they grew) and the
  // component's type, degree and ties
  GeomSlot G;
  memset(&G, 0, sizeof(G));
  G.nv = nv; G.nf = nf; G.comp_pb = 0; G.eval = -1;
  G.fv = (int *)B.fv.p; G.af_off = (int *)B.off.p; G.af_idx = (int *)B.idx.p; G.u = (double *)B.u.p;
  G.a = (double *)B.a.p; G.r = (double *)B.r.p; G.x = (double *)B.x.p; G.vn = (double *)B.vn.p;
  G.fn = (double *)B.fn.p; G.fx = (double *)B.fx.p; G.area = (double *)B.ar.p; G.x32 = (float4 *)B.x32.p;
  G.vn32 = (float4 *)B.vn32.p; G.fn32 = (float4 *)B.fn32.p;
  G.comp_type = ctype; G.nhar = nhar; G.hbasis = nbasis ? (const double *)B.hb.p : nullptr; G.sf_tie = sf_tie;
  ok = ok && hipMemcpy(B.slot.p, &G, sizeof(G), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(B.u.p, u, sizeof(double) * 3 * nv, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(B.a.p, a, sizeof(double) * 3 * nv, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(B.r.p, r, sizeof(double) * nv, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(B.pc.p, pc, sizeof(double) * npc, hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) { (void)hipGetLastError(); B.mod = nullptr; set_err(-3, "realize_mod upload"); return; }
  GeomSlot *ds = (GeomSlot *)B.slot.p;
  int *dflag = (int *)B.flag.p;
  double *dlogf = (double *)((char *)B.flag.p + sizeof(double));
  hipLaunchKernelGGL(k_realize_vertices, dim3((nv + 255) / 256, 1), dim3(256), 0, 0, ds, 1, (const double *)B.pc.p,
                     dflag, dlogf);
  if (ctype == HARMONIC)
    hipLaunchKernelGGL(k_realize_harmonic_bad, dim3(1), dim3(256), 0, 0, ds, (const double *)B.pc.p, dflag, dlogf);
  hipLaunchKernelGGL(k_realize_facets, dim3((nf + 255) / 256, 1), dim3(256), 0, 0, ds, 1);
  hipLaunchKernelGGL(k_realize_vnormals, dim3((nv + 255) / 256, 1), dim3(256), 0, 0, ds, 1);
  // compute_moments (realize_mod.c:757-840): area, volume, COM, inertia tensor
  const int nb = moment_blocks(nf);
  hipLaunchKernelGGL(k_moments_part, dim3(nb, 1), dim3(256), 0, 0, ds, (double *)B.part.p, nb);
  hipLaunchKernelGGL(k_moments_fin, dim3(1), dim3(64), 0, 0, ds, (const double *)B.pc.p, (const double *)B.part.p, nb,
                     (double *)B.mom.p, (const double *)nullptr, 0);
  double mom[MOM_N];
  // results through the thread's pinned stage (plain host memory if it cannot be pinned)
  const size_t nout = 7 * (size_t)nv + 7 * (size_t)nf;
  std::vector<double> plain;
  double *x = host_stage(nout);
  if (!x) { plain.resize(nout); x = plain.data(); }
  double *vn = x + 3 * (size_t)nv, *fn = vn + 3 * (size_t)nv, *fx = fn + 3 * (size_t)nf, *ar = fx + 3 * (size_t)nf;
  double *rr = ar + nf;
  int flag = 0;
  double logf = 0.0;
  ok = hipDeviceSynchronize() == hipSuccess &&
       hipMemcpy(x, B.x.p, sizeof(double) * 3 * nv, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(vn, B.vn.p, sizeof(double) * 3 * nv, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(fn, B.fn.p, sizeof(double) * 3 * nf, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(fx, B.fx.p, sizeof(double) * 3 * nf, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(ar, B.ar.p, sizeof(double) * nf, hipMemcpyDeviceToHost) == hipSuccess &&
       (ctype == VERTEX || hipMemcpy(rr, B.r.p, sizeof(double) * nv, hipMemcpyDeviceToHost) == hipSuccess) &&
       hipMemcpy(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(&logf, dlogf, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(mom, B.mom.p, sizeof(double) * MOM_N, hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) { (void)hipGetLastError(); set_err(-4, "realize_mod kernels"); return; }
  for (int i = 0; i < nv; i++) {
    for (int k = 0; k < 3; k++) { rv->v[i].x[k] = x[3 * i + k]; rv->v[i].n[k] = vn[3 * i + k]; }
    if (ctype != VERTEX) rv->v[i].r.val = rr[i];  // the analytic component's realized distance
  }
  // a harmonic scattering law reads each facet's angular coordinates
  // (realize_mod.c:517-545)
  bool harm_law = false;
  for (int l = 0; l < mod->photo.noptlaws; l++) {
    const int t = mod->photo.opttype[l];
    harm_law = harm_law || t == HARMLAMBERT || t == HARMLOMMEL || t == HARMHAPKE || t == HARMKAAS;
  }
  for (int l = 0; l < mod->photo.nradlaws; l++) harm_law = harm_law || mod->photo.radtype[l] == HARMCOSINE_DIFF;
  for (int f = 0; f < nf; f++) {
    for (int k = 0; k < 3; k++) { rv->f[f].n[k] = fn[3 * f + k]; rv->f[f].x[k] = fx[3 * f + k]; }
    rv->f[f].area = ar[f];
    if (harm_law) {
      const double *c = rv->f[f].x;
      rv->f[f].theta = atan2(sqrt(c[0] * c[0] + c[1] * c[1]), c[2]);
      rv->f[f].phi = atan2(c[1], c[0]);
    }
  }
  for (int j = 0; j < 3; j++) {
    double sf = 1.0;  // ELLIPSE, OVOID
    if (ctype == VERTEX || ctype == HARMONIC) {
      sf = pc[j];
      if (sf <= 0.01) sf = 0.01 / (1 + 0.01 - sf);
    }
    rv->scalefactor[j].val = sf;
  }
  par->baddiam = (unsigned char)flag;
  par->baddiam_logfactor = logf;
  // check_surface for one component (realize_mod.c:580-593): everything active
  for (int i = 0; i < nv; i++) rv->v[i].act = 1;
  for (int f = 0; f < nf; f++) rv->f[f].act = 1;
  if (rv->s)
    for (int k = 0; k < rv->ns; k++) rv->s[k].act = 1;
  // the component's moments equal the model's (one component, every facet active)
  mod->shape.area = comp->area = mom[MOM_AREA];
  mod->shape.volume = comp->volume = mom[MOM_VOL];
  for (int j = 0; j < 3; j++) {
    mod->shape.com[j] = comp->com[j] = mom[MOM_COM + j];
    for (int k = 0; k < 3; k++) mod->shape.inertia[j][k] = comp->inertia[j][k] = mom[MOM_I + 3 * j + k];
  }
}

// ---------------------------------------------------------------------------
// penalties (penalties.c:120): the penalty sum of par's list for the model as
// realize_mod left it (vertex coordinates, facet normals, the moments in
// mod->shape) and the datasets' angle / spin offsets; sets par->pen.base.
// The sides, vertices and principal axes are reduced on the device by the
// batched path's kernels, with the caller's moments in place of their own.
// ---------------------------------------------------------------------------
namespace {
struct PenBufs {
  GrowBuf fv, x, fn, ar, sides, u, r, slot, given, pb, type, weight, base, total, mom, part, ev, evslot, jobs, loc,
      noncos, law;
};
std::mutex g_pen_mu;
PenBufs g_pb;
}  // namespace

extern "C" double penalties(struct par_t *par, struct mod_t *mod, struct dat_t *dat) {
  const int n = par->pen.n;
  if (n <= 0) return 0.0;
  if (mod->shape.ncomp != 1) { set_err(-10, "penalties"); return 0.0; }
  struct comp_t *comp = &mod->shape.comp[0];
  struct vertices_t *rv = &comp->real;
  const int nv = rv->nv, nf = rv->nf, ns = rv->s ? rv->ns : 0;
  std::vector<int> type(n);
  std::vector<double> weight(n);
  int need = 0, law_need = 0;
  for (int i = 0; i < n; i++) {
    type[i] = par->pen.type[i + 1];
    weight[i] = par->pen.weight[i + 1];
    if ((!pen_supported(type[i]) && !pen_supported_laws(type[i])) || (type[i] == RDEV && comp->type != VERTEX)) {
      set_err(-17, "penalties: type outside this path");
      return 0.0;
    }
    need |= pen_needs(type[i]);
    if (pen_supported_laws(type[i])) law_need |= 1 << pen_law_slot(type[i]);
  }
  // the law-reading penalties take the per-facet values and tables the
  // caller's realize_photo left (component 0), one job per law
  std::vector<PhotoJob> jobs;
  std::vector<double> lawloc, noncos;
  if (law_need) {
    int have = 0, nc = 0;
    auto add = [&](int radar, int t, int k, int kind, const struct param_t *first, size_t elem, int cnt) {
      PhotoJob J;
      memset(&J, 0, sizeof(J));
      J.kind = kind; J.type = t; J.k = k; J.n = cnt; J.radar = radar; J.loc = (long long)lawloc.size();
      for (int f = 0; f < cnt; f++) {
        const struct param_t *q = (const struct param_t *)((const char *)first + elem * f);
        for (int j = 0; j < k; j++) lawloc.push_back(q[j].val);
      }
      jobs.push_back(J);
    };
    for (int l = 0; l < mod->photo.noptlaws; l++) {
      union optscat_t *o = &mod->photo.optical[l];
      const int t = mod->photo.opttype[l];
      const int kind = (t == HARMLAMBERT || t == HARMLOMMEL || t == HARMHAPKE || t == HARMKAAS) ? PHOTO_HARM : PHOTO_INHO;
      switch (t) {
      case HARMLAMBERT: case HARMLOMMEL: add(0, t, 1, kind, &o->harmR.local[0][0].R, sizeof(struct R_t), nf); break;
      case INHOLAMBERT: case INHOLOMMEL: add(0, t, 1, kind, &o->inhoR.local[0][0].R, sizeof(struct R_t), nf); break;
      case HARMHAPKE: add(0, t, 5, kind, &o->harmhapke.local[0][0].w, sizeof(struct hapke_t), nf); break;
      case INHOHAPKE: add(0, t, 5, kind, &o->inhohapke.local[0][0].w, sizeof(struct hapke_t), nf); break;
      case HARMKAAS: add(0, t, 5, kind, &o->harmkaas.local[0][0].R, sizeof(struct kaas_t), nf); break;
      case INHOKAAS: add(0, t, 5, kind, &o->inhokaas.local[0][0].R, sizeof(struct kaas_t), nf); break;
      default: continue;
      }
      have |= 1 << 0;
      if (t != HARMHAPKE) have |= 1 << 1;  // optalbvar counts no harmonic-Hapke facet
      if (t == HARMHAPKE || t == INHOHAPKE) have |= (1 << 2) | (1 << 3);
    }
    for (int l = 0; l < mod->photo.nradlaws; l++) {
      union radscat_t *r = &mod->photo.radar[l];
      const int t = mod->photo.radtype[l];
      if (t == HARMCOSINE_DIFF) add(1, t, 2, PHOTO_HARM, &r->harmcosine.local[0][0].R, sizeof(struct RC_t), nf);
      else if (t == INHOCOSINE_DIFF) add(1, t, 2, PHOTO_INHO, &r->inhocosine.local[0][0].R, sizeof(struct RC_t), nf);
      else if (t == TABULARLAW) {
        add(1, t, 1, PHOTO_TABLE, r->tabular.rho, sizeof(struct param_t), r->tabular.n);
        if (!nc) nc = r->tabular.n - 1;
        if (par->pen.x_noncosine && r->tabular.n == nc + 1) have |= 1 << 8;
        continue;
      } else continue;
      have |= (1 << 4) | (1 << 5) | (1 << 6) | (1 << 7);
    }
    if ((law_need & have) != law_need) { set_err(-17, "penalties: no law the penalty reads"); return 0.0; }
    if (law_need >> 8 & 1) {
      for (int i = 0; i < nc; i++) noncos.push_back(par->pen.x_noncosine[i]);
      noncos.push_back(par->pen.xmean_noncosine);
    }
  }
  if ((need & PEN_NEED_SIDES) && ns <= 0) { set_err(-18, "penalties: model has no side list"); return 0.0; }
  // host inputs: corners, coordinates, facet normals, sides, directions, deviations
  std::vector<int> fv(3 * (size_t)nf);
  std::vector<double> x(3 * (size_t)nv), fn(3 * (size_t)nf), u(3 * (size_t)nv), r(nv), ar(nf);
  std::vector<int4> sides(std::max(ns, 1));
  for (int f = 0; f < nf; f++)
    for (int k = 0; k < 3; k++) { fv[3 * f + k] = rv->f[f].v[k]; fn[3 * f + k] = rv->f[f].n[k]; }
  for (int f = 0; f < nf; f++) ar[f] = rv->f[f].area;
  for (int i = 0; i < nv; i++) {
    for (int k = 0; k < 3; k++) { x[3 * i + k] = rv->v[i].x[k]; u[3 * i + k] = rv->v[i].u[k]; }
    r[i] = rv->v[i].r.val;
  }
  for (int q = 0; q < ns; q++) sides[q] = make_int4(rv->s[q].v[0], rv->s[q].v[1], rv->s[q].f[0], rv->s[q].f[1]);
  // the moments and realized scale factors realize_mod left
  double given[17];
  given[0] = mod->shape.area;
  given[1] = mod->shape.volume;
  for (int j = 0; j < 3; j++) given[2 + j] = mod->shape.com[j];
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) given[5 + 3 * j + k] = mod->shape.inertia[j][k];
  for (int j = 0; j < 3; j++) given[14 + j] = rv->scalefactor[j].val;
  // a parameter block with the spin block and one offset block per dataset
  const int nsets = dat->nsets;
  std::vector<double> pb(PB_SPIN_END + (size_t)PB_SET_STRIDE * nsets, 0.0);
  for (int j = 0; j < 3; j++) pb[PB_INERTIA + j] = mod->spin.inertia[j].val;
  pb[PB_NIMP] = mod->spin.n_impulse;
  for (int q = 0; q < mod->spin.n_impulse && q < MAXIMP; q++)
    for (int j = 0; j < 3; j++) pb[PB_IMP + 3 * q + j] = mod->spin.impulse[q][j].val;
  for (int k = 0; k < nsets; k++)
    for (int j = 0; j < 3; j++) {
      pb[PB_SPIN_END + PB_SET_STRIDE * k + PB_SET_ANGLEOFF + j] = dat->set[k].angleoff[j].val;
      pb[PB_SPIN_END + PB_SET_STRIDE * k + PB_SET_OMEGAOFF + j] = dat->set[k].omegaoff[j].val;
    }
  EvalRec E;
  memset(&E, 0, sizeof(E));
  E.pb = 0;
  E.pen_set_off = PB_SPIN_END;
  E.pen_nsets = nsets;
  E.photo_job0 = 0;
  E.photo_njobs = (int)jobs.size();
  const int nb = moment_blocks(nf), zero = 0;
  std::lock_guard<std::mutex> lk(g_pen_mu);
  PenBufs &B = g_pb;
  int rc = B.fv.get(sizeof(int) * fv.size()) | B.x.get(sizeof(double) * x.size()) | B.fn.get(sizeof(double) * fn.size()) |
           B.ar.get(sizeof(double) * ar.size()) | B.sides.get(sizeof(int4) * sides.size()) |
           B.u.get(sizeof(double) * u.size()) | B.r.get(sizeof(double) * r.size()) | B.slot.get(sizeof(GeomSlot)) |
           B.given.get(sizeof(given)) | B.pb.get(sizeof(double) * pb.size()) | B.type.get(sizeof(int) * n) |
           B.weight.get(sizeof(double) * n) | B.base.get(sizeof(double) * n) | B.total.get(sizeof(double)) |
           B.mom.get(sizeof(double) * MOM_N) | B.part.get(sizeof(double) * std::max(MP1_N, MP2_N) * nb) |
           B.ev.get(sizeof(EvalRec)) | B.evslot.get(sizeof(int)) |
           B.jobs.get(sizeof(PhotoJob) * std::max((size_t)1, jobs.size())) |
           B.loc.get(sizeof(double) * std::max((size_t)1, lawloc.size())) |
           B.noncos.get(sizeof(double) * std::max((size_t)1, noncos.size())) | B.law.get(sizeof(double) * PEN_LAW_N);
  if (rc) { set_err(rc, "penalties allocation"); return 0.0; }
  if (law_need &&
      (hipMemcpy(B.jobs.p, jobs.data(), sizeof(PhotoJob) * jobs.size(), hipMemcpyHostToDevice) != hipSuccess ||
       (!lawloc.empty() &&
        hipMemcpy(B.loc.p, lawloc.data(), sizeof(double) * lawloc.size(), hipMemcpyHostToDevice) != hipSuccess) ||
       (!noncos.empty() &&
        hipMemcpy(B.noncos.p, noncos.data(), sizeof(double) * noncos.size(), hipMemcpyHostToDevice) != hipSuccess))) {
    (void)hipGetLastError();
    set_err(-3, "penalties upload");
    return 0.0;
  }
  GeomSlot G;
  memset(&G, 0, sizeof(G));
  G.nv = nv; G.nf = nf; G.eval = -1; G.comp_type = comp->type;
  G.fv = (int *)B.fv.p; G.x = (double *)B.x.p; G.fn = (double *)B.fn.p; G.area = (double *)B.ar.p;
  G.u = (double *)B.u.p; G.r = (double *)B.r.p;
  G.sides = ns > 0 ? (const int4 *)B.sides.p : nullptr;
  G.ns = ns;
  bool ok = hipMemcpy(B.fv.p, fv.data(), sizeof(int) * fv.size(), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.x.p, x.data(), sizeof(double) * x.size(), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.fn.p, fn.data(), sizeof(double) * fn.size(), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.ar.p, ar.data(), sizeof(double) * ar.size(), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.sides.p, sides.data(), sizeof(int4) * sides.size(), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.u.p, u.data(), sizeof(double) * u.size(), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.r.p, r.data(), sizeof(double) * r.size(), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.slot.p, &G, sizeof(G), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.given.p, given, sizeof(given), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.pb.p, pb.data(), sizeof(double) * pb.size(), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.type.p, type.data(), sizeof(int) * n, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.weight.p, weight.data(), sizeof(double) * n, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.ev.p, &E, sizeof(E), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.evslot.p, &zero, sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) { (void)hipGetLastError(); set_err(-3, "penalties upload"); return 0.0; }
  const GeomSlot *ds = (const GeomSlot *)B.slot.p;
  double *part = (double *)B.part.p, *dmom = (double *)B.mom.p;
  const int pa = (need & PEN_NEED_PA) ? 1 : 0;
  hipLaunchKernelGGL(k_moments_part, dim3(nb, 1), dim3(256), 0, 0, ds, part, nb);
  hipLaunchKernelGGL(k_moments_fin, dim3(1), dim3(64), 0, 0, ds, (const double *)B.pb.p, (const double *)part, nb, dmom,
                     (const double *)B.given.p, pa);
  if (need & PEN_NEED_VERTS) {
    hipLaunchKernelGGL(k_pen_vert_part, dim3(nb, 1), dim3(256), 0, 0, ds, (const double *)dmom, part, nb, need);
    hipLaunchKernelGGL(k_pen_vert_fin, dim3(1), dim3(64), 0, 0, ds, (const double *)part, nb, dmom, need);
  }
  PenDev P;
  P.n = n; P.type = (const int *)B.type.p; P.weight = (const double *)B.weight.p; P.base = (double *)B.base.p;
  P.total = (double *)B.total.p; P.mom = dmom; P.law = nullptr;
  if (law_need) {
    hipLaunchKernelGGL(k_pen_laws, dim3(1), dim3(256), 0, 0, (const EvalRec *)B.ev.p, 1, (const int *)B.evslot.p, ds,
                       (const PhotoJob *)B.jobs.p, (const double *)B.loc.p, (const double *)B.noncos.p, law_need,
                       (double *)B.law.p);
    P.law = (const double *)B.law.p;
  }
  hipLaunchKernelGGL(k_pen_eval, dim3(1), dim3(64), 0, 0, (const EvalRec *)B.ev.p, 1, (const int *)B.evslot.p,
                     (const double *)B.pb.p, P);
  std::vector<double> base(n);
  double total = 0.0;
  ok = hipDeviceSynchronize() == hipSuccess &&
       hipMemcpy(base.data(), B.base.p, sizeof(double) * n, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(&total, B.total.p, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) { (void)hipGetLastError(); set_err(-4, "penalties kernels"); return 0.0; }
  for (int i = 0; i < n; i++) par->pen.base[i + 1] = base[i];
  return total;
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
