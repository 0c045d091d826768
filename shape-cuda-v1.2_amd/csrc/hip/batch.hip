// batch.hip -- the batched evaluation interface (include/shape_hip.h, part 2).
//
// A batch packs one or more models (mod_t/dat_t pairs sharing par_t) into the
// device engine once; each evaluation names a model and a set of parameter
// addresses inside it (for instance entries of par->fpntr) with trial values.
// One replay evaluates every evaluation: setup, realization, rasterization,
// synthesis, chi-square and the objective, all on the device.  With
// use_graph the replay is a captured hipGraph; trial values are refreshed
// from a host staging copy as the graph's first node.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "engine_host.h"
#include "shape_hip.h"

using namespace shp;

struct shape_batch {
  Engine E;
  std::vector<int> ntarg;
  int use_graph = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.0f;
};

extern "C" shape_batch *shape_batch_create(struct par_t *par, struct mod_t **mods, struct dat_t **dats, int nmodels,
                                           int nevals, const int *eval_model, const int *n_targets,
                                           double *const *targets, const double *values, int method, int flags) {
  // arguments checked before any device call
  const char *bad = nullptr;
  long long ntot = 0;
  if (!par || !mods || !dats || nmodels < 1) bad = "no models";
  else if (nevals < 1 || !eval_model || !n_targets) bad = "no evaluations";
  else if (method != SHAPE_METHOD_FP64 && method != SHAPE_METHOD_MIXED) bad = "unknown method";
  for (int e = 0; !bad && e < nevals; e++) {
    if (eval_model[e] < 0 || eval_model[e] >= nmodels) bad = "evaluation's model index out of range";
    else if (n_targets[e] < 0) bad = "negative target count";
    else ntot += n_targets[e];
  }
  if (!bad && ntot > 0 && (!targets || !values)) bad = "targets or values missing";
  for (long long q = 0; !bad && q < ntot; q++)
    if (!targets[q]) bad = "null target address";
  if (bad) {
    fprintf(stderr, "shape_batch_create: %s\n", bad);
    return nullptr;
  }
  shape_batch *B = new shape_batch();
  std::vector<EvalSpec> ev(nevals);
  int q = 0;
  B->ntarg.assign(n_targets, n_targets + nevals);
  for (int e = 0; e < nevals; e++) {
    ev[e].model = eval_model[e];
    for (int k = 0; k < n_targets[e]; k++, q++) {
      ev[e].targets.push_back(targets[q]);
      ev[e].values.push_back(values[q]);
    }
  }
  RunOptions o;
  o.method = method == SHAPE_METHOD_MIXED ? METHOD_MIXED : METHOD_FP64;
  o.keep_fit = (flags & SHAPE_BATCH_KEEP_FIT) ? 1 : 0;
  o.global_zbuf = (flags & SHAPE_BATCH_GLOBAL_ZBUF) ? 1 : 0;
  o.facet_map = (flags & SHAPE_BATCH_FACET_MAP) ? 1 : 0;
  o.penalties = 1;
  B->use_graph = (flags & SHAPE_BATCH_GRAPH) ? 1 : 0;
  int rc = B->E.build(par, mods, dats, nmodels, ev, o);
  if (!rc && B->use_graph) rc = B->E.capture_graph();
  if (!rc && (hipEventCreate(&B->ev0) != hipSuccess || hipEventCreate(&B->ev1) != hipSuccess)) rc = -5;
  if (rc) {
    fprintf(stderr, "shape_batch_create failed (code %d)\n", rc);
    delete B;
    return nullptr;
  }
  return B;
}

extern "C" void shape_batch_destroy(shape_batch *B) {
  if (!B) return;
  if (B->ev0) (void)hipEventDestroy(B->ev0);
  if (B->ev1) (void)hipEventDestroy(B->ev1);
  delete B;
}

extern "C" int shape_batch_set_values(shape_batch *B, const double *values) {
  int q = 0;
  for (int e = 0; e < B->E.nevals; e++) {
    std::vector<double> v(values + q, values + q + B->ntarg[e]);
    q += B->ntarg[e];
    int rc = B->E.set_values(e, v);
    if (rc) return rc;
  }
  return 0;
}

// Enqueue one replay (asynchronous); the events bracket it on the engine's stream.
extern "C" int shape_batch_launch(shape_batch *B) {
  if (hipEventRecord(B->ev0, B->E.stream) != hipSuccess) return -6;
  int rc;
  if (B->use_graph) rc = B->E.launch_graph();
  else {
    rc = B->E.upload_params();
    if (!rc) rc = B->E.enqueue();
  }
  if (rc) return rc;
  if (hipEventRecord(B->ev1, B->E.stream) != hipSuccess) return -6;
  return 0;
}

// Wait for the last replay; copy objective values and chi-squares (nevals each, either may be NULL).
extern "C" int shape_batch_wait(shape_batch *B, double *objective, double *chi2v) {
  int rc = B->E.sync();
  if (rc) return rc;
  (void)hipEventElapsedTime(&B->last_ms, B->ev0, B->ev1);
  if (objective || chi2v) {
    std::vector<EvalRec> ev;
    if ((rc = B->E.download_evals(ev))) return rc;
    for (int e = 0; e < B->E.nevals; e++) {
      if (objective) objective[e] = ev[e].objective;
      if (chi2v) chi2v[e] = ev[e].chi2;
    }
  }
  return B->E.error_flags() ? -30 : 0;
}

extern "C" int shape_batch_run(shape_batch *B, double *objective, double *chi2v) {
  int rc = shape_batch_launch(B);
  return rc ? rc : shape_batch_wait(B, objective, chi2v);
}

// One round of shape_batch_line_search: the k trial values are the batch's values.
static int batch_search_eval(void *ctx, int k, const double *x, double *f) {
  shape_batch *B = (shape_batch *)ctx;
  if (k != B->E.nevals) return 1;
  int rc = shape_batch_set_values(B, x);
  return rc ? rc : shape_batch_run(B, f, nullptr);
}

extern "C" int shape_batch_line_search(shape_batch *B, double x0, double step, double tol, double abstol,
                                       int max_rounds, double *xmin, double *fmin, int *rounds) {
  if (!B || B->E.nevals < 1) return -1;
  const auto &t0 = B->E.eval_targets[0];
  for (int e = 0; e < B->E.nevals; e++) {
    const auto &t = B->E.eval_targets[e];
    if (B->ntarg[e] != 1 || t.size() != 1 || t[0].kind != t0[0].kind || t[0].model != t0[0].model ||
        t[0].offset != t0[0].offset)
      return -3;
  }
  return shape_line_search(batch_search_eval, B, B->E.nevals, x0, step, tol, abstol, max_rounds, xmin, fmin, rounds);
}

// Many searches in one batch: evaluations [s k, (s + 1) k) form search s,
// each setting the same single parameter (of any model).
extern "C" int shape_batch_line_searches(shape_batch *B, int nsearch, const double *x0, const double *step,
                                         double tol, double abstol, int max_rounds, double *xmin, double *fmin,
                                         int *status, int *rounds) {
  if (!B || nsearch < 1 || B->E.nevals % nsearch != 0) return -1;
  const int k = B->E.nevals / nsearch;
  for (int s = 0; s < nsearch; s++) {
    const auto &t0 = B->E.eval_targets[(size_t)s * k];
    for (int e = s * k; e < (s + 1) * k; e++) {
      const auto &t = B->E.eval_targets[e];
      if (B->ntarg[e] != 1 || t.size() != 1 || t0.size() != 1 || t[0].kind != t0[0].kind ||
          t[0].model != t0[0].model || t[0].offset != t0[0].offset)
        return -3;
    }
  }
  return shape_line_searches(batch_search_eval, B, nsearch, k, x0, step, tol, abstol, max_rounds, xmin, fmin,
                             status, rounds);
}

// One pass of bestfit's loop over the fitted parameters (bestfit.c:392-470):
// every evaluation of the batch sets the same P parameters in the same
// order; parameter p is searched (k = nevals trial values per round) with
// the others held at x, then x[p] takes the minimum found.
namespace {
struct PassCtx {
  shape_batch *B;
  int P, p;
  const double *x;
  std::vector<double> vals;
};
int pass_eval(void *ctx, int k, const double *xs, double *f) {
  PassCtx &C = *(PassCtx *)ctx;
  if (k != C.B->E.nevals) return 1;
  for (int e = 0; e < k; e++)
    for (int q = 0; q < C.P; q++) C.vals[(size_t)e * C.P + q] = q == C.p ? xs[e] : C.x[q];
  int rc = shape_batch_set_values(C.B, C.vals.data());
  return rc ? rc : shape_batch_run(C.B, f, nullptr);
}
}  // namespace

extern "C" int shape_batch_fit_pass(shape_batch *B, const double *step, const double *tol, const double *abstol,
                                    int max_rounds, double *x, double *fmin, int *rounds) {
  if (!B || !step || !tol || !abstol || !x || B->E.nevals < 3) return -1;
  const int P = B->ntarg[0];
  if (P < 1) return -1;
  const auto &t0 = B->E.eval_targets[0];
  for (int e = 1; e < B->E.nevals; e++) {
    const auto &t = B->E.eval_targets[e];
    if (B->ntarg[e] != P || t.size() != t0.size()) return -3;
    for (int q = 0; q < P; q++)
      if (t[q].kind != t0[q].kind || t[q].model != t0[q].model || t[q].offset != t0[q].offset) return -3;
  }
  PassCtx C{B, P, 0, x, std::vector<double>((size_t)B->E.nevals * P)};
  int worst = 0, total = 0;
  double fm = HUGE_VAL;
  for (int p = 0; p < P; p++) {
    C.p = p;
    double xm = 0.0;
    int r = 0;
    const int rc = shape_line_search(pass_eval, &C, B->E.nevals, x[p], step[p], tol[p], abstol[p], max_rounds, &xm,
                                     &fm, &r);
    if (rc < 0) return rc;
    worst = std::max(worst, rc);
    total += r;
    x[p] = xm;
  }
  if (fmin) *fmin = fm;
  if (rounds) *rounds = total;
  return worst;
}

extern "C" float shape_batch_last_ms(shape_batch *B) { return B->last_ms; }

extern "C" void *shape_batch_stream(shape_batch *B) { return (void *)B->E.stream; }

extern "C" int shape_batch_nframes(shape_batch *B, int e) {
  return (e >= 0 && e < B->E.nevals) ? B->E.h_evals[e].nframes : -1;
}

// per-frame chi2 and calibration of evaluation e, in set/frame order
extern "C" int shape_batch_frame_results(shape_batch *B, int e, double *chi2v, double *cal) {
  std::vector<FrameRec> fr;
  int rc = B->E.download_frames(fr);
  if (rc) return rc;
  const EvalRec &E = B->E.h_evals[e];
  for (int q = 0; q < E.nframes; q++) {
    if (chi2v) chi2v[q] = fr[E.first_frame + q].chi2;
    if (cal) cal[q] = fr[E.first_frame + q].cal;
  }
  return 0;
}

// chi-square terms of every frame of evaluation e: o2, m2, om, overflow_o2,
// overflow_m2, cal, chi2, badradar_logfactor (8 doubles per frame)
extern "C" int shape_batch_frame_terms(shape_batch *B, int e, double *out) {
  std::vector<FrameRec> fr;
  int rc = B->E.download_frames(fr);
  if (rc) return rc;
  const EvalRec &E = B->E.h_evals[e];
  for (int q = 0; q < E.nframes; q++) {
    const FrameRec &F = fr[E.first_frame + q];
    double *o = out + 8 * q;
    o[0] = F.o2; o[1] = F.m2; o[2] = F.om; o[3] = F.overflow_o2; o[4] = F.overflow_m2;
    o[5] = F.cal; o[6] = F.chi2; o[7] = F.badradar_logfactor;
  }
  return 0;
}

// Penalties of evaluation e after the last replay: par->pen.base[1..n] as
// penalties.c leaves them into base[0..n), the weighted sum into *total;
// returns n (0 when par lists none) or < 0.
extern "C" int shape_batch_penalties(shape_batch *B, int e, double *base, double *total) {
  if (e < 0 || e >= B->E.nevals) return -20;
  std::vector<double> b;
  int rc = B->E.download_penalties(e, b, total);
  if (rc) return rc;
  if (base) std::copy(b.begin(), b.end(), base);
  return (int)b.size();
}

// FP64 copy of the fit image of frame q (0-based within evaluation e); needs SHAPE_BATCH_KEEP_FIT
extern "C" int shape_batch_fit(shape_batch *B, int e, int q, double *out) {
  const EvalRec &E = B->E.h_evals[e];
  if (q < 0 || q >= E.nframes) return -20;
  const FrameRec &F = B->E.h_frames[E.first_frame + q];
  long long N = F.type == VT_DELAY ? (long long)F.ndel * F.ndop : (F.type == VT_DOPPLER ? F.ndop : F.lc_n);
  if (hipMemcpy(out, B->E.d_fit + F.off, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess) return -4;
  if (F.nviews > 1 && F.type != VT_LGHTCRV)
    for (long long i = 0; i < N; i++) out[i] /= F.nviews;
  return 0;
}

// owner facets of view v of frame q of evaluation e (needs SHAPE_BATCH_FACET_MAP)
extern "C" int shape_batch_facet_map(shape_batch *B, int e, int q, int v, int *out) {
  Engine &E = B->E;
  if (!E.opt.facet_map) return -21;
  if (e < 0 || e >= E.nevals) return -20;
  const EvalRec &R = E.h_evals[e];
  if (q < 0 || q >= R.nframes) return -20;
  const FrameRec &F = E.h_frames[R.first_frame + q];
  if (F.type == VT_LGHTCRV || v < 0 || v >= F.nviews) return -20;
  const ViewRec &V = E.h_views[F.first_view + v];
  const int *src = E.pd.fmap ? E.pd.fmap : E.pd.f;
  if (!src) return -21;
  const size_t np = (size_t)(2 * V.n + 1) * (2 * V.n + 1);
  return hipMemcpy(out, src + V.pos_off, sizeof(int) * np, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -4;
}

// counters of the last replay: [0] replayed pixels (FP64 method: owner faces
// away; mixed: owner faces away or pixel on a facet edge), [1] error flags,
// [2] mixed: replayed pixels whose candidates the second raster pass
// collected, [3] reserved (0: every footprint width is synthesized)
extern "C" int shape_batch_counters(shape_batch *B, int *out4) {
  return hipMemcpy(out4, B->E.pd.nstale, sizeof(int) * 4, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -4;
}

extern "C" int shape_batch_sizes(shape_batch *B, long long *out8) {
  Engine &E = B->E;
  out8[0] = E.nevals; out8[1] = E.nviews; out8[2] = E.nframes; out8[3] = E.nslots;
  out8[4] = E.total_pix; out8[5] = E.fit_len; out8[6] = E.max_nf; out8[7] = E.max_pix;
  return 0;
}

// ---------------------------------------------------------------------------
// measurement helpers (benchmarks and profiling)
// ---------------------------------------------------------------------------
// Time every kernel of eager launches with HIP events on the batch's stream
// (batches created without SHAPE_BATCH_GRAPH only).  Totals accumulate over
// launches until shape_batch_time_kernels is called again with on = 1.
extern "C" int shape_batch_time_kernels(shape_batch *B, int on) {
  if (B->use_graph) return -21;
  B->E.time_kernels = on != 0;
  if (on) { B->E.kt_names.clear(); B->E.kt_ms.clear(); B->E.kt_count.clear(); }
  return 0;
}

// Entry idx of the kernel totals: name, total milliseconds and launches.
// Returns the number of entries, or -1 when idx is out of range.
extern "C" int shape_batch_kernel_stat(shape_batch *B, int idx, char *name, int namelen, double *total_ms,
                                       long long *launches) {
  int rc = B->E.collect_kernel_times();
  if (rc) return rc;
  int nk = (int)B->E.kt_names.size();
  if (idx < 0 || idx >= nk) return -1;
  if (name && namelen > 0) {
    strncpy(name, B->E.kt_names[idx].c_str(), (size_t)namelen - 1);
    name[namelen - 1] = 0;
  }
  if (total_ms) *total_ms = B->E.kt_ms[idx];
  if (launches) *launches = B->E.kt_count[idx];
  return nk;
}

// Footprint of the last replay: [0] POS pixels inside the views' model
// rectangles, [1] pixels of all views' POS grids, [2] image cells inside the
// frames' occupied windows, [3] image cells of all radar frames,
// [4] vertices and [5] facets summed over the geometry slots, [6] z-buffer
// keys claimed (mixed method).
extern "C" int shape_batch_footprint(shape_batch *B, long long *out7) {
  long long *out6 = out7;
  std::vector<ViewRec> vs;
  std::vector<FrameRec> fr;
  int rc = B->E.download_views(vs);
  if (!rc) rc = B->E.download_frames(fr);
  if (rc) return rc;
  long long bbox = 0, grid = 0, win = 0, cells = 0, nv = 0, nf = 0, claimed = 0;
  for (const ViewRec &V : vs) {
    long long w = 2LL * V.n + 1;
    grid += w * w;
    claimed += V.nclaimed;
    if (V.xlim[1] >= V.xlim[0] && V.ylim[1] >= V.ylim[0])
      bbox += (long long)(V.xlim[1] - V.xlim[0] + 1) * (V.ylim[1] - V.ylim[0] + 1);
  }
  for (const FrameRec &F : fr) {
    if (F.type != VT_DELAY && F.type != VT_DOPPLER) continue;
    int r0 = 1, r1 = F.type == VT_DELAY ? F.ndel : 1, c0 = 0x7fffffff, c1 = -0x7fffffff;
    int d0 = 0x7fffffff, d1 = -0x7fffffff;
    for (int v = 0; v < F.nviews; v++) {
      const ViewRec &V = vs[F.first_view + v];
      c0 = std::min(c0, V.idoplim[0]); c1 = std::max(c1, V.idoplim[1]);
      d0 = std::min(d0, V.idellim[0]); d1 = std::max(d1, V.idellim[1]);
    }
    if (F.type == VT_DELAY) { r0 = std::max(1, d0); r1 = std::min(F.ndel, d1); }
    c0 = std::max(1, c0); c1 = std::min(F.ndop, c1);
    if (r1 >= r0 && c1 >= c0) win += (long long)(r1 - r0 + 1) * (c1 - c0 + 1);
    cells += (long long)(F.type == VT_DELAY ? F.ndel : 1) * F.ndop;
  }
  for (const GeomSlot &G : B->E.h_slots) { nv += G.nv; nf += G.nf; }
  out6[0] = bbox; out6[1] = grid; out6[2] = win; out6[3] = cells; out6[4] = nv; out6[5] = nf;
  out6[6] = claimed;
  return 0;
}

// ---------------------------------------------------------------------------
// Launch-setting tuner.  The mixed method's launch settings (Tune: the tile
// kernel's form -- one workgroup per tile item, or persistent with 2, 4 or 6
// workgroups per compute unit --, the FP32 replay's and the candidate
// collection's grids) change how the work is spread over the device, never
// the results.  shape_batch_tune times replays of the batch under each
// candidate, one setting at a time (tile form, then replay grid, then
// collection grid), keeps the fastest, and caches the choice under the
// batch's problem shape: method, views, geometry slots, facets, tile grid,
// POS pixels, frames, lightcurve and global-raster use, and the device's
// compute units.  With SHAPE_HIP_TUNE_FILE set, choices are read from and
// appended to that file, one line each: "key tile_grid replay_grid
// collect_grid ms".
// ---------------------------------------------------------------------------
namespace {

std::mutex g_tune_mu;
std::map<std::string, std::pair<Tune, double>> g_tune_cache;
bool g_tune_file_read = false;

int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    return 0;
  return cus;
}

std::string shape_key(const Engine &E) {
  char buf[256];
  snprintf(buf, sizeof(buf), "m%d_v%d_s%d_f%d_t%d_p%lld_fr%d_lc%d_g%d_cu%d", E.opt.method, E.nviews, E.nslots, E.max_nf,
           E.max_bin_ntx, E.total_pix, E.nframes, E.any_lc ? 1 : 0, E.any_global ? 1 : 0, device_cus());
  return buf;
}

void read_tune_file() {
  if (g_tune_file_read) return;
  g_tune_file_read = true;
  const char *path = getenv("SHAPE_HIP_TUNE_FILE");
  if (!path) return;
  FILE *f = fopen(path, "r");
  if (!f) return;
  char key[200];
  Tune t;
  double ms;
  while (fscanf(f, "%199s %d %d %d %lf", key, &t.tile_grid, &t.replay_grid, &t.collect_grid, &ms) == 5)
    g_tune_cache[key] = {t, ms};
  fclose(f);
}

void append_tune_file(const std::string &key, const Tune &t, double ms) {
  const char *path = getenv("SHAPE_HIP_TUNE_FILE");
  if (!path) return;
  FILE *f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%s %d %d %d %.6f\n", key.c_str(), t.tile_grid, t.replay_grid, t.collect_grid, ms);
  fclose(f);
}

bool tune_valid(const Tune &t) {
  return t.tile_grid >= 0 && t.tile_grid % NXCD == 0 && t.replay_grid >= 1 && t.collect_grid >= 1;
}

// settings into effect: a captured graph holds launch parameters, so it is
// captured again
int apply_tune(shape_batch *B, const Tune &t) {
  B->E.tune = t;
  return B->use_graph ? B->E.capture_graph() : 0;
}

// median replay time (ms) over `replays` replays after one untimed one
int time_replays(shape_batch *B, int replays, double *ms) {
  int rc = shape_batch_run(B, nullptr, nullptr);
  if (rc) return rc;
  std::vector<double> v;
  for (int r = 0; r < replays; r++) {
    if ((rc = shape_batch_run(B, nullptr, nullptr))) return rc;
    v.push_back(B->last_ms);
  }
  std::sort(v.begin(), v.end());
  *ms = v[v.size() / 2];
  return 0;
}

}  // namespace

static void tune_out(const Tune &t, int *settings) {
  if (!settings) return;
  settings[0] = t.tile_grid;
  settings[1] = t.replay_grid;
  settings[2] = t.collect_grid;
}

extern "C" int shape_batch_get_tune(shape_batch *B, int *settings) {
  tune_out(B->E.tune, settings);
  return 0;
}

extern "C" int shape_batch_set_tune(shape_batch *B, const int *settings) {
  Tune t;
  t.tile_grid = settings[0];
  t.replay_grid = settings[1];
  t.collect_grid = settings[2];
  if (!tune_valid(t) || (t.tile_grid > 0 && !B->E.d_tile_work)) return -1;
  return apply_tune(B, t);
}

extern "C" int shape_batch_tune(shape_batch *B, int replays, int force, int *settings, double *best_ms) {
  if (replays < 1) return -1;
  Engine &E = B->E;
  const std::string key = shape_key(E);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    read_tune_file();
    auto it = g_tune_cache.find(key);
    if (!force && it != g_tune_cache.end()) {
      int rc = apply_tune(B, it->second.first);
      tune_out(E.tune, settings);
      if (best_ms) *best_ms = it->second.second;
      return rc;
    }
  }
  Tune best = E.tune;
  double best_t = 0.0;
  int rc = apply_tune(B, best);
  if (!rc) rc = time_replays(B, replays, &best_t);
  if (rc) return rc;
  auto consider = [&](Tune t) -> int {
    if (!tune_valid(t)) return 0;
    double ms = 0.0;
    int r = apply_tune(B, t);
    if (!r) r = time_replays(B, replays, &ms);
    if (r) return r;
    if (ms < best_t) { best_t = ms; best = t; }
    return 0;
  };
  if (E.opt.method == METHOD_MIXED) {
    const int cus = std::max(NXCD, device_cus());
    if (E.any_binned && E.d_tile_work) {
      for (int per_cu : {2, 4, TILE_WG_PER_CU}) {
        Tune t = best;
        t.tile_grid = std::max(NXCD, cus * per_cu / NXCD * NXCD);
        if ((rc = consider(t))) break;
      }
      Tune t = best;
      t.tile_grid = 0;
      if (!rc) rc = consider(t);
    }
    for (int g : {512, 1024, 2048, 4096, 8192}) {
      if (rc) break;
      Tune t = best;
      t.replay_grid = g;
      rc = consider(t);
    }
    for (int g : {256, 1024, 2048, 4096}) {
      if (rc) break;
      Tune t = best;
      t.collect_grid = g;
      rc = consider(t);
    }
  }
  if (!rc) rc = apply_tune(B, best);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = {best, best_t};
    append_tune_file(key, best, best_t);
  }
  tune_out(best, settings);
  if (best_ms) *best_ms = best_t;
  return 0;
}
