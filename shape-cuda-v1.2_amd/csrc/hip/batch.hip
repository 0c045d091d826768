// batch.hip -- the batched evaluation interface (include/shape_hip.h, part 2).
//
// A batch packs one or more models (mod_t/dat_t pairs sharing par_t) into the
// device engine once; each evaluation names a model and a set of parameter
// addresses inside it (for instance entries of par->fpntr) with trial values.
// One replay evaluates every evaluation: setup, realization, rasterization,
// synthesis, chi-square and the objective, all on the device.  With
// use_graph the replay is a captured hipGraph; trial values are refreshed
// from a host staging copy as the graph's first node.
#include <string.h>

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
  o.method = method == 1 ? METHOD_MIXED : METHOD_FP64;
  o.keep_fit = (flags & SHAPE_BATCH_KEEP_FIT) ? 1 : 0;
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

extern "C" float shape_batch_last_ms(shape_batch *B) { return B->last_ms; }

extern "C" hipStream_t shape_batch_stream(shape_batch *B) { return B->E.stream; }

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

// counters of the last replay: [0] replayed / inherited-normal pixels,
// [1] error flags, [2] mixed-path pixels whose owner faces away, [3] footprints too wide
extern "C" int shape_batch_counters(shape_batch *B, int *out4) {
  return hipMemcpy(out4, B->E.pd.nstale, sizeof(int) * 4, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -4;
}

extern "C" int shape_batch_sizes(shape_batch *B, long long *out8) {
  Engine &E = B->E;
  out8[0] = E.nevals; out8[1] = E.nviews; out8[2] = E.nframes; out8[3] = E.nslots;
  out8[4] = E.total_pix; out8[5] = E.fit_len; out8[6] = E.max_nf; out8[7] = E.max_pix;
  return 0;
}
