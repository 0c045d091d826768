// engine_host.h -- host side of the evaluation engine.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"
#include "kernels.h"
#include "shape_abi.h"

// Run f(k) for k in [0, n) on up to 16 host threads, interleaved so that
// frames of different sizes balance.  The drop-ins copy the reference's
// per-row host arrays to and from flat buffers: memory-bound host work that
// one thread would take most of a call on.  f must write disjoint data.
template <class F> inline void host_parallel_for(int n, F &&f) {
  const int hw = (int)std::thread::hardware_concurrency();
  const int nt = std::max(1, std::min(std::min(n, 16), hw > 0 ? hw : 1));
  if (nt <= 1) {
    for (int k = 0; k < n; k++) f(k);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int t = 0; t < nt; t++)
    th.emplace_back([&f, t, nt, n] {
      for (int k = t; k < n; k += nt) f(k);
    });
  for (auto &x : th) x.join();
}

namespace shp {

// A pinned host staging buffer per host thread, grown on demand and kept
// for the thread's life: the drop-ins stage their flat arrays in it, so the
// copies to and from the device run at full link rate without pinning pages
// on every call.  Returns nullptr if the allocation fails.
double *host_stage(size_t n);

// Where a fittable scalar lives inside an evaluation's parameter block, or
// which vertex deviation it is.
struct ParamTarget {
  int kind;    // 0: parameter block offset, 1: vertex deviation r
  int model;
  int offset;  // block offset or vertex index
};

struct EvalSpec {
  int model;
  std::vector<const double *> targets;  // host addresses (e.g. par->fpntr[p])
  std::vector<double> values;
};

// Launch settings of the mixed method that do not change results, only how
// the work is spread over the device; shape_batch_tune picks them by timing
// replays.
struct Tune {
  int tile_grid = 0;       // tile kernel: 0 = one workgroup per tile item, > 0 = persistent, this many workgroups
  int replay_grid = 2048;  // FP32 replay: workgroups at most (one thread per record, grid-stride)
  int collect_grid = 2048; // global candidate collection: workgroups (grid-stride)
};

struct RunOptions {
  int method = METHOD_FP64;
  int host_spin = 0;        // 1: take ae/spin/deloff/dopoff from the host structs
  int host_geometry = 0;    // 1: upload realized x / normals instead of realizing on device
  int keep_fit = 1;         // 1: leave fit images in place (0: read-and-clear in chi2)
  int with_body = 0;        // allocate body/comp arrays (drop-ins)
  int views_only = 0;       // pack parameters and views only (spin / offset drop-ins)
  int global_zbuf = 0;      // mixed method: rasterize every view through the global z-buffer
  int facet_map = 0;        // keep every POS pixel's owner facet for inspection
  int only_set = -1;        // >= 0: views and frames of this dataset only (calc_deldop & co.)
  int penalties = 0;        // 1: add par's penalties to each objective (the batched interface)
};

class Engine {
 public:
  ~Engine();
  // Pack one scene per model; evaluations reference models.  Returns 0 or an error code.
  int build(struct par_t *par, struct mod_t **mods, struct dat_t **dats, int nmodels,
            const std::vector<EvalSpec> &evals, const RunOptions &opt);
  // Re-pack what a single-model, single-evaluation engine reads from the
  // caller's structs on every call -- views (epochs, spin state, frames of
  // reference), frame scalars, parameter values, realized geometry when
  // opt.host_geometry -- into the existing device arenas.  Returns 1 when
  // the structs no longer have the shape the engine was built for (the
  // caller then builds a new engine), 0 on success, < 0 on a device error.
  int refresh(struct par_t *par, struct mod_t *mod, struct dat_t *dat);
  // Rewrite the trial values of evaluation e (same targets as at build time).
  // Waits for a replay in flight first: the replay's first step reads them.
  int set_values(int e, const std::vector<double> &values);
  int upload_params();
  // Enqueue one full evaluation of every evaluation on the stream.
  int enqueue();
  int capture_graph();
  int launch_graph();
  int sync();
  // results
  int download_evals(std::vector<EvalRec> &out);
  int download_frames(std::vector<FrameRec> &out);
  int download_view_out(std::vector<double> &out);
  int download_views(std::vector<ViewRec> &out);
  // Write fits, frame scalars, view data and flags of evaluation e into the
  // host structs.  accumulate = 0 (calc_fits.c:340-409): seen flags and the
  // par flags and log factors are set; accumulate = 1 (one dataset,
  // calc_deldop / calc_doppler / calc_lghtcrv): seen flags are or-ed in and
  // the flags and un-normalized log factors added, as those routines do.
  int writeback(int e, struct par_t *par, struct mod_t *mod, struct dat_t *dat, int write_pos, int accumulate = 0);
  int error_flags();

  hipStream_t stream = nullptr;
  RunOptions opt;
  // sizes
  int nviews = 0, nframes = 0, nevals = 0, nslots = 0, pb_stride = 0;
  int max_nf = 0, max_pix = 0, nsinc2 = 1, sinc2width = 6;
  double mask_tol = 0.0, mincosine_seen = 0.0;
  int exclude_seen = -1;
  int last_nstale = 0;      // pixels replayed in the last evaluation (diagnostic)
  long long total_pix = 0, fit_len = 0, obs_len = 0, total_tiles = 0;
  long long lcx_len = 0, lct_len = 0;  // lightcurve epochs: calculated points, observed points x views
  bool any_lc = false, any_dd = false, any_dop = false;
  bool any_harmonic = false;  // a model with a HARMONIC component (its per-vertex baddiam guard)
  // mixed method: views taking the binned tile kernel / the global z-buffer
  bool any_binned = false, any_global = false;
  int max_bin_ntx = 0, max_nblk = 0;
  long long total_bins = 0;
  long long lc_rec_total = 0;  // replay records in the lightcurve views' ranges
  int max_lc_rec = 0;          // the largest such range
  std::vector<ViewRec> h_views;
  std::vector<FrameRec> h_frames;
  std::vector<EvalRec> h_evals;
  std::vector<GeomSlot> h_slots;
  std::vector<int> h_eval_slot;
  std::vector<double> h_params;
  double *h_pin = nullptr;  // pinned staging copy of h_params (source of the graph's upload node)
  std::vector<std::vector<ParamTarget>> eval_targets;
  std::vector<std::map<const double *, ParamTarget>> model_params;
  // device
  ViewRec *d_views = nullptr;
  FrameRec *d_frames = nullptr;
  EvalRec *d_evals = nullptr;
  GeomSlot *d_slots = nullptr;
  int *d_eval_slot = nullptr, *d_slot_flags = nullptr;
  double *d_slot_logf = nullptr;
  // realize_photo: jobs, their bad-photometry outputs, the resolved values
  std::vector<PhotoJob> h_photo_jobs;
  PhotoJob *d_photo_jobs = nullptr;
  int *d_job_flag = nullptr;
  double *d_job_log = nullptr;
  double *d_lawloc = nullptr;
  long long lawloc_len = 0;
  int photo_nblk = 1;  // workgroups per realize_photo job (256 facets each)
  // the law-reading penalties: requested slots, noncosine's abscissae, values
  int pen_law_need = 0, noncos_n = 0;
  double *d_noncos = nullptr, *d_pen_law = nullptr;
  double *d_params = nullptr, *d_fit = nullptr, *d_obs = nullptr, *d_oov = nullptr;
  double *d_overflow = nullptr, *d_view_out = nullptr;
  double *d_lc_x = nullptr, *d_lc_t = nullptr, *d_lc_y = nullptr, *d_lc_y2 = nullptr;
  float *d_obs32 = nullptr, *d_oov32 = nullptr;
  int *d_bin_cnt = nullptr;                 // per 32x32 tile of every binned view
  unsigned int *d_bin_list = nullptr;       // BIN_CAP entries per tile: block id << 4 | quarter mask
  int2 *d_out_list = nullptr;               // (view, block) of blocks reaching outside the POS window
  int *d_tile_work = nullptr;               // the persistent tile kernel's work counters (2 NXCD ints)
  // Launch settings a tuner may change between replays (Tune)
  Tune tune;
  // penalties: the list, per-slot moments and terms, per-evaluation results
  PenDev pen{};
  int pen_need = 0, pen_nb = 1;
  double *d_mom = nullptr, *d_mom_part = nullptr;
  int download_penalties(int e, std::vector<double> &base, double *total);
  PosDev pd{};
  std::vector<void *> allocs;
  size_t dev_bytes = 0;     // device memory held by allocs
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  // Optional timing of eager launches: HIP events around every kernel on
  // `stream`; collect_kernel_times() folds the last launches into the totals.
  bool time_kernels = false;
  void tbegin(const char *name);
  void tend();
  int collect_kernel_times();
  std::vector<std::string> kt_names;
  std::vector<double> kt_ms;
  std::vector<long long> kt_count;

 private:
  std::vector<hipEvent_t> t_events;
  std::vector<const char *> t_names;
  int t_used = 0;
  int alloc(void **p, size_t bytes);
  int pack_model(int m, struct par_t *par, struct mod_t *mod, struct dat_t *dat, bool repack = false);
  int upload_host_geometry(int slot);
  int setup_penalties(struct par_t *par);
  void enqueue_penalties();
};

// record a failure of a reference entry point for shape_hip_last_error()
void record_error(int rc, const char *where);

// launch a kernel, bracketed by timing events when the engine times kernels
#define SHP_LAUNCH(E, name, ...)  \
  do {                            \
    (E).tbegin(name);             \
    hipLaunchKernelGGL(__VA_ARGS__); \
    (E).tend();                   \
  } while (0)

}  // namespace shp
