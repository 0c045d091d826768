/*
 * shape_hip.h -- C ABI of libshape_hip.so, the MI355X (gfx950) implementation
 * of SHAPE's model-evaluation path.
 *
 * Part 1 keeps the reference's entry points, names and signatures, so the
 * library links in place of the reference's routines in existing SHAPE host
 * code (prototypes in the reference's src/shape/shape2.h:1443-1599).  Every
 * computation runs in HIP kernels; the caller's structs are read and written
 * exactly where the reference reads and writes them.  The reference reports
 * nothing for GPU failures, so these entries record the failure in
 * shape_hip_last_error() and leave their outputs untouched.
 *
 * Part 2 is the batched evaluation interface (no counterpart in the
 * reference): many trial parameter values, datasets and models evaluated in
 * one device-resident launch sequence replayed as a hipGraph.
 */
#ifndef SHAPE_HIP_H
#define SHAPE_HIP_H

#include "shape_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Part 1: reference entry points ---------------- */

/* realize_mod.c:153  vertex coordinates, facet normals/areas/centroids and
   vertex normals of the model's (single, VERTEX-type) component, its
   active flags (check_surface, :572) and its moments (compute_moments,
   :757: area, volume, centre of mass and inertia tensor of the component and
   of the model); sets par->baddiam and par->baddiam_logfactor. */
void realize_mod(struct par_t *par, struct mod_t *mod);

/* penalties.c:120  sum over par's penalty list of |weight| * base, each base
   the penalty (or its reciprocal for a negative weight) written into
   par->pen.base[i]; reads the realized model (coordinates, facet normals,
   side list, the moments realize_mod left), spin inertia and impulses, and
   the datasets' angle and spin offsets.  Types NONSMOOTH..MAXELLIPDEV,
   BIFURCATION and IMPULSE; others (the harmonic / inhomogeneous /
   tabular-law penalties) record an error and return 0. */
double penalties(struct par_t *par, struct mod_t *mod, struct dat_t *dat);

/* vary_params.c:87  the means behind bestfit's joint adjustments, from the
   first dataset as the reference loops: *deldop_zmax (distance toward Earth
   of the sub-radar point, frames whose delay constant floats),
   *rad_xsec (radar cross section, absolutely calibrated frames),
   *opt_brightness (absolutely calibrated lightcurve at its observation
   epochs), *cos_subradarlat (frames whose Doppler scale floats), each the
   weighted sum over dat's weight sums.  The model and spin states must be
   realized (realize_mod, realize_spin).  Radar frames and lightcurves of
   one view each (an error otherwise); the fit / POS scratch the reference
   leaves in those frames is not written.  action is unused, as in the
   single-process reference. */
void vary_params(struct par_t *par, struct mod_t *mod, struct dat_t *dat, int action, double *deldop_zmax,
                 double *rad_xsec, double *opt_brightness, double *cos_subradarlat);

/* realize_spin.c:41  Euler matrices ae, intrinsic and total spin vectors of
   every frame view and lightcurve point (principal-axis rotators). */
void realize_spin(struct par_t *par, struct mod_t *mod, struct dat_t *dat);

/* realize_delcor.c:50  delay / Doppler COM offsets of every view from the
   delay-correction polynomials (with the delcor0 modes and '=' states). */
void realize_delcor(struct dat_t *dat, double delta_delcor0, int delcor0_mode);

/* calc_fits.c:340  model fits of all delay-Doppler, Doppler and lightcurve
   datasets; writes fit images, per-frame limits, overflow terms, facet
   "seen" flags and the posbnd / badradar flags and log factors in par.
   POS arrays are written back only after shape_hip_set_strict(1). */
void calc_fits(struct par_t *par, struct mod_t *mod, struct dat_t *dat);

/* calc_fits.c:440 / :830 / :1509  the fits of dataset s alone (a
   delay-Doppler, Doppler or lightcurve set), as the HostMultiThreading
   driver calls them (calc_fits_hmt.c:146-155): raises par's posbnd /
   badradar flags and adds the un-normalized log factors, or-s in the
   facets' seen flags.  Disjoint datasets may run from several host threads
   at once. */
void calc_deldop(struct par_t *par, struct mod_t *mod, struct deldop_t *deldop, int s);
void calc_doppler(struct par_t *par, struct mod_t *mod, struct doppler_t *doppler, int s);
void calc_lghtcrv(struct par_t *par, struct mod_t *mod, struct lghtcrv_t *lghtcrv, int s);

/* realize_spin.c:286 / :317  the '=' states of the datasets' angle and spin
   offsets (write_dat.c:123-124 calls them) */
void realize_angleoff(struct dat_t *dat);
void realize_omegaoff(struct dat_t *dat);

/* chi2.c:240  chi-square of the current fits; floats calibration factors,
   writes frame->chi2, cal.val, set chi2 and dat->chi2, returns dat->chi2. */
double chi2(struct par_t *par, struct dat_t *dat, int list_breakdown);

/* chi2 (above) on fits, observations and weights already in device memory,
   allocating no device memory: no counterpart in the reference, for callers
   that cannot allocate (code under hipGraph capture, tight serving loops).
   d_fit / d_obs / d_oov hold every frame's cells in dataset order, each
   frame row-major (delay-Doppler: ndel rows of ndop; Doppler: ndop;
   lightcurve: n points), as chi2 flattens the structs.  Frame shapes,
   calibration states, weights and overflow terms come from dat; results go
   into dat as chi2 writes them, and the total is returned.  Its frame
   records and sums live in a fixed device array per GPU (at most 8192
   frames and 8192 datasets; more is an error): calls that overlap on one
   GPU race. */
double shape_chi2_noalloc(struct par_t *par, struct dat_t *dat, const double *d_fit, const double *d_obs,
                          const double *d_oov);

/* posclr.c:24  blank sky over the whole POS (z = -1e20, cos = b = 0,
   f/body/comp = -1, and the *ill arrays for a bistatic POS); empty rectangles. */
void posclr(struct pos_t *pos);

/* posvis.c:84  rasterize a realized vertex set (x, facet and vertex normals)
   into the observer (src = 0) or source (src = 1) view of the POS with the
   CPU's z-buffer rule; a pixel the POS already owns goes only to a nearer
   facet.  Writes z, cos, f/body/comp and the rectangle; returns outbnd and
   sets pos->posbnd_logfactor. */
int posvis(struct vertices_t *verts, double orbit_offset[3], struct pos_t *pos, int smooth, int src, int body,
           int comp);

/* posmask.c:58  shadowing: zero pos->cose where the source view shows a
   different surface element nearer than tol. */
void posmask(struct pos_t *pos, double tol);

/* radlaw.c:44  radar scattering law ilaw at cos(incidence) cosinc (cosine,
   Gaussian, Hagfors, cosine-QS and the hybrid laws). */
double radlaw(struct photo_t *photo, int ilaw, double cosinc, int c, int f);

/* pos2deldop.c:176  add view v's echo power from frame frm's POS into
   frame->fit (accumulated); sets idellim, idoplim, dellim, doplim, km2Hz,
   the overflow terms and badradar_logfactor; returns badradar. */
int pos2deldop(struct par_t *par, struct photo_t *photo, double orbit_xoff, double orbit_yoff,
               double orbit_dopoff, struct deldop_t *deldop, int body, int set, int frm, int v);

/* pos2doppler.c:149  as pos2deldop for a Doppler spectrum, with the
   overflow spectrum's o2, m2, cross section and mean Doppler. */
int pos2doppler(struct par_t *par, struct photo_t *photo, double orbit_xoff, double orbit_yoff,
                double orbit_dopoff, struct doppler_t *doppler, int body, int set, int frm, int v);

/* apply_photo.c:78  brightness of each lit and illuminated pixel of body into
   pos->b (Lambert, Lommel-Seeliger, geometric, Kaasalainen); returns the
   sum, or 1e-40 when it is zero. */
double apply_photo(struct mod_t *mod, int ilaw, double phase, double intensityfactor, struct pos_t *pos,
                   int body, int s, int frm);

/* ---------------- library control ---------------- */
int shape_hip_last_error(void);       /* 0, or the code of the last failure */
void shape_hip_clear_error(void);
void shape_hip_set_strict(int write_pos); /* 1: calc_fits also writes POS arrays back */
int shape_hip_device_count(void);
/* Self-test of the device's wave-level primitives (no counterpart in the
 * reference): over 256 input ints, per thread: the inclusive prefix sum, min,
 * max and sum over its 64-lane wave, and the FP32 min and max (as bit
 * patterns).  out holds 6 x 256 ints.  0 on success. */
int shape_hip_wave_selftest(const int *in, int *out);
void shape_hip_realize_delcor(struct par_t *par, struct mod_t *mod, struct dat_t *dat,
                              double delta_delcor0, int delcor0_mode);

/* calc_fits and the one-dataset entries keep each call site's device arenas
   (model, dataset) across calls and refresh them from the structs; this
   frees every cached arena not in use (no reference counterpart) */
void shape_hip_release_cache(void);

/* number of POS pixels whose cos(scattering) the last calc_fits inherited
   from an earlier-claiming facet (diagnostic) */
int shape_hip_last_replayed(void);

/* ---------------- Part 2: batched evaluation (no reference counterpart) ----
   Evaluates many trial parameter values of bestfit's objective (bestfit.c:718)
   at once: realization, rendering, synthesis, chi-square, par's penalties
   (penalties.c, the geometric, moment, spin and offset types) and the
   bad-model factors for every evaluation, device-resident.  Each
   evaluation e uses model eval_model[e] (an index into mods/dats) and sets
   n_targets[e] parameters, given as addresses of `val` fields inside that
   model's structs (e.g. par->fpntr entries), to the given trial values.
   targets/values are flattened over evaluations. */
typedef struct shape_batch shape_batch;

#define SHAPE_METHOD_FP64 0   /* plain FP64 (the drop-in's kernels) */
#define SHAPE_METHOD_MIXED 1  /* FP32 bulk, FP64 setup and reductions */

#define SHAPE_BATCH_GRAPH 1     /* capture one replay into a hipGraph */
#define SHAPE_BATCH_KEEP_FIT 2  /* keep fit images for inspection (single replay) */
#define SHAPE_BATCH_GLOBAL_ZBUF 4 /* mixed method: rasterize every view through a global z-buffer
                                     instead of per-tile LDS z-buffers (same results; for checking) */
#define SHAPE_BATCH_FACET_MAP 8   /* keep every POS pixel's owner facet (shape_batch_facet_map) */

shape_batch *shape_batch_create(struct par_t *par, struct mod_t **mods, struct dat_t **dats, int nmodels,
                                int nevals, const int *eval_model, const int *n_targets,
                                double *const *targets, const double *values, int method, int flags);
int shape_batch_set_values(shape_batch *b, const double *values);
int shape_batch_launch(shape_batch *b);                 /* enqueue one replay */
int shape_batch_wait(shape_batch *b, double *objective, double *chi2v);
int shape_batch_run(shape_batch *b, double *objective, double *chi2v);
float shape_batch_last_ms(shape_batch *b);              /* device time of the last replay */
int shape_batch_nframes(shape_batch *b, int e);
int shape_batch_frame_results(shape_batch *b, int e, double *chi2v, double *cal);
int shape_batch_fit(shape_batch *b, int e, int frame, double *out);
/* owner facet of every POS pixel of view v of frame `frame` of evaluation e
   ((2n+1)^2 ints, row i = -n..n, column j = -n..n; -1: no owner), from the
   last replay; needs SHAPE_BATCH_FACET_MAP; radar frames only */
int shape_batch_facet_map(shape_batch *b, int e, int frame, int v, int *out);
/* per frame of evaluation e: o2, m2, om, overflow_o2, overflow_m2, cal, chi2,
   badradar_logfactor (8 doubles per frame, set/frame order) */
int shape_batch_frame_terms(shape_batch *b, int e, double *out);
/* penalties of evaluation e after the last replay: par->pen.base[1..n] into
   base[0..n) and their weighted sum (the term added to chi2/dof) into *total;
   returns n.  shape_batch_create fails when par lists a penalty type outside
   NONSMOOTH..MAXELLIPDEV, BIFURCATION, IMPULSE, or one that reads the side
   list of a model without one, and when par turns on a vary_* joint
   adjustment (evaluate those through the drop-ins: vary_params, realize_*,
   calc_fits, chi2, penalties). */
int shape_batch_penalties(shape_batch *b, int e, double *base, double *total);
/* [0] pixels replayed in FP64, [1] error flags, [2] mixed: replayed pixels
   whose candidate facets a second raster pass collected (the rest come from
   their tile's own list), [3] reserved (0; footprints of every width are
   synthesized) */
int shape_batch_counters(shape_batch *b, int *out4);
/* nevals, nviews, nframes, geometry slots, POS pixels, fit cells, max facets, max POS pixels */
int shape_batch_sizes(shape_batch *b, long long *out8);
void shape_batch_destroy(shape_batch *b);

/* ---------------- Part 3: batched line search (replaces mnbrak + brent_abs) --
   The fit's search over one parameter (bestfit.c:446-468: nr/mnbrak.c:10
   brackets a minimum from x0 and x0 + step, nr/brent_abs.c:13 refines it to
   tol * |x| + max(abstol, 1e-10)) with k trial values per round instead of
   one: the first round spreads x0 + step * {0, +1, -1, +g, -g, ...} (g the
   golden ratio), rounds continue past an end point that is lowest, and with
   a bracket (a, x, b) a round evaluates the parabola's vertex u, u -+ tol1
   and an even grid over (a, b).  Stops on brent_abs's test, so the returned
   minimum carries the same tolerance.  Returns 0 (converged), 1 (max_rounds
   reached: *xmin is the lowest point so far), -1 (bad arguments), -2 (an
   evaluation failed), -3 (shape_batch_line_search: the batch's evaluations
   do not all set the same single parameter). */
#define SHAPE_SEARCH_MAX_K 64
/* evaluate the objective at x[0..k); return 0, or nonzero on failure */
typedef int (*shape_eval_fn)(void *ctx, int k, const double *x, double *f);
int shape_line_search(shape_eval_fn eval, void *ctx, int k, double x0, double step, double tol, double abstol,
                      int max_rounds, double *xmin, double *fmin, int *rounds);
/* the same, each round one replay of b: b's k evaluations each set the same
   single parameter (its trial values); its objectives are the f values */
int shape_batch_line_search(shape_batch *b, double x0, double step, double tol, double abstol, int max_rounds,
                            double *xmin, double *fmin, int *rounds);

/* nsearch independent searches sharing every round (eval gets nsearch * k
   values, search s at [s k, (s + 1) k)); a search that has stopped repeats
   its best point.  status[s]: 0 converged, 1 max_rounds reached.  Returns
   the worst status, or -1 / -2 as above. */
int shape_line_searches(shape_eval_fn eval, void *ctx, int nsearch, int k, const double *x0, const double *step,
                        double tol, double abstol, int max_rounds, double *xmin, double *fmin, int *status,
                        int *rounds);
/* the same over a batch whose evaluations form nsearch groups of k = nevals
   / nsearch consecutive evaluations, each group setting one parameter (of
   any of the batch's models): many models' searches in one replay per
   round.  -3: the groups do not each set one common parameter. */
int shape_batch_line_searches(shape_batch *b, int nsearch, const double *x0, const double *step, double tol,
                              double abstol, int max_rounds, double *xmin, double *fmin, int *status, int *rounds);
/* One pass of bestfit's loop over the fitted parameters (bestfit.c:392-470,
   mnbrak + brent_abs per parameter): every evaluation of b sets the same P
   parameters in the same order; for p = 0 .. P-1 parameter p is searched
   from x[p] with step[p], tol[p], abstol[p] (bestfit's fparstep, fpartol,
   fparabstol), nevals trial values per round, the others held at x, and
   x[p] takes the minimum.  fmin: the objective at the pass's end; rounds:
   replays taken.  Returns 0, 1 (a search reached max_rounds), -1, -2, -3. */
int shape_batch_fit_pass(shape_batch *b, const double *step, const double *tol, const double *abstol, int max_rounds,
                         double *x, double *fmin, int *rounds);

/* ---------------- Part 4: launch-setting tuner (no counterpart in the reference) --
   Settings that change how the mixed method spreads its work over the
   device, never its results: [0] tile kernel form (0: one workgroup per
   32x32 tile item; > 0, a multiple of 8: persistent, that many workgroups
   taking tile items from per-XCD work counters), [1] FP32 replay workgroups
   (at most), [2] candidate-collection workgroups.  shape_batch_tune times
   `replays` replays (median, after one untimed) under each candidate, one
   setting at a time -- tile form (per-tile grid; persistent with 2, 4 or 6
   workgroups per compute unit), then replay grid, then collection grid --
   and keeps the fastest; the choice is cached per problem shape for the
   process (force = 1 tunes again) and, with SHAPE_HIP_TUNE_FILE set, read
   from and appended to that file.  settings (3 ints) and best_ms may be
   NULL.  A batch with SHAPE_BATCH_GRAPH is captured again under the chosen
   settings.  The FP64 method has no such settings (returns its defaults). */
int shape_batch_tune(shape_batch *b, int replays, int force, int *settings, double *best_ms);
int shape_batch_get_tune(shape_batch *b, int *settings);
/* -1: invalid settings (a persistent grid not a multiple of 8, or on a batch
   without binned views) */
int shape_batch_set_tune(shape_batch *b, const int *settings);

/* measurement: the batch's hipStream_t; per-kernel HIP-event timing of eager
   replays (batches created without SHAPE_BATCH_GRAPH); and the last replay's
   footprint: model-rectangle pixels, POS pixels, occupied image cells, image
   cells, vertices, facets, claimed z-buffer keys (7 values) */
void *shape_batch_stream(shape_batch *b);
int shape_batch_time_kernels(shape_batch *b, int on);
int shape_batch_kernel_stat(shape_batch *b, int idx, char *name, int namelen, double *total_ms,
                            long long *launches);
int shape_batch_footprint(shape_batch *b, long long *out7);

#ifdef __cplusplus
}
#endif
#endif
