/*
 * oracle.h -- FP64 CPU oracle for SHAPE's model-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a clean-room restatement, in plain C, of
 * the reference's CPU routines in src/shape, used as the checker for the
 * HIP library.  Only the test suite, the smoke check and the CPU-baseline leg
 * of the benchmark may load it.  The product library never links, loads or
 * falls back to it.
 *
 * Every function takes the same ABI structs as the reference routine it
 * restates (include/shape_abi.h) and has the same side effects, except that
 * nothing here prints, exits, or keeps static state:
 *   - the Doppler overflow spectrum is a per-call buffer instead of the
 *     `static double fit_overflow[]` of pos2doppler.c:166;
 *   - the per-point printf in chi2_lghtcrv (chi2.c:1147) is omitted;
 *   - WRITE-action output is omitted;
 *   - fatal conditions (unsupported law or component type) set
 *     oracle_last_error() and return instead of calling bailout().
 *
 * Pins (tests/test_oracle.py): the reference ships no tests or golden
 * vectors, so each function is checked against closed forms, conservation
 * laws, brute force or an independent library.  Parity unpinned (restated
 * from the reference with no independent check): the tabular and
 * inhomogeneous radar laws (radlaw.c:56-66,118-135), the accumulation order
 * of vertex-normal sums, and libm sin/pow/exp at the ulp level.
 */
#ifndef SHAPE_ORACLE_H
#define SHAPE_ORACLE_H

#include "shape_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *oracle_last_error(void);
void oracle_clear_error(void);

/* model realization: realize_mod.c:153-565 (VERTEX branch, facnrm.c, vertex normals) */
void oracle_realize_mod(struct par_t *par, struct mod_t *mod);
/* spin state per view: realize_spin.c:41-280 + inteuler.c:86-113 (PA rotators) */
void oracle_realize_spin(struct par_t *par, struct mod_t *mod, struct dat_t *dat);
/* inteuler.c:114-392 (oracle_npa.c): an NPA rotator's spin state.  The
   integrator of one interval, its macro step count, the body-frame period of
   w, K(k) from the complementary modulus, and the whole state from the spin
   epoch through impulse epochs t[0..n-1] (method BRUTE or PERIODIC; 0 where
   the period is not finite) */
void oracle_npa_integrate(const double I[3], double t0, double t, double w[3], double m[3][3]);
int oracle_npa_steps(const double I[3], const double w[3], double dt);
double oracle_npa_period(const double inertia[3], const double w[3]);
double oracle_ellip_k(double kc);
int oracle_inteuler_npa(const struct spin_t *spin, const double t[], double imp[][3], int n, int method,
                        double w[3], double m[3][3]);
/* COM offsets: realize_delcor.c:50-151, deldopoffs.c:17-40, dopoffs.c:19-39 */
void oracle_realize_delcor(struct dat_t *dat, double delta_delcor0, int delcor0_mode);

/* moments and penalties: realize_mod.c:757-840 with facmom.c (oracle_realize_mod
   calls it last), diag_inertia.c (Jacobi, nr/jacobi.c), penalties.c:120-1250
   (the geometric, moment, spin and offset penalties; sets par->pen.base) */
void oracle_compute_moments(struct mod_t *mod);
void oracle_jacobi3(double a[3][3], double d[3], double v[3][3]);
void oracle_diag_inertia(double inertia[3][3], double pmoment[3], double ap[3][3]);
double oracle_penalties(struct par_t *par, struct mod_t *mod, struct dat_t *dat);

/* vary_params.c:87-325: the four means of the joint adjustments (first
   dataset only, as the reference loops) */
void oracle_vary_params(struct par_t *par, struct mod_t *mod, struct dat_t *dat, double *deldop_zmax,
                        double *rad_xsec, double *opt_brightness, double *cos_subradarlat);

/* the fit's one-parameter line search: nr/mnbrak.c:10, nr/brent_abs.c:13,
   bestfit.c:446-468 (mnbrak from (x0, x0 + step), then brent_abs) */
typedef double (*oracle_fn)(double);
void oracle_mnbrak(double *a, double *b, double *c, double *fa, double *fb, double *fc, oracle_fn f);
double oracle_brent_abs(double ax, double bx, double cx, oracle_fn f, double tol, double abstol, double *xmin);
double oracle_line_search(oracle_fn f, double x0, double step, double tol, double abstol, double *xmin);

/* plane-of-sky rendering: posclr.c:24, posvis.c:84-364, posmask.c:58-174 */
void oracle_posclr(struct pos_t *pos);
int oracle_posvis(struct vertices_t *verts, double orbit_offset[3], struct pos_t *pos,
                  int smooth, int src, int body, int comp);
void oracle_posmask(struct pos_t *pos, double tol);

/* radar synthesis: radlaw.c:44-148, pos2deldop.c:176-581, pos2doppler.c:149-501 */
double oracle_radlaw(struct photo_t *photo, int ilaw, double cosinc, int c, int f);
int oracle_pos2deldop(struct par_t *par, struct photo_t *photo, double orbit_xoff,
                      double orbit_yoff, double orbit_dopoff, struct deldop_t *deldop,
                      int body, int set, int frm, int v);
int oracle_pos2doppler(struct par_t *par, struct photo_t *photo, double orbit_xoff,
                       double orbit_yoff, double orbit_dopoff, struct doppler_t *doppler,
                       int body, int set, int frm, int v);

/* optical: apply_photo.c:78-309; natural cubic spline nr/spline.c, nr/splint.c */
/* realize_photo.c:122-969 (albedo factors 1, mode 0): '=' states, harmonic
   laws' local values at the facets' angular coordinates, bounds checks into
   par->badphoto */
void oracle_realize_photo(struct par_t *par, struct mod_t *mod);
/* associated Legendre function P_l^m with the Condon-Shortley phase */
double oracle_plm(int l, int m, double x);
/* astro/hapke.c:41-176: Hapke's bidirectional reflectance */
double oracle_hapke(double cosi, double cose, double phase, double w, double h, double B0, double g,
                    double theta);
double oracle_apply_photo(struct mod_t *mod, int ilaw, double phase, double intensityfactor,
                          struct pos_t *pos, int body, int s, int frm);
void oracle_spline(double *x, double *y, int n, double yp1, double ypn, double *y2);
void oracle_splint(double *xa, double *ya, double *y2a, int n, double x, double *y);

/* forward model and reduction: calc_fits.c:340-1959 (fit-time), chi2.c:240-1233 */
void oracle_calc_fits(struct par_t *par, struct mod_t *mod, struct dat_t *dat);
double oracle_chi2(struct par_t *par, struct dat_t *dat, int list_breakdown);
/* objective() of bestfit.c:718-871 with no penalties: chi2/dof times the
   bad-model factors.  Realization is the caller's job, as in the reference. */
double oracle_objective_tail(struct par_t *par, struct dat_t *dat);
/* objective() of bestfit.c:718-871 in full: chi2/dof + penalties, times the
   bad-model factors (the model realized by the caller) */
double oracle_objective(struct par_t *par, struct mod_t *mod, struct dat_t *dat);

/* HMT-style CPU baseline (HostMultiThreading/calc_fits_hmt.c): datasets are
   dealt round-robin to nthreads threads (inputnode = s mod T); requires
   pos_scope LOCAL so frames do not share POS memory. */
void oracle_calc_fits_threads(struct par_t *par, struct mod_t *mod, struct dat_t *dat,
                              int nthreads);

#ifdef __cplusplus
}
#endif
#endif
