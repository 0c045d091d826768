/*
 * scene.h -- synthetic SHAPE scenes built directly as ABI structs.
 *
 * The reference builds par_t/mod_t/dat_t from text files (read_par.c,
 * read_mod.c, read_dat.c).  This module builds the same structs from a
 * numeric specification instead, so that benchmarks and parity tests can run
 * without data files.  Allocation follows the reference's conventions
 * (Numerical Recipes offset matrices, one POS block shared by every frame
 * when pos_scope is GLOBAL, read_dat.c:1239-1290).
 */
#ifndef SHAPE_SCENE_H
#define SHAPE_SCENE_H

#include <stddef.h>

#include "shape_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { SHP_SHAPE_ELLIPSOID_TESS = 0, SHP_SHAPE_ICOSPHERE = 1 };

/* Free-parameter kinds, numbered as mkparlist.c numbers fpartype. */
enum { SHP_SIZEPAR = 1, SHP_SHAPEPAR = 2, SHP_PHOTOPAR = 3, SHP_SPINPAR = 4,
       SHP_DELCORPAR = 5, SHP_DOPSCALEPAR = 6 };

typedef struct shp_spec {
  /* geometry */
  int shape_kind;          /* SHP_SHAPE_* */
  int theta_steps;         /* ellipsoid tessellation resolution (read_mod.c:2179) */
  int ico_level;           /* icosphere subdivision level */
  double semi_axes[3];     /* km */
  double rdev_sigma;       /* vertex radial deviation: r *= 1 + sigma*N(0,1) */
  unsigned long long seed_geom;
  unsigned long long seed_frames;
  /* spin state (principal-axis rotator) */
  double spin_period_h;
  double pole_lambda_deg, pole_beta_deg, phase0_deg;
  double t0;               /* spin epoch (JD) */
  double epoch_span_days;  /* frame epochs are spread over [t0, t0+span] */
  /* plane of sky */
  int pos_n;               /* POS is (2n+1)^2 pixels */
  double km_per_pixel;
  int pos_scope;           /* GLOBAL or LOCAL */
  int pos_smooth;
  /* radar */
  double Ftx;              /* MHz */
  int codemethod, spb, stride, dopfftlen;
  double R_min, R_max, C_min, C_max;   /* cosine law per set */
  int one_set_per_frame;   /* 1: every radar frame is its own set */
  /* delay-Doppler frames */
  int n_deldop;
  int ndel_min, ndel_max, ndop_min, ndop_max;
  double del_per_pixel;
  double dop_per_pixel;    /* <= 0: choose per frame so the echo fits */
  /* Doppler-only frames */
  int n_doppler;
  int nbin_min, nbin_max;
  double dop_per_bin_min, dop_per_bin_max;
  /* lightcurves */
  int n_lightcurves;
  int lc_points;           /* observed points per lightcurve */
  int lc_ncalc;            /* calculated points per lightcurve */
  double phase_min_deg, phase_max_deg;
  double kaas_wt;          /* Lambert weight of the Kaasalainen law */
  /* delay correction polynomial (usec, usec/day) */
  double delcor_a0, delcor_a1;
  /* viewing geometry drift: sub-observer latitude range (deg) */
  double obs_lat_min_deg, obs_lat_max_deg;
  /* smearing: views per radar frame, centred on the frame epoch (v0 = nviews / 2) */
  int radar_nviews;
  double view_interval_days;
  int lc_nviews;           /* smearing of lightcurve points (same interval) */
  /* Doppler response: sinc^2 evaluated at nsinc2 x nsinc2 points per POS
     pixel over sinc2width bins (par_t nsinc2 / sinc2width) */
  int nsinc2, sinc2width;
  /* component type (read_mod.c:289-420): VERTEX (0, the default), or the
     analytic ELLIPSE (1), OVOID (2) and HARMONIC (3) components realized from
     their parameters on the ellipsoid tessellation's directions
     (theta_steps); ovoid asymmetry k; harmonic degree and the rms of the
     random l >= 1 coefficients (relative to the mean radius) */
  int comp_type;
  double ovoid_k;
  int nhar;
  double har_sigma;
} shp_spec;

typedef struct shp_scene shp_scene;

void shp_spec_defaults(shp_spec *spec);
shp_scene *shp_scene_create(const shp_spec *spec);
void shp_scene_destroy(shp_scene *sc);
struct par_t *shp_scene_par(shp_scene *sc);
struct mod_t *shp_scene_mod(shp_scene *sc);
struct dat_t *shp_scene_dat(shp_scene *sc);
/* number of vertices / facets actually generated */
int shp_scene_nv(shp_scene *sc);
int shp_scene_nf(shp_scene *sc);

/* Turn the current fit images into observations: obs = A*fit + N(0,sigma),
   A chosen per frame so the brightest model pixel has signal-to-noise snr.
   Lightcurves get obs = fit*(1 + rel_sigma*N(0,1)).  Must be called after a
   calc_fits on the truth model. */
void shp_scene_synthesize_obs(shp_scene *sc, double snr, double lc_rel_sigma,
                              unsigned long long seed);

/* NR-style allocation (row pointers into one zeroed block). */
double **shp_matrix(int nrl, int nrh, int ncl, int nch);
int **shp_imatrix(int nrl, int nrh, int ncl, int nch);
double *shp_vector(int nl, int nh);
void shp_free_matrix(double **m, int nrl, int ncl);
void shp_free_imatrix(int **m, int nrl, int ncl);
void shp_free_vector(double *v, int nl);

/* ---- accessors used by tests and benchmarks ---- */
int shp_set_type(struct dat_t *dat, int s);
int shp_dat_nsets(struct dat_t *dat);
int shp_set_nframes(struct dat_t *dat, int s);      /* frames, or ncalc for lightcurves */
void shp_frame_dims(struct dat_t *dat, int s, int f, int out[2]); /* ndel, ndop (or n, ncalc) */
/* Copy a fit/obs/oneovervar image row-major into out.  which: 0 fit, 1 obs, 2 oneovervar */
void shp_frame_image(struct dat_t *dat, int s, int f, int which, double *out);
void shp_frame_set_image(struct dat_t *dat, int s, int f, int which, const double *in);
/* Per-frame scalars: chi2, cal, overflow_o2, overflow_m2, badradar_logfactor,
   dof, idellim[2], idoplim[2], dellim[2], doplim[2], km2Hz(v0), overflow_xsec,
   overflow_dopmean, weight -> 18 doubles */
void shp_frame_scalars(struct dat_t *dat, int s, int f, double *out);
struct pos_t *shp_frame_pos(struct dat_t *dat, int s, int f);
/* Copy a POS array (2n+1)^2 row-major (x major, y minor) into out.
   which: 0 b, 1 cose, 2 cosi, 3 z, 4 cosill, 5 zill  (doubles)
          10 f, 11 body, 12 comp, 13 fill               (ints, written as doubles) */
void shp_pos_array(struct pos_t *pos, int which, double *out);
void shp_pos_limits(struct pos_t *pos, int out[8]); /* xlim, ylim, xlim2, ylim2 */
/* Copy every array, rectangle and matrix of src into dst (same n) */
void shp_pos_copy(struct pos_t *dst, const struct pos_t *src);
/* posbnd_logfactor, bistatic, km_per_pixel, n */
void shp_pos_scalars(struct pos_t *pos, double out[4]);
/* the realized vertex set of component 0, and a set's deldop_t / doppler_t */
struct vertices_t *shp_mod_vertices(struct mod_t *mod);
void *shp_set_desc(struct dat_t *dat, int s);
/* view data: ae[9], oe[9], spin[3], intspin[3], deloff, dopoff -> 29 doubles */
void shp_view_data(struct dat_t *dat, int s, int f, int v, double *out);
void shp_set_view_data(struct dat_t *dat, int s, int f, int v, const double *in);
/* realized geometry: x[3*nv], vertex n[3*nv], facet n[3*nf], area[nf] */
void shp_geometry(struct mod_t *mod, double *vx, double *vn, double *fn, double *area);
void shp_facet_seen(struct mod_t *mod, unsigned char *out);
void shp_set_vertex_normals(struct mod_t *mod, const double *vn);
void shp_fits_bracket(struct par_t *par, struct mod_t *mod, struct dat_t *dat, int start);
void shp_set_offset(struct dat_t *dat, int s, int which, int j, int state, double val);
double shp_get_offset(struct dat_t *dat, int s, int which, int j);
/* par flags: posbnd, badradar, baddiam; logfactors posbnd, badradar, baddiam */
void shp_par_flags(struct par_t *par, double out[6]);
void shp_frame_consts(struct dat_t *dat, int s, int f, double out[9]);
void shp_par_set(struct par_t *par, int which, double v);
struct photo_t *shp_mod_photo(struct mod_t *mod);
int shp_vertex_facets(struct mod_t *mod, int v, int *out, int cap);
void shp_radar_law(struct mod_t *mod, struct dat_t *dat, int s, double out[2]);
void shp_pos_set_array(struct pos_t *pos, int which, const double *in);
void shp_pos_set_frames(struct pos_t *pos, const double *oe, const double *se);
int shp_set_radar_law(struct mod_t *mod, int l, int type, const double *p);
double *shp_param_ptr(shp_scene *sc, int i);
/* penalties (par->pen, 1-based as penalties.c reads it) and spin quantities */
int shp_scene_set_penalties(shp_scene *sc, int n, const int *type, const double *weight);
void shp_pen_base(struct par_t *par, double *out);
void shp_set_spin_inertia(struct mod_t *mod, const double *in);
int shp_set_impulses(struct mod_t *mod, int n, const double *t, const double *imp);
void shp_moments(struct mod_t *mod, int mode, double out[14]);
int shp_sides(struct mod_t *mod, int *out, int cap);
void shp_facets(struct mod_t *mod, int *out);
int shp_harmonic(struct mod_t *mod, int which, double *a, double *b);
void shp_set_vertex_r(struct mod_t *mod, const double *r);
void shp_set_comp_transform(struct mod_t *mod, const double *rot, const double *off);
/* vary_params: par's vary flags, calibration / Doppler-scale / delay-constant
   states, and the weight sums read_dat forms */
void shp_set_vary(struct par_t *par, int delcor0, int radalb, int optalb, int dopscale);
void shp_set_state(struct dat_t *dat, int s, int f, int which, int state, double val);
void shp_dat_vary_weights(struct dat_t *dat);
void shp_vertex_dirs(struct mod_t *mod, double *u, double *r);
int shp_param_count(shp_scene *sc);
int shp_param_type(shp_scene *sc, int i);
/* structs allocated by the caller (the input-file readers fill them):
   sizes (0 par_t, 1 mod_t, 2 dat_t), the mod_t / dat_t file name, component
   0's realized counts, vertex side lists, facet side triples, and a view's
   epoch with its plane-of-sky spin contribution (t, orbspin[3]; for a
   lightcurve, calculated point f's light-time corrected epoch) */
size_t shp_struct_size(int which);
void shp_set_name(int which, void *p, const char *name);
int shp_mod_nv(struct mod_t *mod);
int shp_mod_nf(struct mod_t *mod);
int shp_vertex_sides(struct mod_t *mod, int *count, int *out, int cap);
void shp_facet_sides(struct mod_t *mod, int *out);
void shp_view_epoch(struct dat_t *dat, int s, int f, int v, double out[4]);
double *shp_mod_param_ptr(struct mod_t *mod, int i);
/* non-principal-axis spin (w rad/day, moments I, par's int_method), and the
   spin epoch / angles / spin vector (7 doubles) */
void shp_set_npa(struct par_t *par, struct mod_t *mod, const double w[3], const double I[3], int method);
void shp_spin_state(struct mod_t *mod, double out[7]);
/* Deep copy of the parts of a scene the forward model writes (fit images,
   frame scalars, view data) from src into dst; both built from one spec. */

int shp_law_local(struct mod_t *mod, int radar, int ilaw, int f, double *out);
int shp_law_table(struct mod_t *mod, int ilaw, double *out, int cap);
void shp_facet_geom(struct mod_t *mod, int f, double out[5]);
void shp_badphoto(struct par_t *par, double out[2]);
double *shp_law_param_ptr(struct mod_t *mod, int radar, int ilaw, int which, int f);

#ifdef __cplusplus
}
#endif
#endif
