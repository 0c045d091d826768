/*
 * shape_abi.h -- the SHAPE data structures as a binary interface.
 *
 * Every struct below has the same members, in the same order and with the same
 * C types, as the corresponding declaration in the reference's
 * src/shape/shape2.h:411-1436 (par_t :431, vertex_t :723, facet_t :740,
 * vertices_t :804, comp_t :819, photo_t :994, spin_t :1007, mod_t :1024,
 * pos_t :1061, deldopview_t :1097, deldopfrm_t :1114, deldop_t :1184,
 * dopfrm_t :1230, doppler_t :1266, crvrend_t :1344, lghtcrv_t :1358,
 * set_t :1404, dat_t :1420).  Layout identity is what lets this library
 * replace the reference's routines inside existing SHAPE host code, so the
 * member lists are an interface, not an implementation choice.  Constants are
 * those of src/shape/const.h and src/macros/func.h.
 *
 * Matrices follow the Numerical Recipes convention of the reference
 * (src/nr/nrutil.c): an array of row pointers, offset so that the first valid
 * index is 1 (images) or -n (plane-of-sky frames).
 *
 * Pure C; no GPU types appear anywhere in this header.
 */
#ifndef SHAPE_ABI_H
#define SHAPE_ABI_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (src/shape/const.h) ---------------------------------- */
#define WRITE 1
#define COVAR 2
#define FORMAT 3
#define FACETS 4
#define SAMPLE 5
#define MIRROR 6
#define FIT 7
#define WAVEFRONT 8
#define REFSHAPE 9
#define MOMENTS 10
#define DELCORINIT 11
#define CONVEXHULL 12
#define MAP 13
#define VIEW 14
#define AREA 15
#define PHOTOFACETS 16
#define PHOTOHARM 17
#define SLICE 18
#define ORBIT 19
#define SPLIT 20

#define NONE 0
#define GLOBAL 1
#define LOCAL 2
#define BRUTE 1
#define PERIODIC 2

/* image resampling (dd_scaling, poset_scaling, poset_resample) */
#define BLOCK 1
#define BILINEAR 2
#define BICUBIC 3
#define CUBICCONV 4
/* byte order of binary data files */
#define BIG_ENDIAN_DATA 1
#define LITTLE_ENDIAN_DATA 2
/* scattering law of rendered views */
#define OPTICALVIEW 1
#define RADARVIEW 2
#define LAMBERTVIEW 3
/* joint variation of calibration-like parameters with size or spin */
#define VARY_NONE 0
#define VARY_SIZE 1
#define VARY_SPIN 2
#define VARY_ALL 3
/* pa_ephfile formats and the scalefitobs options */
#define HORIZONS 0
#define DATATAKING 1
#define SCALE_SEPARATELY 0
#define SCALE_MAXFITOBS 1
#define SCALE_FIT 2
#define SCALE_OBS 3
/* smearing: the frame epoch is the centre view or the first view */
#define SMEAR_CENTER 0
#define SMEAR_FIRST 1

#define MAPMODE_DELDOP 0
#define MAPMODE_POS 1
#define MAPMODE_FACETS 2

#define D2R 0.01745329251994330
#define R2D 57.29577951308232
#define PIE 3.141592653589793
#define PIOVER2 1.570796326794897
#define TWOPI 6.283185307179586
#define HUGENUMBER 1.0e20
#define AU 1.4959787061e8
#define SMALLVAL 1.0e-10
#define TINYCALFACT 1.0e-40
#define LN10 2.302585092994046
#define HALFSECOND 5.8e-6
#define DAYSPERAU 5.775518328e-3
#define GRAVCONST 4.98232e-10
#define SOLAR_VMAG -26.762

#define SHORT 0
#define LONG_ORIG 1
#define LONG_MOD 2

#define MAXOVERFLOW 2000
#define KM2US 6.67128190396304
#define KM2HZFACT 7.72139109254982e-5
#define MAXBINS 100

/* penalty types, par->pen.type[1..n] (const.h:102-126) */
#define OPTALBDEL 1
#define OPTALBVAR 2
#define THETAVAR 3
#define THETADEL 4
#define RADALBDEL 5
#define RADALBVAR 6
#define RAD_C_DEL 7
#define RAD_C_VAR 8
#define NONSMOOTH 9
#define CONCAVITY 10
#define RDEV 11
#define VOLUME 12
#define COMDEV 13
#define INERTIADEV 14
#define INERTIADEV_UNI 15
#define PA3TILT 16
#define NONPA 17
#define NONPA_UNI 18
#define EULEROFFS 19
#define FLATTENING 20
#define MAXRDEV 21
#define MAXELLIPDEV 22
#define NONCOSINE 23
#define BIFURCATION 24
#define IMPULSE 25

/* ---- par_t and its penalty block (shape2.h:415-716) ------------------ */
struct pen_t {
  int n;
  int *type;
  double *weight;
  double *base;
  double *x_noncosine;
  double *y_noncosine;
  double xmean_noncosine;
};

#define MAXDELCORPAR 10
#define MAXSUNMAGS 10
#define MAXCHOSENFACETS 20

struct par_t {
  unsigned char action, speckle, pos_smooth, pos_scope, dd_scaling, int_method,
      lcrv_pos, posbnd, badphoto, baddiam, badposet, badradar, baddopscale;
  double bad_objfactor;
  double posbnd_logfactor, baddiam_logfactor, badphoto_logfactor,
      badposet_logfactor, badradar_logfactor, baddopscale_logfactor;
  unsigned char term_badmodel, avoid_badpos, warn_badradar, showstate, read_node;
  int theta_steps, ver_samps, pos_pixels;
  double pos_width, dd_gamma, mask_tol;
  double length_step, length_tol, length_abstol;
  double ratio_step, ratio_tol, ratio_abstol;
  double angle_step, angle_tol, angle_abstol;
  double spin_step, spin_tol, spin_abstol;
  double spindot_step, spindot_tol, spindot_abstol;
  double lib_amp_step, lib_amp_tol, lib_amp_abstol;
  double lib_freq_step, lib_freq_tol, lib_freq_abstol;
  double lib_phase_step, lib_phase_tol, lib_phase_abstol;
  double photo_step, photo_tol, photo_abstol;
  double inertia_step, inertia_tol, inertia_abstol;
  double delcor_step[MAXDELCORPAR];
  double delcor_tol;
  double delcor_abstol[MAXDELCORPAR];
  double xyoff_step, xyoff_tol, xyoff_abstol;
  double term_prec;
  int term_maxiter;
  double dd_resid, optposmax, radposmax, radfitmin, radfitmax, radobsmin, radobsmax;
  char pa_ephfile[255];
  unsigned char pa_ephformat;
  double pa_startepoch, pa_stopepoch;
  double pa_angleoff[3];
  double pa_omegaoff[3];
  int nsinc2, sinc2width;
  unsigned char scalefitobs, listres;
  int dd_clockwiserot, dd_xflip, dd_yflip;
  double delcorthresh, jdoffset;
  unsigned char perform_ltc;
  char convex_file[255];
  struct pen_t pen;
  unsigned char plot_spinvec, plot_subradar, plot_com;
  unsigned char plot_pa[3];
  unsigned char plot_angmom, pierce_spinvec, pierce_angmom, write_obs, delcor_read;
  char delcor_file[255];
  unsigned char delcor_verbose;
  int first_fitpar;
  double view_long, view_lat;
  unsigned char view_shadows;
  double view_sunlong, view_sunlat;
  unsigned char listfit;
  char listfit_path[255];
  double area_latincr, area_longincr;
  unsigned char poset_resample, image_rebin, poset_scaling;
  double poset_resid;
  unsigned char list_posetcal;
  double delta_rotphase;
  unsigned char endian, mark_unseen;
  double mincosine_seen;
  unsigned char lcrv_writeall;
  int exclude_seen;
  unsigned char pa_scatlaw, view_scatlaw, sky_radlaw, sky_optlaw, radfacets;
  char rad_global_R_state, rad_global_C_state, rad_local_R_state, rad_local_C_state;
  unsigned char optfacets;
  char opt_global_R_state, opt_global_wt_state, opt_global_A0_state,
      opt_global_D_state, opt_global_k_state, opt_global_w_state,
      opt_global_h_state, opt_global_B0_state, opt_global_g_state,
      opt_global_theta_state;
  char opt_local_R_state, opt_local_wt_state, opt_local_A0_state,
      opt_local_D_state, opt_local_k_state, opt_local_w_state,
      opt_local_h_state, opt_local_B0_state, opt_local_g_state,
      opt_local_theta_state;
  unsigned char radharm;
  int rad_R_nharm, rad_C_nharm;
  unsigned char optharm;
  int opt_R_nharm, opt_wt_nharm, opt_A0_nharm, opt_D_nharm, opt_k_nharm,
      opt_w_nharm, opt_h_nharm, opt_B0_nharm, opt_g_nharm, opt_theta_nharm;
  double rad_R_min, rad_R_max, rad_C_min, rad_C_max, rad_rho_min, rad_rho_max;
  double opt_R_min, opt_R_max, opt_wt_min, opt_wt_max, opt_A0_min, opt_A0_max,
      opt_D_min, opt_D_max, opt_k_min, opt_k_max, opt_w_min, opt_w_max,
      opt_h_min, opt_h_max, opt_B0_min, opt_B0_max, opt_g_min, opt_g_max,
      opt_theta_min, opt_theta_max;
  double sun_appmag[MAXSUNMAGS];
  unsigned char vary_delcor0, vary_radalb, vary_optalb, vary_dopscale;
  double slice_long, slice_lat, slice_offset, slice_viewlong, slice_viewlat,
      slice_sunlong, slice_sunlat;
  unsigned char slice_scatlaw, slice_shadows;
  double slice_planefrac, slice_skyfactor, slice_dimfactor;
  unsigned char slice_read, slice_dointerp;
  double semimajor_axis[2];
  double r_pericenter[2];
  double eccentricity[2];
  double t_pericenter[2];
  double long_asc_node[2];
  double inclination[2];
  double arg_pericenter[2];
  double binary_gravparam[2];
  double orbit_period[2];
  unsigned char orbit_reflex;
  int is_triple;
  double orbit_posfactor[3];
  double write_posfactor, slice_posfactor, view_posfactor, map_posfactor;
  char maskdir[80];
  unsigned char write_chi2fit0;
  double chi2fit0_thresh;
  int npar_update;
  double split_norm[3];
  double split_const;
  int map_set, map_frame;
  unsigned char map_mode;
  int map_dellim[2];
  int map_doplim[2];
  int map_xposlim[2];
  int map_yposlim[2];
  int map_comp[MAXCHOSENFACETS];
  int map_facet[MAXCHOSENFACETS];
  double map_fitcutoff, map_poscutoff;
  unsigned char map_verbose, map_overflow;
  double int_abstol;
  unsigned char pa_bodycoords, pa_highlight;
  int pa_comp[MAXCHOSENFACETS];
  int pa_facet[MAXCHOSENFACETS];
  unsigned char view_highlight;
  int view_comp[MAXCHOSENFACETS];
  int view_facet[MAXCHOSENFACETS];
  unsigned char write_highlight;
  int write_comp[MAXCHOSENFACETS];
  int write_facet[MAXCHOSENFACETS];
  double dopscale_min, dopscale_max;
  double objfunc_start;
  unsigned char listpos_deldop, listpos_opt;
  char listpos_path[255];
  int nfpar;
  double **fpntr;
  double *fparstep;
  double *fpartol;
  double *fparabstol;
  int *fpartype;
};

/* ---- mod_t and its parts (shape2.h:719-1031) --------------------------- */
struct param_t {
  char state; /* 'f' free, 'c' constant, '=' tied */
  double val;
};

struct vertex_t {
  double a[3];
  double u[3];
  struct param_t r;
  double x[3];
  double n[3];
  int naf;
  int *af;
  int nas;
  int *as;
  unsigned char act;
  double **afactor;
  double **bfactor;
  int v_mirror;
};

struct facet_t {
  int v[3];
  int s[3];
  double n[3];
  unsigned char act;
  unsigned char seen;
  double theta;
  double phi;
  double area;
  double x[3];
};

struct side_t {
  int v[2];
  int f[2];
  unsigned char act;
};

struct ellipse_t {
  int ntheta;
  struct param_t two_a, a_over_b, b_over_c;
};

struct ovoid_t {
  int ntheta;
  struct param_t two_a, a_over_b, b_over_c, k;
};

struct harmonic_t {
  int ntheta;
  int nhar;
  struct param_t scalefactor[3];
  struct param_t **a;
  struct param_t **b;
  double **a_save;
  double **b_save;
};

struct vertices_t {
  struct param_t scalefactor[3];
  int nv;
  struct vertex_t *v;
  int nf;
  struct facet_t *f;
  int ns;
  struct side_t *s;
};

#define ELLIPSE 1
#define OVOID 2
#define HARMONIC 3
#define VERTEX 4
struct comp_t {
  unsigned char type;
  struct param_t off[3];
  struct param_t rot[3];
  double m[3][3];
  union desc_t {
    struct ellipse_t ell;
    struct ovoid_t ovoid;
    struct harmonic_t har;
    struct vertices_t ver;
  } desc;
  struct vertices_t real;
  double inertia[3][3];
  double com[3];
  double volume;
  double area;
};

struct shape_t {
  int ncomp;
  struct comp_t *comp;
  double inertia[3][3];
  double com[3];
  double volume;
  double area;
};

/* radar scattering laws */
#define NOLAW 0
#define COSINELAW_DIFF 1
struct RC_t {
  struct param_t R, C;
  double R_save;
};
#define GAUSSIANLAW 2
#define HAGFORSLAW 3
#define COSINELAW_QS 4
struct quasispec_t {
  struct param_t R, C;
  double cos_cutoff;
  double R_save;
};
#define GAUSSIAN_COSINE 5
#define HAGFORS_COSINE 6
#define COSINE_COSINE 7
struct hybridradar_t {
  struct quasispec_t qs;
  struct RC_t diff;
};
#define HARMCOSINE_DIFF 8
struct harmcosine_t {
  struct harmonic_t R, C;
  struct RC_t **local;
};
#define INHOCOSINE_DIFF 9
struct inhocosine_t {
  struct RC_t global;
  struct RC_t **local;
};
#define TABULARLAW 10
struct tabular_t {
  int n;
  struct param_t *rho;
  double *rho_save;
};
union radscat_t {
  struct RC_t RC;
  struct tabular_t tabular;
  struct quasispec_t quasispec;
  struct hybridradar_t hybrid;
  struct harmcosine_t harmcosine;
  struct inhocosine_t inhocosine;
};

/* optical scattering laws */
#define GEOMETRICAL 1
#define LAMBERTLAW 2
#define LOMMEL 3
struct R_t {
  struct param_t R;
  double R_save;
};
#define HARMLAMBERT 4
#define HARMLOMMEL 5
struct harmR_t {
  struct harmonic_t R;
  struct R_t **local;
};
#define INHOLAMBERT 6
#define INHOLOMMEL 7
struct inhoR_t {
  struct R_t global;
  struct R_t **local;
};
#define HAPKE 8
struct hapke_t {
  struct param_t w, h, B0, g, theta;
  double w_save;
};
#define HARMHAPKE 9
struct harmhapke_t {
  struct harmonic_t w, h, B0, g, theta;
  struct hapke_t **local;
};
#define INHOHAPKE 10
struct inhohapke_t {
  struct hapke_t global;
  struct hapke_t **local;
};
#define KAASALAINEN 11
struct kaas_t {
  struct param_t R, wt, A0, D, k;
  double R_save;
};
#define HARMKAAS 12
struct harmkaas_t {
  struct harmonic_t R, wt, A0, D, k;
  struct kaas_t **local;
};
#define INHOKAAS 13
struct inhokaas_t {
  struct kaas_t global;
  struct kaas_t **local;
};
union optscat_t {
  struct R_t R;
  struct harmR_t harmR;
  struct inhoR_t inhoR;
  struct hapke_t hapke;
  struct harmhapke_t harmhapke;
  struct inhohapke_t inhohapke;
  struct kaas_t kaas;
  struct harmkaas_t harmkaas;
  struct inhokaas_t inhokaas;
};

struct photo_t {
  int nradlaws;
  unsigned char *radtype;
  union radscat_t *radar;
  int noptlaws;
  unsigned char *opttype;
  union optscat_t *optical;
};

#define MAXIMP 10
struct spin_t {
  unsigned char pa;
  double t0;
  struct param_t angle[3];
  struct param_t omega[3];
  struct param_t omegadot[3];
  int n_impulse;
  double t_impulse[MAXIMP];
  struct param_t impulse[MAXIMP][3];
  struct param_t inertia[3];
  struct param_t lib_amp;
  struct param_t lib_freq;
  struct param_t lib_phase;
};

struct mod_t {
  int nfpar;
  struct shape_t shape;
  struct photo_t photo;
  struct spin_t spin;
  char name[80];
};

/* ---- dat_t and its parts (shape2.h:1038-1436) -------------------------- */
struct delcor_t {
  double t0;
  int n;
  struct param_t *a;
  double delcor0_save;
  int *equals_set;
};

struct ephpnt_t {
  double t, ra, dec, dist;
};

struct ephem_t {
  int n;
  struct ephpnt_t *pnt;
};

struct pos_facet_t; /* never defined by the reference; pointer only */

struct pos_t {
  int n;                /* frame is (2n+1)^2 pixels, indices -n..n */
  double km_per_pixel;
  double **b;
  double **cose;
  double **cosi;
  double **z;
  float *b_s;
  double *b_d;
  float *cose_s;
  float *cosi_s;
  float *z_s;
  int **body;
  int **comp;
  int **f;
  int xlim[2];
  int ylim[2];
  int bistatic;
  double **cosill;
  double **zill;
  float *cosill_s;
  float *zill_s;
  int **bodyill;
  int **compill;
  int **fill;
  int xlim2[2];
  int ylim2[2];
  double ae[3][3];
  double oe[3][3];
  double se[3][3];
  double posbnd_logfactor;
  int nf;
  struct pos_facet_t *facet;
};

struct deldopview_t {
  double t;
  double ae[3][3];
  double oe[3][3];
  double intspin[3];
  double orbspin[3];
  double spin[3];
  int idellim[2];
  int idoplim[2];
  double dellim[2];
  double doplim[2];
  double deloff;
  double dopoff;
  double km2Hz;
};

struct deldopfrm_t {
  double t0;
  int n_integrate;
  double t_integrate[MAXIMP + 2];
  double impulse[MAXIMP + 2][3];
  double nlooks;
  double weight;
  int pixels_weighted;
  double sdev;
  double delcom;
  struct param_t cal;
  int ndel;
  int ndop;
  double delcom_vig;
  double dopcom_vig;
  double dopDC_vig;
  double **obs;
  double **fit;
  float *fit_s;
  double **oneovervar;
  struct pos_t pos;
  struct deldopview_t *view;
  char name[40];
  int idellim[2];
  int idoplim[2];
  double dellim[2];
  double doplim[2];
  double idelvig[2];
  double idopvig[2];
  double chi2;
  double overflow_o2, overflow_m2, overflow_xsec, overflow_delmean, overflow_dopmean;
  float overflow_o232, overflow_m232, overflow_xsec32, overflow_delmean32,
      overflow_dopmean32;
  double dof;
  double badradar_logfactor;
  double **map_fit;
  double **map_pos;
  double **map_facet_power;
  struct param_t angleoff[3];
  struct param_t omegaoff[3];
  struct delcor_t delcor;
  double Ftx;
  double del_per_pixel;
  double dop_per_pixel;
  double dopDC;
  double dopfftlen;
  double dopcom;
  struct param_t dopscale;
  double dopscale_save;
  int nviews;
  double view_interval;
  unsigned char smearing_mode;
  int v0;
  int codemethod;
  int spb;
  int stride;
};

struct deldop_t {
  int iradlaw;
  struct ephem_t astephem;
  double Ftx;
  struct delcor_t delcor;
  int ndel;
  int ndop;
  double del_per_pixel;
  double dop_per_pixel;
  double dopcom;
  double dopDC;
  int codemethod;
  int spb;
  int stride;
  int dopfftlen;
  struct param_t dopscale;
  double dopscale_save;
  int nviews;
  double view_interval;
  unsigned char smearing_mode;
  int v0;
  int nframes;
  struct deldopfrm_t *frame;
  char dir[80];
  double dof;
  double sum_deldop_zmax_weights;
  double sum_rad_xsec_weights;
  double sum_cos_subradarlat_weights;
  double fit_overflow[MAXOVERFLOW][MAXOVERFLOW]; /* 30.5 MiB, embedded */
};

struct dopview_t {
  double t;
  double ae[3][3];
  double oe[3][3];
  double intspin[3];
  double orbspin[3];
  double spin[3];
  int idoplim[2];
  double doplim[2];
  double dopoff;
  double km2Hz;
};

struct dopfrm_t {
  double t0;
  int n_integrate;
  double t_integrate[MAXIMP + 2];
  double impulse[MAXIMP + 2][3];
  double nlooks;
  double weight;
  int pixels_weighted;
  double sdev;
  struct param_t cal;
  int ndop;
  double dopcom_vig;
  double *obs;
  double *fit;
  double *oneovervar;
  float *fit_s;
  struct pos_t pos;
  struct dopview_t *view;
  char name[40];
  int idoplim[2];
  double doplim[2];
  double idopvig[2];
  double chi2;
  double overflow_o2, overflow_m2, overflow_xsec, overflow_dopmean;
  double dof;
  double badradar_logfactor;
  double *map_fit;
  double **map_pos;
  double **map_facet_power;
};

struct doppler_t {
  int iradlaw;
  struct ephem_t astephem;
  double Ftx;
  struct delcor_t delcor;
  int ndop;
  double dop_per_bin;
  double dopcom;
  struct param_t dopscale;
  double dopscale_save;
  int nviews;
  double view_interval;
  unsigned char smearing_mode;
  int v0;
  int nframes;
  struct dopfrm_t *frame;
  char dir[80];
  double dof;
  double sum_rad_xsec_weights;
  double sum_cos_subradarlat_weights;
  double fit_overflow[MAXOVERFLOW];
};

struct posetview_t {
  double t;
  double ae[3][3];
  double oe[3][3];
  double se[3][3];
  double intspin[3];
  double orbspin[3];
  double spin[3];
  double solar_phase;
  double solar_azimuth;
};

struct posetfrm_t {
  double t0;
  int n_integrate;
  double t_integrate[MAXIMP + 2];
  double impulse[MAXIMP + 2][3];
  double weight;
  int pixels_weighted;
  double northangle;
  struct param_t cal;
  int nrow;
  int ncol;
  double rowcom_vig;
  double colcom_vig;
  struct pos_t obs;
  struct pos_t fit;
  struct pos_t pos;
  double **oneovervar;
  struct posetview_t *view;
  char name[40];
  struct param_t off[2];
  double chi2;
  double dof;
};

struct poset_t {
  int ioptlaw;
  struct ephem_t astephem;
  struct ephem_t solephem;
  double angle_per_pixel;
  int nviews;
  double view_interval;
  unsigned char smearing_mode;
  int v0;
  int nframes;
  struct posetfrm_t *frame;
  char dir[80];
  double dof;
};

struct crvrend_t {
  struct pos_t pos;
  double ae[3][3];
  double oe[3][3];
  double se[3][3];
  double intspin[3];
  double orbspin[3];
  double spin[3];
  int n_integrate;
  double t_integrate[MAXIMP + 2];
  double impulse[MAXIMP + 2][3];
};

struct lghtcrv_t {
  int ioptlaw;
  struct ephem_t astephem;
  struct ephem_t solephem;
  struct ephem_t astephem_sun;
  struct ephem_t obsephem_sun;
  int nviews;
  double view_interval;
  unsigned char smearing_mode;
  int v0;
  int n;
  double weight;
  double *t0;
  double **t;
  double *obs;
  double *oneovervar;
  double *fit;
  double *rotphase_calc;
  double *rotphase_obs;
  struct param_t cal;
  char name[80];
  int ncalc;
  double *x0;
  double *x;
  double *y;
  double *y2;
  double *solar_phase;
  double *solar_azimuth;
  struct crvrend_t *rend;
  double dof;
  double sum_opt_brightness_weights;
  int ncalc_obsfile;
  double jdstart;
  double jdstop;
  double jdinterval;
};

#define DELAY 0
#define DOPPLER 1
#define POS 2
#define LGHTCRV 3
struct set_t {
  unsigned char type;
  union setdesc_t {
    struct deldop_t deldop;
    struct doppler_t doppler;
    struct poset_t poset;
    struct lghtcrv_t lghtcrv;
  } desc;
  double chi2;
  double dof;
  int inputnode;
  struct param_t angleoff[3];
  struct param_t omegaoff[3];
};

struct dat_t {
  struct pos_t pos;
  int nsets;
  struct set_t *set;
  double chi2;
  double chi2_variance;
  double dof;
  double dof_deldop;
  double dof_doppler;
  double dof_poset;
  double dof_lghtcrv;
  double sum_deldop_zmax_weights;
  double sum_rad_xsec_weights;
  double sum_opt_brightness_weights;
  double sum_cos_subradarlat_weights;
  char name[40];
};

#ifdef __cplusplus
}
#endif
#endif /* SHAPE_ABI_H */
