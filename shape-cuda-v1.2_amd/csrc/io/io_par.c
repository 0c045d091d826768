/*
 * io_par.c -- the parameter file (read_par.c:372-2475): keyword / value pairs
 * up to the PENALTIES keyword, then the penalty list.
 *
 * Where the reference tests each keyword in a chain of string comparisons,
 * this reader looks keywords up in one table that gives, per keyword, the
 * value's kind (yes/no, option word, state, integer, number, bounded number,
 * string or a vector of these on the rest of the line), the par_t member it
 * sets, a unit conversion and any follow-up rule.  The keywords, option words
 * and conversions are the par-file format (read_par.c:404-1352); the
 * defaults are initialize_par's (read_par.c:1365-1605) and the checks and
 * derived defaults adjust_par's (read_par.c:1915-2301).
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "io_internal.h"

enum { PK_YESNO, PK_YESNOVEC, PK_STATE, PK_OPTION, PK_STRING, PK_INT, PK_INTVEC, PK_DOUBLE,
       PK_DOUBLEVEC, PK_MINMAX, PK_SPECIAL };
enum { CV_NONE, CV_D2R, CV_GRAV };
/* follow-up rules for vectors (read_par.c:497-511, 552-557, 611-614, 656-694, 1156-1177) */
enum { PP_NONE, PP_FILL, PP_EXACT, PP_TRIPLE, PP_POSFACTOR, PP_LIMITS, PP_MASKDIR };

typedef struct {
  const char *const *word;
  const unsigned char *val;
  int n;
} optset;

typedef struct {
  const char *name;
  unsigned char kind;
  int arg;              /* vector length, state '=' allowed, or 0 min / 1 max */
  unsigned char conv;
  size_t off;
  const optset *opt;
  unsigned char post;
} parkw;

/* ---- the keyword table ---- */
static const char *const o_action_w[] = {"fit", "write", "format", "mirror", "facets", "sample", "covar", "wavefront", "refshape", "moments", "delcorinit", "convexhull", "map", "view", "area", "photofacets", "photoharm", "slice", "orbit", "split"};
static const unsigned char o_action_v[] = {FIT, WRITE, FORMAT, MIRROR, FACETS, SAMPLE, COVAR, WAVEFRONT, REFSHAPE, MOMENTS, DELCORINIT, CONVEXHULL, MAP, VIEW, AREA, PHOTOFACETS, PHOTOHARM, SLICE, ORBIT, SPLIT};
static const optset o_action = {o_action_w, o_action_v, 20};
static const char *const o_scope_w[] = {"global", "local"};
static const unsigned char o_scope_v[] = {GLOBAL, LOCAL};
static const optset o_scope = {o_scope_w, o_scope_v, 2};
static const char *const o_scaling_w[] = {"none", "block", "bilinear", "bicubic", "cubicconv"};
static const unsigned char o_scaling_v[] = {NONE, BLOCK, BILINEAR, BICUBIC, CUBICCONV};
static const optset o_scaling = {o_scaling_w, o_scaling_v, 5};
static const char *const o_intmethod_w[] = {"brute", "periodic"};
static const unsigned char o_intmethod_v[] = {BRUTE, PERIODIC};
static const optset o_intmethod = {o_intmethod_w, o_intmethod_v, 2};
static const char *const o_ephformat_w[] = {"horizons", "datataking"};
static const unsigned char o_ephformat_v[] = {HORIZONS, DATATAKING};
static const optset o_ephformat = {o_ephformat_w, o_ephformat_v, 2};
static const char *const o_scalefitobs_w[] = {"separate", "maxfitobs", "fit", "obs"};
static const unsigned char o_scalefitobs_v[] = {SCALE_SEPARATELY, SCALE_MAXFITOBS, SCALE_FIT, SCALE_OBS};
static const optset o_scalefitobs = {o_scalefitobs_w, o_scalefitobs_v, 4};
static const char *const o_resample_w[] = {"block", "bilinear", "bicubic", "cubicconv"};
static const unsigned char o_resample_v[] = {BLOCK, BILINEAR, BICUBIC, CUBICCONV};
static const optset o_resample = {o_resample_w, o_resample_v, 4};
static const char *const o_endian_w[] = {"big", "little"};
static const unsigned char o_endian_v[] = {BIG_ENDIAN_DATA, LITTLE_ENDIAN_DATA};
static const optset o_endian = {o_endian_w, o_endian_v, 2};
static const char *const o_viewlaw_w[] = {"optical", "radar", "lambertian"};
static const unsigned char o_viewlaw_v[] = {OPTICALVIEW, RADARVIEW, LAMBERTVIEW};
static const optset o_viewlaw = {o_viewlaw_w, o_viewlaw_v, 3};
static const char *const o_skyrad_w[] = {"radar", "lambertian"};
static const unsigned char o_skyrad_v[] = {RADARVIEW, LAMBERTVIEW};
static const optset o_skyrad = {o_skyrad_w, o_skyrad_v, 2};
static const char *const o_skyopt_w[] = {"optical", "lambertian"};
static const unsigned char o_skyopt_v[] = {OPTICALVIEW, LAMBERTVIEW};
static const optset o_skyopt = {o_skyopt_w, o_skyopt_v, 2};
static const char *const o_vary_w[] = {"none", "no", "size", "all", "yes"};
static const unsigned char o_vary_v[] = {VARY_NONE, VARY_NONE, VARY_SIZE, VARY_ALL, VARY_ALL};
static const optset o_vary = {o_vary_w, o_vary_v, 5};
static const char *const o_mapmode_w[] = {"deldop", "pos", "facets"};
static const unsigned char o_mapmode_v[] = {MAPMODE_DELDOP, MAPMODE_POS, MAPMODE_FACETS};
static const optset o_mapmode = {o_mapmode_w, o_mapmode_v, 3};
static const char *const o_varydop_w[] = {"none", "no", "spin", "all", "yes"};
static const unsigned char o_varydop_v[] = {VARY_NONE, VARY_NONE, VARY_SPIN, VARY_ALL, VARY_ALL};
static const optset o_varydop = {o_varydop_w, o_varydop_v, 5};

static const parkw kwtab[] = {
  {"action", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, action), &o_action, PP_NONE},
  {"speckle", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, speckle), NULL, PP_NONE},
  {"pos_smooth", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, pos_smooth), NULL, PP_NONE},
  {"read_node", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, read_node), NULL, PP_NONE},
  {"pos_scope", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, pos_scope), &o_scope, PP_NONE},
  {"dd_scaling", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, dd_scaling), &o_scaling, PP_NONE},
  {"int_method", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, int_method), &o_intmethod, PP_NONE},
  {"lcrv_pos", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, lcrv_pos), NULL, PP_NONE},
  {"theta_steps", PK_INT, 0, CV_NONE, offsetof(struct par_t, theta_steps), NULL, PP_NONE},
  {"ver_samps", PK_INT, 0, CV_NONE, offsetof(struct par_t, ver_samps), NULL, PP_NONE},
  {"pos_pixels", PK_INT, 0, CV_NONE, offsetof(struct par_t, pos_pixels), NULL, PP_NONE},
  {"pos_width", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, pos_width), NULL, PP_NONE},
  {"dd_gamma", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, dd_gamma), NULL, PP_NONE},
  {"mask_tol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, mask_tol), NULL, PP_NONE},
  {"length_step", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, length_step), NULL, PP_NONE},
  {"length_tol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, length_tol), NULL, PP_NONE},
  {"length_abstol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, length_abstol), NULL, PP_NONE},
  {"angle_step", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, angle_step), NULL, PP_NONE},
  {"angle_tol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, angle_tol), NULL, PP_NONE},
  {"angle_abstol", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, angle_abstol), NULL, PP_NONE},
  {"spin_step", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, spin_step), NULL, PP_NONE},
  {"spin_tol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, spin_tol), NULL, PP_NONE},
  {"spin_abstol", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, spin_abstol), NULL, PP_NONE},
  {"spindot_step", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, spindot_step), NULL, PP_NONE},
  {"spindot_tol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, spindot_tol), NULL, PP_NONE},
  {"spindot_abstol", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, spindot_abstol), NULL, PP_NONE},
  {"photo_step", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, photo_step), NULL, PP_NONE},
  {"photo_tol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, photo_tol), NULL, PP_NONE},
  {"photo_abstol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, photo_abstol), NULL, PP_NONE},
  {"inertia_step", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, inertia_step), NULL, PP_NONE},
  {"inertia_tol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, inertia_tol), NULL, PP_NONE},
  {"inertia_abstol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, inertia_abstol), NULL, PP_NONE},
  {"delcor_step", PK_DOUBLEVEC, MAXDELCORPAR, CV_NONE, offsetof(struct par_t, delcor_step), NULL, PP_FILL},
  {"delcor_tol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, delcor_tol), NULL, PP_NONE},
  {"delcor_abstol", PK_DOUBLEVEC, MAXDELCORPAR, CV_NONE, offsetof(struct par_t, delcor_abstol), NULL, PP_FILL},
  {"xyoff_step", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, xyoff_step), NULL, PP_NONE},
  {"xyoff_tol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, xyoff_tol), NULL, PP_NONE},
  {"xyoff_abstol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, xyoff_abstol), NULL, PP_NONE},
  {"term_prec", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, term_prec), NULL, PP_NONE},
  {"dd_resid", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, dd_resid), NULL, PP_NONE},
  {"optposmax", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, optposmax), NULL, PP_NONE},
  {"radposmax", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, radposmax), NULL, PP_NONE},
  {"pa_ephfile", PK_STRING, 0, CV_NONE, offsetof(struct par_t, pa_ephfile), NULL, PP_NONE},
  {"pa_ephformat", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, pa_ephformat), &o_ephformat, PP_NONE},
  {"pa_startepoch", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, pa_startepoch), NULL, PP_NONE},
  {"pa_stopepoch", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, pa_stopepoch), NULL, PP_NONE},
  {"pa_angleoff", PK_DOUBLEVEC, 3, CV_NONE, offsetof(struct par_t, pa_angleoff), NULL, PP_EXACT},
  {"pa_omegaoff", PK_DOUBLEVEC, 3, CV_NONE, offsetof(struct par_t, pa_omegaoff), NULL, PP_EXACT},
  {"nsinc2", PK_INT, 0, CV_NONE, offsetof(struct par_t, nsinc2), NULL, PP_NONE},
  {"sinc2width", PK_INT, 0, CV_NONE, offsetof(struct par_t, sinc2width), NULL, PP_NONE},
  {"scalefitobs", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, scalefitobs), &o_scalefitobs, PP_NONE},
  {"listres", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, listres), NULL, PP_NONE},
  {"dd_maxsides", PK_SPECIAL, 0, CV_NONE, 0, NULL, PP_NONE},
  {"delcorthresh", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, delcorthresh), NULL, PP_NONE},
  {"jdoffset", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, jdoffset), NULL, PP_NONE},
  {"perform_ltc", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, perform_ltc), NULL, PP_NONE},
  {"convex_file", PK_STRING, 0, CV_NONE, offsetof(struct par_t, convex_file), NULL, PP_NONE},
  {"plot_spinvec", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, plot_spinvec), NULL, PP_NONE},
  {"plot_subradar", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, plot_subradar), NULL, PP_NONE},
  {"plot_com", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, plot_com), NULL, PP_NONE},
  {"plot_pa", PK_YESNOVEC, 3, CV_NONE, offsetof(struct par_t, plot_pa), NULL, PP_EXACT},
  {"write_obs", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, write_obs), NULL, PP_NONE},
  {"delcor_read", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, delcor_read), NULL, PP_NONE},
  {"delcor_file", PK_STRING, 0, CV_NONE, offsetof(struct par_t, delcor_file), NULL, PP_NONE},
  {"first_fitpar", PK_INT, 0, CV_NONE, offsetof(struct par_t, first_fitpar), NULL, PP_NONE},
  {"view_long", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, view_long), NULL, PP_NONE},
  {"view_lat", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, view_lat), NULL, PP_NONE},
  {"listfit", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, listfit), NULL, PP_NONE},
  {"listfit_path", PK_STRING, 0, CV_NONE, offsetof(struct par_t, listfit_path), NULL, PP_NONE},
  {"area_latincr", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, area_latincr), NULL, PP_NONE},
  {"area_longincr", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, area_longincr), NULL, PP_NONE},
  {"poset_scaling", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, poset_scaling), &o_scaling, PP_NONE},
  {"poset_resample", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, poset_resample), &o_resample, PP_NONE},
  {"image_rebin", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, image_rebin), NULL, PP_NONE},
  {"poset_resid", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, poset_resid), NULL, PP_NONE},
  {"list_posetcal", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, list_posetcal), NULL, PP_NONE},
  {"delta_rotphase", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, delta_rotphase), NULL, PP_NONE},
  {"endian", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, endian), &o_endian, PP_NONE},
  {"mark_unseen", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, mark_unseen), NULL, PP_NONE},
  {"maxangle_seen", PK_SPECIAL, 0, CV_NONE, 0, NULL, PP_NONE},
  {"lcrv_writeall", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, lcrv_writeall), NULL, PP_NONE},
  {"exclude_seen", PK_INT, 0, CV_NONE, offsetof(struct par_t, exclude_seen), NULL, PP_NONE},
  {"pa_scatlaw", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, pa_scatlaw), &o_viewlaw, PP_NONE},
  {"view_scatlaw", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, view_scatlaw), &o_viewlaw, PP_NONE},
  {"sky_radlaw", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, sky_radlaw), &o_skyrad, PP_NONE},
  {"sky_optlaw", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, sky_optlaw), &o_skyopt, PP_NONE},
  {"radfacets", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, radfacets), NULL, PP_NONE},
  {"rad_global_R_state", PK_STATE, 0, CV_NONE, offsetof(struct par_t, rad_global_R_state), NULL, PP_NONE},
  {"rad_global_C_state", PK_STATE, 0, CV_NONE, offsetof(struct par_t, rad_global_C_state), NULL, PP_NONE},
  {"rad_local_R_state", PK_STATE, 1, CV_NONE, offsetof(struct par_t, rad_local_R_state), NULL, PP_NONE},
  {"rad_local_C_state", PK_STATE, 1, CV_NONE, offsetof(struct par_t, rad_local_C_state), NULL, PP_NONE},
  {"optfacets", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, optfacets), NULL, PP_NONE},
  {"opt_global_R_state", PK_STATE, 0, CV_NONE, offsetof(struct par_t, opt_global_R_state), NULL, PP_NONE},
  {"opt_global_wt_state", PK_STATE, 0, CV_NONE, offsetof(struct par_t, opt_global_wt_state), NULL, PP_NONE},
  {"opt_global_A0_state", PK_STATE, 0, CV_NONE, offsetof(struct par_t, opt_global_A0_state), NULL, PP_NONE},
  {"opt_global_D_state", PK_STATE, 0, CV_NONE, offsetof(struct par_t, opt_global_D_state), NULL, PP_NONE},
  {"opt_global_k_state", PK_STATE, 0, CV_NONE, offsetof(struct par_t, opt_global_k_state), NULL, PP_NONE},
  {"opt_global_w_state", PK_STATE, 0, CV_NONE, offsetof(struct par_t, opt_global_w_state), NULL, PP_NONE},
  {"opt_global_h_state", PK_STATE, 0, CV_NONE, offsetof(struct par_t, opt_global_h_state), NULL, PP_NONE},
  {"opt_global_B0_state", PK_STATE, 0, CV_NONE, offsetof(struct par_t, opt_global_B0_state), NULL, PP_NONE},
  {"opt_global_g_state", PK_STATE, 0, CV_NONE, offsetof(struct par_t, opt_global_g_state), NULL, PP_NONE},
  {"opt_global_theta_state", PK_STATE, 0, CV_NONE, offsetof(struct par_t, opt_global_theta_state), NULL, PP_NONE},
  {"opt_local_R_state", PK_STATE, 1, CV_NONE, offsetof(struct par_t, opt_local_R_state), NULL, PP_NONE},
  {"opt_local_wt_state", PK_STATE, 1, CV_NONE, offsetof(struct par_t, opt_local_wt_state), NULL, PP_NONE},
  {"opt_local_A0_state", PK_STATE, 1, CV_NONE, offsetof(struct par_t, opt_local_A0_state), NULL, PP_NONE},
  {"opt_local_D_state", PK_STATE, 1, CV_NONE, offsetof(struct par_t, opt_local_D_state), NULL, PP_NONE},
  {"opt_local_k_state", PK_STATE, 1, CV_NONE, offsetof(struct par_t, opt_local_k_state), NULL, PP_NONE},
  {"opt_local_w_state", PK_STATE, 1, CV_NONE, offsetof(struct par_t, opt_local_w_state), NULL, PP_NONE},
  {"opt_local_h_state", PK_STATE, 1, CV_NONE, offsetof(struct par_t, opt_local_h_state), NULL, PP_NONE},
  {"opt_local_B0_state", PK_STATE, 1, CV_NONE, offsetof(struct par_t, opt_local_B0_state), NULL, PP_NONE},
  {"opt_local_g_state", PK_STATE, 1, CV_NONE, offsetof(struct par_t, opt_local_g_state), NULL, PP_NONE},
  {"opt_local_theta_state", PK_STATE, 1, CV_NONE, offsetof(struct par_t, opt_local_theta_state), NULL, PP_NONE},
  {"radharm", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, radharm), NULL, PP_NONE},
  {"rad_R_nharm", PK_INT, 0, CV_NONE, offsetof(struct par_t, rad_R_nharm), NULL, PP_NONE},
  {"rad_C_nharm", PK_INT, 0, CV_NONE, offsetof(struct par_t, rad_C_nharm), NULL, PP_NONE},
  {"optharm", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, optharm), NULL, PP_NONE},
  {"opt_R_nharm", PK_INT, 0, CV_NONE, offsetof(struct par_t, opt_R_nharm), NULL, PP_NONE},
  {"opt_wt_nharm", PK_INT, 0, CV_NONE, offsetof(struct par_t, opt_wt_nharm), NULL, PP_NONE},
  {"opt_A0_nharm", PK_INT, 0, CV_NONE, offsetof(struct par_t, opt_A0_nharm), NULL, PP_NONE},
  {"opt_D_nharm", PK_INT, 0, CV_NONE, offsetof(struct par_t, opt_D_nharm), NULL, PP_NONE},
  {"opt_k_nharm", PK_INT, 0, CV_NONE, offsetof(struct par_t, opt_k_nharm), NULL, PP_NONE},
  {"opt_w_nharm", PK_INT, 0, CV_NONE, offsetof(struct par_t, opt_w_nharm), NULL, PP_NONE},
  {"opt_h_nharm", PK_INT, 0, CV_NONE, offsetof(struct par_t, opt_h_nharm), NULL, PP_NONE},
  {"opt_B0_nharm", PK_INT, 0, CV_NONE, offsetof(struct par_t, opt_B0_nharm), NULL, PP_NONE},
  {"opt_g_nharm", PK_INT, 0, CV_NONE, offsetof(struct par_t, opt_g_nharm), NULL, PP_NONE},
  {"opt_theta_nharm", PK_INT, 0, CV_NONE, offsetof(struct par_t, opt_theta_nharm), NULL, PP_NONE},
  {"view_shadows", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, view_shadows), NULL, PP_NONE},
  {"view_sunlong", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, view_sunlong), NULL, PP_NONE},
  {"view_sunlat", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, view_sunlat), NULL, PP_NONE},
  {"sun_appmag", PK_DOUBLEVEC, MAXSUNMAGS, CV_NONE, offsetof(struct par_t, sun_appmag), NULL, PP_FILL},
  {"rad_R_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, rad_R_min), NULL, PP_NONE},
  {"rad_R_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, rad_R_max), NULL, PP_NONE},
  {"rad_C_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, rad_C_min), NULL, PP_NONE},
  {"rad_C_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, rad_C_max), NULL, PP_NONE},
  {"opt_R_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, opt_R_min), NULL, PP_NONE},
  {"opt_R_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, opt_R_max), NULL, PP_NONE},
  {"opt_wt_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, opt_wt_min), NULL, PP_NONE},
  {"opt_wt_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, opt_wt_max), NULL, PP_NONE},
  {"opt_A0_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, opt_A0_min), NULL, PP_NONE},
  {"opt_A0_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, opt_A0_max), NULL, PP_NONE},
  {"opt_D_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, opt_D_min), NULL, PP_NONE},
  {"opt_D_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, opt_D_max), NULL, PP_NONE},
  {"opt_k_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, opt_k_min), NULL, PP_NONE},
  {"opt_k_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, opt_k_max), NULL, PP_NONE},
  {"opt_w_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, opt_w_min), NULL, PP_NONE},
  {"opt_w_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, opt_w_max), NULL, PP_NONE},
  {"opt_h_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, opt_h_min), NULL, PP_NONE},
  {"opt_h_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, opt_h_max), NULL, PP_NONE},
  {"opt_B0_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, opt_B0_min), NULL, PP_NONE},
  {"opt_B0_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, opt_B0_max), NULL, PP_NONE},
  {"opt_g_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, opt_g_min), NULL, PP_NONE},
  {"opt_g_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, opt_g_max), NULL, PP_NONE},
  {"opt_theta_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, opt_theta_min), NULL, PP_NONE},
  {"opt_theta_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, opt_theta_max), NULL, PP_NONE},
  {"ratio_step", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, ratio_step), NULL, PP_NONE},
  {"ratio_tol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, ratio_tol), NULL, PP_NONE},
  {"ratio_abstol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, ratio_abstol), NULL, PP_NONE},
  {"vary_delcor0", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, vary_delcor0), &o_vary, PP_NONE},
  {"vary_radalb", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, vary_radalb), &o_vary, PP_NONE},
  {"vary_optalb", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, vary_optalb), &o_vary, PP_NONE},
  {"slice_long", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, slice_long), NULL, PP_NONE},
  {"slice_lat", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, slice_lat), NULL, PP_NONE},
  {"slice_offset", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, slice_offset), NULL, PP_NONE},
  {"slice_viewlong", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, slice_viewlong), NULL, PP_NONE},
  {"slice_viewlat", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, slice_viewlat), NULL, PP_NONE},
  {"slice_sunlong", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, slice_sunlong), NULL, PP_NONE},
  {"slice_sunlat", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, slice_sunlat), NULL, PP_NONE},
  {"slice_scatlaw", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, slice_scatlaw), &o_viewlaw, PP_NONE},
  {"slice_shadows", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, slice_shadows), NULL, PP_NONE},
  {"slice_planefrac", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, slice_planefrac), NULL, PP_NONE},
  {"slice_skyfactor", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, slice_skyfactor), NULL, PP_NONE},
  {"slice_dimfactor", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, slice_dimfactor), NULL, PP_NONE},
  {"slice_read", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, slice_read), NULL, PP_NONE},
  {"slice_dointerp", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, slice_dointerp), NULL, PP_NONE},
  {"semimajor_axis", PK_DOUBLEVEC, 2, CV_NONE, offsetof(struct par_t, semimajor_axis), NULL, PP_TRIPLE},
  {"r_pericenter", PK_DOUBLEVEC, 2, CV_NONE, offsetof(struct par_t, r_pericenter), NULL, PP_TRIPLE},
  {"eccentricity", PK_DOUBLEVEC, 2, CV_NONE, offsetof(struct par_t, eccentricity), NULL, PP_TRIPLE},
  {"t_pericenter", PK_DOUBLEVEC, 2, CV_NONE, offsetof(struct par_t, t_pericenter), NULL, PP_TRIPLE},
  {"long_asc_node", PK_DOUBLEVEC, 2, CV_D2R, offsetof(struct par_t, long_asc_node), NULL, PP_TRIPLE},
  {"inclination", PK_DOUBLEVEC, 2, CV_D2R, offsetof(struct par_t, inclination), NULL, PP_TRIPLE},
  {"arg_pericenter", PK_DOUBLEVEC, 2, CV_D2R, offsetof(struct par_t, arg_pericenter), NULL, PP_TRIPLE},
  {"binary_totmass", PK_DOUBLEVEC, 2, CV_GRAV, offsetof(struct par_t, binary_gravparam), NULL, PP_TRIPLE},
  {"orbit_period", PK_DOUBLEVEC, 2, CV_NONE, offsetof(struct par_t, orbit_period), NULL, PP_TRIPLE},
  {"orbit_reflex", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, orbit_reflex), NULL, PP_NONE},
  {"orbit_posfactor", PK_DOUBLEVEC, 3, CV_NONE, offsetof(struct par_t, orbit_posfactor), NULL, PP_POSFACTOR},
  {"write_posfactor", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, write_posfactor), NULL, PP_NONE},
  {"radfitmax", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, radfitmax), NULL, PP_NONE},
  {"radobsmax", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, radobsmax), NULL, PP_NONE},
  {"term_badmodel", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, term_badmodel), NULL, PP_NONE},
  {"avoid_badpos", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, avoid_badpos), NULL, PP_NONE},
  {"bad_objfactor", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, bad_objfactor), NULL, PP_NONE},
  {"maskdir", PK_STRING, 0, CV_NONE, offsetof(struct par_t, maskdir), NULL, PP_MASKDIR},
  {"write_chi2fit0", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, write_chi2fit0), NULL, PP_NONE},
  {"warn_badradar", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, warn_badradar), NULL, PP_NONE},
  {"plot_angmom", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, plot_angmom), NULL, PP_NONE},
  {"pierce_spinvec", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, pierce_spinvec), NULL, PP_NONE},
  {"pierce_angmom", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, pierce_angmom), NULL, PP_NONE},
  {"slice_posfactor", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, slice_posfactor), NULL, PP_NONE},
  {"view_posfactor", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, view_posfactor), NULL, PP_NONE},
  {"npar_update", PK_INT, 0, CV_NONE, offsetof(struct par_t, npar_update), NULL, PP_NONE},
  {"rad_rho_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, rad_rho_min), NULL, PP_NONE},
  {"rad_rho_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, rad_rho_max), NULL, PP_NONE},
  {"split_plane", PK_SPECIAL, 0, CV_NONE, 0, NULL, PP_NONE},
  {"map_set", PK_INT, 0, CV_NONE, offsetof(struct par_t, map_set), NULL, PP_NONE},
  {"map_frame", PK_INT, 0, CV_NONE, offsetof(struct par_t, map_frame), NULL, PP_NONE},
  {"map_mode", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, map_mode), &o_mapmode, PP_NONE},
  {"map_dellim", PK_INTVEC, 2, CV_NONE, offsetof(struct par_t, map_dellim), NULL, PP_LIMITS},
  {"map_doplim", PK_INTVEC, 2, CV_NONE, offsetof(struct par_t, map_doplim), NULL, PP_LIMITS},
  {"map_xposlim", PK_INTVEC, 2, CV_NONE, offsetof(struct par_t, map_xposlim), NULL, PP_LIMITS},
  {"map_yposlim", PK_INTVEC, 2, CV_NONE, offsetof(struct par_t, map_yposlim), NULL, PP_LIMITS},
  {"map_posfactor", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, map_posfactor), NULL, PP_NONE},
  {"map_fitcutoff", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, map_fitcutoff), NULL, PP_NONE},
  {"map_poscutoff", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, map_poscutoff), NULL, PP_NONE},
  {"map_comp", PK_INTVEC, MAXCHOSENFACETS, CV_NONE, offsetof(struct par_t, map_comp), NULL, PP_NONE},
  {"map_facet", PK_INTVEC, MAXCHOSENFACETS, CV_NONE, offsetof(struct par_t, map_facet), NULL, PP_NONE},
  {"map_verbose", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, map_verbose), NULL, PP_NONE},
  {"radfitmin", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, radfitmin), NULL, PP_NONE},
  {"radobsmin", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, radobsmin), NULL, PP_NONE},
  {"int_abstol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, int_abstol), NULL, PP_NONE},
  {"pa_bodycoords", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, pa_bodycoords), NULL, PP_NONE},
  {"pa_highlight", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, pa_highlight), NULL, PP_NONE},
  {"pa_comp", PK_INTVEC, MAXCHOSENFACETS, CV_NONE, offsetof(struct par_t, pa_comp), NULL, PP_NONE},
  {"pa_facet", PK_INTVEC, MAXCHOSENFACETS, CV_NONE, offsetof(struct par_t, pa_facet), NULL, PP_NONE},
  {"view_highlight", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, view_highlight), NULL, PP_NONE},
  {"view_comp", PK_INTVEC, MAXCHOSENFACETS, CV_NONE, offsetof(struct par_t, view_comp), NULL, PP_NONE},
  {"view_facet", PK_INTVEC, MAXCHOSENFACETS, CV_NONE, offsetof(struct par_t, view_facet), NULL, PP_NONE},
  {"vary_dopscale", PK_OPTION, 0, CV_NONE, offsetof(struct par_t, vary_dopscale), &o_varydop, PP_NONE},
  {"dopscale_min", PK_MINMAX, 0, CV_NONE, offsetof(struct par_t, dopscale_min), NULL, PP_NONE},
  {"dopscale_max", PK_MINMAX, 1, CV_NONE, offsetof(struct par_t, dopscale_max), NULL, PP_NONE},
  {"objfunc_start", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, objfunc_start), NULL, PP_NONE},
  {"chi2fit0_thresh", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, chi2fit0_thresh), NULL, PP_NONE},
  {"listpos_deldop", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, listpos_deldop), NULL, PP_NONE},
  {"listpos_opt", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, listpos_opt), NULL, PP_NONE},
  {"listpos_path", PK_STRING, 0, CV_NONE, offsetof(struct par_t, listpos_path), NULL, PP_NONE},
  {"term_maxiter", PK_INT, 0, CV_NONE, offsetof(struct par_t, term_maxiter), NULL, PP_NONE},
  {"write_highlight", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, write_highlight), NULL, PP_NONE},
  {"write_comp", PK_INTVEC, MAXCHOSENFACETS, CV_NONE, offsetof(struct par_t, write_comp), NULL, PP_NONE},
  {"write_facet", PK_INTVEC, MAXCHOSENFACETS, CV_NONE, offsetof(struct par_t, write_facet), NULL, PP_NONE},
  {"delcor_verbose", PK_YESNO, 0, CV_NONE, offsetof(struct par_t, delcor_verbose), NULL, PP_NONE},
  {"lib_amp_step", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, lib_amp_step), NULL, PP_NONE},
  {"lib_amp_tol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, lib_amp_tol), NULL, PP_NONE},
  {"lib_amp_abstol", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, lib_amp_abstol), NULL, PP_NONE},
  {"lib_freq_step", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, lib_freq_step), NULL, PP_NONE},
  {"lib_freq_tol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, lib_freq_tol), NULL, PP_NONE},
  {"lib_freq_abstol", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, lib_freq_abstol), NULL, PP_NONE},
  {"lib_phase_step", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, lib_phase_step), NULL, PP_NONE},
  {"lib_phase_tol", PK_DOUBLE, 0, CV_NONE, offsetof(struct par_t, lib_phase_tol), NULL, PP_NONE},
  {"lib_phase_abstol", PK_DOUBLE, 0, CV_D2R, offsetof(struct par_t, lib_phase_abstol), NULL, PP_NONE},
};

#define NKW ((int)(sizeof kwtab / sizeof kwtab[0]))

/* initialize_par (read_par.c:1365-1605): every default, and the sentinels
   (-9.99, -1, -HUGENUMBER) adjust_par tests to see what the file set */
static void par_defaults(struct par_t *p)
{
  memset(p, 0, sizeof *p);
  p->pos_smooth = 1;
  p->pos_scope = GLOBAL;
  p->dd_scaling = NONE;
  p->int_method = PERIODIC;
  p->dd_gamma = 1.0;
  p->dd_resid = 5.0;
  p->pa_ephformat = HORIZONS;
  p->pa_startepoch = -1.0;
  p->pa_stopepoch = HUGENUMBER;
  p->scalefitobs = SCALE_SEPARATELY;
  p->perform_ltc = 1;
  p->view_lat = PIE / 2;
  p->view_sunlat = PIE / 2;
  p->area_latincr = 5.0 * D2R;
  p->area_longincr = 5.0 * D2R;
  p->poset_scaling = NONE;
  p->poset_resample = BILINEAR;
  p->poset_resid = 5.0;
  p->endian = BIG_ENDIAN_DATA;
  p->exclude_seen = -1;
  p->pa_scatlaw = p->view_scatlaw = p->sky_radlaw = LAMBERTVIEW;
  p->sky_optlaw = OPTICALVIEW;
  p->radfacets = p->optfacets = 1;
  p->rad_global_R_state = p->rad_global_C_state = 'c';
  p->rad_local_R_state = p->rad_local_C_state = 'f';
  p->opt_global_R_state = p->opt_global_wt_state = p->opt_global_A0_state = 'c';
  p->opt_global_D_state = p->opt_global_k_state = p->opt_global_w_state = 'c';
  p->opt_global_h_state = p->opt_global_B0_state = p->opt_global_g_state = 'c';
  p->opt_global_theta_state = 'c';
  p->opt_local_R_state = p->opt_local_wt_state = p->opt_local_A0_state = 'f';
  p->opt_local_D_state = p->opt_local_k_state = p->opt_local_w_state = 'f';
  p->opt_local_h_state = p->opt_local_B0_state = p->opt_local_g_state = 'f';
  p->opt_local_theta_state = 'f';
  p->radharm = p->optharm = 1;
  for (int i = 0; i < MAXSUNMAGS; i++) p->sun_appmag[i] = SOLAR_VMAG;
  p->rad_R_max = p->rad_C_max = p->rad_rho_max = HUGENUMBER;
  p->opt_R_max = p->opt_A0_max = p->opt_D_max = p->opt_h_max = p->opt_B0_max = HUGENUMBER;
  p->opt_wt_max = 1.0;
  p->opt_k_min = -HUGENUMBER;
  p->opt_w_max = 1.0;
  p->opt_g_min = -1.0;
  p->opt_g_max = 1.0;
  p->opt_theta_max = PIE / 2;
  p->vary_delcor0 = p->vary_radalb = p->vary_optalb = VARY_NONE;
  p->slice_lat = PIE / 2;
  p->slice_sunlat = PIE / 2;
  p->slice_scatlaw = LAMBERTVIEW;
  p->slice_planefrac = 0.85;
  p->slice_skyfactor = 0.2;
  p->slice_dimfactor = 0.2;
  p->slice_dointerp = 1;
  for (int i = 0; i < 3; i++) p->orbit_posfactor[i] = 1.0;
  p->write_posfactor = 1.0;
  p->bad_objfactor = 2.0;
  p->pierce_spinvec = p->pierce_angmom = 1;
  p->slice_posfactor = p->view_posfactor = 1.0;
  p->npar_update = 20;
  p->map_mode = MAPMODE_DELDOP;
  p->map_posfactor = 1.0;
  p->map_fitcutoff = p->map_poscutoff = -9.99;
  p->int_abstol = 1.0e-7;
  p->vary_dopscale = VARY_NONE;
  p->dopscale_max = HUGENUMBER;
  /* sentinels */
  p->ver_samps = -1;
  p->pos_pixels = -1;
  p->pos_width = -9.99;
  p->mask_tol = -9.99;
  double *sent[] = {&p->length_step, &p->length_tol, &p->length_abstol, &p->ratio_step, &p->ratio_tol,
                    &p->ratio_abstol, &p->angle_step, &p->angle_tol, &p->angle_abstol, &p->spin_step,
                    &p->spin_tol, &p->spin_abstol, &p->spindot_step, &p->spindot_tol, &p->spindot_abstol,
                    &p->lib_amp_step, &p->lib_amp_tol, &p->lib_amp_abstol, &p->lib_freq_step,
                    &p->lib_freq_tol, &p->lib_freq_abstol, &p->lib_phase_step, &p->lib_phase_tol,
                    &p->lib_phase_abstol, &p->photo_step, &p->photo_tol, &p->photo_abstol,
                    &p->inertia_step, &p->inertia_tol, &p->inertia_abstol, &p->delcor_tol,
                    &p->xyoff_step, &p->xyoff_tol, &p->xyoff_abstol, &p->term_prec, &p->delcorthresh,
                    &p->objfunc_start};
  for (size_t i = 0; i < sizeof sent / sizeof sent[0]; i++) *sent[i] = -9.99;
  for (int i = 0; i < MAXDELCORPAR; i++) p->delcor_step[i] = p->delcor_abstol[i] = -9.99;
  p->nsinc2 = -1;
  p->sinc2width = -1;
  p->dd_clockwiserot = -1;
  p->slice_viewlong = p->slice_viewlat = -HUGENUMBER;
  for (int i = 0; i < 2; i++) {
    p->semimajor_axis[i] = p->r_pericenter[i] = p->eccentricity[i] = p->t_pericenter[i] = -9.99;
    p->long_asc_node[i] = p->inclination[i] = p->arg_pericenter[i] = -HUGENUMBER;
    p->binary_gravparam[i] = p->orbit_period[i] = -9.99;
  }
  p->is_triple = -1;
  p->split_const = -HUGENUMBER;
  p->map_set = p->map_frame = -1;
  p->map_dellim[0] = p->map_doplim[0] = p->map_xposlim[0] = p->map_yposlim[0] = 999;
  p->map_dellim[1] = p->map_doplim[1] = p->map_xposlim[1] = p->map_yposlim[1] = -999;
  for (int i = 0; i < MAXCHOSENFACETS; i++)
    p->map_facet[i] = p->pa_facet[i] = p->view_facet[i] = p->write_facet[i] = -1;
}

/* the option words of dd_maxsides: image corner of max delay and Doppler
   -> clockwise rotation, x flip, y flip (read_par.c:518-553) */
static const struct { const char *w; int rot, xf, yf; } maxsides[] = {
    {"right-top", 0, 0, 0},    {"right-bottom", 0, 0, 1}, {"left-top", 0, 1, 0},
    {"left-bottom", 0, 1, 1},  {"top-right", 1, 0, 1},    {"top-left", 1, 1, 1},
    {"bottom-right", 1, 0, 0}, {"bottom-left", 1, 1, 0}};

static void check_triple(io_ctx *cx, struct par_t *p, int n_read, int n_max)
{
  if (p->is_triple == -1) p->is_triple = (n_read == n_max);
  else if ((n_read == n_max) != (p->is_triple == 1))
    io_fail(cx, "read_par: 'orbit' action parameters are mixed on binary vs. triple system");
}

/* a vector on the rest of the line: up to n numbers; returns how many */
static int read_vec(io_ctx *cx, io_file *f, const char *name, int kind, void *dst, int n)
{
  char line[4096], *save = NULL;
  io_restofline(f, line, sizeof line);
  int k = 0;
  for (char *w = strtok_r(line, " \t\r\n", &save); w; w = strtok_r(NULL, " \t\r\n", &save)) {
    if (k == n) io_fail(cx, "read_par: %s should have no more than %d components", name, n);
    if (kind == PK_INTVEC) ((int *)dst)[k] = io_str_to_int(cx, w);
    else if (kind == PK_DOUBLEVEC) ((double *)dst)[k] = io_str_to_double(cx, w);
    else if (!strcmp(w, "yes")) ((unsigned char *)dst)[k] = 1;
    else if (!strcmp(w, "no")) ((unsigned char *)dst)[k] = 0;
    else io_fail(cx, "read_par: unrecognized value for %s component", name);
    k++;
  }
  return k;
}

static void read_keyword(io_ctx *cx, io_file *f, struct par_t *p, const parkw *k)
{
  char w[IO_MAXLEN + 1];
  char *base = (char *)p + k->off;
  switch (k->kind) {
  case PK_YESNO:
    io_word(f, w, sizeof w);
    if (!strcmp(w, "yes")) *(unsigned char *)base = 1;
    else if (!strcmp(w, "no")) *(unsigned char *)base = 0;
    else io_fail(cx, "read_par: unrecognized value for %s", k->name);
    return;
  case PK_STATE:
    io_word(f, w, sizeof w);
    if (!strcmp(w, "c") || !strcmp(w, "f") || (k->arg && !strcmp(w, "="))) *base = w[0];
    else io_fail(cx, "read_par: unrecognized value for %s", k->name);
    return;
  case PK_OPTION:
    io_word(f, w, sizeof w);
    for (int i = 0; i < k->opt->n; i++)
      if (!strcmp(w, k->opt->word[i])) { *(unsigned char *)base = k->opt->val[i]; return; }
    io_fail(cx, "read_par: unrecognized value for %s", k->name);
  case PK_STRING:
    io_word(f, w, sizeof w);
    if (strlen(w) >= 80) io_fail(cx, "read_par: value of %s longer than 79 characters", k->name);
    strcpy(base, w);
    if (k->post == PP_MASKDIR && (!strcmp(base, "\"\"") || !strcmp(base, "''"))) base[0] = '\0';
    return;
  case PK_INT:
    *(int *)base = io_int(cx, f);
    return;
  case PK_DOUBLE:
    *(double *)base = io_double(cx, f);
    if (k->conv == CV_D2R) *(double *)base *= D2R;
    return;
  case PK_MINMAX: {
    /* a bound only tightens the default (read_par.c:1807-1838) */
    const double v = io_double(cx, f);
    double *dst = (double *)base;
    if ((k->arg == 0 && v >= *dst) || (k->arg == 1 && v <= *dst)) *dst = v;
    return;
  }
  case PK_YESNOVEC:
  case PK_INTVEC:
  case PK_DOUBLEVEC: {
    const int n = read_vec(cx, f, k->name, k->kind, base, k->arg);
    if (k->kind == PK_DOUBLEVEC && k->conv != CV_NONE)
      for (int i = 0; i < n; i++) ((double *)base)[i] *= (k->conv == CV_D2R ? D2R : GRAVCONST);
    switch (k->post) {
    case PP_FILL:
      for (int i = n; i < k->arg; i++) ((double *)base)[i] = ((double *)base)[0];
      break;
    case PP_EXACT:
      if (n != k->arg) io_fail(cx, "read_par: %s must have %d components", k->name, k->arg);
      break;
    case PP_TRIPLE:
      check_triple(cx, p, n, 2);
      break;
    case PP_POSFACTOR:
      if (n == 1) io_fail(cx, "read_par: %s must have 2 or 3 components", k->name);
      check_triple(cx, p, n, 3);
      break;
    case PP_LIMITS:
      if (n < 1 || n > 2) io_fail(cx, "read_par: %s must have 1 or 2 components", k->name);
      if (n == 1) ((int *)base)[1] = ((int *)base)[0];
      break;
    }
    return;
  }
  }
}

static void read_special(io_ctx *cx, io_file *f, struct par_t *p, const char *name)
{
  char w[IO_MAXLEN + 1];
  if (!strcmp(name, "dd_maxsides")) {
    io_word(f, w, sizeof w);
    for (size_t i = 0; i < sizeof maxsides / sizeof maxsides[0]; i++)
      if (!strcmp(w, maxsides[i].w)) {
        p->dd_clockwiserot = maxsides[i].rot;
        p->dd_xflip = maxsides[i].xf;
        p->dd_yflip = maxsides[i].yf;
        return;
      }
    io_fail(cx, "read_par: unrecognized value for dd_maxsides");
  } else if (!strcmp(name, "maxangle_seen")) {
    const double a = io_double(cx, f);
    if (a < 0.0 || a > 90.0) io_fail(cx, "read_par: must have 0.0 <= maxangle_seen <= 90.0 deg");
    p->mincosine_seen = cos(a * D2R);
  } else { /* split_plane: a plane normalized by its normal's length */
    double sp[4];
    if (read_vec(cx, f, name, PK_DOUBLEVEC, sp, 4) != 4)
      io_fail(cx, "read_par: split_plane must have 4 components");
    const double nc = sqrt(sp[0] * sp[0] + sp[1] * sp[1] + sp[2] * sp[2]);
    for (int i = 0; i < 3; i++) p->split_norm[i] = sp[i] / nc;
    p->split_const = sp[3] / nc;
  }
}

#define TOLPAIR(t, a) \
  do { if (p->t < 0 && p->a < 0) io_fail(cx, "read_par: must specify " #t " or " #a " or both"); \
       p->t = fmax(p->t, 0.0); p->a = fmax(p->a, 0.0); } while (0)
#define TOLDEFAULT(t, a, dt, da) \
  do { if (p->t < 0 && p->a < 0) { p->t = (dt); p->a = (da); } \
       else { p->t = fmax(p->t, 0.0); p->a = fmax(p->a, 0.0); } } while (0)

/* adjust_par (read_par.c:1915-2301): required parameters per action and the
   defaults that depend on what the file gave */
static void par_adjust(io_ctx *cx, struct par_t *p)
{
  const int fitlike = p->action == FIT || p->action == COVAR;
  if (fitlike) {
    if (p->length_step < 0) io_fail(cx, "read_par: must specify length_step");
    if (p->ratio_step < 0) io_fail(cx, "read_par: must specify ratio_step");
    if (p->angle_step < 0) io_fail(cx, "read_par: must specify angle_step");
    if (p->spin_step < 0) io_fail(cx, "read_par: must specify spin_step");
    if (p->photo_step < 0) p->photo_step = p->length_step;
    if (p->inertia_step < 0) p->inertia_step = p->spin_step / D2R;
    if (p->delcor_step[0] < 0)
      for (int i = 0; i < MAXDELCORPAR; i++) p->delcor_step[i] = p->spin_step / D2R;
    if (p->xyoff_step < 0) p->xyoff_step = 1.0;
    if (p->spindot_step < 0) {
      p->spindot_step = p->spin_step / 10000.0;
      /* the libration steps get defaults only on this branch, as in the
         reference (read_par.c:1985-2004) */
      if (p->lib_amp_step < 0) p->lib_amp_step = 0.01 / D2R;
      if (p->lib_freq_step < 0) p->lib_freq_step = p->spin_step / D2R;
      if (p->lib_phase_step < 0) p->lib_phase_step = 1 / D2R;
    }
  }
  if (p->action == FIT) {
    TOLPAIR(length_tol, length_abstol);
    TOLPAIR(ratio_tol, ratio_abstol);
    TOLPAIR(angle_tol, angle_abstol);
    TOLPAIR(spin_tol, spin_abstol);
    TOLDEFAULT(photo_tol, photo_abstol, p->length_tol, p->length_abstol);
    TOLDEFAULT(inertia_tol, inertia_abstol, p->spin_tol, p->spin_abstol);
    if (p->delcor_tol < 0 && p->delcor_abstol[0] < 0) {
      p->delcor_tol = p->spin_tol;
      for (int i = 0; i < MAXDELCORPAR; i++) p->delcor_abstol[i] = p->spin_abstol;
    } else {
      p->delcor_tol = fmax(p->delcor_tol, 0.0);
      for (int i = 0; i < MAXDELCORPAR; i++) p->delcor_abstol[i] = fmax(p->delcor_abstol[i], 0.0);
    }
    TOLDEFAULT(xyoff_tol, xyoff_abstol, 0.0, 0.01);
    TOLDEFAULT(spindot_tol, spindot_abstol, p->spin_tol, (p->spin_abstol / D2R) / 10000.0);
    TOLDEFAULT(lib_amp_tol, lib_amp_abstol, 0.1, 0.1 / D2R);
    TOLDEFAULT(lib_freq_tol, lib_freq_abstol, 0.1, 0.1);
    TOLDEFAULT(lib_phase_tol, lib_phase_abstol, 0.1, 0.1);
  }
  const int a = p->action;
  if (a == FIT || a == WRITE || a == MOMENTS || a == FORMAT || a == COVAR || a == VIEW || a == AREA ||
      a == DELCORINIT || a == SLICE || a == ORBIT || a == MAP) {
    if (p->pos_pixels < 0) io_fail(cx, "read_par: must specify pos_pixels");
    if (p->pos_width < 0.0) io_fail(cx, "read_par: must specify pos_width");
  }
  if (a == SAMPLE && p->ver_samps < 0) io_fail(cx, "read_par: must specify ver_samps for sample action");
  if (a == FIT && p->term_prec < 0) io_fail(cx, "read_par: must specify term_prec for fit action");
  if (a == MOMENTS) {
    if (!p->pa_ephfile[0]) io_fail(cx, "read_par: must specify pa_ephfile for moments action");
    if (p->pa_highlight && p->pa_facet[0] < 0) io_fail(cx, "read_par: must specify pa_facet if pa_highlight = yes");
  }
  if (a == CONVEXHULL && !p->convex_file[0]) io_fail(cx, "read_par: must specify convex_file for convexhull action");
  if (a == DELCORINIT && p->delcor_read && !p->delcor_file[0])
    io_fail(cx, "read_par: must specify delcor_file for delcorinit action if delcor_read = yes");
  if (a == WRITE && p->mark_unseen && p->pos_scope == GLOBAL)
    io_fail(cx, "read_par: must set pos_scope = local for write action if mark_unseen = yes");
  if (a == SPLIT && p->split_const == -HUGENUMBER) io_fail(cx, "read_par: must specify split_plane for split action");
  if (a == MAP) {
    if (p->map_mode != MAPMODE_FACETS && (p->map_set < 0 || p->map_frame < 0))
      io_fail(cx, "read_par: must specify map_set and map_frame if map_mode = pos or deldop");
    else if (p->map_mode == MAPMODE_POS &&
             (p->map_xposlim[0] > p->map_xposlim[1] || p->map_yposlim[0] > p->map_yposlim[1]))
      io_fail(cx, "read_par: must specify map_xposlim and map_yposlim if map_mode = pos");
    else if (p->map_mode == MAPMODE_DELDOP &&
             (p->map_dellim[0] > p->map_dellim[1] || p->map_doplim[0] > p->map_doplim[1]))
      io_fail(cx, "read_par: must specify map_dellim and map_doplim if map_mode = deldop");
    else if (p->map_mode == MAPMODE_FACETS && p->map_facet[0] < 0)
      io_fail(cx, "read_par: must specify map_facet if map_mode = facets");
  }
  if (a == VIEW && p->view_highlight && p->view_facet[0] < 0)
    io_fail(cx, "read_par: must specify view_facet if view_highlight = yes");
  if (p->mask_tol < 0)
    p->mask_tol = (p->pos_pixels > 0 && p->pos_width > 0.0) ? p->pos_width / (p->pos_pixels - 1) : 0.0;
  if (p->nsinc2 < 0) p->nsinc2 = 1;
  if (p->sinc2width < 0) p->sinc2width = 6;
  if (p->dd_clockwiserot < 0) p->dd_clockwiserot = p->dd_xflip = p->dd_yflip = 0;
  if (a == DELCORINIT && p->delcorthresh < 0) p->delcorthresh = 2.0;
  if (a == SLICE) {
    if (p->slice_viewlong == -HUGENUMBER) p->slice_viewlong = p->slice_long;
    if (p->slice_viewlat == -HUGENUMBER) p->slice_viewlat = p->slice_lat;
  }
  if (a == ORBIT) {
    const int nsat = (p->is_triple == 1) ? 2 : 1;
    for (int i = 0; i < nsat; i++) {
      if (p->eccentricity[0] < 0.0) p->eccentricity[i] = 0.0;
      if (p->t_pericenter[0] < 0.0) p->t_pericenter[i] = 2451545.0;
      if (p->long_asc_node[0] == -HUGENUMBER) p->long_asc_node[i] = 0.0;
      if (p->inclination[0] == -HUGENUMBER) p->inclination[i] = 0.0;
      if (p->arg_pericenter[0] == -HUGENUMBER) p->arg_pericenter[i] = 0.0;
    }
    const double twopi = 2 * PIE;
    for (int i = 0; i < nsat; i++) {
      p->inclination[i] -= twopi * floor(p->inclination[i] / twopi);
      if (p->inclination[i] > PIE) {
        p->inclination[i] = twopi - p->inclination[i];
        p->long_asc_node[i] += PIE;
        p->arg_pericenter[i] += PIE;
      }
      p->long_asc_node[i] -= twopi * floor(p->long_asc_node[i] / twopi);
      p->arg_pericenter[i] -= twopi * floor(p->arg_pericenter[i] / twopi);
    }
    if (p->r_pericenter[0] > 0.0 && p->semimajor_axis[0] > 0.0) {
      for (int i = 0; i < nsat; i++) p->semimajor_axis[i] = -HUGENUMBER;
    } else if (p->semimajor_axis[0] > 0.0) {
      for (int i = 0; i < nsat; i++) p->r_pericenter[i] = p->semimajor_axis[i] * (1 - p->eccentricity[i]);
    } else if (p->r_pericenter[0] < 0.0) {
      io_fail(cx, "read_par: must specify r_pericenter or semimajor_axis for orbit action");
    }
    if (p->binary_gravparam[0] > 0.0 && p->orbit_period[0] > 0.0) {
      for (int i = 0; i < nsat; i++) p->orbit_period[i] = -HUGENUMBER;
    } else if (p->orbit_period[0] > 0.0) {
      if (p->eccentricity[0] >= 1.0) io_fail(cx, "read_par: must have eccentricity < 1.0 if orbit_period is specified");
      for (int i = 0; i < nsat; i++)
        p->binary_gravparam[i] = pow(2 * PIE / p->orbit_period[i], 2.0) *
                                 pow(p->r_pericenter[i] / (1 - p->eccentricity[i]), 3.0);
    } else if (p->binary_gravparam[0] < 0.0) {
      io_fail(cx, "read_par: must specify binary_totmass or orbit_period for orbit action");
    }
  }
}

/* penalty names and their type codes (read_par.c:2336-2416) */
static const struct { const char *name; int type; } pentab[] = {
    {"optalbdel", OPTALBDEL},   {"optalbvar", OPTALBVAR},     {"thetadel", THETADEL},
    {"thetavar", THETAVAR},     {"radalbdel", RADALBDEL},     {"radalbvar", RADALBVAR},
    {"rad_c_del", RAD_C_DEL},   {"rad_c_var", RAD_C_VAR},     {"nonsmooth", NONSMOOTH},
    {"concavity", CONCAVITY},   {"rdev", RDEV},               {"comdev", COMDEV},
    {"volume", VOLUME},         {"inertiadev", INERTIADEV},   {"inertiadev_uni", INERTIADEV_UNI},
    {"pa3tilt", PA3TILT},       {"nonpa", NONPA},             {"nonpa_uni", NONPA_UNI},
    {"euleroffs", EULEROFFS},   {"flattening", FLATTENING},   {"maxrdev", MAXRDEV},
    {"maxellipdev", MAXELLIPDEV}, {"noncosine", NONCOSINE},   {"bifurcation", BIFURCATION},
    {"impulse", IMPULSE}};
#define NPEN ((int)(sizeof pentab / sizeof pentab[0]))

/* read_penalties (read_par.c:2303-2422): a count, then name / weight pairs;
   the arrays are 1-based, as penalties.c reads them */
static void read_penalties(io_ctx *cx, io_file *f, struct par_t *p)
{
  char w[IO_MAXLEN + 1];
  int seen[NPEN];
  memset(seen, 0, sizeof seen);
  p->pen.n = io_int(cx, f);
  if (p->pen.n < 0) io_fail(cx, "read_par: negative number of penalties");
  if (p->pen.n > 0) {
    p->pen.type = io_ivector(cx, 1, p->pen.n);
    p->pen.weight = io_vector(cx, 1, p->pen.n);
    p->pen.base = io_vector(cx, 1, p->pen.n);
  }
  for (int i = 1; i <= p->pen.n; i++) {
    io_word(f, w, sizeof w);
    p->pen.weight[i] = io_double(cx, f);
    int j = 0;
    while (j < NPEN && strcmp(w, pentab[j].name)) j++;
    if (j == NPEN) io_fail(cx, "read_par: don't recognize penalty '%s'", w);
    if (seen[j]) io_fail(cx, "read_par: penalty '%s' is specified more than once", w);
    seen[j] = 1;
    p->pen.type[i] = pentab[j].type;
  }
}

static int par_body(io_ctx *cx, const char *name, struct par_t *par)
{
  io_file f;
  char w[IO_MAXLEN + 1];
  unsigned char given[NKW];
  memset(given, 0, sizeof given);
  par_defaults(par);
  io_open(cx, &f, name);
  if (!io_word(&f, w, sizeof w) || strcmp(w, "PARAMETERS"))
    io_fail(cx, "read_par: %s: expected keyword PARAMETERS", name);
  for (;;) {
    if (!io_word(&f, w, sizeof w)) break;
    if (!strcmp(w, "PENALTIES")) break;
    int k = 0;
    while (k < NKW && strcmp(w, kwtab[k].name)) k++;
    if (k == NKW) io_fail(cx, "read_par: %s: don't recognize parameter '%s'", name, w);
    if (given[k]) io_fail(cx, "read_par: '%s' is specified more than once in the parameter file", w);
    given[k] = 1;
    if (kwtab[k].kind == PK_SPECIAL) read_special(cx, &f, par, kwtab[k].name);
    else read_keyword(cx, &f, par, &kwtab[k]);
  }
  par_adjust(cx, par);
  if (strcmp(w, "PENALTIES")) io_fail(cx, "read_par: %s: expected keyword PENALTIES", name);
  read_penalties(cx, &f, par);
  io_close(cx, &f);
  return 0;
}

int shape_read_par(const char *name, struct par_t *par)
{
  io_ctx cx;
  memset(&cx, 0, sizeof cx);
  if (setjmp(cx.jb)) return io_unwind(&cx);
  shape_io_free_par(par);
  par_body(&cx, name, par);
  io_arena_commit(&cx, par);
  return 0;
}

void read_par(char *name, struct par_t *par)
{
  if (shape_read_par(name, par) < 0) {
    printf("%s\n", shape_io_error());
    fprintf(stderr, "ERROR: %s\n", shape_io_error());
    exit(2);
  }
}
