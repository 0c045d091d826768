/*
 * oracle_fits.c -- forward-model driver, chi-square and the HMT-style baseline.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

static void copy33(double d[3][3], double s[3][3]) { memcpy(d, s, sizeof(double) * 9); }

/* calc_fits.c:440-602 (fit-time path) */
void oracle_calc_deldop(struct par_t *par, struct mod_t *mod, struct deldop_t *deldop, int s,
                        struct o_flags *fl)
{
  double orbit_offset[3] = {0.0, 0.0, 0.0};
  for (int f = 0; f < deldop->nframes; f++) {
    struct deldopfrm_t *frame = &deldop->frame[f];
    struct pos_t *pos = &frame->pos;
    double **fit_store = NULL;
    double ov_o2 = 0, ov_m2 = 0, ov_xs = 0, ov_dm = 0, ov_pm = 0;
    int ndel = frame->ndel, ndop = frame->ndop;
    if (deldop->nviews > 1) {
      fit_store = (double **)malloc(sizeof(double *) * (ndel + 1));
      for (int i = 1; i <= ndel; i++) fit_store[i] = (double *)calloc(ndop + 1, sizeof(double));
    }
    for (int v2 = deldop->v0 + 1; v2 <= deldop->v0 + deldop->nviews; v2++) {
      int v = v2 % deldop->nviews;
      copy33(pos->ae, frame->view[v].ae);
      copy33(pos->oe, frame->view[v].oe);
      pos->bistatic = 0;
      oracle_posclr(pos);
      for (int c = 0; c < mod->shape.ncomp; c++)
        if (oracle_posvis(&mod->shape.comp[c].real, orbit_offset, pos, (int)par->pos_smooth, 0, 0, c) &&
            v == deldop->v0) {
          fl->posbnd = 1;
          fl->posbnd_logfactor += frame->dof * pos->posbnd_logfactor;
        }
      if (s != par->exclude_seen && v == deldop->v0) {
        for (int k = pos->xlim[0]; k <= pos->xlim[1]; k++)
          for (int l = pos->ylim[0]; l <= pos->ylim[1]; l++)
            if (pos->cose[k][l] > par->mincosine_seen && pos->f[k][l] >= 0)
              mod->shape.comp[pos->comp[k][l]].real.f[pos->f[k][l]].seen = 1;
      }
      for (int i = 1; i <= ndel; i++)
        for (int j = 1; j <= ndop; j++) frame->fit[i][j] = 0.0;
      if (oracle_pos2deldop(par, &mod->photo, 0.0, 0.0, 0.0, deldop, 0, s, f, v)) {
        fl->badradar = 1;
        fl->badradar_logfactor += frame->dof * frame->badradar_logfactor / deldop->nviews;
      }
      if (deldop->nviews > 1) {
        for (int i = 1; i <= ndel; i++)
          for (int j = 1; j <= ndop; j++) fit_store[i][j] += frame->fit[i][j];
        ov_o2 += frame->overflow_o2;
        ov_m2 += frame->overflow_m2;
        ov_xs += frame->overflow_xsec;
        ov_dm += frame->overflow_delmean;
        ov_pm += frame->overflow_dopmean;
      }
    }
    if (deldop->nviews > 1) {
      for (int i = 1; i <= ndel; i++) {
        for (int j = 1; j <= ndop; j++) frame->fit[i][j] = fit_store[i][j] / deldop->nviews;
        free(fit_store[i]);
      }
      free(fit_store);
      frame->overflow_o2 = ov_o2 / deldop->nviews;
      frame->overflow_m2 = ov_m2 / deldop->nviews;
      frame->overflow_xsec = ov_xs / deldop->nviews;
      frame->overflow_delmean = ov_dm / deldop->nviews;
      frame->overflow_dopmean = ov_pm / deldop->nviews;
    }
    if (par->dd_gamma != 1.0) { /* gamma_trans.c */
      for (int i = 1; i <= ndel; i++)
        for (int j = 1; j <= ndop; j++)
          if (frame->fit[i][j] > 0.0) frame->fit[i][j] = pow(frame->fit[i][j], 1 / par->dd_gamma);
    }
  }
}

/* calc_fits.c:830-962 (fit-time path) */
void oracle_calc_doppler(struct par_t *par, struct mod_t *mod, struct doppler_t *doppler, int s,
                         struct o_flags *fl)
{
  double orbit_offset[3] = {0.0, 0.0, 0.0};
  for (int f = 0; f < doppler->nframes; f++) {
    struct dopfrm_t *frame = &doppler->frame[f];
    struct pos_t *pos = &frame->pos;
    double *fit_store = NULL;
    double ov_o2 = 0, ov_m2 = 0, ov_xs = 0, ov_pm = 0;
    int ndop = frame->ndop;
    if (doppler->nviews > 1) fit_store = (double *)calloc(ndop + 1, sizeof(double));
    for (int v2 = doppler->v0 + 1; v2 <= doppler->v0 + doppler->nviews; v2++) {
      int v = v2 % doppler->nviews;
      copy33(pos->ae, frame->view[v].ae);
      copy33(pos->oe, frame->view[v].oe);
      pos->bistatic = 0;
      oracle_posclr(pos);
      for (int c = 0; c < mod->shape.ncomp; c++)
        if (oracle_posvis(&mod->shape.comp[c].real, orbit_offset, pos, (int)par->pos_smooth, 0, 0, c) &&
            v == doppler->v0) {
          fl->posbnd = 1;
          fl->posbnd_logfactor += frame->dof * pos->posbnd_logfactor;
        }
      if (s != par->exclude_seen && v == doppler->v0) {
        for (int k = pos->xlim[0]; k <= pos->xlim[1]; k++)
          for (int l = pos->ylim[0]; l <= pos->ylim[1]; l++)
            if (pos->cose[k][l] > par->mincosine_seen && pos->f[k][l] >= 0)
              mod->shape.comp[pos->comp[k][l]].real.f[pos->f[k][l]].seen = 1;
      }
      for (int j = 1; j <= ndop; j++) frame->fit[j] = 0.0;
      if (oracle_pos2doppler(par, &mod->photo, 0.0, 0.0, 0.0, doppler, 0, s, f, v)) {
        fl->badradar = 1;
        fl->badradar_logfactor += frame->dof * frame->badradar_logfactor / doppler->nviews;
      }
      if (doppler->nviews > 1) {
        for (int j = 1; j <= ndop; j++) fit_store[j] += frame->fit[j];
        ov_o2 += frame->overflow_o2;
        ov_m2 += frame->overflow_m2;
        ov_xs += frame->overflow_xsec;
        ov_pm += frame->overflow_dopmean;
      }
    }
    if (doppler->nviews > 1) {
      for (int j = 1; j <= ndop; j++) frame->fit[j] = fit_store[j] / doppler->nviews;
      free(fit_store);
      frame->overflow_o2 = ov_o2 / doppler->nviews;
      frame->overflow_m2 = ov_m2 / doppler->nviews;
      frame->overflow_xsec = ov_xs / doppler->nviews;
      frame->overflow_dopmean = ov_pm / doppler->nviews;
    }
  }
}

/* calc_fits.c:1509-1959 (fit-time path) */
void oracle_calc_lghtcrv(struct par_t *par, struct mod_t *mod, struct lghtcrv_t *lghtcrv, int s,
                         struct o_flags *fl)
{
  double orbit_offset[3] = {0.0, 0.0, 0.0};
  double posbnd_logfactor = 0.0;
  int n = lghtcrv->n, ncalc = lghtcrv->ncalc;
  for (int i = 1; i <= ncalc; i++) {
    struct crvrend_t *rend = &lghtcrv->rend[i];
    struct pos_t *pos = &rend->pos;
    copy33(pos->ae, rend->ae);
    copy33(pos->oe, rend->oe);
    copy33(pos->se, rend->se);
    pos->bistatic = 1;
    oracle_posclr(pos);
    for (int c = 0; c < mod->shape.ncomp; c++)
      if (oracle_posvis(&mod->shape.comp[c].real, orbit_offset, pos, (int)par->pos_smooth, 0, 0, c)) {
        fl->posbnd = 1;
        posbnd_logfactor += pos->bistatic ? 0.5 * pos->posbnd_logfactor : pos->posbnd_logfactor;
      }
    if (pos->bistatic) {
      for (int c = 0; c < mod->shape.ncomp; c++)
        if (oracle_posvis(&mod->shape.comp[c].real, orbit_offset, pos, 0, 1, 0, c)) {
          fl->posbnd = 1;
          posbnd_logfactor += 0.5 * pos->posbnd_logfactor;
        }
      oracle_posmask(pos, par->mask_tol);
    }
    if (s != par->exclude_seen) {
      for (int k = pos->xlim[0]; k <= pos->xlim[1]; k++)
        for (int l = pos->ylim[0]; l <= pos->ylim[1]; l++)
          if (pos->cose[k][l] > par->mincosine_seen && pos->cosi[k][l] > par->mincosine_seen &&
              pos->f[k][l] >= 0)
            mod->shape.comp[pos->comp[k][l]].real.f[pos->f[k][l]].seen = 1;
    }
    double intensityfactor = pow(pos->km_per_pixel / AU, 2.0);
    lghtcrv->y[i] = oracle_apply_photo(mod, lghtcrv->ioptlaw, lghtcrv->solar_phase[i], intensityfactor,
                                       pos, 0, s, i);
  }
  oracle_spline(lghtcrv->x, lghtcrv->y, ncalc, 2.0e30, 2.0e30, lghtcrv->y2);
  for (int i = 1; i <= n; i++) {
    lghtcrv->fit[i] = 0.0;
    for (int v = 0; v < lghtcrv->nviews; v++) {
      double interp;
      oracle_splint(lghtcrv->x, lghtcrv->y, lghtcrv->y2, ncalc, lghtcrv->t[i][v], &interp);
      lghtcrv->fit[i] += interp;
    }
    lghtcrv->fit[i] /= lghtcrv->nviews;
  }
  fl->posbnd_logfactor += lghtcrv->dof * (posbnd_logfactor / ncalc);
}

static void reset_flags(struct par_t *par, struct mod_t *mod)
{
  par->posbnd = 0;
  par->badposet = 0;
  par->badradar = 0;
  par->posbnd_logfactor = 0.0;
  par->badposet_logfactor = 0.0;
  par->badradar_logfactor = 0.0;
  for (int c = 0; c < mod->shape.ncomp; c++)
    for (int f = 0; f < mod->shape.comp[c].real.nf; f++) mod->shape.comp[c].real.f[f].seen = 0;
}

static void finish_flags(struct par_t *par, struct dat_t *dat)
{
  par->posbnd_logfactor /= dat->dof;
  par->badposet_logfactor /= dat->dof_poset;
  par->badradar_logfactor /= (dat->dof_deldop + dat->dof_doppler);
}

static void run_set(struct par_t *par, struct mod_t *mod, struct dat_t *dat, int s, struct o_flags *fl)
{
  switch (dat->set[s].type) {
  case DELAY: oracle_calc_deldop(par, mod, &dat->set[s].desc.deldop, s, fl); break;
  case DOPPLER: oracle_calc_doppler(par, mod, &dat->set[s].desc.doppler, s, fl); break;
  case LGHTCRV: oracle_calc_lghtcrv(par, mod, &dat->set[s].desc.lghtcrv, s, fl); break;
  default: oracle_set_error("oracle_calc_fits: plane-of-sky sets are not restated");
  }
}

static void merge_flags(struct par_t *par, const struct o_flags *fl)
{
  if (fl->posbnd) par->posbnd = 1;
  if (fl->badradar) par->badradar = 1;
  par->posbnd_logfactor += fl->posbnd_logfactor;
  par->badradar_logfactor += fl->badradar_logfactor;
}

/* calc_fits.c:340-410 */
void oracle_calc_fits(struct par_t *par, struct mod_t *mod, struct dat_t *dat)
{
  reset_flags(par, mod);
  for (int s = 0; s < dat->nsets; s++) {
    struct o_flags fl = {0, 0, 0.0, 0.0};
    run_set(par, mod, dat, s, &fl);
    merge_flags(par, &fl);
  }
  finish_flags(par, dat);
}

/* ---- HMT-style threaded driver -------------------------------------- */
struct o_thread_arg {
  struct par_t *par;
  struct mod_t *mod;
  struct dat_t *dat;
  int tid, nthreads;
  struct o_flags *flags; /* one per set */
};

static void *o_thread_main(void *p)
{
  struct o_thread_arg *a = (struct o_thread_arg *)p;
  for (int s = a->tid; s < a->dat->nsets; s += a->nthreads) run_set(a->par, a->mod, a->dat, s, &a->flags[s]);
  return NULL;
}

void oracle_calc_fits_threads(struct par_t *par, struct mod_t *mod, struct dat_t *dat, int nthreads)
{
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 1 && par->pos_scope != LOCAL) {
    /* frames share one POS under GLOBAL scope: threads would race on it */
    oracle_set_error("oracle_calc_fits_threads: threads need pos_scope LOCAL");
    return;
  }
  reset_flags(par, mod);
  struct o_flags *flags = (struct o_flags *)calloc(dat->nsets, sizeof(struct o_flags));
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  struct o_thread_arg *args = (struct o_thread_arg *)malloc(sizeof(struct o_thread_arg) * nthreads);
  for (int t = 0; t < nthreads; t++) {
    args[t] = (struct o_thread_arg){par, mod, dat, t, nthreads, flags};
    if (t > 0) pthread_create(&th[t], NULL, o_thread_main, &args[t]);
  }
  o_thread_main(&args[0]);
  for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
  for (int s = 0; s < dat->nsets; s++) merge_flags(par, &flags[s]); /* fixed set order */
  finish_flags(par, dat);
  free(flags);
  free(th);
  free(args);
}

/* ---- chi-square: chi2.c:240-300, :395-466, :722-788, :1125-1170 ---- */
static double o_frame_err(double o2, double m2, double om, struct param_t *cal, double weight)
{
  if (cal->state == 'f') cal->val = (om > 0.0) ? om / m2 : TINYCALFACT;
  double c = cal->val;
  return weight * (o2 - 2 * c * om + c * c * m2);
}

double oracle_chi2(struct par_t *par, struct dat_t *dat, int list_breakdown)
{
  (void)par; (void)list_breakdown;
  dat->chi2 = 0.0;
  for (int s = 0; s < dat->nsets; s++) {
    struct set_t *set = &dat->set[s];
    double chi2_set = 0.0;
    if (set->type == DELAY) {
      struct deldop_t *dd = &set->desc.deldop;
      for (int f = 0; f < dd->nframes; f++) {
        struct deldopfrm_t *fr = &dd->frame[f];
        double o2 = fr->overflow_o2, m2 = fr->overflow_m2, om = 0.0;
        for (int i = 1; i <= fr->ndel; i++)
          for (int j = 1; j <= fr->ndop; j++) {
            o2 += fr->obs[i][j] * fr->obs[i][j] * fr->oneovervar[i][j];
            m2 += fr->fit[i][j] * fr->fit[i][j] * fr->oneovervar[i][j];
            om += fr->fit[i][j] * fr->obs[i][j] * fr->oneovervar[i][j];
          }
        fr->chi2 = o_frame_err(o2, m2, om, &fr->cal, fr->weight);
        chi2_set += fr->chi2;
      }
    } else if (set->type == DOPPLER) {
      struct doppler_t *dp = &set->desc.doppler;
      for (int f = 0; f < dp->nframes; f++) {
        struct dopfrm_t *fr = &dp->frame[f];
        double o2 = fr->overflow_o2, m2 = fr->overflow_m2, om = 0.0;
        for (int j = 1; j <= fr->ndop; j++) {
          o2 += fr->obs[j] * fr->obs[j] * fr->oneovervar[j];
          m2 += fr->fit[j] * fr->fit[j] * fr->oneovervar[j];
          om += fr->fit[j] * fr->obs[j] * fr->oneovervar[j];
        }
        fr->chi2 = o_frame_err(o2, m2, om, &fr->cal, fr->weight);
        chi2_set += fr->chi2;
      }
    } else if (set->type == LGHTCRV) {
      struct lghtcrv_t *lc = &set->desc.lghtcrv;
      double o2 = 0.0, m2 = 0.0, om = 0.0;
      for (int i = 1; i <= lc->n; i++) {
        o2 += lc->obs[i] * lc->obs[i] * lc->oneovervar[i];
        m2 += lc->fit[i] * lc->fit[i] * lc->oneovervar[i];
        om += lc->fit[i] * lc->obs[i] * lc->oneovervar[i];
      }
      chi2_set = o_frame_err(o2, m2, om, &lc->cal, lc->weight);
    } else {
      oracle_set_error("oracle_chi2: plane-of-sky sets are not restated");
    }
    set->chi2 = chi2_set;
    dat->chi2 += chi2_set;
  }
  return dat->chi2;
}

static double o_bad_factors(struct par_t *par, double err)
{
  if (par->baddiam) err *= par->bad_objfactor * exp(par->baddiam_logfactor);
  if (par->badphoto) err *= par->bad_objfactor * exp(par->badphoto_logfactor);
  if (par->posbnd) err *= par->bad_objfactor * exp(par->posbnd_logfactor);
  if (par->badposet) err *= par->bad_objfactor * exp(par->badposet_logfactor);
  if (par->badradar) err *= par->bad_objfactor * exp(par->badradar_logfactor);
  if (par->baddopscale) err *= par->bad_objfactor * exp(par->baddopscale_logfactor);
  return err;
}

/* bestfit.c:789-866 with penalties = 0 */
double oracle_objective_tail(struct par_t *par, struct dat_t *dat)
{
  return o_bad_factors(par, oracle_chi2(par, dat, 0) / dat->dof);
}

/* bestfit.c:789-866: err = chi2/dof (:790-793), err += penalties (:807-811),
   then the bad-model factors (:813-866) */
double oracle_objective(struct par_t *par, struct mod_t *mod, struct dat_t *dat)
{
  double err = oracle_chi2(par, dat, 0) / dat->dof;
  err += oracle_penalties(par, mod, dat);
  return o_bad_factors(par, err);
}
