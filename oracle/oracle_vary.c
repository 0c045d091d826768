/*
 * oracle_vary.c -- the means vary_params feeds to bestfit's joint
 * adjustments (delay constant, Doppler scales, radar and optical albedos).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * oracle_vary_params restates vary_params.c:87-325 step by step: the first
 * dataset only (the reference's loop runs s < 1, vary_params.c:108), each
 * radar frame's view v0 rendered afresh (posclr, posvis) and synthesized
 * with view index 0 (pos2deldop / pos2doppler), the lightcurve's points
 * rendered from the observer and the sun with the shadow mask and the
 * photometry, the spline to the observation epochs at view v0; sums
 * weighted by frame (or lightcurve) weight and divided by the weight sums
 * read_dat leaves in dat (zero where a sum is zero).  Fits and POS arrays
 * are written as the reference writes them.
 */
#include "oracle_internal.h"

static void o_copy33(double a[3][3], double b[3][3])
{
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = b[i][j];
}

/* cos(sub-radar latitude) of a frame view: the target-to-Earth vector in body
   coordinates is the third row of oe ae^T (vary_params.c:177-184) */
static double o_cos_delta(double ae[3][3], double oe[3][3])
{
  double oa[3][3];
  o_mtrnsps(oa, ae);
  o_mmmul(oa, oe, oa);
  return sqrt(oa[2][0] * oa[2][0] + oa[2][1] * oa[2][1]);
}

void oracle_vary_params(struct par_t *par, struct mod_t *mod, struct dat_t *dat, double *deldop_zmax,
                        double *rad_xsec, double *opt_brightness, double *cos_subradarlat)
{
  double orbit_offset[3] = {0.0, 0.0, 0.0};
  double s_zmax = 0.0, s_xsec = 0.0, s_bright = 0.0, s_cos = 0.0;
  for (int s = 0; s < 1 && s < dat->nsets; s++) {
    struct set_t *set = &dat->set[s];
    if (set->type == DELAY) {
      struct deldop_t *deldop = &set->desc.deldop;
      for (int f = 0; f < deldop->nframes; f++) {
        struct deldopfrm_t *frame = &deldop->frame[f];
        const int do_zmax = par->vary_delcor0 != 0 && deldop->delcor.a[0].state != 'c';
        const int do_xsec = par->vary_radalb != 0 && frame->cal.state == 'c';
        if (do_zmax || do_xsec) {
          const double weight = frame->weight;
          struct pos_t *pos = &frame->pos;
          o_copy33(pos->ae, frame->view[deldop->v0].ae);
          o_copy33(pos->oe, frame->view[deldop->v0].oe);
          pos->bistatic = 0;
          oracle_posclr(pos);
          for (int c = 0; c < mod->shape.ncomp; c++)
            oracle_posvis(&mod->shape.comp[c].real, orbit_offset, pos, (int)par->pos_smooth, 0, 0, c);
          for (int i = 1; i <= frame->ndel; i++)
            for (int j = 1; j <= frame->ndop; j++) frame->fit[i][j] = 0.0;
          oracle_pos2deldop(par, &mod->photo, 0.0, 0.0, 0.0, deldop, 0, s, f, 0);
          if (do_zmax) {
            double zmax = -HUGENUMBER;
            for (int x = pos->xlim[0]; x <= pos->xlim[1]; x++)
              for (int y = pos->ylim[0]; y <= pos->ylim[1]; y++)
                if (pos->cose[x][y] > 0.0) zmax = O_MAX(zmax, pos->z[x][y]);
            s_zmax += zmax * weight;
          }
          if (do_xsec) {
            double xsec = frame->overflow_xsec;
            for (int i = 1; i <= frame->ndel; i++)
              for (int j = 1; j <= frame->ndop; j++) xsec += frame->fit[i][j];
            xsec *= frame->cal.val;
            s_xsec += xsec * weight;
          }
        }
        if (par->vary_dopscale != 0 && deldop->dopscale.state != 'c')
          s_cos += o_cos_delta(frame->view[deldop->v0].ae, frame->view[deldop->v0].oe) * frame->weight;
      }
    } else if (set->type == DOPPLER) {
      struct doppler_t *doppler = &set->desc.doppler;
      for (int f = 0; f < doppler->nframes; f++) {
        struct dopfrm_t *frame = &doppler->frame[f];
        if (par->vary_radalb != 0 && frame->cal.state == 'c') {
          const double weight = frame->weight;
          struct pos_t *pos = &frame->pos;
          o_copy33(pos->ae, frame->view[doppler->v0].ae);
          o_copy33(pos->oe, frame->view[doppler->v0].oe);
          pos->bistatic = 0;
          oracle_posclr(pos);
          for (int c = 0; c < mod->shape.ncomp; c++)
            oracle_posvis(&mod->shape.comp[c].real, orbit_offset, pos, (int)par->pos_smooth, 0, 0, c);
          for (int j = 1; j <= frame->ndop; j++) frame->fit[j] = 0.0;
          oracle_pos2doppler(par, &mod->photo, 0.0, 0.0, 0.0, doppler, 0, s, f, 0);
          double xsec = frame->overflow_xsec;
          for (int j = 1; j <= frame->ndop; j++) xsec += frame->fit[j];
          xsec *= frame->cal.val;
          s_xsec += xsec * weight;
        }
        if (par->vary_dopscale != 0 && doppler->dopscale.state != 'c')
          s_cos += o_cos_delta(frame->view[doppler->v0].ae, frame->view[doppler->v0].oe) * frame->weight;
      }
    } else if (set->type == LGHTCRV) {
      struct lghtcrv_t *lc = &set->desc.lghtcrv;
      if (par->vary_optalb != 0 && lc->cal.state == 'c') {
        const int n = lc->n, ncalc = lc->ncalc;
        const double weight = lc->weight;
        for (int i = 1; i <= ncalc; i++) {
          struct pos_t *pos = &lc->rend[i].pos;
          o_copy33(pos->ae, lc->rend[i].ae);
          o_copy33(pos->oe, lc->rend[i].oe);
          o_copy33(pos->se, lc->rend[i].se);
          pos->bistatic = 1;
          oracle_posclr(pos);
          for (int c = 0; c < mod->shape.ncomp; c++)
            oracle_posvis(&mod->shape.comp[c].real, orbit_offset, pos, (int)par->pos_smooth, 0, 0, c);
          for (int c = 0; c < mod->shape.ncomp; c++)
            oracle_posvis(&mod->shape.comp[c].real, orbit_offset, pos, 0, 1, 0, c);
          oracle_posmask(pos, par->mask_tol);
          const double intensityfactor = pow(pos->km_per_pixel / AU, 2.0);
          lc->y[i] = oracle_apply_photo(mod, lc->ioptlaw, lc->solar_phase[i], intensityfactor, pos, 0, s, i);
        }
        oracle_spline(lc->x, lc->y, ncalc, 2.0e30, 2.0e30, lc->y2);
        for (int i = 1; i <= n; i++) {
          oracle_splint(lc->x, lc->y, lc->y2, ncalc, lc->t[i][lc->v0], &lc->fit[i]);
          s_bright += lc->fit[i] * weight;
        }
      }
    } else if (set->type != POS) {
      oracle_set_error("oracle_vary_params: unknown dataset type");
      return;
    }
  }
  *deldop_zmax = dat->sum_deldop_zmax_weights > 0.0 ? s_zmax / dat->sum_deldop_zmax_weights : 0.0;
  *rad_xsec = dat->sum_rad_xsec_weights > 0.0 ? s_xsec / dat->sum_rad_xsec_weights : 0.0;
  *opt_brightness = dat->sum_opt_brightness_weights > 0.0 ? s_bright / dat->sum_opt_brightness_weights : 0.0;
  *cos_subradarlat = dat->sum_cos_subradarlat_weights > 0.0 ? s_cos / dat->sum_cos_subradarlat_weights : 0.0;
}
