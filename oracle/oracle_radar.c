/*
 * oracle_radar.c -- radar scattering laws and delay-Doppler / Doppler synthesis.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

/* radlaw.c:44-148 */
double oracle_radlaw(struct photo_t *photo, int ilaw, double cosinc, int c, int f)
{
  union radscat_t *r = &photo->radar[ilaw];
  double d = -9.99;
  switch (photo->radtype[ilaw]) {
  case COSINELAW_DIFF:
    d = r->RC.R.val * (r->RC.C.val + 1) * pow(cosinc, 2 * r->RC.C.val - 1);
    break;
  case TABULARLAW:
    if (f < 0) d = 0.0;
    else {
      double inc = acos(cosinc), angres = (PIE / 2) / (r->tabular.n - 1), ratio = inc / angres;
      int i = (int)floor(ratio);
      double rho1 = r->tabular.rho[i].val, rho2 = r->tabular.rho[i + 1].val;
      d = (rho1 + (ratio - i) * (rho2 - rho1)) / cosinc;
    }
    break;
  case GAUSSIANLAW:
    if (cosinc >= r->quasispec.cos_cutoff) {
      double tan2 = 1 / (cosinc * cosinc) - 1;
      d = r->quasispec.R.val * r->quasispec.C.val * exp(-r->quasispec.C.val * tan2) /
          (cosinc * cosinc * cosinc * cosinc * cosinc);
    } else d = 0.0;
    break;
  case HAGFORSLAW:
    if (cosinc >= r->quasispec.cos_cutoff) {
      double sin2 = 1 - cosinc * cosinc;
      double arg = cosinc * cosinc * cosinc * cosinc + r->quasispec.C.val * sin2;
      d = 0.5 * r->quasispec.R.val * r->quasispec.C.val * pow(arg, -1.5) / cosinc;
    } else d = 0.0;
    break;
  case COSINELAW_QS:
    if (cosinc >= r->quasispec.cos_cutoff)
      d = r->quasispec.R.val * (r->quasispec.C.val + 1) * pow(cosinc, 2 * r->quasispec.C.val - 1);
    else d = 0.0;
    break;
  case GAUSSIAN_COSINE:
    d = r->hybrid.diff.R.val * (r->hybrid.diff.C.val + 1) * pow(cosinc, 2 * r->hybrid.diff.C.val - 1);
    if (cosinc >= r->hybrid.qs.cos_cutoff) {
      double tan2 = 1 / (cosinc * cosinc) - 1;
      d += r->hybrid.qs.R.val * r->hybrid.qs.C.val * exp(-r->hybrid.qs.C.val * tan2) /
           (cosinc * cosinc * cosinc * cosinc * cosinc);
    }
    break;
  case HAGFORS_COSINE:
    d = r->hybrid.diff.R.val * (r->hybrid.diff.C.val + 1) * pow(cosinc, 2 * r->hybrid.diff.C.val - 1);
    if (cosinc >= r->hybrid.qs.cos_cutoff) {
      double sin2 = 1 - cosinc * cosinc;
      double arg = cosinc * cosinc * cosinc * cosinc + r->hybrid.qs.C.val * sin2;
      d += 0.5 * r->hybrid.qs.R.val * r->hybrid.qs.C.val * pow(arg, -1.5) / cosinc;
    }
    break;
  case COSINE_COSINE:
    d = r->hybrid.diff.R.val * (r->hybrid.diff.C.val + 1) * pow(cosinc, 2 * r->hybrid.diff.C.val - 1);
    if (cosinc >= r->hybrid.qs.cos_cutoff)
      d += r->hybrid.qs.R.val * (r->hybrid.qs.C.val + 1) * pow(cosinc, 2 * r->hybrid.qs.C.val - 1);
    break;
  case HARMCOSINE_DIFF:
    if (f < 0) d = 0.0;
    else {
      struct RC_t *l = &r->harmcosine.local[c][f];
      d = l->R.val * (l->C.val + 1) * pow(cosinc, 2 * l->C.val - 1);
    }
    break;
  case INHOCOSINE_DIFF:
    if (f < 0) d = 0.0;
    else {
      struct RC_t *l = &r->inhocosine.local[c][f];
      d = l->R.val * (l->C.val + 1) * pow(cosinc, 2 * l->C.val - 1);
    }
    break;
  default:
    oracle_set_error("oracle_radlaw: unsupported radar scattering law");
  }
  return d;
}

/* ---- shared pieces of the two synthesis routines -------------------- */
struct o_dopgeom {
  double ax, ay, xincr, yincr, dopdiff_bl, dopdiff_max, dop_extra, km2Hz;
};

/* pos2deldop.c:243-288 / pos2doppler.c:179-227: Doppler bins per POS pixel */
static int o_dopgeom(struct par_t *par, double oe[3][3], double spin[3], double dopscale, double Ftx,
                     double dop_per_bin, double kmpp, struct o_dopgeom *g, double *badlog)
{
  double w[3];
  int badradar = 0;
  o_cotrans(w, oe, spin, 1);
  double dopfact = dopscale * KM2HZFACT * kmpp * Ftx / dop_per_bin;
  g->ax = -w[1] * dopfact;
  g->ay = w[0] * dopfact;
  g->km2Hz = sqrt(g->ax * g->ax + g->ay * g->ay) * dop_per_bin / kmpp;
  if (w[0] != 0.0 || w[1] != 0.0)
    g->dop_extra = g->km2Hz * 0.5 * kmpp * sqrt(w[0] * w[0] + w[1] * w[1]) / O_MAX(fabs(w[0]), fabs(w[1]));
  else
    g->dop_extra = 0.0;
  g->xincr = g->ax / par->nsinc2;
  g->yincr = g->ay / par->nsinc2;
  g->dopdiff_bl = -(par->nsinc2 - 1) * (g->xincr + g->yincr) / 2;
  g->dopdiff_max = (par->nsinc2 - 1) * (fabs(g->xincr) + fabs(g->yincr)) / 2;
  if (2 * g->dopdiff_max + par->sinc2width + 1 > MAXBINS) {
    badradar = 1;
    *badlog += log((2 * g->dopdiff_max + par->sinc2width + 1) / MAXBINS);
  }
  return badradar;
}

/* mean sinc^2 over nsinc2^2 points of a pixel (pos2deldop.c:432-457) */
static double o_sinc2_mean(const struct par_t *par, const struct o_dopgeom *g, double dopPOS, int idop)
{
  if (par->nsinc2 == 1) return o_sinc2(dopPOS - idop);
  if (par->nsinc2 == 2) {
    double bl = dopPOS + g->dopdiff_bl - idop;
    return (o_sinc2(bl) + o_sinc2(bl + g->xincr) + o_sinc2(bl + g->yincr) + o_sinc2(bl + g->xincr + g->yincr)) / 4;
  }
  double arg_left = dopPOS + g->dopdiff_bl - idop, sum = 0.0;
  for (int i = 0; i < par->nsinc2; i++) {
    double a = arg_left;
    for (int j = 0; j < par->nsinc2; j++) {
      sum += o_sinc2(a);
      a += g->xincr;
    }
    arg_left += g->yincr;
  }
  return sum / (par->nsinc2 * par->nsinc2);
}

/* pos2deldop.c:176-581 (overflow accounting disabled as in the reference, :550-578) */
int oracle_pos2deldop(struct par_t *par, struct photo_t *photo, double orbit_xoff, double orbit_yoff,
                      double orbit_dopoff, struct deldop_t *deldop, int body, int set, int frm, int v)
{
  (void)set;
  struct deldopfrm_t *frame = &deldop->frame[frm];
  struct pos_t *pos = &frame->pos;
  int ndel = frame->ndel, ndop = frame->ndop;
  double del_c[MAXBINS + 1], dop_c[MAXBINS + 1];

  frame->idellim[0] = ndel + 999999;
  frame->idellim[1] = -999999;
  frame->idoplim[0] = ndop + 999999;
  frame->idoplim[1] = -999999;
  frame->dellim[0] = HUGENUMBER;
  frame->dellim[1] = -HUGENUMBER;
  frame->doplim[0] = HUGENUMBER;
  frame->doplim[1] = -HUGENUMBER;
  frame->badradar_logfactor = 0.0;

  int codemethod = deldop->codemethod, spb = deldop->spb, stride = deldop->stride;
  int dopfftlen = deldop->dopfftlen;
  double dopDC_vig = frame->dopDC_vig;
  int spb_over_stride = spb / stride;
  double one_over_spb = 1.0 / spb;
  int spb_sq = spb * spb;
  double const1, const2;
  if (codemethod != LONG_ORIG) {
    const1 = (3 * spb - 1) / (2.0 * stride);
    const2 = (spb - 1) / (2.0 * stride);
  } else {
    const1 = (double)spb_over_stride;
    const2 = 0.0;
  }
  double delfact = -KM2US / deldop->del_per_pixel;
  struct o_dopgeom g;
  int badradar = o_dopgeom(par, frame->view[v].oe, frame->view[v].spin, deldop->dopscale.val, deldop->Ftx,
                           deldop->dop_per_pixel, pos->km_per_pixel, &g, &frame->badradar_logfactor);
  frame->view[v].km2Hz = g.km2Hz;
  double delshift = frame->delcom_vig + frame->view[v].deloff;
  double dopshift = frame->dopcom_vig + frame->view[v].dopoff + orbit_dopoff;
  double kmpp2 = pos->km_per_pixel * pos->km_per_pixel;

  for (int x = pos->xlim[0]; x <= pos->xlim[1]; x++)
    for (int y = pos->ylim[0]; y <= pos->ylim[1]; y++) {
      if (!(pos->cose[x][y] > 0.0 && pos->body[x][y] == body && pos->z[x][y] != -1e20)) continue;
      double delPOS = pos->z[x][y] * delfact + delshift;
      int idel_min = (int)floor(delPOS - const1) + 1;
      int idel_max = (int)ceil(delPOS + const1) - 1;
      double dopPOS = g.ax * (x - orbit_xoff) + g.ay * (y - orbit_yoff) + dopshift;
      int idop_min = (int)floor(dopPOS - g.dopdiff_max + 1 - par->sinc2width / 2.0);
      int idop_max = (int)floor(dopPOS + g.dopdiff_max + par->sinc2width / 2.0);
      double codefactor = (codemethod == SHORT) ? o_sinc2((dopPOS - dopDC_vig) / dopfftlen) : 1.0;

      if (idel_min < frame->idellim[0]) frame->idellim[0] = idel_min;
      if (idel_max > frame->idellim[1]) frame->idellim[1] = idel_max;
      if (idop_min < frame->idoplim[0]) frame->idoplim[0] = idop_min;
      if (idop_max > frame->idoplim[1]) frame->idoplim[1] = idop_max;
      frame->dellim[0] = O_MIN(frame->dellim[0], delPOS);
      frame->dellim[1] = O_MAX(frame->dellim[1], delPOS);
      frame->doplim[0] = O_MIN(frame->doplim[0], dopPOS);
      frame->doplim[1] = O_MAX(frame->doplim[1], dopPOS);
      int in_bounds = (idel_min >= 1) && (idel_max <= ndel) && (idop_min >= 1) && (idop_max <= ndop);

      for (int idel = idel_min; idel <= idel_max; idel++) {
        double dc;
        if (codemethod != LONG_ORIG) {
          if (spb == 1) {
            dc = o_tri2(delPOS - idel);
          } else if (spb == 2) {
            double arg = (delPOS - (idel - const2)) / spb_over_stride;
            dc = o_tri(arg) + o_tri(arg - 0.5);
            dc *= dc / 4;
          } else {
            dc = 0.0;
            int m_min = O_MAX((int)floor((delPOS - idel - const2) * stride), 0);
            int m_max = O_MIN((int)ceil((delPOS - idel + const1) * stride), spb) - 1;
            double arg = (delPOS - (idel - const2)) / spb_over_stride - m_min * one_over_spb;
            for (int m = m_min; m <= m_max; m++) {
              dc += o_tri(arg);
              arg -= one_over_spb;
            }
            dc *= dc / spb_sq;
          }
        } else {
          dc = o_tri2((delPOS - idel) / spb_over_stride);
        }
        del_c[idel - idel_min] = dc;
      }
      for (int idop = idop_min; idop <= idop_max; idop++) {
        int k = O_MIN(idop - idop_min, MAXBINS);
        dop_c[k] = o_sinc2_mean(par, &g, dopPOS, idop);
      }
      double sumweights = 0.0;
      for (int idel = idel_min; idel <= idel_max; idel++)
        for (int idop = idop_min; idop <= idop_max; idop++) {
          int k = O_MIN(idop - idop_min, MAXBINS);
          sumweights += del_c[idel - idel_min] * dop_c[k];
        }
      double amp = oracle_radlaw(photo, deldop->iradlaw, pos->cose[x][y], pos->comp[x][y], pos->f[x][y]) *
                   pos->km_per_pixel * pos->km_per_pixel * codefactor / sumweights;
      (void)kmpp2;
      if (in_bounds) {
        for (int idel = idel_min; idel <= idel_max; idel++)
          for (int idop = idop_min; idop <= idop_max; idop++) {
            int k = O_MIN(idop - idop_min, MAXBINS);
            frame->fit[idel][idop] += amp * del_c[idel - idel_min] * dop_c[k];
          }
      }
    }

  frame->dellim[0] = (frame->dellim[0] - delshift) * deldop->del_per_pixel;
  frame->dellim[1] = (frame->dellim[1] - delshift) * deldop->del_per_pixel;
  frame->doplim[0] = (frame->doplim[0] - dopshift) * deldop->dop_per_pixel - g.dop_extra;
  frame->doplim[1] = (frame->doplim[1] - dopshift) * deldop->dop_per_pixel + g.dop_extra;
  frame->overflow_o2 = 0.0;
  frame->overflow_m2 = 0.0;
  frame->overflow_xsec = 0.0;
  frame->overflow_delmean = 0.0;
  frame->overflow_dopmean = 0.0;
  return badradar;
}

/* pos2doppler.c:149-501, with the overflow spectrum held per call */
int oracle_pos2doppler(struct par_t *par, struct photo_t *photo, double orbit_xoff, double orbit_yoff,
                       double orbit_dopoff, struct doppler_t *doppler, int body, int set, int frm, int v)
{
  (void)set;
  struct dopfrm_t *frame = &doppler->frame[frm];
  struct pos_t *pos = &frame->pos;
  int ndop = frame->ndop, any_overflow = 0, idop0 = 0;
  double dop_c[MAXBINS + 1];
  double *fit_overflow = (double *)calloc(MAXOVERFLOW, sizeof(double));

  frame->idoplim[0] = ndop + 999999;
  frame->idoplim[1] = -999999;
  frame->doplim[0] = HUGENUMBER;
  frame->doplim[1] = -HUGENUMBER;
  frame->badradar_logfactor = 0.0;
  struct o_dopgeom g;
  int badradar = o_dopgeom(par, frame->view[v].oe, frame->view[v].spin, doppler->dopscale.val, doppler->Ftx,
                           doppler->dop_per_bin, pos->km_per_pixel, &g, &frame->badradar_logfactor);
  frame->view[v].km2Hz = g.km2Hz;
  double dopshift = frame->dopcom_vig + frame->view[v].dopoff + orbit_dopoff;

  for (int x = pos->xlim[0]; x <= pos->xlim[1]; x++)
    for (int y = pos->ylim[0]; y <= pos->ylim[1]; y++) {
      if (!(pos->cose[x][y] > 0.0 && pos->body[x][y] == body)) continue;
      double dopPOS = g.ax * (x - orbit_xoff) + g.ay * (y - orbit_yoff) + dopshift;
      int idop_min = (int)floor(dopPOS - g.dopdiff_max + 1 - par->sinc2width / 2.0);
      int idop_max = (int)floor(dopPOS + g.dopdiff_max + par->sinc2width / 2.0);
      if (idop_min < frame->idoplim[0]) frame->idoplim[0] = idop_min;
      if (idop_max > frame->idoplim[1]) frame->idoplim[1] = idop_max;
      frame->doplim[0] = O_MIN(frame->doplim[0], dopPOS);
      frame->doplim[1] = O_MAX(frame->doplim[1], dopPOS);
      int in_bounds = (idop_min >= 1) && (idop_max <= ndop);
      if (!in_bounds && !any_overflow) {
        any_overflow = 1;
        memset(fit_overflow, 0, sizeof(double) * MAXOVERFLOW);
        idop0 = MAXOVERFLOW / 2 - (int)floor(dopshift + 0.5);
      }
      for (int idop = idop_min; idop <= idop_max; idop++) {
        int k = O_MIN(idop - idop_min, MAXBINS);
        dop_c[k] = o_sinc2_mean(par, &g, dopPOS, idop);
      }
      double sumweights = 0.0;
      for (int idop = idop_min; idop <= idop_max; idop++) sumweights += dop_c[O_MIN(idop - idop_min, MAXBINS)];
      double amp = oracle_radlaw(photo, doppler->iradlaw, pos->cose[x][y], pos->comp[x][y], pos->f[x][y]) *
                   pos->km_per_pixel * pos->km_per_pixel / sumweights;
      if (in_bounds) {
        for (int idop = idop_min; idop <= idop_max; idop++)
          frame->fit[idop] += amp * dop_c[O_MIN(idop - idop_min, MAXBINS)];
      } else {
        int idop1 = O_MAX(idop_min, -idop0), idop2 = O_MIN(idop_max, -idop0 + MAXOVERFLOW - 1);
        for (int idop = idop1; idop <= idop2; idop++)
          fit_overflow[idop + idop0] += amp * dop_c[O_MIN(idop - idop_min, MAXBINS)];
      }
    }

  frame->doplim[0] = (frame->doplim[0] - dopshift) * doppler->dop_per_bin - g.dop_extra;
  frame->doplim[1] = (frame->doplim[1] - dopshift) * doppler->dop_per_bin + g.dop_extra;
  frame->overflow_o2 = 0.0;
  frame->overflow_m2 = 0.0;
  frame->overflow_xsec = 0.0;
  frame->overflow_dopmean = 0.0;
  double sdev_sq = frame->sdev * frame->sdev, variance = sdev_sq;
  double lookfact = (frame->nlooks > 0.0) ? 1.0 / frame->nlooks : 0.0;
  if (any_overflow) {
    int j1 = O_MAX(frame->idoplim[0] + idop0, 0), j2 = O_MIN(frame->idoplim[1] + idop0, MAXOVERFLOW - 1);
    for (int j = j1; j <= j2; j++) {
      if (fit_overflow[j] != 0.0) {
        if (par->speckle) variance = sdev_sq + lookfact * fit_overflow[j] * fit_overflow[j];
        frame->overflow_o2 += 1.0;
        frame->overflow_m2 += fit_overflow[j] * fit_overflow[j] / variance;
        frame->overflow_xsec += fit_overflow[j];
        frame->overflow_dopmean += (j - idop0) * fit_overflow[j];
      }
    }
    if (frame->overflow_xsec != 0.0) frame->overflow_dopmean /= frame->overflow_xsec;
    if ((frame->idoplim[0] + idop0) < 0 || (frame->idoplim[1] + idop0) >= MAXOVERFLOW) {
      badradar = 1;
      double dopfactor = (O_MAX(frame->idoplim[1] + idop0, MAXOVERFLOW) - O_MIN(frame->idoplim[0] + idop0, 0)) /
                         (1.0 * MAXOVERFLOW);
      frame->badradar_logfactor += log(dopfactor);
    }
  }
  free(fit_overflow);
  return badradar;
}
