/*
 * oracle_photo.c -- optical brightness and the natural cubic spline.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>

#include "oracle_internal.h"

#define O_TINY 1.0e-40

/* apply_photo.c:78-309.  Hapke laws need astro/hapke.c and are not restated. */
double oracle_apply_photo(struct mod_t *mod, int ilaw, double phase, double intensityfactor,
                          struct pos_t *pos, int body, int s, int frm)
{
  (void)s; (void)frm;
  union optscat_t *o = &mod->photo.optical[ilaw];
  double sum = 0.0;
  int type = mod->photo.opttype[ilaw];
  double scale = 0, scale_ls = 0, scale_l = 0;
  switch (type) {
  case LAMBERTLAW: scale = o->R.R.val / PIE; break;
  case LOMMEL: scale = o->R.R.val / (4 * PIE); break;
  case GEOMETRICAL: break;
  case KAASALAINEN: {
    double pf = o->kaas.A0.val * exp(-phase / o->kaas.D.val) + o->kaas.k.val * phase + 1;
    scale_ls = (1 - o->kaas.wt.val) * pf * o->kaas.R.val / (4 * PIE);
    scale_l = o->kaas.wt.val * pf * o->kaas.R.val / PIE;
    break;
  }
  case HARMLAMBERT: case INHOLAMBERT: case HARMLOMMEL: case INHOLOMMEL:
  case HARMKAAS: case INHOKAAS:
    break;
  default:
    oracle_set_error("oracle_apply_photo: unsupported optical scattering law");
    return O_TINY;
  }
  for (int i = pos->xlim[0]; i <= pos->xlim[1]; i++)
    for (int j = pos->ylim[0]; j <= pos->ylim[1]; j++) {
      if (!(pos->cose[i][j] > 0.0 && pos->cosi[i][j] > 0.0 && pos->body[i][j] == body)) continue;
      double ci = pos->cosi[i][j], ce = pos->cose[i][j];
      int c = pos->comp[i][j], f = pos->f[i][j];
      switch (type) {
      case LAMBERTLAW: pos->b[i][j] = intensityfactor * scale * ci; break;
      case LOMMEL: pos->b[i][j] = intensityfactor * scale * ci / (ci + ce); break;
      case GEOMETRICAL: pos->b[i][j] = intensityfactor * o->R.R.val; break;
      case KAASALAINEN: pos->b[i][j] = intensityfactor * ci * (scale_ls / (ci + ce) + scale_l); break;
      case HARMLAMBERT: case INHOLAMBERT: {
        if (f < 0) continue;
        struct R_t *l = type == HARMLAMBERT ? &o->harmR.local[c][f] : &o->inhoR.local[c][f];
        pos->b[i][j] = intensityfactor * (l->R.val / PIE) * ci;
        break;
      }
      case HARMLOMMEL: case INHOLOMMEL: {
        if (f < 0) continue;
        struct R_t *l = type == HARMLOMMEL ? &o->harmR.local[c][f] : &o->inhoR.local[c][f];
        pos->b[i][j] = intensityfactor * (l->R.val / (4 * PIE)) * ci / (ci + ce);
        break;
      }
      case HARMKAAS: case INHOKAAS: {
        if (f < 0) continue;
        struct kaas_t *l = type == HARMKAAS ? &o->harmkaas.local[c][f] : &o->inhokaas.local[c][f];
        double pf = l->A0.val * exp(-phase / l->D.val) + l->k.val * phase + 1;
        double sls = (1 - l->wt.val) * pf * l->R.val / (4 * PIE);
        double sl = l->wt.val * pf * l->R.val / PIE;
        pos->b[i][j] = intensityfactor * ci * (sls / (ci + ce) + sl);
        break;
      }
      }
      sum += pos->b[i][j];
    }
  if (sum == 0.0) sum = O_TINY;
  return sum;
}

/* nr/spline.c:3-37: second derivatives of the interpolating cubic spline
   (yp1, ypn > 0.99e30 select the natural end condition) */
void oracle_spline(double *x, double *y, int n, double yp1, double ypn, double *y2)
{
  double *u = (double *)malloc(sizeof(double) * (n + 1));
  double qn, un;
  if (yp1 > 0.99e30) y2[1] = u[1] = 0.0;
  else {
    y2[1] = -0.5;
    u[1] = (3.0 / (x[2] - x[1])) * ((y[2] - y[1]) / (x[2] - x[1]) - yp1);
  }
  for (int i = 2; i <= n - 1; i++) {
    double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    double p = sig * y2[i - 1] + 2.0;
    y2[i] = (sig - 1.0) / p;
    u[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
    u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
  }
  if (ypn > 0.99e30) qn = un = 0.0;
  else {
    qn = 0.5;
    un = (3.0 / (x[n] - x[n - 1])) * (ypn - (y[n] - y[n - 1]) / (x[n] - x[n - 1]));
  }
  y2[n] = (un - qn * u[n - 1]) / (qn * y2[n - 1] + 1.0);
  for (int k = n - 1; k >= 1; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
  free(u);
}

/* nr/splint.c:3-23 */
void oracle_splint(double *xa, double *ya, double *y2a, int n, double x, double *y)
{
  int klo = 1, khi = n;
  while (khi - klo > 1) {
    int k = (khi + klo) >> 1;
    if (xa[k] > x) khi = k;
    else klo = k;
  }
  double h = xa[khi] - xa[klo];
  if (h == 0.0) {
    oracle_set_error("oracle_splint: bad xa input");
    *y = NAN;
    return;
  }
  double a = (xa[khi] - x) / h, b = (x - xa[klo]) / h;
  *y = a * ya[klo] + b * ya[khi] + ((a * a * a - a) * y2a[klo] + (b * b * b - b) * y2a[khi]) * (h * h) / 6.0;
}
