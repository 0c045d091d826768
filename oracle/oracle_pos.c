/*
 * oracle_pos.c -- plane-of-sky clear, rasterization with z-buffer, shadow mask.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <limits.h>
#include <stdlib.h>

#include "oracle_internal.h"

/* posclr.c:24-67 */
void oracle_posclr(struct pos_t *pos)
{
  int n = pos->n;
  for (int i = -n; i <= n; i++)
    for (int j = -n; j <= n; j++) {
      pos->b[i][j] = pos->cose[i][j] = pos->cosi[i][j] = 0.0;
      pos->z[i][j] = -HUGENUMBER;
      pos->body[i][j] = pos->comp[i][j] = pos->f[i][j] = -1;
    }
  pos->xlim[0] = pos->ylim[0] = n;
  pos->xlim[1] = pos->ylim[1] = -n;
  if (pos->bistatic) {
    for (int i = -n; i <= n; i++)
      for (int j = -n; j <= n; j++) {
        pos->cosill[i][j] = 0.0;
        pos->zill[i][j] = -HUGENUMBER;
        pos->bodyill[i][j] = pos->compill[i][j] = pos->fill[i][j] = -1;
      }
    pos->xlim2[0] = pos->ylim2[0] = n;
    pos->xlim2[1] = pos->ylim2[1] = -n;
  }
}

static int clamp_int(double v)
{
  if (v < INT_MIN) return INT_MIN;
  if (v > INT_MAX) return INT_MAX;
  return (int)v;
}

/* posvis.c:322-364: grow the overall and in-frame rectangles */
static void o_posrect(struct pos_t *pos, int src, double imin_d, double imax_d, double jmin_d,
                      double jmax_d, double ov[4])
{
  ov[0] = O_MIN(ov[0], imin_d);
  ov[1] = O_MAX(ov[1], imax_d);
  ov[2] = O_MIN(ov[2], jmin_d);
  ov[3] = O_MAX(ov[3], jmax_d);
  int imin = imin_d < INT_MIN ? INT_MIN : (int)imin_d;
  int imax = imax_d > INT_MAX ? INT_MAX : (int)imax_d;
  int jmin = jmin_d < INT_MIN ? INT_MIN : (int)jmin_d;
  int jmax = jmax_d > INT_MAX ? INT_MAX : (int)jmax_d;
  int n = pos->n;
  int *xl = src ? pos->xlim2 : pos->xlim, *yl = src ? pos->ylim2 : pos->ylim;
  if (imin < xl[0]) xl[0] = O_MAX(imin, -n);
  if (imax > xl[1]) xl[1] = O_MIN(imax, n);
  if (jmin < yl[0]) yl[0] = O_MAX(jmin, -n);
  if (jmax > yl[1]) yl[1] = O_MIN(jmax, n);
}

/* posvis.c:84-317 */
int oracle_posvis(struct vertices_t *verts, double orbit_offset[3], struct pos_t *pos, int smooth,
                  int src, int body, int comp)
{
  double oa[3][3], usrc[3] = {0, 0, 1};
  double **cosa, **cosb, **zz;
  int **bod, **cmp, **fac;
  int outbnd = 0, n = pos->n;
  double ov[4] = {HUGENUMBER, -HUGENUMBER, HUGENUMBER, -HUGENUMBER};
  double kmpp = pos->km_per_pixel;

  pos->posbnd_logfactor = 0.0;
  o_mtrnsps(oa, pos->ae);
  if (src) {
    o_mmmul(oa, pos->se, oa);
    cosa = pos->cosill; cosb = NULL; zz = pos->zill;
    bod = pos->bodyill; cmp = pos->compill; fac = pos->fill;
  } else {
    o_mmmul(oa, pos->oe, oa);
    cosa = pos->cose; cosb = pos->cosi; zz = pos->z;
    bod = pos->body; cmp = pos->comp; fac = pos->f;
    if (pos->bistatic) {
      usrc[0] = usrc[1] = 0.0;
      usrc[2] = 1.0;
      o_cotrans(usrc, pos->se, usrc, -1);
      o_cotrans(usrc, pos->oe, usrc, 1);
    }
  }

  for (int f = 0; f < verts->nf; f++) {
    double nrm[3];
    for (int i = 0; i <= 2; i++) nrm[i] = verts->f[f].n[i];
    o_cotrans(nrm, oa, nrm, 1);
    if (!(nrm[2] > 0.0)) continue;

    double v0[3], v1[3], v2[3];
    const int *fv = verts->f[f].v;
    o_cotrans(v0, oa, verts->v[fv[0]].x, 1);
    o_cotrans(v1, oa, verts->v[fv[1]].x, 1);
    o_cotrans(v2, oa, verts->v[fv[2]].x, 1);
    for (int i = 0; i <= 2; i++) {
      v0[i] += orbit_offset[i];
      v1[i] += orbit_offset[i];
      v2[i] += orbit_offset[i];
    }
    double imin_d = floor(O_MIN(v0[0], O_MIN(v1[0], v2[0])) / kmpp - SMALLVAL + 0.5);
    double imax_d = floor(O_MAX(v0[0], O_MAX(v1[0], v2[0])) / kmpp + SMALLVAL + 0.5);
    double jmin_d = floor(O_MIN(v0[1], O_MIN(v1[1], v2[1])) / kmpp - SMALLVAL + 0.5);
    double jmax_d = floor(O_MAX(v0[1], O_MAX(v1[1], v2[1])) / kmpp + SMALLVAL + 0.5);
    int imin = imin_d < INT_MIN ? INT_MIN : (int)imin_d;
    int imax = imax_d > INT_MAX ? INT_MAX : (int)imax_d;
    int jmin = jmin_d < INT_MIN ? INT_MIN : (int)jmin_d;
    int jmax = jmax_d > INT_MAX ? INT_MAX : (int)jmax_d;
    if (imin < -n || imax > n || jmin < -n || jmax > n) outbnd = 1;
    int i1 = O_MAX(imin, -n), i2 = O_MIN(imax, n), j1 = O_MAX(jmin, -n), j2 = O_MIN(jmax, n);

    if (i1 > n || i2 < -n || j1 > n || j2 < -n) {
      o_posrect(pos, src, imin_d, imax_d, jmin_d, jmax_d, ov);
      continue;
    }
    for (int i = i1; i <= i2; i++) {
      double x0 = i * kmpp;
      for (int j = j1; j <= j2; j++) {
        double x1 = j * kmpp, den;
        double s = ((x0 - v0[0]) * (v2[1] - v1[1]) - (v2[0] - v1[0]) * (x1 - v0[1])) *
                   (den = 1 / ((v1[0] - v0[0]) * (v2[1] - v1[1]) - (v2[0] - v1[0]) * (v1[1] - v0[1])));
        if (!((s >= -SMALLVAL) && (s <= 1.0 + SMALLVAL))) continue;
        double t = ((v1[0] - v0[0]) * (x1 - v0[1]) - (x0 - v0[0]) * (v1[1] - v0[1])) * den;
        if (!((t >= -SMALLVAL) && (t <= s + SMALLVAL))) continue;
        double z = v0[2] + s * (v1[2] - v0[2]) + t * (v2[2] - v1[2]);
        if (!((z > zz[i][j]) || (fac[i][j] < 0))) continue;
        zz[i][j] = z;
        if (smooth) {
          const struct vertex_t *a = &verts->v[fv[0]], *b = &verts->v[fv[1]], *c = &verts->v[fv[2]];
          for (int k = 0; k <= 2; k++)
            nrm[k] = a->n[k] + s * (b->n[k] - a->n[k]) + t * (c->n[k] - b->n[k]);
          o_cotrans(nrm, oa, nrm, 1);
          o_normalize(nrm);
        }
        if (nrm[2] > 0.0) {
          cosa[i][j] = nrm[2];
          if (!src && pos->bistatic) {
            cosb[i][j] = o_dot(nrm, usrc);
            if (cosb[i][j] <= 0.0) cosa[i][j] = 0.0;
          }
        }
        o_posrect(pos, src, (double)i, (double)i, (double)j, (double)j, ov);
        bod[i][j] = body;
        cmp[i][j] = comp;
        fac[i][j] = f;
      }
    }
  }
  if (outbnd) {
    double xfactor = (O_MAX(ov[1], n) - O_MIN(ov[0], -n) + 1) / (2 * n + 1);
    double yfactor = (O_MAX(ov[3], n) - O_MIN(ov[2], -n) + 1) / (2 * n + 1);
    pos->posbnd_logfactor = log(xfactor * yfactor);
  }
  (void)clamp_int;
  return outbnd;
}

/* posmask.c:58-174 */
void oracle_posmask(struct pos_t *pos, double tol)
{
  double so[3][3], x[3];
  double bignum = 0.99 * HUGENUMBER;
  o_mtrnsps(so, pos->oe);
  o_mmmul(so, pos->se, so);
  double ppk = 1 / pos->km_per_pixel;
  int n = pos->n;
  for (int i = -n; i <= n; i++)
    for (int j = -n; j <= n; j++) {
      if (pos->cose[i][j] == 0.0) continue;
      x[0] = i * pos->km_per_pixel;
      x[1] = j * pos->km_per_pixel;
      x[2] = pos->z[i][j];
      o_cotrans(x, so, x, 1);
      double i0 = x[0] * ppk, j0 = x[1] * ppk;
      int im = o_iround(i0), jm = o_iround(j0);
      if (!(fabs(i0) < n && fabs(j0) < n && pos->zill[im][jm] > -bignum &&
            (pos->f[i][j] != pos->fill[im][jm] || pos->comp[i][j] != pos->compill[im][jm] ||
             pos->body[i][j] != pos->bodyill[im][jm])))
        continue;
      int i1 = (int)floor(i0), j1 = (int)floor(j0);
      double zill;
      if (pos->zill[i1][j1] > -bignum && pos->zill[i1 + 1][j1] > -bignum &&
          pos->zill[i1][j1 + 1] > -bignum && pos->zill[i1 + 1][j1 + 1] > -bignum) {
        double t = i0 - i1, u = j0 - j1;
        zill = (1 - t) * (1 - u) * pos->zill[i1][j1] + t * (1 - u) * pos->zill[i1 + 1][j1] +
               t * u * pos->zill[i1 + 1][j1 + 1] + (1 - t) * u * pos->zill[i1][j1 + 1];
      } else {
        zill = pos->zill[im][jm];
        int is = (i0 >= im) ? 1 : -1;
        int i2 = im + is;
        if (abs(i2) <= n && pos->zill[i2][jm] > -bignum) {
          zill += fabs(i0 - im) * (pos->zill[i2][jm] - pos->zill[im][jm]);
        } else {
          i2 = im - is;
          if (abs(i2) <= n && pos->zill[i2][jm] > -bignum)
            zill -= fabs(i0 - im) * (pos->zill[i2][jm] - pos->zill[im][jm]);
        }
        int js = (j0 >= jm) ? 1 : -1;
        int j2 = jm + js;
        if (abs(j2) <= n && pos->zill[im][j2] > -bignum) {
          zill += fabs(j0 - jm) * (pos->zill[im][j2] - pos->zill[im][jm]);
        } else {
          j2 = jm - js;
          if (abs(j2) <= n && pos->zill[im][j2] > -bignum)
            zill -= fabs(j0 - jm) * (pos->zill[im][j2] - pos->zill[im][jm]);
        }
      }
      if (zill - x[2] > tol) pos->cose[i][j] = 0.0;
    }
}
