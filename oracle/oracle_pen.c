/*
 * oracle_pen.c -- model moments (area, volume, centre of mass, inertia
 * tensor) and the penalty functions of the objective.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * oracle_compute_moments restates realize_mod.c:757-840 (compute_moments)
 * with facmom.c:44-160 for each facet: the volume between the facet and the
 * coordinate plane, its first moment and its inertia tensor, uniform
 * density, each coordinate axis in turn by cyclic permutation.
 *
 * oracle_diag_inertia restates diag_inertia.c:11-60: the principal moments
 * and axes of the inertia tensor by the cyclic Jacobi method with the
 * threshold strategy of Numerical Recipes in C, 2nd ed., section 11.1
 * (nr/jacobi.c), eigenvalues left in the order the rotations leave them.
 *
 * oracle_penalties restates penalties.c:120-1250 for every penalty that
 * depends on the model's geometry, moments, spin state and dataset offsets:
 * nonsmooth (:597), concavity (:635), rdev (:689), maxrdev (:729),
 * maxellipdev (:765), volume (:846), comdev (:853), inertiadev (:877),
 * inertiadev_uni (:926), pa3tilt (:958), nonpa (:984), nonpa_uni (:1003),
 * euleroffs (:1034), flattening (:1053), bifurcation (:1083) and impulse
 * (:1211), with the weights' rule (:1228-1237).  The albedo, Hapke-angle and
 * radar-law penalties (optalbdel ... rad_c_var, noncosine) read the
 * harmonic, inhomogeneous and tabular scattering laws' per-facet parameters,
 * which are outside this path: they set the error and return 0.
 */
#include <stdio.h>

#include "oracle_internal.h"

#define O_PIE 3.141592653589793
#define O_R2D 57.29577951308232
#define O_TINYPEN 1.0e-10
#define O_NBIF 20            /* penalties.c:116-117: NBIFURCATION, zones = NBIFURCATION - 1 */
#define O_NZONES (O_NBIF - 1)
#define O_HUGE 1.0e20
#define O_SMALLVAL 1.0e-10

/* facmom.c: one facet's contribution to the volume, the first moment and the
   inertia tensor of the region it sweeps out parallel to each axis */
static void o_facmom(const double fv0[3], const double fv1[3], const double fv2[3], const double fn[3], double *dv,
                     double dvr[3], double dI[3][3])
{
  double v0[3], v1[3], v2[3], n[3], tmp;
  for (int i = 0; i <= 2; i++) { v0[i] = fv0[i]; v1[i] = fv1[i]; v2[i] = fv2[i]; n[i] = fn[i]; }
  int sign = n[2] < 0.0 ? -1 : 1;
  double jac = fabs((v1[0] - v0[0]) * (v2[1] - v1[1]) - (v2[0] - v1[0]) * (v1[1] - v0[1]));
  *dv = sign * jac * (v0[2] + v1[2] + v2[2]) / 6.0;
  for (int i = 0; i <= 2; i++) {
    double r = sign * jac * (v0[2] * (v0[2] + v1[2] + v2[2]) + v1[2] * (v1[2] + v2[2]) + v2[2] * v2[2]) / 24.0;
    double s = -(sign * jac / 120.0) *
               ((v0[1] + v1[1] + v2[1]) * (v0[2] * v0[2] + v1[2] * v1[2] + v2[2] * v2[2]) +
                2.0 * (v0[1] * v0[2] + v1[1] * v1[2] + v2[1] * v2[2]) * (v0[2] + v1[2] + v2[2]) +
                (v0[1] * v1[2] * v2[2] + v1[1] * v0[2] * v2[2] + v2[1] * v0[2] * v1[2]));
    double t = (sign * jac / 60.0) *
               ((v0[2] * v0[2] + v1[2] * v1[2] + v2[2] * v2[2]) * (v0[2] + v1[2] + v2[2]) + v0[2] * v1[2] * v2[2]);
    switch (i) {
    case 0: dvr[2] = r; dI[1][2] = s; dI[0][0] = t; dI[1][1] = t; break;
    case 1: dvr[1] = r; dI[0][1] = s; dI[0][0] += t; dI[2][2] = t; break;
    case 2: dvr[0] = r; dI[0][2] = s; dI[1][1] += t; dI[2][2] += t; break;
    }
    if (i < 2) {  /* (x, y, z) -> (z, x, y): the next axis plays z */
      tmp = v0[2]; v0[2] = v0[1]; v0[1] = v0[0]; v0[0] = tmp;
      tmp = v1[2]; v1[2] = v1[1]; v1[1] = v1[0]; v1[0] = tmp;
      tmp = v2[2]; v2[2] = v2[1]; v2[1] = v2[0]; v2[0] = tmp;
      tmp = n[2]; n[2] = n[1]; n[1] = n[0]; n[0] = tmp;
      jac = fabs((v1[0] - v0[0]) * (v2[1] - v1[1]) - (v2[0] - v1[0]) * (v1[1] - v0[1]));
      sign = n[2] < 0.0 ? -1 : 1;
    }
  }
  dI[1][0] = dI[0][1];
  dI[2][0] = dI[0][2];
  dI[2][1] = dI[1][2];
}

void oracle_compute_moments(struct mod_t *mod)
{
  double dI[3][3], dcom[3], dv;
  mod->shape.area = 0.0;
  mod->shape.volume = 0.0;
  for (int k = 0; k <= 2; k++) {
    mod->shape.com[k] = 0.0;
    for (int j = 0; j <= 2; j++) mod->shape.inertia[k][j] = 0.0;
  }
  for (int c = 0; c < mod->shape.ncomp; c++) {
    struct comp_t *comp = &mod->shape.comp[c];
    struct vertices_t *real = &comp->real;
    comp->area = 0.0;
    for (int f = 0; f < real->nf; f++) {
      comp->area += real->f[f].area;
      if (real->f[f].act) mod->shape.area += real->f[f].area;
    }
    comp->volume = 0.0;
    for (int k = 0; k <= 2; k++) {
      comp->com[k] = 0.0;
      for (int j = 0; j <= 2; j++) comp->inertia[k][j] = 0.0;
    }
    for (int f = 0; f < real->nf; f++) {
      const struct facet_t *fa = &real->f[f];
      o_facmom(real->v[fa->v[0]].x, real->v[fa->v[1]].x, real->v[fa->v[2]].x, fa->n, &dv, dcom, dI);
      comp->volume += dv;
      for (int j = 0; j <= 2; j++) {
        comp->com[j] += dcom[j];
        for (int k = 0; k <= 2; k++) comp->inertia[j][k] += dI[j][k];
      }
      if (fa->act) {
        mod->shape.volume += dv;
        for (int j = 0; j <= 2; j++) {
          mod->shape.com[j] += dcom[j];
          for (int k = 0; k <= 2; k++) mod->shape.inertia[j][k] += dI[j][k];
        }
      }
    }
    for (int j = 0; j <= 2; j++) comp->com[j] /= comp->volume;
  }
  for (int j = 0; j <= 2; j++) mod->shape.com[j] /= mod->shape.volume;
}

/* Eigen-decomposition of a symmetric 3x3 matrix by cyclic Jacobi rotations:
   sweeps over the upper triangle in row order; in the first three sweeps an
   element is rotated away only above 0.2 * (sum of |off-diagonal|) / 9,
   after the fourth an element negligible beside both diagonal entries is
   zeroed; the diagonal is refreshed from the accumulated updates after each
   sweep.  d gets the eigenvalues, the columns of v the eigenvectors. */
void oracle_jacobi3(double a_in[3][3], double d[3], double v[3][3])
{
  double a[3][3], b[3], z[3];
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++) { a[p][q] = a_in[p][q]; v[p][q] = p == q ? 1.0 : 0.0; }
  for (int p = 0; p < 3; p++) { b[p] = d[p] = a[p][p]; z[p] = 0.0; }
  for (int sweep = 1; sweep <= 50; sweep++) {
    double sm = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
    if (sm == 0.0) return;
    double tresh = sweep < 4 ? 0.2 * sm / 9 : 0.0;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        double g = 100.0 * fabs(a[p][q]);
        if (sweep > 4 && fabs(d[p]) + g == fabs(d[p]) && fabs(d[q]) + g == fabs(d[q])) {
          a[p][q] = 0.0;
        } else if (fabs(a[p][q]) > tresh) {
          double h = d[q] - d[p], t;
          if (fabs(h) + g == fabs(h)) {
            t = a[p][q] / h;
          } else {
            double theta = 0.5 * h / a[p][q];
            t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
            if (theta < 0.0) t = -t;
          }
          double c = 1.0 / sqrt(1 + t * t), s = t * c, tau = s / (1.0 + c);
          h = t * a[p][q];
          z[p] -= h; z[q] += h; d[p] -= h; d[q] += h;
          a[p][q] = 0.0;
          /* rotate the other upper-triangle entries of rows/columns p and q:
             (j, p)-(j, q) for j < p, (p, j)-(j, q) for p < j < q, (p, j)-(q, j) for j > q */
          for (int j = 0; j < 3; j++) {
            double *x, *y;
            if (j < p) { x = &a[j][p]; y = &a[j][q]; }
            else if (j > p && j < q) { x = &a[p][j]; y = &a[j][q]; }
            else if (j > q) { x = &a[p][j]; y = &a[q][j]; }
            else continue;
            double gg = *x, hh = *y;
            *x = gg - s * (hh + gg * tau);
            *y = hh + s * (gg - hh * tau);
          }
          for (int j = 0; j < 3; j++) {
            double gg = v[j][p], hh = v[j][q];
            v[j][p] = gg - s * (hh + gg * tau);
            v[j][q] = hh + s * (gg - hh * tau);
          }
        }
      }
    for (int p = 0; p < 3; p++) { b[p] += z[p]; d[p] = b[p]; z[p] = 0.0; }
  }
  oracle_set_error("oracle_jacobi3: no convergence in 50 sweeps");
}

void oracle_diag_inertia(double inertia[3][3], double pmoment[3], double ap[3][3])
{
  oracle_jacobi3(inertia, pmoment, ap);
}

/* the model's effective radius: the radius of the sphere of its volume */
static double o_reff(const struct mod_t *mod) { return pow(3 * mod->shape.volume / (4 * O_PIE), 1.0 / 3.0); }

/* The per-facet value a law-reading penalty takes from law ilaw at facet f
   of component c (penalties.c:150-522): kind 0 the optical albedo (R, or
   Hapke's w), 1 Hapke's theta, 2 the radar R, 3 the radar C; NULL where the
   law carries none. */
static const double *o_law_value(struct mod_t *mod, int kind, int ilaw, int c, int f)
{
  if (kind >= 2) {
    union radscat_t *r = &mod->photo.radar[ilaw];
    const struct RC_t *q = mod->photo.radtype[ilaw] == HARMCOSINE_DIFF   ? &r->harmcosine.local[c][f]
                           : mod->photo.radtype[ilaw] == INHOCOSINE_DIFF ? &r->inhocosine.local[c][f]
                                                                         : NULL;
    return q ? (kind == 2 ? &q->R.val : &q->C.val) : NULL;
  }
  union optscat_t *o = &mod->photo.optical[ilaw];
  switch (mod->photo.opttype[ilaw]) {
  case HARMLAMBERT: case HARMLOMMEL: return kind == 0 ? &o->harmR.local[c][f].R.val : NULL;
  case INHOLAMBERT: case INHOLOMMEL: return kind == 0 ? &o->inhoR.local[c][f].R.val : NULL;
  case HARMHAPKE: return kind == 0 ? &o->harmhapke.local[c][f].w.val : &o->harmhapke.local[c][f].theta.val;
  case INHOHAPKE: return kind == 0 ? &o->inhohapke.local[c][f].w.val : &o->inhohapke.local[c][f].theta.val;
  case HARMKAAS: return kind == 0 ? &o->harmkaas.local[c][f].R.val : NULL;
  case INHOKAAS: return kind == 0 ? &o->inhokaas.local[c][f].R.val : NULL;
  }
  return NULL;
}

/* the "del" penalties: over sides whose facets are both active, the
   length-weighted sum of |x - y| over that of |x + y| (penalties.c:150-217) */
static double o_law_del(struct mod_t *mod, int kind)
{
  double a = 0.0, b = 0.0;
  const int nlaws = kind >= 2 ? mod->photo.nradlaws : mod->photo.noptlaws;
  for (int c = 0; c < mod->shape.ncomp; c++) {
    struct vertices_t *r = &mod->shape.comp[c].real;
    for (int s = 0; s < r->ns; s++) {
      const int f1 = r->s[s].f[0], f2 = r->s[s].f[1];
      if (!(r->f[f1].act && r->f[f2].act)) continue;
      const double *x1 = r->v[r->s[s].v[0]].x, *x2 = r->v[r->s[s].v[1]].x;
      const double len = sqrt((x1[0] - x2[0]) * (x1[0] - x2[0]) + (x1[1] - x2[1]) * (x1[1] - x2[1]) +
                              (x1[2] - x2[2]) * (x1[2] - x2[2]));
      for (int l = 0; l < nlaws; l++) {
        const double *x = o_law_value(mod, kind, l, c, f1), *y = o_law_value(mod, kind, l, c, f2);
        if (!x) continue;
        a += len * fabs(*x + *y);
        b += len * fabs(*x - *y);
      }
    }
  }
  if (a == 0.0) {
    oracle_set_error("oracle_penalties: a 'del' penalty without a law it reads");
    return 0.0;
  }
  return b / a;
}

/* the "var" penalties: the fractional variance over active facets,
   ntot (sum x^2) / (sum x)^2 - 1, clamped at 0 (penalties.c:219-277).  The
   reference's optalbvar does not count harmonic-Hapke facets in ntot; kept. */
static double o_law_var(struct mod_t *mod, int kind, int albvar)
{
  double av = 0.0, av2 = 0.0;
  int ntot = 0;
  const int nlaws = kind >= 2 ? mod->photo.nradlaws : mod->photo.noptlaws;
  for (int c = 0; c < mod->shape.ncomp; c++) {
    struct vertices_t *r = &mod->shape.comp[c].real;
    for (int f = 0; f < r->nf; f++) {
      if (!r->f[f].act) continue;
      for (int l = 0; l < nlaws; l++) {
        const double *x = o_law_value(mod, kind, l, c, f);
        if (!x) continue;
        av += *x;
        av2 += *x * *x;
        if (!(albvar && mod->photo.opttype[l] == HARMHAPKE)) ++ntot;
      }
    }
  }
  if (ntot == 0) {
    oracle_set_error("oracle_penalties: a 'var' penalty without a law it reads");
    return 0.0;
  }
  const double pen = ntot * (av2 / (av * av)) - 1.0;
  return pen < 0.0 ? 0.0 : pen;
}

/* log10 of a tabular rho, continued as a straight line below 1e-20 */
static double o_logrho(double rho)
{
  return rho > 1.0e-20 ? log10(rho) : log10(1.0e-20) + (rho / 1.0e-20 - 1) / LN10;
}

/* noncosine (penalties.c:523-588): the mean squared residual of an
   unweighted line fit of log10 rho against log10 cos(theta) over the table's
   points below 90 deg, the 90 deg point adding its residual about the fit at
   the point before it when it is the larger */
static double o_noncosine(struct par_t *par, struct mod_t *mod)
{
  int ntot = 0;
  double resid2sum = 0.0;
  for (int l = 0; l < mod->photo.nradlaws; l++) {
    if (mod->photo.radtype[l] != TABULARLAW) continue;
    const struct tabular_t *t = &mod->photo.radar[l].tabular;
    const int n = t->n;
    ntot += n;
    double ymean = 0.0;
    for (int j = 0; j < n - 1; j++) {
      par->pen.y_noncosine[j] = o_logrho(t->rho[j].val);
      ymean += par->pen.y_noncosine[j];
    }
    ymean /= (n - 1);
    double x2 = 0.0, y2 = 0.0, xy = 0.0;
    for (int j = 0; j < n - 1; j++) {
      const double x = par->pen.x_noncosine[j] - par->pen.xmean_noncosine, y = par->pen.y_noncosine[j] - ymean;
      x2 += x * x;
      y2 += y * y;
      xy += x * y;
    }
    const double slope = xy / x2, intercept = ymean - slope * par->pen.xmean_noncosine;
    resid2sum += y2 - xy * xy / x2;
    if (t->rho[n - 1].val > t->rho[n - 2].val) {
      const double y = o_logrho(t->rho[n - 1].val) - (slope * par->pen.x_noncosine[n - 2] + intercept);
      resid2sum += y * y;
    }
  }
  if (ntot == 0) {
    oracle_set_error("oracle_penalties: noncosine without a tabular law");
    return 0.0;
  }
  return O_MAX(resid2sum / ntot, 0.0);
}

/* one penalty's value (before its weight); sets the error for a type not restated */
static double o_penalty(int type, struct par_t *par, struct mod_t *mod, struct dat_t *dat, double pm[3],
                        double ap[3][3])
{
  double pen = 0.0;
  int ntot = 0;
  switch (type) {
  case OPTALBDEL: return o_law_del(mod, 0);
  case THETADEL: return o_law_del(mod, 1);
  case RADALBDEL: return o_law_del(mod, 2);
  case RAD_C_DEL: return o_law_del(mod, 3);
  case OPTALBVAR: return o_law_var(mod, 0, 1);
  case THETAVAR: return o_law_var(mod, 1, 0);
  case RADALBVAR: return o_law_var(mod, 2, 0);
  case RAD_C_VAR: return o_law_var(mod, 3, 0);
  case NONCOSINE: return o_noncosine(par, mod);
  case NONSMOOTH:
    for (int c = 0; c < mod->shape.ncomp; c++) {
      struct vertices_t *r = &mod->shape.comp[c].real;
      for (int s = 0; s < r->ns; s++)
        if (r->s[s].act) {
          double x = 1 - o_dot(r->f[r->s[s].f[0]].n, r->f[r->s[s].f[1]].n);
          pen += x * x * x * x;
          ntot++;
        }
    }
    return pen / ntot;
  case CONCAVITY:
    for (int c = 0; c < mod->shape.ncomp; c++) {
      struct vertices_t *r = &mod->shape.comp[c].real;
      for (int s = 0; s < r->ns; s++)
        if (r->s[s].act) {
          int f1 = r->s[s].f[0], f2 = r->s[s].f[1], v1 = r->s[s].v[0], v2 = r->s[s].v[1], v3 = 0;
          for (int j = 0; j <= 2; j++)
            if (r->f[f1].v[j] != v1 && r->f[f1].v[j] != v2) v3 = r->f[f1].v[j];
          double disp[3];
          for (int j = 0; j <= 2; j++) disp[j] = r->v[v3].x[j] - r->v[v1].x[j];
          if (o_dot(disp, r->f[f2].n) > 0.0) {
            double x = 1 - o_dot(r->f[f1].n, r->f[f2].n);
            pen += x * x;
          }
          ntot++;
        }
    }
    return pen / ntot;
  case RDEV:
  case MAXRDEV: {
    double reff = o_reff(mod);
    for (int c = 0; c < mod->shape.ncomp; c++) {
      if (mod->shape.comp[c].type != VERTEX) continue;
      struct vertices_t *r = &mod->shape.comp[c].real;
      for (int v = 0; v < r->nv; v++)
        if (r->v[v].act) {
          double scale = 0.0;
          for (int j = 0; j <= 2; j++) {
            double y = r->v[v].u[j] * r->scalefactor[j].val;
            scale += y * y;
          }
          scale = sqrt(scale);
          double x = scale * r->v[v].r.val / reff;
          if (type == RDEV) {
            pen += x * x;
            ntot++;
          } else {
            x *= x;
            pen = O_MAX(pen, x);
          }
        }
    }
    if (type == RDEV) {
      if (ntot == 0) {
        oracle_set_error("oracle_penalties: rdev needs a vertex component");
        return 0.0;
      }
      pen /= ntot;
    }
    return pen;
  }
  case MAXELLIPDEV: {
    double reff = o_reff(mod), volume = mod->shape.volume, radius[3], deeve[3];
    radius[0] = sqrt((5 / volume) * (-pm[0] + pm[1] + pm[2]) / 2);
    radius[1] = sqrt((5 / volume) * (pm[0] - pm[1] + pm[2]) / 2);
    radius[2] = sqrt((5 / volume) * (pm[0] + pm[1] - pm[2]) / 2);
    double scale = pow(3 * volume / (4 * O_PIE * radius[0] * radius[1] * radius[2]), 1.0 / 3.0);
    for (int j = 0; j <= 2; j++) deeve[j] = scale * radius[j];
    for (int c = 0; c < mod->shape.ncomp; c++) {
      struct vertices_t *r = &mod->shape.comp[c].real;
      for (int v = 0; v < r->nv; v++)
        if (r->v[v].act) {
          double w[3];
          o_cotrans(w, ap, r->v[v].x, -1);
          double rr = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
          double theta = atan2(sqrt(w[0] * w[0] + w[1] * w[1]), w[2]), phi = atan2(w[1], w[0]);
          double xt = sin(theta) * cos(phi) / deeve[0], yt = sin(theta) * sin(phi) / deeve[1],
                 zt = cos(theta) / deeve[2];
          double rd = 1 / sqrt(xt * xt + yt * yt + zt * zt);
          double x = (rr - rd) / reff;
          x *= x;
          pen = O_MAX(pen, x);
        }
    }
    return pen;
  }
  case VOLUME:
    return mod->shape.volume;
  case COMDEV: {
    double reff = o_reff(mod);
    for (int k = 0; k <= 2; k++) pen += mod->shape.com[k] * mod->shape.com[k];
    return pen / (reff * reff);
  }
  case INERTIADEV: {
    double a = 0.0, b = 0.0, adotb = 0.0;
    for (int j = 0; j <= 2; j++) {
      a += mod->spin.inertia[j].val * mod->spin.inertia[j].val;
      adotb += mod->spin.inertia[j].val * mod->shape.inertia[j][j];
      for (int k = 0; k <= 2; k++) b += mod->shape.inertia[j][k] * mod->shape.inertia[j][k];
    }
    return 1 - adotb / sqrt(a * b);
  }
  case INERTIADEV_UNI: {
    double a = 0.0, b = 0.0, adotb = 0.0;
    for (int j = 0; j <= 2; j++) {
      a += pm[j] * pm[j];
      adotb += pm[j] * mod->shape.inertia[j][j];
      for (int k = 0; k <= 2; k++) b += mod->shape.inertia[j][k] * mod->shape.inertia[j][k];
    }
    return 1 - adotb / sqrt(a * b);
  }
  case PA3TILT:
    return O_MAX(0.0, 1 - ap[2][2] * ap[2][2]);
  case NONPA_UNI:
    pen = O_MAX((pm[0] / pm[2]) - 1, (pm[1] / pm[2]) - 1) + 0.01;
    return pen < 0.0 ? 0.0 : pen;
  case FLATTENING: {
    double boc = sqrt(O_MIN((-pm[0] + pm[1] + pm[2]), (pm[0] - pm[1] + pm[2])) / (pm[0] + pm[1] - pm[2]));
    if (boc > 1.0) {
      double x = boc - 1;
      return x * x * x * x;
    }
    return 0.0;
  }
  case BIFURCATION: {
    int jmax;
    if (pm[0] <= pm[1] && pm[0] <= pm[2]) jmax = 0;
    else if (pm[1] <= pm[0] && pm[1] <= pm[2]) jmax = 1;
    else jmax = 2;
    int j1 = (jmax + 1) % 3, j2 = (jmax + 2) % 3;
    double lo = O_HUGE, hi = -O_HUGE, sum[O_NZONES], cnt[O_NZONES], mean[O_NZONES];
    for (int c = 0; c < mod->shape.ncomp; c++) {
      struct vertices_t *r = &mod->shape.comp[c].real;
      for (int v = 0; v < r->nv; v++) {
        lo = O_MIN(lo, r->v[v].x[jmax]);
        hi = O_MAX(hi, r->v[v].x[jmax]);
      }
    }
    double incr = (hi - lo + O_SMALLVAL) / O_NBIF;
    for (int k = 0; k < O_NZONES; k++) sum[k] = cnt[k] = 0.0;
    for (int c = 0; c < mod->shape.ncomp; c++) {
      struct vertices_t *r = &mod->shape.comp[c].real;
      for (int v = 0; v < r->nv; v++)
        if (r->v[v].act) {
          double rho2 = r->v[v].x[j1] * r->v[v].x[j1] + r->v[v].x[j2] * r->v[v].x[j2];
          double x = (r->v[v].x[jmax] - lo) / incr, w1, w2;
          int k = (int)floor(x), k1, k2;
          double y = x - k;
          if (k == 0) { k1 = k2 = 0; w1 = 0.0; w2 = 1.0; }
          else if (k == O_NZONES) { k1 = k2 = O_NZONES - 1; w1 = 1.0; w2 = 0.0; }
          else { k1 = k - 1; k2 = k; w1 = 1 - pow(y, 6); w2 = 1 - pow(1 - y, 6); }
          sum[k1] += w1 * rho2; sum[k2] += w2 * rho2;
          cnt[k1] += w1; cnt[k2] += w2;
        }
    }
    for (int k = 0; k < O_NZONES; k++) mean[k] = sum[k] / cnt[k];
    double deepest = O_HUGE;
    for (int k = 1; k < O_NZONES - 1; k++)
      for (int k1 = 0; k1 < k; k1++)
        if (mean[k1] > mean[k])
          for (int k2 = k + 1; k2 < O_NZONES; k2++)
            if (mean[k2] > mean[k]) deepest = O_MIN(deepest, mean[k] / ((mean[k1] + mean[k2]) / 2));
    return deepest < 1.0 ? (1 / deepest - 1) * (1 / deepest - 1) : 0.0;
  }
  case NONPA:
    pen = O_MAX((mod->spin.inertia[0].val / mod->spin.inertia[2].val) - 1,
                (mod->spin.inertia[1].val / mod->spin.inertia[2].val) - 1) + 0.01;
    return pen < 0.0 ? 0.0 : pen;
  case EULEROFFS:
    for (int k = 0; k < dat->nsets; k++)
      for (int j = 0; j <= 2; j++) {
        pen += dat->set[k].angleoff[j].val * dat->set[k].angleoff[j].val;
        pen += dat->set[k].omegaoff[j].val * dat->set[k].omegaoff[j].val;
        ntot += 2;
      }
    return pen * (O_R2D * O_R2D / ntot);
  case IMPULSE:
    for (int n = 0; n < mod->spin.n_impulse; n++)
      for (int j = 0; j <= 2; j++) pen += mod->spin.impulse[n][j].val * mod->spin.impulse[n][j].val;
    return pen * (O_R2D * O_R2D / (3 * mod->spin.n_impulse));
  default:
    oracle_set_error("oracle_penalties: penalty type not restated");
    return 0.0;
  }
}

double oracle_penalties(struct par_t *par, struct mod_t *mod, struct dat_t *dat)
{
  double sum = 0.0, pm[3] = {0.0, 0.0, 0.0}, ap[3][3] = {{0.0}};
  int got_pa = 0;
  for (int i = 1; i <= par->pen.n; i++) {
    const int t = par->pen.type[i];
    if (!got_pa && (t == MAXELLIPDEV || t == INERTIADEV_UNI || t == PA3TILT ||
                    t == NONPA_UNI || t == FLATTENING || t == BIFURCATION)) {
      oracle_diag_inertia(mod->shape.inertia, pm, ap);
      got_pa = 1;
    }
    double pen = o_penalty(t, par, mod, dat, pm, ap);
    par->pen.base[i] = par->pen.weight[i] >= 0.0 ? pen : 1.0 / O_MAX(pen, O_TINYPEN);
    sum += fabs(par->pen.weight[i]) * par->pen.base[i];
  }
  return sum;
}
