/*
 * oracle_realize.c -- model realization, spin state and COM offsets.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdio.h>
#include <string.h>

#include "oracle_internal.h"

#define O_SMALLRATIO 0.01 /* realize_mod.c:137-145 */
#define O_HAIRWIDTH 1.0e-7
#define O_SMALLOVOIDK1 0.01
#define O_SMALLOVOIDK2 1.0e-6
#define O_OVOIDTOL 1.0e-6
#define O_SMALLCOEFF3 1.0e-5
#define O_HALFSECOND 5.8e-6

static char g_err[256];
void oracle_set_error(const char *msg) { snprintf(g_err, sizeof g_err, "%s", msg); }
const char *oracle_last_error(void) { return g_err; }
void oracle_clear_error(void) { g_err[0] = 0; }

/* facet normal and area: facnrm.c:13-25 */
static double o_facnrm(struct vertices_t *vt, int fi)
{
  double a[3], b[3];
  struct facet_t *f = &vt->f[fi];
  for (int i = 0; i < 3; i++) {
    a[i] = vt->v[f->v[1]].x[i] - vt->v[f->v[0]].x[i];
    b[i] = vt->v[f->v[2]].x[i] - vt->v[f->v[1]].x[i];
  }
  double area = 0.5 * o_cross(f->n, a, b);
  o_normalize(f->n);
  return area;
}

/* scale factors of a VERTEX or HARMONIC component: '=' ties to the previous
   factor, and factors at or below SMALLRATIO made positive with the baddiam
   flag and log factor (realize_mod.c:425-441, :455-466) */
static void o_scale_factors(struct par_t *par, struct param_t sf[3], struct vertices_t *real)
{
  for (int j = 0; j <= 2; j++) {
    if (j > 0 && sf[j].state == '=') sf[j].val = sf[j - 1].val;
    real->scalefactor[j].val = sf[j].val;
    if (real->scalefactor[j].val <= O_SMALLRATIO) {
      par->baddiam = 1;
      par->baddiam_logfactor += log(1 + O_SMALLRATIO - real->scalefactor[j].val);
      real->scalefactor[j].val = O_SMALLRATIO / (1 + O_SMALLRATIO - real->scalefactor[j].val);
    }
  }
}

/* ELLIPSE (realize_mod.c:217-284): inverse squared radii from the diameter and
   the two axis ratios, each kept positive (HAIRWIDTH / SMALLRATIO) with the
   baddiam flag; each vertex's distance along its direction cosines */
static void o_realize_ellipse(struct par_t *par, struct comp_t *comp)
{
  struct vertices_t *real = &comp->real;
  double a[3], diam, ratio;
  diam = comp->desc.ell.two_a.val;
  if (diam > O_HAIRWIDTH) {
    a[0] = 2.0 / diam;
  } else {
    a[0] = (2.0 / O_HAIRWIDTH) * (1 + O_HAIRWIDTH - diam);
    par->baddiam = 1;
    par->baddiam_logfactor += log(1 + O_HAIRWIDTH - diam);
  }
  diam = (2.0 / a[0]);
  ratio = comp->desc.ell.a_over_b.val;
  if (ratio > O_SMALLRATIO) {
    a[1] = 2.0 / (diam / ratio);
  } else {
    a[1] = (2.0 / (diam / O_SMALLRATIO)) / (1 + O_SMALLRATIO - ratio);
    par->baddiam = 1;
    par->baddiam_logfactor += log(1 + O_SMALLRATIO - ratio);
  }
  diam = (2.0 / a[1]);
  ratio = comp->desc.ell.b_over_c.val;
  if (ratio > O_SMALLRATIO) {
    a[2] = 2.0 / (diam / ratio);
  } else {
    a[2] = (2.0 / (diam / O_SMALLRATIO)) / (1 + O_SMALLRATIO - ratio);
    par->baddiam = 1;
    par->baddiam_logfactor += log(1 + O_SMALLRATIO - ratio);
  }
  for (int i = 0; i <= 2; i++) a[i] *= a[i];
  for (int i = 0; i < real->nv; i++) {
    double den = 0.0;
    for (int j = 0; j <= 2; j++) den += a[j] * (real->v[i].u[j] * real->v[i].u[j]);
    real->v[i].r.val = 1 / sqrt(den);
  }
  for (int j = 0; j <= 2; j++) real->scalefactor[j].val = 1.0;
}

/* Real roots of coeff[3] x^3 + coeff[2] x^2 + coeff[1] x + coeff[0]
   (realize_mod.c:852-941): the quadratic / linear cases below a small cubic
   coefficient, otherwise the trigonometric (three roots) or Cardano (one root)
   solution of Numerical Recipes in C, 2nd ed., section 5.6, with r^2 - q^3
   written out in the coefficients.  Returns the number of roots, or -1 where
   the reference bails out. */
static int o_cubic_roots(const double coeff[4], double root[3])
{
  root[0] = root[1] = root[2] = -HUGENUMBER;
  if (fabs(coeff[3]) < O_SMALLCOEFF3) {
    const double a = coeff[2], b = coeff[1], c = coeff[0];
    if (fabs(a) < SMALLVAL) {
      if (fabs(b) < SMALLVAL) return -1;
      root[0] = -c / b;
      return 1;
    }
    const double disc = b * b - 4 * a * c;
    if (disc < 0.0) return -1;
    if (fabs(b) < SMALLVAL) {
      root[0] = sqrt(disc) / (2 * a);
      root[1] = -root[0];
    } else {
      const int bsign = (b < 0.0) ? -1 : 1;
      const double q = -0.5 * (b + bsign * sqrt(disc));
      root[0] = q / a;
      root[1] = c / q;
    }
    return 2;
  }
  const double a = coeff[2] / coeff[3], b = coeff[1] / coeff[3], c = coeff[0] / coeff[3];
  const double q = (a * a - 3 * b) / 9;
  const double r = (2 * a * a * a - 9 * a * b + 27 * c) / 54;
  const double d = (4 * a * a * a * c - a * a * b * b - 18 * a * b * c + 27 * c * c + 4 * b * b * b) / 108;
  if (d >= 0.0) {
    const double rsign = (r < 0.0) ? -1 : 1;
    const double s = -rsign * pow(fabs(r) + sqrt(d), 1.0 / 3);
    const double t = (fabs(s) >= SMALLVAL) ? q / s : 0.0;
    root[0] = s + t - a / 3;
    return 1;
  }
  const double qs = sqrt(q), theta = acos(r / (q * qs));
  root[0] = -2 * qs * cos(theta / 3) - a / 3;
  root[1] = -2 * qs * cos((theta + 2 * PIE) / 3) - a / 3;
  root[2] = -2 * qs * cos((theta - 2 * PIE) / 3) - a / 3;
  return 3;
}

/* OVOID (realize_mod.c:285-394): parameters kept in bounds with the baddiam
   flag; x0 puts the centre of mass at the origin (series for small |k|);
   each vertex's r / a is the largest admissible non-negative root of the
   ovoid's cubic along its direction, or for tiny |k| the ellipsoid's value
   with one Newton step.  Returns -1 where the reference bails out. */
static int o_realize_ovoid(struct par_t *par, struct comp_t *comp)
{
  struct vertices_t *real = &comp->real;
  double ar = comp->desc.ovoid.two_a.val / 2;
  if (ar <= O_HAIRWIDTH / 2) {
    par->baddiam = 1;
    par->baddiam_logfactor += log(1 + O_HAIRWIDTH - 2 * ar);
    ar = (O_HAIRWIDTH / 2) / (1 + O_HAIRWIDTH - 2 * ar);
  }
  double ab = comp->desc.ovoid.a_over_b.val;
  if (ab <= O_SMALLRATIO) {
    par->baddiam = 1;
    par->baddiam_logfactor += log(1 + O_SMALLRATIO - ab);
    ab = O_SMALLRATIO / (1 + O_SMALLRATIO - ab);
  }
  double bc = comp->desc.ovoid.b_over_c.val;
  if (bc <= O_SMALLRATIO) {
    par->baddiam = 1;
    par->baddiam_logfactor += log(1 + O_SMALLRATIO - bc);
    bc = O_SMALLRATIO / (1 + O_SMALLRATIO - bc);
  }
  double k = comp->desc.ovoid.k.val;
  if (fabs(k) > 1 - SMALLVAL) {
    par->baddiam = 1;
    par->baddiam_logfactor += log(fabs(k) + SMALLVAL);
    k = k > 0.0 ? 1 - SMALLVAL * (1 - SMALLVAL) / k : -1 - SMALLVAL * (1 - SMALLVAL) / k;
  }
  double x0;
  if (fabs(k) > O_SMALLOVOIDK1) {
    const double x0term = 3 * (1 - k * k) * log((1 + k) / (1 - k));
    const double numer = 2.0 * k * (3 - 2 * k * k) - x0term;
    const double denom = 2.0 * k * (3 - k * k) - x0term;
    x0 = (ar / k) * (numer / denom);
  } else {
    x0 = 0.4 * k * ar;
  }
  for (int i = 0; i < real->nv; i++) {
    const double ac = ab * bc;
    const double h = ab * ab * real->v[i].u[1] * real->v[i].u[1] + ac * ac * real->v[i].u[2] * real->v[i].u[2];
    const double alpha0 = x0 / ar, ux = real->v[i].u[0];
    double coeff[4], root[3], good;
    coeff[3] = (h - ux * ux) * k * ux;
    coeff[2] = (1 + 3 * k * alpha0) * ux * ux + h * (1 - k * alpha0);
    coeff[1] = (k - (2 + 3 * k * alpha0) * alpha0) * ux;
    coeff[0] = -(1 - alpha0 * alpha0) * (1 + k * alpha0);
    if (fabs(k) <= O_SMALLOVOIDK2) {
      good = 1 / sqrt(ux * ux + h);
      good -= (coeff[3] * good * good * good + coeff[1] * good) /
              (3 * coeff[3] * good * good + 2 * coeff[2] * good + coeff[1]);
    } else {
      const int nr = o_cubic_roots(coeff, root);
      if (nr < 0) {
        oracle_set_error("oracle_realize_mod: ovoid cubic has no admissible form");
        return -1;
      }
      good = -HUGENUMBER;
      for (int q = 0; q < nr; q++)
        if (root[q] >= 0.0 && fabs(root[q] * ux - alpha0) - 1 < O_OVOIDTOL) good = O_MAX(good, root[q]);
    }
    if (good < 0.0) {
      oracle_set_error("oracle_realize_mod: no ovoid vertex displacement");
      return -1;
    }
    real->v[i].r.val = good * ar;
  }
  for (int j = 0; j <= 2; j++) real->scalefactor[j].val = 1.0;
  return 0;
}

/* HARMONIC (realize_mod.c:395-442): each vertex's distance is the
   coefficients times the vertex's spherical-harmonic factors (afactor,
   bfactor, set up with the vertex directions), kept above HAIRWIDTH / 2 with
   the baddiam flag; then the scale factors */
static void o_realize_harmonic(struct par_t *par, struct comp_t *comp)
{
  struct vertices_t *real = &comp->real;
  const int L = comp->desc.har.nhar;
  for (int i = 0; i < real->nv; i++) {
    double r = 0.0;
    for (int l = 0; l <= L; l++) {
      r += comp->desc.har.a[l][0].val * real->v[i].afactor[l][0];
      for (int m = 1; m <= l; m++)
        r += comp->desc.har.a[l][m].val * real->v[i].afactor[l][m] +
             comp->desc.har.b[l][m].val * real->v[i].bfactor[l][m];
    }
    if (r > O_HAIRWIDTH / 2) {
      real->v[i].r.val = r;
    } else {
      par->baddiam = 1;
      par->baddiam_logfactor += log(1 + O_HAIRWIDTH - 2 * r) / ((L + 1) * (L + 1));
      real->v[i].r.val = (O_HAIRWIDTH / 2) / (1 + O_HAIRWIDTH - 2 * r);
    }
  }
  o_scale_factors(par, comp->desc.har.scalefactor, real);
}

/* realize_mod.c:153-177 -> realize_coordinates :183, the ELLIPSE / OVOID /
   HARMONIC branches :217-442 (above), VERTEX branch :444-467,
   coordinates :475-479, component rotation/offset :484-512, facets :532-546,
   vertex normals :552-563, check_surface for one component (:580-593: every
   vertex, facet and side active), compute_moments :757 (oracle_pen.c). */
void oracle_realize_mod(struct par_t *par, struct mod_t *mod)
{
  par->baddiam = 0;
  par->baddiam_logfactor = 0.0;
  int harmonic_law = 0;
  for (int l = 0; l < mod->photo.noptlaws; l++) {
    const int t = mod->photo.opttype[l];
    if (t == HARMLAMBERT || t == HARMLOMMEL || t == HARMHAPKE || t == HARMKAAS) harmonic_law = 1;
  }
  for (int l = 0; l < mod->photo.nradlaws; l++)
    if (mod->photo.radtype[l] == HARMCOSINE_DIFF) harmonic_law = 1;
  for (int c = 0; c < mod->shape.ncomp; c++) {
    struct comp_t *comp = &mod->shape.comp[c];
    struct vertices_t *real = &comp->real;
    int nv = real->nv, nf = real->nf;
    if (comp->type == ELLIPSE) {
      o_realize_ellipse(par, comp);
    } else if (comp->type == OVOID) {
      if (o_realize_ovoid(par, comp)) return;
    } else if (comp->type == HARMONIC) {
      o_realize_harmonic(par, comp);
    } else if (comp->type == VERTEX) {
      for (int i = 0; i < nv; i++)
        if (real->v[i].r.state == '=') real->v[i].r.val = real->v[real->v[i].v_mirror].r.val;
      o_scale_factors(par, comp->desc.ver.scalefactor, real);
    } else {
      oracle_set_error("oracle_realize_mod: unknown component type");
      return;
    }
    for (int i = 0; i < nv; i++)
      for (int j = 0; j <= 2; j++)
        real->v[i].x[j] = real->scalefactor[j].val * (real->v[i].u[j] * real->v[i].r.val + real->v[i].a[j]);

    o_euler2mat(comp->m, comp->rot[0].val, comp->rot[1].val, comp->rot[2].val);
    if (!(comp->rot[0].val == 0 && comp->rot[1].val == 0 && comp->rot[2].val == 0)) {
      for (int i = 0; i < nv; i++) {
        double x[3];
        for (int j = 0; j <= 2; j++) {
          x[j] = 0.0;
          for (int k = 0; k <= 2; k++) x[j] += comp->m[j][k] * real->v[i].x[k];
        }
        for (int j = 0; j <= 2; j++) real->v[i].x[j] = x[j];
      }
    }
    if (!(comp->off[0].val == 0.0 && comp->off[1].val == 0.0 && comp->off[2].val == 0.0)) {
      for (int i = 0; i < nv; i++)
        for (int j = 0; j <= 2; j++) real->v[i].x[j] += comp->off[j].val;
    }
    for (int f = 0; f < nf; f++) {
      real->f[f].area = o_facnrm(real, f);
      for (int j = 0; j <= 2; j++)
        real->f[f].x[j] = (real->v[real->f[f].v[0]].x[j] + real->v[real->f[f].v[1]].x[j] +
                           real->v[real->f[f].v[2]].x[j]) / 3;
      if (harmonic_law) { /* realize_mod.c:517-545 */
        const double *x = real->f[f].x;
        real->f[f].theta = atan2(sqrt(x[0] * x[0] + x[1] * x[1]), x[2]);
        real->f[f].phi = atan2(x[1], x[0]);
      }
    }
    for (int i = 0; i < nv; i++) {
      double n[3] = {0.0, 0.0, 0.0};
      for (int j = 0; j < real->v[i].naf; j++) {
        int f = real->v[i].af[j];
        for (int k = 0; k <= 2; k++) n[k] += real->f[f].n[k];
      }
      o_normalize(n);
      for (int k = 0; k <= 2; k++) real->v[i].n[k] = n[k];
    }
    for (int i = 0; i < nv; i++) real->v[i].act = 1;
    for (int f = 0; f < nf; f++) real->f[f].act = 1;
    for (int k = 0; k < real->ns; k++) real->s[k].act = 1;
  }
  oracle_compute_moments(mod);
}

/* realize_spin.c:350-398: epochs and impulses met between t0 and t */
static void o_realize_impulse(const struct spin_t *spin, double t, double t_int[], double imp[][3],
                              int *n_int)
{
  int k = 0;
  t_int[0] = spin->t0;
  imp[0][0] = imp[0][1] = imp[0][2] = 0.0;
  if (t >= spin->t0) {
    for (int n = 0; n < spin->n_impulse; n++)
      if (spin->t_impulse[n] > spin->t0 && spin->t_impulse[n] <= t) {
        k++;
        t_int[k] = spin->t_impulse[n];
        for (int j = 0; j <= 2; j++) imp[k][j] = spin->impulse[n][j].val;
      }
    if (t_int[k] < t) {
      k++;
      t_int[k] = t;
      imp[k][0] = imp[k][1] = imp[k][2] = 0.0;
    }
  } else {
    for (int n = spin->n_impulse - 1; n >= 0; n--)
      if (spin->t_impulse[n] < spin->t0 && spin->t_impulse[n] >= t) {
        k++;
        t_int[k] = spin->t_impulse[n];
        for (int j = 0; j <= 2; j++) imp[k][j] = -spin->impulse[n][j].val;
      }
    if (t_int[k] > t) {
      k++;
      t_int[k] = t;
      imp[k][0] = imp[k][1] = imp[k][2] = 0.0;
    }
  }
  *n_int = k + 1;
}

/* inteuler.c:93-113: principal-axis rotator kinematics; NPA rotators by
   Euler's equations (oracle_npa.c) */
static int o_inteuler(const struct spin_t *spin, const double t[], double imp[][3], int n, int method,
                      double w[3], double m[3][3])
{
  if (!spin->pa) {
    if (!oracle_inteuler_npa(spin, t, imp, n, method, w, m)) {
      oracle_set_error("oracle_realize_spin: the NPA spin state has no finite period");
      return 0;
    }
    return 1;
  }
  double dt = t[n - 1] - t[0];
  double psi = spin->angle[2].val + spin->omega[2].val * dt + 0.5 * spin->omegadot[2].val * dt * dt +
               spin->lib_amp.val * sin(spin->lib_freq.val * dt + spin->lib_phase.val);
  for (int i = 0; i <= 1; i++) w[i] = spin->omega[i].val + spin->omegadot[i].val * dt;
  w[2] = spin->omega[2].val + spin->omegadot[2].val * dt +
         spin->lib_amp.val * spin->lib_freq.val * cos(spin->lib_freq.val * dt + spin->lib_phase.val);
  for (int k = 0; k < n; k++) {
    psi += imp[k][2] * (t[n - 1] - t[k]);
    w[2] += imp[k][2];
  }
  o_euler2mat(m, spin->angle[0].val, spin->angle[1].val, psi);
  return 1;
}

/* one view: impulses, kinematics, spin offsets, body->ecliptic, add orbital spin
   (realize_spin.c:87-128) */
static void o_spin_view(struct par_t *par, struct mod_t *mod, struct set_t *set, double t, double t_int[],
                        double imp[][3], int *n_int, double intspin[3], double ae[3][3],
                        const double orbspin[3], double spin[3])
{
  o_realize_impulse(&mod->spin, t, t_int, imp, n_int);
  o_inteuler(&mod->spin, t_int, imp, *n_int, par->int_method, intspin, ae);
  for (int j = 0; j <= 2; j++) intspin[j] += set->omegaoff[j].val;
  o_cotrans(intspin, ae, intspin, -1);
  for (int j = 0; j <= 2; j++) spin[j] = orbspin[j] + intspin[j];
}

/* realize_spin.c:41-280 (angle/omega offsets realized with '=' states, :283-342) */
void oracle_realize_spin(struct par_t *par, struct mod_t *mod, struct dat_t *dat)
{
  for (int j = 0; j <= 2; j++) {
    int sa = -1, so = -1;
    for (int s = 0; s < dat->nsets; s++) {
      if (dat->set[s].angleoff[j].state != '=') sa = s;
      else if (sa >= 0) dat->set[s].angleoff[j].val = dat->set[sa].angleoff[j].val;
      if (dat->set[s].omegaoff[j].state != '=') so = s;
      else if (so >= 0) dat->set[s].omegaoff[j].val = dat->set[so].omegaoff[j].val;
    }
  }
  for (int s = 0; s < dat->nsets; s++) {
    struct set_t *set = &dat->set[s];
    double anglesave[3];
    for (int j = 0; j <= 2; j++) {
      anglesave[j] = mod->spin.angle[j].val;
      mod->spin.angle[j].val += set->angleoff[j].val;
    }
    if (set->type == DELAY) {
      struct deldop_t *dd = &set->desc.deldop;
      for (int f = 0; f < dd->nframes; f++)
        for (int k = 0; k < dd->nviews; k++) {
          struct deldopfrm_t *fr = &dd->frame[f];
          struct deldopview_t *v = &fr->view[k];
          o_spin_view(par, mod, set, v->t, fr->t_integrate, fr->impulse, &fr->n_integrate, v->intspin,
                      v->ae, v->orbspin, v->spin);
        }
    } else if (set->type == DOPPLER) {
      struct doppler_t *dp = &set->desc.doppler;
      for (int f = 0; f < dp->nframes; f++)
        for (int k = 0; k < dp->nviews; k++) {
          struct dopfrm_t *fr = &dp->frame[f];
          struct dopview_t *v = &fr->view[k];
          o_spin_view(par, mod, set, v->t, fr->t_integrate, fr->impulse, &fr->n_integrate, v->intspin,
                      v->ae, v->orbspin, v->spin);
        }
    } else if (set->type == LGHTCRV) {
      struct lghtcrv_t *lc = &set->desc.lghtcrv;
      for (int i = 1; i <= lc->ncalc; i++) {
        struct crvrend_t *r = &lc->rend[i];
        o_spin_view(par, mod, set, lc->x[i], r->t_integrate, r->impulse, &r->n_integrate, r->intspin,
                    r->ae, r->orbspin, r->spin);
      }
    } else {
      oracle_set_error("oracle_realize_spin: plane-of-sky sets are not restated");
    }
    for (int j = 0; j <= 2; j++) mod->spin.angle[j].val = anglesave[j];
  }
}

/* deldopoffs.c:17-40 */
static void o_deldopoffs(struct deldop_t *dd)
{
  for (int f = 0; f < dd->nframes; f++)
    for (int k = 0; k < dd->nviews; k++) {
      double x = 1.0, dop = 0.0, del = dd->delcor.a[0].val;
      double arg = dd->frame[f].view[k].t - dd->delcor.t0;
      for (int n = 1; n <= dd->delcor.n; n++) {
        dop += n * dd->delcor.a[n].val * x;
        del += dd->delcor.a[n].val * (x *= arg);
      }
      dd->frame[f].view[k].deloff = del / dd->del_per_pixel;
      dd->frame[f].view[k].dopoff = -dop * dd->Ftx / (dd->dop_per_pixel * 86400.0);
    }
}

/* dopoffs.c:19-39 */
static void o_dopoffs(struct doppler_t *dp)
{
  for (int f = 0; f < dp->nframes; f++)
    for (int k = 0; k < dp->nviews; k++) {
      double x = 1.0, dop = 0.0;
      double arg = dp->frame[f].view[k].t - dp->delcor.t0;
      for (int n = 1; n <= dp->delcor.n; n++) {
        dop += n * dp->delcor.a[n].val * x;
        x *= arg;
      }
      dp->frame[f].view[k].dopoff = -dop * dp->Ftx / (dp->dop_per_bin * 86400.0);
    }
}

/* realize_delcor.c:50-151: resolve '=' polynomials, then per-view offsets */
void oracle_realize_delcor(struct dat_t *dat, double delta_delcor0, int delcor0_mode)
{
  int s_dc = -1, type_dc = -1, n_dc = -1;
  double t0_dc = -1.0;
  for (int s = 0; s < dat->nsets; s++) {
    struct set_t *set = &dat->set[s];
    if (set->type != DELAY && set->type != DOPPLER) continue;
    struct delcor_t *dc = set->type == DELAY ? &set->desc.deldop.delcor : &set->desc.doppler.delcor;
    int n = dc->n;
    double t0 = dc->t0;
    if (dc->a[0].state != '=') {
      s_dc = s;
      type_dc = set->type;
      n_dc = n;
      t0_dc = t0;
      if (set->type == DELAY && dc->a[0].state == 'f') {
        if (delcor0_mode != 0) dc->a[0].val = dc->delcor0_save + delta_delcor0;
        if (delcor0_mode != 1) dc->delcor0_save = dc->a[0].val;
      }
    } else if (s_dc < 0 || n != n_dc || fabs(t0 - t0_dc) > O_HALFSECOND) {
      oracle_set_error("oracle_realize_delcor: invalid '=' delay polynomial");
      return;
    } else {
      struct set_t *src = &dat->set[s_dc];
      struct delcor_t *sdc = type_dc == DELAY ? &src->desc.deldop.delcor : &src->desc.doppler.delcor;
      dc->t0 = t0_dc;
      for (int i = 0; i <= n; i++) dc->a[i].val = sdc->a[i].val;
    }
    if (set->type == DELAY) o_deldopoffs(&set->desc.deldop);
    else o_dopoffs(&set->desc.doppler);
  }
}
