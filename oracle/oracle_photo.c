/*
 * oracle_photo.c -- optical brightness and the natural cubic spline.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

#define O_TINY 1.0e-40

/* Hapke's bidirectional reflectance of a rough particulate surface
   (astro/hapke.c:41-176): Hapke 1984 (Icarus 59, 41) with the Henyey-Greenstein
   phase function, the opposition surge B0 / (1 + tan(phase/2) / h), the
   macroscopic-roughness correction of the mean slope theta (shadowing S and
   the effective cosines mu0', mu'), and Hapke 2002's (Icarus 157, 523)
   approximation to Chandrasekhar's H function.  Out-of-range parameters give
   0, as the reference does. */
double oracle_hapke(double cosi, double cose, double phase, double w, double h, double B0, double g,
                    double theta)
{
  const double pi = 4 * atan(1.0);
  if (w <= 0.0 || w >= 1.0 || h <= 0.0 || B0 < 0.0 || fabs(g) >= 1.0 || theta < 0.0 || theta >= pi / 2 ||
      phase < 0.0 || phase > pi)
    return 0.0;
  const double mu0 = cosi, mu = cose;
  const double cosphase = (g != 0.0 || theta != 0.0) ? cos(phase) : -999.99;
  const double s1w = sqrt(1 - w), r0 = (1 - s1w) / (1 + s1w);
  const double P = (g == 0.0) ? 1.0 : (1 - g * g) / pow(1 + 2 * g * cosphase + g * g, 1.5);
  const double B = (B0 == 0.0) ? 0.0 : B0 / (1 + tan(phase / 2.0) / h);
  double mu0p, mup, S;
  if (theta == 0.0) {
    mu0p = mu0;
    mup = mu;
    S = 1.0;
  } else {
    const double i = acos(cosi), sini = sin(i), e = acos(cose), sine = sin(e);
    double cospsi = 1.0;
    if (sini > 0.0 && sine > 0.0) {
      cospsi = (cosphase - cosi * cose) / (sini * sine);
      if (fabs(cospsi) > 1.0) cospsi /= fabs(cospsi);
    }
    const double psi = acos(cospsi), hs = sin(psi / 2.0), s2p2 = hs * hs;
    const double tt = tan(theta), ct = 1 / tt;
    const double f = (psi < pi) ? exp(-2.0 * tan(psi / 2.0)) : 0.0;
    const double C = 1 / sqrt(1 + pi * tt * tt);
    double E1e, E2e, E1i, E2i;
    if (e >= pi / 2) E1e = E2e = 1.0;
    else if (e > 0.0) { const double a = ct / tan(e); E1e = exp(-(2 / pi) * a); E2e = exp(-(1 / pi) * a * a); }
    else E1e = E2e = 0.0;
    if (i >= pi / 2) E1i = E2i = 1.0;
    else if (i > 0.0) { const double a = ct / tan(i); E1i = exp(-(2 / pi) * a); E2i = exp(-(1 / pi) * a * a); }
    else E1i = E2i = 0.0;
    const double mu0p0 = C * (mu0 + sini * tt * E2i / (2 - E1i));
    const double mup0 = C * (mu + sine * tt * E2e / (2 - E1e));
    if (i < e) {
      mu0p = C * (mu0 + sini * tt * (cospsi * E2e + s2p2 * E2i) / (2 - E1e - (psi / pi) * E1i));
      mup = C * (mu + sine * tt * (E2e - s2p2 * E2i) / (2 - E1e - (psi / pi) * E1i));
      S = ((mup * mu0) / (mup0 * mu0p0)) * C / (1 - f * (1 - C * mu0 / mu0p0));
    } else {
      mu0p = C * (mu0 + sini * tt * (E2i - s2p2 * E2e) / (2 - E1i - (psi / pi) * E1e));
      mup = C * (mu + sine * tt * (cospsi * E2i + s2p2 * E2e) / (2 - E1i - (psi / pi) * E1e));
      S = ((mup * mu0) / (mup0 * mu0p0)) * C / (1 - f * (1 - C * mu / mup0));
    }
  }
  const double H0 = mu0p > O_TINY ? 1 / (1 - w * mu0p * (r0 + 0.5 * (1 - 2 * r0 * mu0p) * log(1 + 1 / mu0p))) : 1.0;
  const double H1 = mup > O_TINY ? 1 / (1 - w * mup * (r0 + 0.5 * (1 - 2 * r0 * mup) * log(1 + 1 / mup))) : 1.0;
  const double r = ((w * mu0p) / (4 * pi * (mu0p + mup))) * ((1 + B) * P + H0 * H1 - 1.0) * S;
  return r > 0.0 ? r : 0.0;
}

/* apply_photo.c:78-309 */
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
  case HARMKAAS: case INHOKAAS: case HAPKE: case HARMHAPKE: case INHOHAPKE:
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
      case HAPKE: {
        const struct hapke_t *l = &o->hapke;
        pos->b[i][j] = intensityfactor * oracle_hapke(ci, ce, phase, l->w.val, l->h.val, l->B0.val, l->g.val,
                                                      l->theta.val);
        break;
      }
      case HARMHAPKE: case INHOHAPKE: {
        if (f < 0) continue;
        const struct hapke_t *l = type == HARMHAPKE ? &o->harmhapke.local[c][f] : &o->inhohapke.local[c][f];
        pos->b[i][j] = intensityfactor * oracle_hapke(ci, ce, phase, l->w.val, l->h.val, l->B0.val, l->g.val,
                                                      l->theta.val);
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

/* ---- realize_photo (realize_photo.c:122-969) with albedo_mode 0 -------- */

/* checkphotopar (realize_photo.c:965-1004): mode bit 0 makes the lower bound
   exclusive (modes 1, 3), mode bit 1 the upper (modes 2, 3) */
static void o_checkphotopar(struct par_t *par, double v, double lo, double hi, int mode)
{
  if ((mode == 0 || mode == 2) ? v < lo : v <= lo) {
    par->badphoto = 1;
    par->badphoto_logfactor += log(1 + lo - v);
  }
  if ((mode == 0 || mode == 1) ? v > hi : v >= hi) {
    par->badphoto = 1;
    par->badphoto_logfactor += log(1 + v - hi);
  }
}

/* associated Legendre function P_l^m(x) with the Condon-Shortley phase
   (the three-term recurrence in l from P_m^m = (-1)^m (2m-1)!! (1-x^2)^(m/2)) */
double oracle_plm(int l, int m, double x)
{
  double pmm = 1.0;
  if (m > 0) {
    const double s = sqrt((1.0 - x) * (1.0 + x));
    double fact = 1.0;
    for (int i = 1; i <= m; i++) {
      pmm *= -fact * s;
      fact += 2.0;
    }
  }
  if (l == m) return pmm;
  double pm1 = x * (2 * m + 1) * pmm;
  if (l == m + 1) return pm1;
  double pl = 0.0;
  for (int ll = m + 2; ll <= l; ll++) {
    pl = (x * (2 * ll - 1) * pm1 - (ll + m - 1) * pmm) / (ll - m);
    pmm = pm1;
    pm1 = pl;
  }
  return pl;
}

/* sqrt((2l+1) (l-m)! / (l+m)!), the factorial ratio as a product */
static double o_nlm(int l, int m)
{
  double r = 1.0;
  for (int k = l - m + 1; k <= l + m; k++) r /= k;
  return sqrt((2 * l + 1) * r);
}

/* an expansion's value at a facet: sum of a_lm cos(m phi) P~_lm + b_lm sin(m phi) P~_lm */
static double o_harm_value(const struct harmonic_t *h, double costheta, double phi)
{
  double v = 0.0;
  for (int l = 0; l <= h->nhar; l++) {
    v += h->a[l][0].val * (o_nlm(l, 0) * oracle_plm(l, 0, costheta));
    for (int m = 1; m <= l; m++) {
      const double plm = o_nlm(l, m) * oracle_plm(l, m, costheta);
      v += h->a[l][m].val * (cos(m * phi) * plm) + h->b[l][m].val * (sin(m * phi) * plm);
    }
  }
  return v;
}

/* Per-facet law parameters as an array: the local structs (R_t, RC_t, hapke_t,
   kaas_t) and the expansion structs (harmR_t ...) hold their k parameters as
   consecutive members.  bounds: {min, max, mode} per parameter. */
struct o_lawdesc {
  int k;
  double lo[5], hi[5];
  int mode[5];
};

static void o_desc(struct par_t *par, int radar, int type, struct o_lawdesc *d)
{
  memset(d, 0, sizeof(*d));
  if (radar) {
    d->k = 2;
    d->lo[0] = par->rad_R_min; d->hi[0] = par->rad_R_max; d->mode[0] = 3;
    d->lo[1] = par->rad_C_min; d->hi[1] = par->rad_C_max; d->mode[1] = 3;
    return;
  }
  switch (type) {
  case HAPKE: case HARMHAPKE: case INHOHAPKE: {
    const double lo[5] = {par->opt_w_min, par->opt_h_min, par->opt_B0_min, par->opt_g_min, par->opt_theta_min};
    const double hi[5] = {par->opt_w_max, par->opt_h_max, par->opt_B0_max, par->opt_g_max, par->opt_theta_max};
    const int md[5] = {3, 3, 2, 3, 2};
    d->k = 5;
    for (int j = 0; j < 5; j++) { d->lo[j] = lo[j]; d->hi[j] = hi[j]; d->mode[j] = md[j]; }
    break;
  }
  case KAASALAINEN: case HARMKAAS: case INHOKAAS: {
    const double lo[5] = {par->opt_R_min, par->opt_wt_min, par->opt_A0_min, par->opt_D_min, par->opt_k_min};
    const double hi[5] = {par->opt_R_max, par->opt_wt_max, par->opt_A0_max, par->opt_D_max, par->opt_k_max};
    const int md[5] = {3, 0, 2, 3, 1};
    d->k = 5;
    for (int j = 0; j < 5; j++) { d->lo[j] = lo[j]; d->hi[j] = hi[j]; d->mode[j] = md[j]; }
    break;
  }
  default:
    d->k = 1;
    d->lo[0] = par->opt_R_min; d->hi[0] = par->opt_R_max; d->mode[0] = 3;
  }
}

/* one law's local structs: base pointer and size of one facet's struct */
static void o_law_arrays(struct mod_t *mod, int radar, int ilaw, int *kind, struct harmonic_t **harm,
                         struct param_t **global, void ***local, size_t *elem)
{
  *harm = NULL; *global = NULL; *local = NULL; *elem = 0; *kind = 0;
  if (radar) {
    union radscat_t *r = &mod->photo.radar[ilaw];
    switch (mod->photo.radtype[ilaw]) {
    case HARMCOSINE_DIFF: *kind = 1; *harm = &r->harmcosine.R; *local = (void **)r->harmcosine.local; *elem = sizeof(struct RC_t); break;
    case INHOCOSINE_DIFF: *kind = 2; *global = &r->inhocosine.global.R; *local = (void **)r->inhocosine.local; *elem = sizeof(struct RC_t); break;
    }
    return;
  }
  union optscat_t *o = &mod->photo.optical[ilaw];
  switch (mod->photo.opttype[ilaw]) {
  case HARMLAMBERT: case HARMLOMMEL: *kind = 1; *harm = &o->harmR.R; *local = (void **)o->harmR.local; *elem = sizeof(struct R_t); break;
  case INHOLAMBERT: case INHOLOMMEL: *kind = 2; *global = &o->inhoR.global.R; *local = (void **)o->inhoR.local; *elem = sizeof(struct R_t); break;
  case HARMHAPKE: *kind = 1; *harm = &o->harmhapke.w; *local = (void **)o->harmhapke.local; *elem = sizeof(struct hapke_t); break;
  case INHOHAPKE: *kind = 2; *global = &o->inhohapke.global.w; *local = (void **)o->inhohapke.local; *elem = sizeof(struct hapke_t); break;
  case HARMKAAS: *kind = 1; *harm = &o->harmkaas.R; *local = (void **)o->harmkaas.local; *elem = sizeof(struct kaas_t); break;
  case INHOKAAS: *kind = 2; *global = &o->inhokaas.global.R; *local = (void **)o->inhokaas.local; *elem = sizeof(struct kaas_t); break;
  }
}

static void o_realize_law(struct par_t *par, struct mod_t *mod, int radar, int ilaw)
{
  const int type = radar ? mod->photo.radtype[ilaw] : mod->photo.opttype[ilaw];
  struct o_lawdesc d;
  o_desc(par, radar, type, &d);
  if (radar) {
    union radscat_t *r = &mod->photo.radar[ilaw];
    switch (type) {
    case NOLAW: return;
    case COSINELAW_DIFF:
      o_checkphotopar(par, r->RC.R.val, par->rad_R_min, par->rad_R_max, 3);
      o_checkphotopar(par, r->RC.C.val, par->rad_C_min, par->rad_C_max, 3);
      return;
    case TABULARLAW:
      for (int i = 0; i < r->tabular.n; i++) {
        if (r->tabular.rho[i].state == '=') r->tabular.rho[i].val = r->tabular.rho[i - 1].val;
        o_checkphotopar(par, r->tabular.rho[i].val, par->rad_rho_min, par->rad_rho_max, 3);
      }
      return;
    case GAUSSIANLAW: case HAGFORSLAW: case COSINELAW_QS:
      o_checkphotopar(par, r->quasispec.R.val, par->rad_R_min, par->rad_R_max, 3);
      o_checkphotopar(par, r->quasispec.C.val, par->rad_C_min, par->rad_C_max, 3);
      return;
    case GAUSSIAN_COSINE: case HAGFORS_COSINE: case COSINE_COSINE:
      o_checkphotopar(par, r->hybrid.qs.R.val, par->rad_R_min, par->rad_R_max, 3);
      o_checkphotopar(par, r->hybrid.qs.C.val, par->rad_C_min, par->rad_C_max, 3);
      o_checkphotopar(par, r->hybrid.diff.R.val, par->rad_R_min, par->rad_R_max, 3);
      o_checkphotopar(par, r->hybrid.diff.C.val, par->rad_C_min, par->rad_C_max, 3);
      return;
    }
  } else {
    union optscat_t *o = &mod->photo.optical[ilaw];
    switch (type) {
    case NOLAW: return;
    case GEOMETRICAL: case LAMBERTLAW: case LOMMEL:
      o_checkphotopar(par, o->R.R.val, par->opt_R_min, par->opt_R_max, 3);
      return;
    case HAPKE: case KAASALAINEN: {
      struct param_t *p = type == HAPKE ? &o->hapke.w : &o->kaas.R;
      for (int j = 0; j < d.k; j++) o_checkphotopar(par, p[j].val, d.lo[j], d.hi[j], d.mode[j]);
      return;
    }
    }
  }
  int kind;
  struct harmonic_t *harm;
  struct param_t *global;
  void **local;
  size_t elem;
  o_law_arrays(mod, radar, ilaw, &kind, &harm, &global, &local, &elem);
  if (!kind) {
    oracle_set_error("oracle_realize_photo: unknown scattering law");
    return;
  }
  if (kind == 2)
    for (int j = 0; j < d.k; j++) o_checkphotopar(par, global[j].val, d.lo[j], d.hi[j], d.mode[j]);
  for (int c = 0; c < mod->shape.ncomp; c++) {
    struct vertices_t *real = &mod->shape.comp[c].real;
    for (int f = 0; f < real->nf; f++) {
      struct param_t *p = (struct param_t *)((char *)local[c] + elem * f);
      for (int j = 0; j < d.k; j++) {
        if (kind == 1) p[j].val = o_harm_value(&harm[j], cos(real->f[f].theta), real->f[f].phi);
        else if (p[j].state == '=') p[j].val = global[j].val;
      }
      for (int j = 0; j < d.k; j++) o_checkphotopar(par, p[j].val, d.lo[j], d.hi[j], d.mode[j]);
    }
  }
}

/* realize_photo.c:122-960 as the objective calls it (albedo factors 1, mode
   0): '=' states resolved (tabular rho to the previous value, an
   inhomogeneous law's local values to its global ones), harmonic laws'
   local values from their expansions at each facet's angular coordinates
   (set by oracle_realize_mod), every value bounds-checked into
   par->badphoto / badphoto_logfactor; optical laws first, then radar. */
void oracle_realize_photo(struct par_t *par, struct mod_t *mod)
{
  par->badphoto = 0;
  par->badphoto_logfactor = 0.0;
  for (int l = 0; l < mod->photo.noptlaws; l++) o_realize_law(par, mod, 0, l);
  for (int l = 0; l < mod->photo.nradlaws; l++) o_realize_law(par, mod, 1, l);
}
