/*
 * io_mod.c -- the model file (read_mod.c:204-2596): shape components, radar
 * and optical scattering laws, and the spin state; then the realization
 * meshes of the analytic components and every component's sides and
 * attached-facet lists.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "io_internal.h"

typedef struct {
  io_ctx *cx;
  io_file *f;
  struct par_t *par;
  struct mod_t *mod;
  int nfpar;
} modrd;

/* readparam, counting free parameters; end of file is an error here */
static void rp(modrd *r, struct param_t *p, const char *what)
{
  const int k = io_param(r->cx, r->f, p);
  if (k < 0) io_fail(r->cx, "read_mod: %s: end of file where %s was expected", r->f->name, what);
  r->nfpar += k;
}

static int fitting(modrd *r, unsigned char vary) { return r->par->action == FIT && vary; }

/* ---- shape (read_mod.c:289-542) ---------------------------------------- */
static int theta_steps(modrd *r, int ntheta)
{
  /* theta_steps in the parameter file overrides the model's (read_mod.c:334-341) */
  return (r->par->theta_steps <= 0 || r->par->theta_steps == ntheta) ? ntheta : r->par->theta_steps;
}

/* a spherical-harmonic expansion: degree L, then a[l][m] (and b[l][m], m > 0)
   for l = 0..L, m = 0..l; a_save / b_save kept for the albedo-like
   expansions; conv scales each value after it is read */
static void read_expansion(modrd *r, struct harmonic_t *h, int keep_save, double conv, int inv,
                           unsigned char vary, const char *what)
{
  io_ctx *cx = r->cx;
  h->nhar = io_int(cx, r->f);
  const int L = h->nhar;
  if (L < 0) io_fail(cx, "read_mod: negative harmonic degree for %s", what);
  h->a = (struct param_t **)io_calloc(cx, (size_t)L + 1, sizeof(struct param_t *));
  h->b = (struct param_t **)io_calloc(cx, (size_t)L + 1, sizeof(struct param_t *));
  if (keep_save) {
    h->a_save = (double **)io_calloc(cx, (size_t)L + 1, sizeof(double *));
    h->b_save = (double **)io_calloc(cx, (size_t)L + 1, sizeof(double *));
  }
  for (int l = 0; l <= L; l++) {
    h->a[l] = (struct param_t *)io_calloc(cx, (size_t)l + 1, sizeof(struct param_t));
    h->b[l] = (struct param_t *)io_calloc(cx, (size_t)l + 1, sizeof(struct param_t));
    if (keep_save) {
      h->a_save[l] = (double *)io_calloc(cx, (size_t)l + 1, sizeof(double));
      h->b_save[l] = (double *)io_calloc(cx, (size_t)l + 1, sizeof(double));
    }
    for (int m = 0; m <= l; m++) {
      for (int ab = 0; ab < (m > 0 ? 2 : 1); ab++) {
        struct param_t *p = ab ? &h->b[l][m] : &h->a[l][m];
        rp(r, p, what);
        if (keep_save && fitting(r, vary) && p->state == 'c' && p->val != 0.0)
          io_fail(cx, "read_mod: must let all nonzero %s coefficients float when varying albedos", what);
        if (conv != 1.0) p->val = inv ? p->val / conv : p->val * conv;
        if (keep_save) (ab ? h->b_save : h->a_save)[l][m] = p->val;
      }
    }
  }
  h->ntheta = -1;
}

static void read_shape(modrd *r)
{
  io_ctx *cx = r->cx;
  io_file *f = r->f;
  struct mod_t *mod = r->mod;
  char type[IO_MAXLEN + 1];
  mod->shape.ncomp = io_int(cx, f);
  if (mod->shape.ncomp < 1) io_fail(cx, "read_mod: a model needs at least one component");
  mod->shape.comp = (struct comp_t *)io_calloc(cx, (size_t)mod->shape.ncomp, sizeof(struct comp_t));
  for (int i = 0; i < mod->shape.ncomp; i++) {
    struct comp_t *c = &mod->shape.comp[i];
    for (int j = 0; j < 3; j++) rp(r, &c->off[j], "a linear offset");
    for (int j = 0; j < 3; j++) {
      rp(r, &c->rot[j], "an angular offset");
      c->rot[j].val += 360.0 * floor((180.0 - c->rot[j].val) / 360.0);
      c->rot[j].val *= D2R;
    }
    io_word(f, type, sizeof type);
    if (!strcmp(type, "ellipse")) {
      c->type = ELLIPSE;
      rp(r, &c->desc.ell.two_a, "2a");
      rp(r, &c->desc.ell.a_over_b, "a/b");
      rp(r, &c->desc.ell.b_over_c, "b/c");
      c->desc.ell.ntheta = theta_steps(r, io_int(cx, f));
    } else if (!strcmp(type, "ovoid")) {
      c->type = OVOID;
      rp(r, &c->desc.ovoid.two_a, "2a");
      rp(r, &c->desc.ovoid.a_over_b, "a/b");
      rp(r, &c->desc.ovoid.b_over_c, "b/c");
      rp(r, &c->desc.ovoid.k, "k");
      c->desc.ovoid.ntheta = theta_steps(r, io_int(cx, f));
    } else if (!strcmp(type, "harmonic")) {
      struct harmonic_t *h = &c->desc.har;
      c->type = HARMONIC;
      h->nhar = io_int(cx, f);
      const int L = h->nhar;
      if (L < 0) io_fail(cx, "read_mod: negative harmonic degree");
      h->a = (struct param_t **)io_calloc(cx, (size_t)L + 1, sizeof(struct param_t *));
      h->b = (struct param_t **)io_calloc(cx, (size_t)L + 1, sizeof(struct param_t *));
      for (int j = 0; j < 3; j++) rp(r, &h->scalefactor[j], "a scale factor");
      if (h->scalefactor[0].state == '=') io_fail(cx, "read_mod: can't use '=' state for 'scale factor 0'");
      for (int l = 0; l <= L; l++) {
        h->a[l] = (struct param_t *)io_calloc(cx, (size_t)l + 1, sizeof(struct param_t));
        h->b[l] = (struct param_t *)io_calloc(cx, (size_t)l + 1, sizeof(struct param_t));
        for (int m = 0; m <= l; m++) {
          rp(r, &h->a[l][m], "a harmonic coefficient");
          if (m > 0) rp(r, &h->b[l][m], "a harmonic coefficient");
        }
      }
      h->ntheta = theta_steps(r, io_int(cx, f));
    } else if (!strcmp(type, "vertex")) {
      struct vertices_t *vt = &c->desc.ver;
      c->type = VERTEX;
      vt->nv = io_int(cx, f);
      if (vt->nv < 4) io_fail(cx, "read_mod: a vertex component needs at least 4 vertices");
      vt->v = (struct vertex_t *)io_calloc(cx, (size_t)vt->nv, sizeof(struct vertex_t));
      for (int j = 0; j < 3; j++) rp(r, &vt->scalefactor[j], "a scale factor");
      if (vt->scalefactor[0].state == '=') io_fail(cx, "read_mod: can't use '=' state for 'scale factor 0'");
      int mirror = 0;
      for (int j = 0; j < vt->nv; j++) {
        struct vertex_t *v = &vt->v[j];
        rp(r, &v->r, "a vertex deviation");
        for (int k = 0; k < 3; k++) v->u[k] = io_double(cx, f);
        for (int k = 0; k < 3; k++) v->a[k] = io_double(cx, f);
        if (v->r.state == '=') mirror = 1;
        else v->v_mirror = -1;
      }
      /* an '=' deviation ties a southern vertex to its northern mirror: the
         vertex with the same base point and direction reflected in z
         (read_mod.c:476-505) */
      if (mirror)
        for (int j = 0; j < vt->nv; j++) {
          struct vertex_t *v = &vt->v[j];
          if (v->r.state != '=') continue;
          int m = 0;
          for (; m < vt->nv; m++) {
            const struct vertex_t *w = &vt->v[m];
            if (m != j && fabs(w->a[0] - v->a[0]) < SMALLVAL && fabs(w->a[1] - v->a[1]) < SMALLVAL &&
                fabs(w->a[2] + v->a[2]) < SMALLVAL && fabs(w->u[0] - v->u[0]) < SMALLVAL &&
                fabs(w->u[1] - v->u[1]) < SMALLVAL && fabs(w->u[2] + v->u[2]) < SMALLVAL)
              break;
          }
          if (m == vt->nv) io_fail(cx, "read_mod: no mirror vertex found for component %d vertex %d", i, j);
          v->v_mirror = m;
        }
      vt->nf = io_int(cx, f);
      if (vt->nf < 4) io_fail(cx, "read_mod: a vertex component needs at least 4 facets");
      vt->f = (struct facet_t *)io_calloc(cx, (size_t)vt->nf, sizeof(struct facet_t));
      for (int j = 0; j < vt->nf; j++)
        for (int k = 0; k < 3; k++) vt->f[j].v[k] = io_int(cx, f);
    } else {
      io_fail(cx, "read_mod: invalid shape description '%s'", type);
    }
  }
}

/* ---- realization meshes of the analytic components --------------------- */
/* N_lm P_lm(x) for 0 <= m <= l <= L into p[l*(L+1)+m], with
   N_lm = sqrt((2l+1) (l-m)!/(l+m)!) and P_lm the associated Legendre
   function with the Condon-Shortley phase: P_mm by its product formula, then
   the upward recurrence in l.  The factorial ratio is formed exactly (the
   reference takes it from exp(gammln), read_mod.c:2202-2204). */
static void legendre_normalized(int L, double x, double *p)
{
  const double s = sqrt((1.0 - x) * (1.0 + x));
  for (int m = 0; m <= L; m++) {
    double pmm = 1.0, fact = 1.0;
    for (int i = 1; i <= m; i++) { pmm *= -fact * s; fact += 2.0; }
    double prev = 0.0, cur = pmm;
    for (int l = m; l <= L; l++) {
      if (l == m + 1) { prev = cur; cur = x * (2 * m + 1) * pmm; }
      else if (l > m + 1) {
        const double nxt = (x * (2 * l - 1) * cur - (l + m - 1) * prev) / (l - m);
        prev = cur;
        cur = nxt;
      }
      double ratio = 1.0;
      for (int i = l - m + 1; i <= l + m; i++) ratio /= i;
      p[l * (L + 1) + m] = sqrt((2 * l + 1) * ratio) * cur;
    }
  }
}

/* Rings of vertices at theta_t = t pi / nt, t = 0..nt, np_t =
   round(2 nt sin theta_t) (at least 1) vertices per ring at phi = 2 pi p / np_t;
   then each band between rings t and t+1 zipped into triangles by walking
   both rings eastward and advancing whichever ring's next vertex comes first
   in longitude (the top ring on ties).  The order of vertices and facets is
   the reference's (setuprealver / setuprealfac, read_mod.c:2179-2384), so the
   realized component matches its own in every index. */
static void realize_mesh(modrd *r, struct comp_t *c, int nt)
{
  io_ctx *cx = r->cx;
  if (nt < 1) io_fail(cx, "read_mod: an analytic component needs theta steps >= 1");
  const double pi = 4 * atan(1.0);
  int *np = (int *)malloc(((size_t)nt + 1) * sizeof(int));
  int *first = (int *)malloc(((size_t)nt + 1) * sizeof(int));
  if (!np || !first) { free(np); free(first); io_fail(cx, "out of host memory"); }
  int nv = 0;
  for (int t = 0; t <= nt; t++) {
    const double theta = (t * pi) / nt;
    const double x = 2 * nt * sin(theta);
    np[t] = (int)(x < 0 ? x - 0.5 : x + 0.5);
    if (np[t] == 0) np[t] = 1;
    first[t] = nv;
    nv += np[t];
  }
  struct vertices_t *rv = &c->real;
  rv->nv = nv;
  rv->v = (struct vertex_t *)io_calloc(cx, (size_t)nv, sizeof(struct vertex_t));
  const int L = c->type == HARMONIC ? c->desc.har.nhar : -1;
  double *plm = L >= 0 ? (double *)malloc(((size_t)L + 1) * (L + 1) * sizeof(double)) : NULL;
  if (L >= 0 && !plm) { free(np); free(first); io_fail(cx, "out of host memory"); }
  int v = 0;
  for (int t = 0; t <= nt; t++) {
    const double theta = (t * pi) / nt;
    for (int p = 0; p < np[t]; p++, v++) {
      const double phi = (p * 2 * pi) / np[t];
      struct vertex_t *x = &rv->v[v];
      x->u[0] = sin(theta) * cos(phi);
      x->u[1] = sin(theta) * sin(phi);
      x->u[2] = cos(theta);
      x->r.state = 'f';
      x->r.val = 0.0;
      if (L >= 0) {
        legendre_normalized(L, x->u[2], plm);
        x->afactor = io_matrix(cx, 0, L, 0, L);
        if (L > 0) x->bfactor = io_matrix(cx, 1, L, 1, L);
        for (int l = 0; l <= L; l++) {
          x->afactor[l][0] = plm[l * (L + 1)];
          for (int m = 1; m <= l; m++) {
            x->afactor[l][m] = cos(m * phi) * plm[l * (L + 1) + m];
            x->bfactor[l][m] = sin(m * phi) * plm[l * (L + 1) + m];
          }
        }
      }
    }
  }
  free(plm);
  /* two passes over the bands: count, then fill */
  int nf = 0;
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      rv->nf = nf;
      rv->f = (struct facet_t *)io_calloc(cx, (size_t)nf, sizeof(struct facet_t));
      nf = 0;
    }
    for (int t = 0; t < nt; t++) {
      const double dtop = (2 * pi) / np[t], dbot = (2 * pi) / np[t + 1];
      int ptop = 0, pbot = 0, wtop = 0, wbot = 0;
      int a = first[t], b = first[t + 1];
      do {
        const int nt_top = ptop + 1, nt_bot = pbot + 1;
        const double phi_top = (nt_top + wtop) * dtop, phi_bot = (nt_bot + wbot) * dbot;
        int nxt;
        if (phi_top <= phi_bot && t != 0) {
          int q = nt_top;
          if (q == np[t]) { q = 0; wtop = np[t]; }
          ptop = q;
          nxt = first[t] + q;
          if (pass) { rv->f[nf].v[0] = a; rv->f[nf].v[1] = b; rv->f[nf].v[2] = nxt; }
          a = nxt;
        } else {
          int q = nt_bot;
          if (q == np[t + 1]) { q = 0; wbot = np[t + 1]; }
          pbot = q;
          nxt = first[t + 1] + q;
          if (pass) { rv->f[nf].v[0] = a; rv->f[nf].v[1] = b; rv->f[nf].v[2] = nxt; }
          b = nxt;
        }
        nf++;
      } while (ptop != 0 || pbot != 0);
    }
  }
  free(np);
  free(first);
}

/* setupreal (read_mod.c:2152-2177): a VERTEX component's realization shares
   its description's arrays; the others get a tessellation */
static void setup_real(modrd *r)
{
  struct mod_t *mod = r->mod;
  for (int i = 0; i < mod->shape.ncomp; i++) {
    struct comp_t *c = &mod->shape.comp[i];
    if (c->type == VERTEX) c->real = c->desc.ver;
    else if (c->type == ELLIPSE) realize_mesh(r, c, c->desc.ell.ntheta);
    else if (c->type == OVOID) realize_mesh(r, c, c->desc.ovoid.ntheta);
    else realize_mesh(r, c, c->desc.har.ntheta);
  }
}

/* ---- photometry (read_mod.c:544-1950) ---------------------------------- */
static double cutoff_cosine(modrd *r)
{
  const double a = io_double(r->cx, r->f);
  if (a <= 0.0 || a > 90.0) io_fail(r->cx, "read_mod: must have 0.0 < cutoff_angle <= 90.0 deg");
  return cos(a * D2R);
}

static void albedo_check(modrd *r, const struct param_t *p, unsigned char vary, int nonzero_only,
                         const char *what)
{
  if (fitting(r, vary) && p->state == 'c' && (!nonzero_only || p->val != 0.0))
    io_fail(r->cx, "read_mod: must let %s float when varying albedos", what);
}

/* per-component, per-facet arrays of an inhomogeneous law: nc must be the
   component count and each nf the realized facet count */
static int inho_count(modrd *r, int c, int want_nc)
{
  io_ctx *cx = r->cx;
  if (c < 0) {
    const int nc = io_int(cx, r->f);
    if (nc != want_nc) io_fail(cx, "read_mod: %d components for an inhomogeneous law != %d model components", nc, want_nc);
    return nc;
  }
  const int nf = io_int(cx, r->f);
  if (nf != r->mod->shape.comp[c].real.nf)
    io_fail(cx, "read_mod: %d facets for inhomogeneous law component %d != %d realized facets", nf, c,
            r->mod->shape.comp[c].real.nf);
  return nf;
}

static void **local_arrays(modrd *r, size_t elem)
{
  const int nc = r->mod->shape.ncomp;
  void **loc = (void **)io_calloc(r->cx, (size_t)nc, sizeof(void *));
  for (int c = 0; c < nc; c++) loc[c] = io_calloc(r->cx, (size_t)r->mod->shape.comp[c].real.nf, elem);
  return loc;
}

static void read_radar_law(modrd *r, int ilaw)
{
  io_ctx *cx = r->cx;
  struct photo_t *ph = &r->mod->photo;
  union radscat_t *L = &ph->radar[ilaw];
  const unsigned char vary = r->par->vary_radalb;
  char w[IO_MAXLEN + 1];
  io_word(r->f, w, sizeof w);
  if (!strcmp(w, "none")) {
    ph->radtype[ilaw] = NOLAW;
  } else if (!strcmp(w, "cosine")) {
    ph->radtype[ilaw] = COSINELAW_DIFF;
    rp(r, &L->RC.R, "R");
    albedo_check(r, &L->RC.R, vary, 0, "R");
    L->RC.R_save = L->RC.R.val;
    rp(r, &L->RC.C, "C");
  } else if (!strcmp(w, "tabular")) {
    ph->radtype[ilaw] = TABULARLAW;
    const int n = L->tabular.n = io_int(cx, r->f);
    if (n < 2) io_fail(cx, "read_mod: a tabular law needs at least two values");
    L->tabular.rho = (struct param_t *)io_calloc(cx, (size_t)n, sizeof(struct param_t));
    L->tabular.rho_save = io_vector(cx, 0, n - 1);
    for (int i = 0; i < n; i++) {
      rp(r, &L->tabular.rho[i], "rho");
      albedo_check(r, &L->tabular.rho[i], vary, 1, "every nonzero rho");
      L->tabular.rho_save[i] = L->tabular.rho[i].val;
    }
    if (L->tabular.rho[0].state == '=')
      io_fail(cx, "read_mod: can't use '=' state for coefficient 0 of the 'tabular' scattering law");
    /* the noncosine penalty's abscissae, log10 cos of the table's angles,
       set up by the first tabular law (read_mod.c:623-640) */
    if (!r->par->pen.x_noncosine) {
      const double step = (PIE / 2) / (n - 1);
      r->par->pen.x_noncosine = io_vector(cx, 0, n - 2);
      r->par->pen.y_noncosine = io_vector(cx, 0, n - 2);
      r->par->pen.xmean_noncosine = 0.0;
      for (int i = 0; i < n - 1; i++) {
        r->par->pen.x_noncosine[i] = log10(cos(i * step));
        r->par->pen.xmean_noncosine += r->par->pen.x_noncosine[i];
      }
      r->par->pen.xmean_noncosine /= (n - 1);
    }
  } else if (!strcmp(w, "gaussian") || !strcmp(w, "hagfors") || !strcmp(w, "cosine_qs")) {
    ph->radtype[ilaw] = !strcmp(w, "gaussian") ? GAUSSIANLAW : !strcmp(w, "hagfors") ? HAGFORSLAW : COSINELAW_QS;
    rp(r, &L->quasispec.R, "R");
    albedo_check(r, &L->quasispec.R, vary, 0, "R");
    L->quasispec.R_save = L->quasispec.R.val;
    rp(r, &L->quasispec.C, "C");
    L->quasispec.cos_cutoff = cutoff_cosine(r);
  } else if (!strcmp(w, "gauss+cosine") || !strcmp(w, "hagfors+cosine") || !strcmp(w, "cosine+cosine")) {
    ph->radtype[ilaw] = !strcmp(w, "gauss+cosine") ? GAUSSIAN_COSINE
                        : !strcmp(w, "hagfors+cosine") ? HAGFORS_COSINE : COSINE_COSINE;
    rp(r, &L->hybrid.qs.R, "R");
    albedo_check(r, &L->hybrid.qs.R, vary, 1, "a nonzero R");
    L->hybrid.qs.R_save = L->hybrid.qs.R.val;
    rp(r, &L->hybrid.qs.C, "C");
    L->hybrid.qs.cos_cutoff = cutoff_cosine(r);
    rp(r, &L->hybrid.diff.R, "R");
    albedo_check(r, &L->hybrid.diff.R, vary, 1, "a nonzero R");
    L->hybrid.diff.R_save = L->hybrid.diff.R.val;
    rp(r, &L->hybrid.diff.C, "C");
  } else if (!strcmp(w, "harmcosine")) {
    ph->radtype[ilaw] = HARMCOSINE_DIFF;
    read_expansion(r, &L->harmcosine.R, 1, 1.0, 0, vary, "R");
    read_expansion(r, &L->harmcosine.C, 0, 1.0, 0, vary, "C");
    L->harmcosine.local = (struct RC_t **)local_arrays(r, sizeof(struct RC_t));
  } else if (!strcmp(w, "inhocosine")) {
    ph->radtype[ilaw] = INHOCOSINE_DIFF;
    struct inhocosine_t *h = &L->inhocosine;
    rp(r, &h->global.R, "R");
    const int globalflag = h->global.R.state == 'c' && h->global.R.val != 0.0;
    h->global.R_save = h->global.R.val;
    rp(r, &h->global.C, "C");
    const int nc = inho_count(r, -1, r->mod->shape.ncomp);
    h->local = (struct RC_t **)io_calloc(cx, (size_t)nc, sizeof(struct RC_t *));
    for (int c = 0; c < nc; c++) {
      const int nf = inho_count(r, c, 0);
      h->local[c] = (struct RC_t *)io_calloc(cx, (size_t)nf, sizeof(struct RC_t));
      for (int f = 0; f < nf; f++) {
        struct RC_t *x = &h->local[c][f];
        rp(r, &x->R, "R");
        if (fitting(r, vary) && (x->R.state == 'c' || (x->R.state == '=' && globalflag)))
          io_fail(cx, "read_mod: must let R float for all facets when varying radar albedos");
        x->R_save = x->R.val;
        rp(r, &x->C, "C");
      }
    }
  } else {
    io_fail(cx, "read_mod: invalid radar scattering law '%s'", w);
  }
}

static void read_hapke(modrd *r, struct hapke_t *h, int keep_save)
{
  rp(r, &h->w, "w");
  if (keep_save) h->w_save = h->w.val;
  rp(r, &h->h, "h");
  rp(r, &h->B0, "B0");
  rp(r, &h->g, "g");
  rp(r, &h->theta, "theta");
  h->theta.val *= D2R;
}

static void read_kaas(modrd *r, struct kaas_t *k)
{
  rp(r, &k->R, "R");
  k->R_save = k->R.val;
  rp(r, &k->wt, "wt");
  rp(r, &k->A0, "A0");
  rp(r, &k->D, "D");
  rp(r, &k->k, "k");
  k->D.val *= D2R;
  k->k.val /= D2R;
}

static void read_optical_law(modrd *r, int ilaw)
{
  io_ctx *cx = r->cx;
  struct photo_t *ph = &r->mod->photo;
  union optscat_t *L = &ph->optical[ilaw];
  const unsigned char vary = r->par->vary_optalb;
  char w[IO_MAXLEN + 1];
  io_word(r->f, w, sizeof w);
  if (!strcmp(w, "none")) {
    ph->opttype[ilaw] = NOLAW;
  } else if (!strcmp(w, "geometrical") || !strcmp(w, "lambertian") || !strcmp(w, "lommel")) {
    ph->opttype[ilaw] = !strcmp(w, "geometrical") ? GEOMETRICAL : !strcmp(w, "lambertian") ? LAMBERTLAW : LOMMEL;
    rp(r, &L->R.R, "R");
    albedo_check(r, &L->R.R, vary, 0, "R");
    L->R.R_save = L->R.R.val;
  } else if (!strcmp(w, "harmlambert") || !strcmp(w, "harmlommel")) {
    ph->opttype[ilaw] = !strcmp(w, "harmlambert") ? HARMLAMBERT : HARMLOMMEL;
    read_expansion(r, &L->harmR.R, 1, 1.0, 0, vary, "R");
    L->harmR.local = (struct R_t **)local_arrays(r, sizeof(struct R_t));
  } else if (!strcmp(w, "inholambert") || !strcmp(w, "inholommel")) {
    ph->opttype[ilaw] = !strcmp(w, "inholambert") ? INHOLAMBERT : INHOLOMMEL;
    struct inhoR_t *h = &L->inhoR;
    rp(r, &h->global.R, "R");
    const int globalflag = h->global.R.state == 'c' && h->global.R.val != 0.0;
    h->global.R_save = h->global.R.val;
    const int nc = inho_count(r, -1, r->mod->shape.ncomp);
    h->local = (struct R_t **)io_calloc(cx, (size_t)nc, sizeof(struct R_t *));
    for (int c = 0; c < nc; c++) {
      const int nf = inho_count(r, c, 0);
      h->local[c] = (struct R_t *)io_calloc(cx, (size_t)nf, sizeof(struct R_t));
      for (int f = 0; f < nf; f++) {
        rp(r, &h->local[c][f].R, "R");
        if (fitting(r, vary) && (h->local[c][f].R.state == 'c' || (h->local[c][f].R.state == '=' && globalflag)))
          io_fail(cx, "read_mod: must let R float for all facets when varying optical albedos");
        h->local[c][f].R_save = h->local[c][f].R.val;
      }
    }
  } else if (!strcmp(w, "hapke")) {
    ph->opttype[ilaw] = HAPKE;
    read_hapke(r, &L->hapke, 1);
    albedo_check(r, &L->hapke.w, vary, 0, "w");
  } else if (!strcmp(w, "harmhapke")) {
    ph->opttype[ilaw] = HARMHAPKE;
    read_expansion(r, &L->harmhapke.w, 1, 1.0, 0, vary, "w");
    read_expansion(r, &L->harmhapke.h, 0, 1.0, 0, vary, "h");
    read_expansion(r, &L->harmhapke.B0, 0, 1.0, 0, vary, "B0");
    read_expansion(r, &L->harmhapke.g, 0, 1.0, 0, vary, "g");
    read_expansion(r, &L->harmhapke.theta, 0, D2R, 0, vary, "theta");
    L->harmhapke.local = (struct hapke_t **)local_arrays(r, sizeof(struct hapke_t));
  } else if (!strcmp(w, "inhohapke")) {
    ph->opttype[ilaw] = INHOHAPKE;
    struct inhohapke_t *h = &L->inhohapke;
    read_hapke(r, &h->global, 1);
    const int globalflag = h->global.w.state == 'c' && h->global.w.val != 0.0;
    const int nc = inho_count(r, -1, r->mod->shape.ncomp);
    h->local = (struct hapke_t **)io_calloc(cx, (size_t)nc, sizeof(struct hapke_t *));
    for (int c = 0; c < nc; c++) {
      const int nf = inho_count(r, c, 0);
      h->local[c] = (struct hapke_t *)io_calloc(cx, (size_t)nf, sizeof(struct hapke_t));
      for (int f = 0; f < nf; f++) {
        read_hapke(r, &h->local[c][f], 1);
        if (fitting(r, vary) && (h->local[c][f].w.state == 'c' || (h->local[c][f].w.state == '=' && globalflag)))
          io_fail(cx, "read_mod: must let w float for all facets when varying optical albedos");
      }
    }
  } else if (!strcmp(w, "kaasalainen")) {
    ph->opttype[ilaw] = KAASALAINEN;
    read_kaas(r, &L->kaas);
    albedo_check(r, &L->kaas.R, vary, 0, "R");
  } else if (!strcmp(w, "harmkaas")) {
    ph->opttype[ilaw] = HARMKAAS;
    read_expansion(r, &L->harmkaas.R, 1, 1.0, 0, vary, "R");
    read_expansion(r, &L->harmkaas.wt, 0, 1.0, 0, vary, "wt");
    read_expansion(r, &L->harmkaas.A0, 0, 1.0, 0, vary, "A0");
    read_expansion(r, &L->harmkaas.D, 0, D2R, 0, vary, "D");
    read_expansion(r, &L->harmkaas.k, 0, D2R, 1, vary, "k");
    L->harmkaas.local = (struct kaas_t **)local_arrays(r, sizeof(struct kaas_t));
  } else if (!strcmp(w, "inhokaas")) {
    ph->opttype[ilaw] = INHOKAAS;
    struct inhokaas_t *h = &L->inhokaas;
    read_kaas(r, &h->global);
    const int globalflag = h->global.R.state == 'c' && h->global.R.val != 0.0;
    const int nc = inho_count(r, -1, r->mod->shape.ncomp);
    h->local = (struct kaas_t **)io_calloc(cx, (size_t)nc, sizeof(struct kaas_t *));
    for (int c = 0; c < nc; c++) {
      const int nf = inho_count(r, c, 0);
      h->local[c] = (struct kaas_t *)io_calloc(cx, (size_t)nf, sizeof(struct kaas_t));
      for (int f = 0; f < nf; f++) {
        read_kaas(r, &h->local[c][f]);
        if (fitting(r, vary) && (h->local[c][f].R.state == 'c' || (h->local[c][f].R.state == '=' && globalflag)))
          io_fail(cx, "read_mod: must let R float for all facets when varying optical albedos");
      }
    }
  } else {
    io_fail(cx, "read_mod: invalid optical scattering law '%s'", w);
  }
}

static void read_photo(modrd *r)
{
  io_ctx *cx = r->cx;
  struct photo_t *ph = &r->mod->photo;
  ph->nradlaws = io_int(cx, r->f);
  if (ph->nradlaws < 0) io_fail(cx, "read_mod: negative number of radar laws");
  ph->radtype = (unsigned char *)io_calloc(cx, (size_t)ph->nradlaws + 1, 1);
  ph->radar = (union radscat_t *)io_calloc(cx, (size_t)ph->nradlaws, sizeof(union radscat_t));
  for (int i = 0; i < ph->nradlaws; i++) read_radar_law(r, i);
  ph->noptlaws = io_int(cx, r->f);
  if (ph->noptlaws < 0) io_fail(cx, "read_mod: negative number of optical laws");
  ph->opttype = (unsigned char *)io_calloc(cx, (size_t)ph->noptlaws + 2, 1);
  ph->optical = (union optscat_t *)io_calloc(cx, (size_t)ph->noptlaws, sizeof(union optscat_t));
  for (int i = 0; i < ph->noptlaws; i++) read_optical_law(r, i);
}

/* ---- spin state (read_mod.c:1952-2150) --------------------------------- */
static void read_spin(modrd *r)
{
  io_ctx *cx = r->cx;
  struct spin_t *s = &r->mod->spin;
  char w[IO_MAXLEN + 1];
  s->t0 = io_caldate(cx, r->f);
  for (int i = 0; i < 3; i++) rp(r, &s->angle[i], "an Euler angle");
  for (int i = 0; i < 3; i++) rp(r, &s->omega[i], "a spin component");
  for (int i = 0; i < 3; i++) rp(r, &s->inertia[i], "a principal moment of inertia");
  for (int i = 0; i < 3; i++) rp(r, &s->omegadot[i], "a YORP component");
  rp(r, &s->lib_amp, "the libration amplitude");
  rp(r, &s->lib_freq, "the libration frequency");
  rp(r, &s->lib_phase, "the libration phase");
  if (!io_word(r->f, w, sizeof w)) io_fail(cx, "read_mod: end of file where the number of spin impulses was expected");
  s->n_impulse = io_str_to_int(cx, w);
  if (s->n_impulse < 0 || s->n_impulse > MAXIMP) io_fail(cx, "read_mod: %d spin impulses (at most %d)", s->n_impulse, MAXIMP);
  for (int n = 0; n < s->n_impulse; n++) {
    s->t_impulse[n] = io_caldate(cx, r->f);
    if (n > 0 && s->t_impulse[n] <= s->t_impulse[n - 1])
      io_fail(cx, "read_mod: spin impulses must be given in increasing chronological order");
    for (int i = 0; i < 3; i++) rp(r, &s->impulse[n][i], "a spin impulse");
  }
  for (int n = s->n_impulse; n < MAXIMP; n++) {
    s->t_impulse[n] = -HUGENUMBER;
    for (int i = 0; i < 3; i++) { s->impulse[n][i].state = 'c'; s->impulse[n][i].val = 0.0; }
  }
  /* principal-axis rotation iff the first two spin components, YORP terms and
     impulse components are held at zero (read_mod.c:2077-2110) */
  s->pa = 0;
  if (s->omega[0].val == 0.0 && s->omega[0].state == 'c' && s->omega[1].val == 0.0 && s->omega[1].state == 'c') {
    s->pa = 1;
    for (int n = 0; n < s->n_impulse && s->pa; n++)
      if (s->impulse[n][0].val != 0.0 || s->impulse[n][0].state == 'f' || s->impulse[n][1].val != 0.0 ||
          s->impulse[n][1].state == 'f')
        s->pa = 0;
  }
  const int yorp01 = s->omegadot[0].val != 0 || s->omegadot[0].state == 'f' || s->omegadot[1].val != 0 ||
                     s->omegadot[1].state == 'f';
  if (s->pa && yorp01) io_fail(cx, "read_mod: for a PA rotator, YORP can only be applied to {spin 2}");
  if (!s->pa && (yorp01 || s->omegadot[2].val != 0 || s->omegadot[2].state == 'f' || s->lib_amp.val != 0 ||
                 s->lib_amp.state == 'f'))
    io_fail(cx, "read_mod: cannot apply NPA rotation and YORP/librations together");
  /* a pole "over the pole" folds into theta in [0, 180] deg (read_mod.c:2129-2136) */
  s->angle[1].val -= 360.0 * floor(s->angle[1].val / 360.0);
  if (s->angle[1].val > 180.0) {
    s->angle[1].val = 360.0 - s->angle[1].val;
    s->angle[0].val += 180.0;
    s->angle[2].val += 180.0;
  }
  for (int i = 0; i < 3; i++) {
    s->angle[i].val -= 360.0 * floor(s->angle[i].val / 360.0);
    s->angle[i].val *= D2R;
    s->omega[i].val *= D2R;
    s->omegadot[i].val *= D2R;
    for (int n = 0; n < s->n_impulse; n++) s->impulse[n][i].val *= D2R;
  }
  s->lib_amp.val *= D2R;
  s->lib_freq.val *= D2R;
  s->lib_phase.val *= D2R;
}

static int mod_body(io_ctx *cx, struct par_t *par, struct mod_t *mod)
{
  io_file f;
  char name[sizeof mod->name];
  memcpy(name, mod->name, sizeof name);
  name[sizeof name - 1] = '\0';
  /* a fresh model: everything but the name */
  memset(mod, 0, sizeof *mod);
  memcpy(mod->name, name, sizeof name);
  io_open(cx, &f, name);
  modrd r = {cx, &f, par, mod, 0};
  read_shape(&r);
  setup_real(&r);
  read_photo(&r);
  read_spin(&r);
  io_close(cx, &f);
  for (int c = 0; c < mod->shape.ncomp; c++) {
    io_setup_sides(cx, &mod->shape.comp[c].real);
    io_setup_vertices(cx, &mod->shape.comp[c].real);
  }
  return r.nfpar;
}

int shape_read_mod(struct par_t *par, struct mod_t *mod)
{
  io_ctx cx;
  memset(&cx, 0, sizeof cx);
  if (setjmp(cx.jb)) return io_unwind(&cx);
  shape_io_free_mod(mod);
  const int n = mod_body(&cx, par, mod);
  io_arena_commit(&cx, mod);
  return n;
}

int read_mod(struct par_t *par, struct mod_t *mod)
{
  const int n = shape_read_mod(par, mod);
  if (n < 0) {
    printf("%s\n", shape_io_error());
    fprintf(stderr, "ERROR: %s\n", shape_io_error());
    exit(2);
  }
  return n;
}
