/*
 * scene.c -- build synthetic SHAPE scenes as ABI structs (see scene.h).
 *
 * What is generated and why it looks like the reference:
 *  - Ellipsoid tessellation: rings of vertices at colatitude t*pi/T with
 *    max(1, round(2T sin theta)) vertices each, zipped into triangles ring by
 *    ring; this reproduces the vertex/facet counts of read_mod.c:2179-2380
 *    (T = 30 gives 1148 vertices, 2292 facets), which is what the
 *    ellipsoid-converted-to-vertex models of the reference contain.
 *  - Icosphere: 4-way subdivision of an icosahedron, V = 10*4^k + 2.
 *  - Vertex components hold base point a = 0, direction u (unit), radius r
 *    along u (vertex_t, shape2.h:723), so realize_mod's VERTEX branch
 *    x = s*(u*r + a) (realize_mod.c:475-479) yields the deformed shape.
 *  - vertex->af lists attached facets in ascending facet order, as
 *    setupvertices (read_mod.c:2504-2590) builds them.
 *  - POS blocks are shared by every frame for pos_scope GLOBAL and separate
 *    per frame for LOCAL (read_dat.c:1239-1330); km_per_pixel = width/(2n).
 */
#include "scene.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* NR-style allocators: row pointers into one zeroed block             */
/* ------------------------------------------------------------------ */
double **shp_matrix(int nrl, int nrh, int ncl, int nch)
{
  size_t nr = (size_t)(nrh - nrl + 1), nc = (size_t)(nch - ncl + 1);
  double **m = (double **)malloc(nr * sizeof(double *));
  double *blk = (double *)calloc(nr * nc, sizeof(double));
  if (!m || !blk) abort();
  for (size_t i = 0; i < nr; i++) m[i] = blk + i * nc - ncl;
  return m - nrl;
}

int **shp_imatrix(int nrl, int nrh, int ncl, int nch)
{
  size_t nr = (size_t)(nrh - nrl + 1), nc = (size_t)(nch - ncl + 1);
  int **m = (int **)malloc(nr * sizeof(int *));
  int *blk = (int *)calloc(nr * nc, sizeof(int));
  if (!m || !blk) abort();
  for (size_t i = 0; i < nr; i++) m[i] = blk + i * nc - ncl;
  return m - nrl;
}

double *shp_vector(int nl, int nh)
{
  double *v = (double *)calloc((size_t)(nh - nl + 1), sizeof(double));
  if (!v) abort();
  return v - nl;
}

void shp_free_matrix(double **m, int nrl, int ncl)
{
  if (!m) return;
  free(m[nrl] + ncl);
  free(m + nrl);
}

void shp_free_imatrix(int **m, int nrl, int ncl)
{
  if (!m) return;
  free(m[nrl] + ncl);
  free(m + nrl);
}

void shp_free_vector(double *v, int nl)
{
  if (v) free(v + nl);
}

/* ------------------------------------------------------------------ */
/* deterministic random numbers (splitmix64 + Box-Muller)              */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t s; int have; double spare; } rng_t;

static uint64_t rng_next(rng_t *r)
{
  uint64_t z = (r->s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double rng_uniform(rng_t *r) { return (rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
static double rng_range(rng_t *r, double a, double b) { return a + (b - a) * rng_uniform(r); }
static int rng_int(rng_t *r, int a, int b) { return a + (int)floor(rng_uniform(r) * (b - a + 1)); }
static double rng_normal(rng_t *r)
{
  if (r->have) { r->have = 0; return r->spare; }
  double u1 = rng_uniform(r), u2 = rng_uniform(r);
  if (u1 < 1e-300) u1 = 1e-300;
  double rad = sqrt(-2.0 * log(u1));
  r->spare = rad * sin(TWOPI * u2);
  r->have = 1;
  return rad * cos(TWOPI * u2);
}
static rng_t rng_make(uint64_t seed) { rng_t r; r.s = seed * 0x2545F4914F6CDD1DULL + 12345; r.have = 0; r.spare = 0; return r; }

/* ------------------------------------------------------------------ */
/* mesh builders                                                      */
/* ------------------------------------------------------------------ */
typedef struct { int nv, nf; double *u; int *fv; } mesh_t; /* u: 3*nv unit dirs, fv: 3*nf */

static int round_half_away(double x) { return x < 0 ? (int)(x - 0.5) : (int)(x + 0.5); }

static void mesh_push_tri(mesh_t *m, int *cap, int a, int b, int c)
{
  if (m->nf == *cap) { *cap = *cap ? 2 * *cap : 1024; m->fv = (int *)realloc(m->fv, sizeof(int) * 3 * (size_t)*cap); }
  m->fv[3 * m->nf] = a; m->fv[3 * m->nf + 1] = b; m->fv[3 * m->nf + 2] = c;
  m->nf++;
}

/* Ring-zipper tessellation of the unit sphere with T colatitude steps. */
static mesh_t mesh_ellipsoid_tess(int T)
{
  mesh_t m = {0, 0, NULL, NULL};
  int *np = (int *)malloc(sizeof(int) * (T + 1));
  int *first = (int *)malloc(sizeof(int) * (T + 1));
  for (int t = 0; t <= T; t++) {
    int k = round_half_away(2 * T * sin(t * PIE / T));
    np[t] = k < 1 ? 1 : k;
    first[t] = m.nv;
    m.nv += np[t];
  }
  m.u = (double *)malloc(sizeof(double) * 3 * m.nv);
  for (int t = 0, v = 0; t <= T; t++) {
    double th = t * PIE / T;
    for (int p = 0; p < np[t]; p++, v++) {
      double ph = p * TWOPI / np[t];
      m.u[3 * v] = sin(th) * cos(ph);
      m.u[3 * v + 1] = sin(th) * sin(ph);
      m.u[3 * v + 2] = cos(th);
    }
  }
  int cap = 0;
  for (int t = 0; t < T; t++) {
    int na = np[t], nb = np[t + 1], a = 0, b = 0;
    /* walk both rings eastward; the step whose next vertex has the smaller
       azimuth goes first (top ring first on ties); poles only fan. */
    while (a < na || b < nb) {
      double pa = (double)(a + 1) / na, pb = (double)(b + 1) / nb;
      int step_top;
      if (na == 1) step_top = 0;
      else if (nb == 1) step_top = 1;
      else if (a >= na) step_top = 0;
      else if (b >= nb) step_top = 1;
      else step_top = (pa <= pb);
      int va = first[t] + (a % na), vb = first[t + 1] + (b % nb);
      if (step_top) {
        a++;
        mesh_push_tri(&m, &cap, va, vb, first[t] + (a % na));
      } else {
        b++;
        mesh_push_tri(&m, &cap, va, vb, first[t + 1] + (b % nb));
      }
      if (na == 1 && b >= nb) a = na;
      if (nb == 1 && a >= na) b = nb;
    }
  }
  free(np);
  free(first);
  return m;
}

/* icosphere: subdivide an icosahedron `level` times, projecting to the sphere */
typedef struct { uint64_t key; int v; } edge_ent;

static int edge_mid(edge_ent *tab, size_t tabsz, mesh_t *m, int *vcap, int a, int b)
{
  int lo = a < b ? a : b, hi = a < b ? b : a;
  uint64_t key = ((uint64_t)lo << 32) | (uint64_t)hi;
  size_t h = (size_t)((key * 0x9E3779B97F4A7C15ULL) >> 20) & (tabsz - 1);
  while (tab[h].key != UINT64_MAX) {
    if (tab[h].key == key) return tab[h].v;
    h = (h + 1) & (tabsz - 1);
  }
  if (m->nv == *vcap) { *vcap *= 2; m->u = (double *)realloc(m->u, sizeof(double) * 3 * (size_t)*vcap); }
  double x = m->u[3 * a] + m->u[3 * b], y = m->u[3 * a + 1] + m->u[3 * b + 1], z = m->u[3 * a + 2] + m->u[3 * b + 2];
  double nn = sqrt(x * x + y * y + z * z);
  m->u[3 * m->nv] = x / nn; m->u[3 * m->nv + 1] = y / nn; m->u[3 * m->nv + 2] = z / nn;
  tab[h].key = key;
  tab[h].v = m->nv;
  return m->nv++;
}

static mesh_t mesh_icosphere(int level)
{
  const double p = (1.0 + sqrt(5.0)) / 2.0;
  static const int F0[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
                                {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
                                {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
  double V0[12][3] = {{-1, p, 0}, {1, p, 0}, {-1, -p, 0}, {1, -p, 0}, {0, -1, p}, {0, 1, p},
                      {0, -1, -p}, {0, 1, -p}, {p, 0, -1}, {p, 0, 1}, {-p, 0, -1}, {-p, 0, 1}};
  mesh_t m = {0, 0, NULL, NULL};
  int vcap = 12;
  m.u = (double *)malloc(sizeof(double) * 3 * vcap);
  for (int i = 0; i < 12; i++) {
    double nn = sqrt(V0[i][0] * V0[i][0] + V0[i][1] * V0[i][1] + V0[i][2] * V0[i][2]);
    for (int k = 0; k < 3; k++) m.u[3 * i + k] = V0[i][k] / nn;
  }
  m.nv = 12;
  int cap = 0;
  for (int f = 0; f < 20; f++) mesh_push_tri(&m, &cap, F0[f][0], F0[f][1], F0[f][2]);
  for (int l = 0; l < level; l++) {
    size_t tabsz = 1;
    while (tabsz < (size_t)m.nf * 4) tabsz <<= 1;
    edge_ent *tab = (edge_ent *)malloc(sizeof(edge_ent) * tabsz);
    for (size_t i = 0; i < tabsz; i++) tab[i].key = UINT64_MAX;
    mesh_t nm = {m.nv, 0, m.u, NULL};
    int ncap = 0;
    for (int f = 0; f < m.nf; f++) {
      int a = m.fv[3 * f], b = m.fv[3 * f + 1], c = m.fv[3 * f + 2];
      int ab = edge_mid(tab, tabsz, &nm, &vcap, a, b);
      int bc = edge_mid(tab, tabsz, &nm, &vcap, b, c);
      int ca = edge_mid(tab, tabsz, &nm, &vcap, c, a);
      mesh_push_tri(&nm, &ncap, a, ab, ca);
      mesh_push_tri(&nm, &ncap, b, bc, ab);
      mesh_push_tri(&nm, &ncap, c, ca, bc);
      mesh_push_tri(&nm, &ncap, ab, bc, ca);
    }
    free(tab);
    free(m.fv);
    m = nm;
  }
  return m;
}

/* orient every facet outward (normal points away from the origin) */
static void mesh_orient(mesh_t *m)
{
  for (int f = 0; f < m->nf; f++) {
    int *v = m->fv + 3 * f;
    const double *p0 = m->u + 3 * v[0], *p1 = m->u + 3 * v[1], *p2 = m->u + 3 * v[2];
    double a[3], b[3], n[3], c[3];
    for (int k = 0; k < 3; k++) { a[k] = p1[k] - p0[k]; b[k] = p2[k] - p1[k]; c[k] = p0[k] + p1[k] + p2[k]; }
    n[0] = a[1] * b[2] - a[2] * b[1];
    n[1] = a[2] * b[0] - a[0] * b[2];
    n[2] = a[0] * b[1] - a[1] * b[0];
    if (n[0] * c[0] + n[1] * c[1] + n[2] * c[2] < 0) { int t = v[1]; v[1] = v[2]; v[2] = t; }
  }
}

/* ------------------------------------------------------------------ */
/* scene                                                              */
/* ------------------------------------------------------------------ */
struct shp_scene {
  shp_spec spec;
  struct par_t par;
  struct mod_t mod;
  struct dat_t dat;
  int nparams;
  double *params[32];
  int ptype[32];
  /* allocations to release */
  void **allocs;
  int nalloc, capalloc;
};

static void *track(shp_scene *sc, void *p)
{
  if (sc->nalloc == sc->capalloc) {
    sc->capalloc = sc->capalloc ? 2 * sc->capalloc : 256;
    sc->allocs = (void **)realloc(sc->allocs, sizeof(void *) * sc->capalloc);
  }
  sc->allocs[sc->nalloc++] = p;
  return p;
}
static void *tcalloc(shp_scene *sc, size_t n, size_t sz) { void *p = calloc(n ? n : 1, sz); if (!p) abort(); return track(sc, p); }

/* matrix whose block and row array are both tracked */
static double **t_matrix(shp_scene *sc, int nrl, int nrh, int ncl, int nch)
{
  double **m = shp_matrix(nrl, nrh, ncl, nch);
  track(sc, m[nrl] + ncl);
  track(sc, m + nrl);
  return m;
}
static int **t_imatrix(shp_scene *sc, int nrl, int nrh, int ncl, int nch)
{
  int **m = shp_imatrix(nrl, nrh, ncl, nch);
  track(sc, m[nrl] + ncl);
  track(sc, m + nrl);
  return m;
}
static double *t_vector(shp_scene *sc, int nl, int nh)
{
  double *v = shp_vector(nl, nh);
  track(sc, v + nl);
  return v;
}

void shp_spec_defaults(shp_spec *s)
{
  memset(s, 0, sizeof(*s));
  s->shape_kind = SHP_SHAPE_ELLIPSOID_TESS;
  s->theta_steps = 30;
  s->semi_axes[0] = 1.0; s->semi_axes[1] = 0.8; s->semi_axes[2] = 0.6;
  s->rdev_sigma = 0.02;
  s->comp_type = 0;
  s->ovoid_k = 0.2;
  s->nhar = 3;
  s->har_sigma = 0.05;
  s->seed_geom = 1001;
  s->seed_frames = 1002;
  s->spin_period_h = 6.0;
  s->pole_lambda_deg = 30.0;
  s->pole_beta_deg = 60.0;
  s->phase0_deg = 0.0;
  s->t0 = 2458000.5;
  s->epoch_span_days = 0.0;
  s->pos_n = 75;
  s->km_per_pixel = 0.015;
  s->pos_scope = GLOBAL;
  s->pos_smooth = 1;
  s->Ftx = 8560.0;
  s->codemethod = LONG_MOD;
  s->spb = 1;
  s->stride = 1;
  s->dopfftlen = 64;
  s->R_min = s->R_max = 0.2;
  s->C_min = s->C_max = 1.5;
  s->one_set_per_frame = 1;
  s->n_deldop = 1;
  s->ndel_min = s->ndel_max = 100;
  s->ndop_min = s->ndop_max = 80;
  s->del_per_pixel = 0.5;
  s->dop_per_pixel = 0.5;
  s->n_doppler = 0;
  s->nbin_min = 100; s->nbin_max = 400;
  s->dop_per_bin_min = 0.25; s->dop_per_bin_max = 2.0;
  s->n_lightcurves = 0;
  s->lc_points = 100;
  s->lc_ncalc = 40;
  s->phase_min_deg = 5.0; s->phase_max_deg = 60.0;
  s->kaas_wt = 0.1;
  s->delcor_a0 = 0.0;
  s->delcor_a1 = 0.0;
  s->obs_lat_min_deg = 10.0;
  s->obs_lat_max_deg = 10.0;
  s->radar_nviews = 1;
  s->view_interval_days = 0.0;
  s->lc_nviews = 1;
  s->nsinc2 = 1;
  s->sinc2width = 6;
}

/* Euler-angle matrix with the reference's convention (util/euler.c). */
static void euler_mat(double m[3][3], double phi, double theta, double psi)
{
  double ct = cos(theta), st = sin(theta), cp = cos(phi), sp = sin(phi), cs = cos(psi), ss = sin(psi);
  m[0][0] = cs * cp - ct * sp * ss;
  m[1][0] = -ss * cp - ct * sp * cs;
  m[2][0] = st * sp;
  m[0][1] = cs * sp + ct * cp * ss;
  m[1][1] = -ss * sp + ct * cp * cs;
  m[2][1] = -st * cp;
  m[0][2] = ss * st;
  m[1][2] = cs * st;
  m[2][2] = ct;
}

/* Frame whose z axis points along dir (ecliptic), y toward ecliptic north. */
static void frame_from_dir(double m[3][3], const double dir[3])
{
  double z[3] = {dir[0], dir[1], dir[2]}, y[3], x[3];
  double nn = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
  for (int k = 0; k < 3; k++) z[k] /= nn;
  double up[3] = {0, 0, 1};
  double d = up[0] * z[0] + up[1] * z[1] + up[2] * z[2];
  if (fabs(d) > 0.999) { up[0] = 1; up[2] = 0; d = z[0]; }
  for (int k = 0; k < 3; k++) y[k] = up[k] - d * z[k];
  nn = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
  for (int k = 0; k < 3; k++) y[k] /= nn;
  x[0] = y[1] * z[2] - y[2] * z[1];
  x[1] = y[2] * z[0] - y[0] * z[2];
  x[2] = y[0] * z[1] - y[1] * z[0];
  for (int k = 0; k < 3; k++) { m[0][k] = x[k]; m[1][k] = y[k]; m[2][k] = z[k]; }
}

static void los_dir(const shp_spec *sp, double frac, double dir[3])
{
  double lat = (sp->obs_lat_min_deg + (sp->obs_lat_max_deg - sp->obs_lat_min_deg) * frac) * D2R;
  double lon = (20.0 + 40.0 * frac) * D2R;
  dir[0] = cos(lat) * cos(lon);
  dir[1] = cos(lat) * sin(lon);
  dir[2] = sin(lat);
}

static void alloc_pos(shp_scene *sc, struct pos_t *pos, int n, double kmpp, int bistatic)
{
  memset(pos, 0, sizeof(*pos));
  pos->n = n;
  pos->km_per_pixel = kmpp;
  pos->b = t_matrix(sc, -n, n, -n, n);
  pos->cose = t_matrix(sc, -n, n, -n, n);
  pos->cosi = t_matrix(sc, -n, n, -n, n);
  pos->z = t_matrix(sc, -n, n, -n, n);
  pos->body = t_imatrix(sc, -n, n, -n, n);
  pos->comp = t_imatrix(sc, -n, n, -n, n);
  pos->f = t_imatrix(sc, -n, n, -n, n);
  pos->cosill = t_matrix(sc, -n, n, -n, n);
  pos->zill = t_matrix(sc, -n, n, -n, n);
  pos->bodyill = t_imatrix(sc, -n, n, -n, n);
  pos->compill = t_imatrix(sc, -n, n, -n, n);
  pos->fill = t_imatrix(sc, -n, n, -n, n);
  pos->bistatic = bistatic;
}

/* Give a frame its POS: shared block (GLOBAL) or its own (LOCAL). */
static void frame_pos(shp_scene *sc, struct pos_t *dst, int bistatic)
{
  if (sc->par.pos_scope == GLOBAL) {
    *dst = sc->dat.pos;
    dst->bistatic = bistatic;
  } else {
    alloc_pos(sc, dst, sc->dat.pos.n, sc->dat.pos.km_per_pixel, bistatic);
  }
}

static void add_param(shp_scene *sc, double *p, int type)
{
  if (sc->nparams < 32) { sc->params[sc->nparams] = p; sc->ptype[sc->nparams] = type; sc->nparams++; }
}

/* The side (edge) list of a closed triangle mesh in read_mod.c's setupsides
   order (:2386-2501): sides numbered by first appearance over the facets in
   order, each facet's corners taken sorted (a < b < c) as (a,b), (a,c),
   (b,c); s.v ascending; s.f[0] the lower-numbered facet, s.f[1] the other;
   each facet's f.s[] in ascending side number.  A hash of the vertex pair
   replaces the reference's linear scan of the list. */
typedef struct { long long key; int s; } side_ent;

static int side_find(side_ent *tab, size_t mask, long long key, int add, int *ns)
{
  size_t h = (size_t)((unsigned long long)key * 0x9E3779B97F4A7C15ULL >> 20) & mask;
  while (tab[h].s >= 0 && tab[h].key != key) h = (h + 1) & mask;
  if (tab[h].s < 0) {
    if (!add) return -1;
    tab[h].key = key;
    tab[h].s = (*ns)++;
  }
  return tab[h].s;
}

static void build_sides(shp_scene *sc, struct vertices_t *vt)
{
  const int nf = vt->nf;
  vt->ns = (3 * nf) / 2;
  vt->s = (struct side_t *)tcalloc(sc, vt->ns, sizeof(struct side_t));
  size_t cap = 1;
  while (cap < 4 * (size_t)vt->ns) cap <<= 1;
  side_ent *tab = (side_ent *)malloc(cap * sizeof(side_ent));
  for (size_t i = 0; i < cap; i++) tab[i].s = -1;
  int ns = 0;
  for (int s = 0; s < vt->ns; s++) vt->s[s].f[0] = vt->s[s].f[1] = -1;
  for (int f = 0; f < nf; f++) {
    int v[3] = {vt->f[f].v[0], vt->f[f].v[1], vt->f[f].v[2]};
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2 - i; j++)
        if (v[j] > v[j + 1]) { int t = v[j]; v[j] = v[j + 1]; v[j + 1] = t; }
    const int pa[3] = {v[0], v[0], v[1]}, pb[3] = {v[1], v[2], v[2]};
    int got[3];
    for (int k = 0; k < 3; k++) {
      const int before = ns;
      int s = side_find(tab, cap - 1, (long long)pa[k] * vt->nv + pb[k], 1, &ns);
      if (s >= vt->ns) abort(); /* not a closed two-manifold mesh */
      if (ns != before) { vt->s[s].v[0] = pa[k]; vt->s[s].v[1] = pb[k]; }
      if (vt->s[s].f[0] < 0) vt->s[s].f[0] = f;
      else vt->s[s].f[1] = f;
      got[k] = s;
    }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2 - i; j++)
        if (got[j] > got[j + 1]) { int t = got[j]; got[j] = got[j + 1]; got[j + 1] = t; }
    for (int k = 0; k < 3; k++) vt->f[f].s[k] = got[k];
  }
  free(tab);
  for (int s = 0; s < vt->ns; s++) vt->s[s].act = 1;
}

/* Normalized associated Legendre functions N_lm P_lm(x), 0 <= m <= l <= L,
   N_lm = sqrt((2l+1) (l-m)!/(l+m)!), into p[l*(L+1)+m]: P_mm by the product
   formula, then the upward recurrence in l (Condon-Shortley phase included,
   as the Numerical Recipes routine the reference calls has it) */
static void legendre_norm(int L, double x, double *p)
{
  const double s = sqrt((1.0 - x) * (1.0 + x));
  for (int m = 0; m <= L; m++) {
    double pmm = 1.0, fact = 1.0;
    for (int i = 1; i <= m; i++) { pmm *= -fact * s; fact += 2.0; }
    double prev = 0.0, cur = pmm;
    for (int l = m; l <= L; l++) {
      if (l == m + 1) { prev = cur; cur = x * (2 * m + 1) * pmm; }
      else if (l > m + 1) { double nxt = (x * (2 * l - 1) * cur - (l + m - 1) * prev) / (l - m); prev = cur; cur = nxt; }
      double ratio = 1.0; /* (l-m)!/(l+m)! */
      for (int i = l - m + 1; i <= l + m; i++) ratio /= i;
      p[l * (L + 1) + m] = sqrt((2 * l + 1) * ratio) * cur;
    }
  }
}

static void build_spin(shp_scene *sc)
{
  const shp_spec *sp = &sc->spec;
  struct mod_t *mod = &sc->mod;
  struct spin_t *s = &mod->spin;
  s->pa = 1;
  s->t0 = sp->t0;
  s->angle[0].val = (sp->pole_lambda_deg + 90.0) * D2R;
  s->angle[1].val = (90.0 - sp->pole_beta_deg) * D2R;
  s->angle[2].val = sp->phase0_deg * D2R;
  s->omega[2].val = TWOPI / (sp->spin_period_h / 24.0);
  for (int k = 0; k < 3; k++) {
    s->angle[k].state = 'f'; s->omega[k].state = k == 2 ? 'f' : 'c'; s->omegadot[k].state = 'c';
    s->inertia[k].state = 'c'; s->inertia[k].val = 1.0;
  }
  s->lib_amp.state = s->lib_freq.state = s->lib_phase.state = 'c';
}

/* An analytic component on the ellipsoid tessellation's directions
   (setuprealver's rings, read_mod.c:2179-2290): base points zero, r to be
   realized; the harmonic factors afactor[l][m] = N_lm P_lm(cos theta)
   cos(m phi), bfactor = ... sin(m phi); the parameters from the spec */
static void build_analytic(shp_scene *sc, struct comp_t *c, const mesh_t *m, rng_t *rg)
{
  const shp_spec *sp = &sc->spec;
  struct vertices_t *vt = &c->real;
  vt->nv = m->nv;
  vt->nf = m->nf;
  vt->v = (struct vertex_t *)tcalloc(sc, m->nv, sizeof(struct vertex_t));
  vt->f = (struct facet_t *)tcalloc(sc, m->nf, sizeof(struct facet_t));
  for (int k = 0; k < 3; k++) { vt->scalefactor[k].state = 'f'; vt->scalefactor[k].val = 1.0; }
  const int L = sp->nhar;
  double *plm = (double *)malloc(sizeof(double) * (L + 1) * (L + 1));
  for (int i = 0; i < m->nv; i++) {
    struct vertex_t *v = &vt->v[i];
    for (int k = 0; k < 3; k++) { v->u[k] = m->u[3 * i + k]; v->a[k] = 0.0; }
    v->r.state = 'f';
    v->act = 1;
    v->v_mirror = i;
    if (sp->comp_type == 3) {
      const double phi = atan2(v->u[1], v->u[0]);
      legendre_norm(L, v->u[2], plm);
      v->afactor = t_matrix(sc, 0, L, 0, L);
      if (L > 0) v->bfactor = t_matrix(sc, 1, L, 1, L);
      for (int l = 0; l <= L; l++) {
        v->afactor[l][0] = plm[l * (L + 1)];
        for (int mm = 1; mm <= l; mm++) {
          v->afactor[l][mm] = cos(mm * phi) * plm[l * (L + 1) + mm];
          v->bfactor[l][mm] = sin(mm * phi) * plm[l * (L + 1) + mm];
        }
      }
    }
  }
  free(plm);
  for (int f = 0; f < m->nf; f++) {
    for (int k = 0; k < 3; k++) vt->f[f].v[k] = m->fv[3 * f + k];
    vt->f[f].act = 1;
  }
  const double *ax = sp->semi_axes;
  if (sp->comp_type == 1) {
    c->type = ELLIPSE;
    c->desc.ell.ntheta = sp->theta_steps;
    c->desc.ell.two_a.state = c->desc.ell.a_over_b.state = c->desc.ell.b_over_c.state = 'f';
    c->desc.ell.two_a.val = 2 * ax[0];
    c->desc.ell.a_over_b.val = ax[0] / ax[1];
    c->desc.ell.b_over_c.val = ax[1] / ax[2];
  } else if (sp->comp_type == 2) {
    c->type = OVOID;
    c->desc.ovoid.ntheta = sp->theta_steps;
    c->desc.ovoid.two_a.state = c->desc.ovoid.a_over_b.state = c->desc.ovoid.b_over_c.state = 'f';
    c->desc.ovoid.k.state = 'f';
    c->desc.ovoid.two_a.val = 2 * ax[0];
    c->desc.ovoid.a_over_b.val = ax[0] / ax[1];
    c->desc.ovoid.b_over_c.val = ax[1] / ax[2];
    c->desc.ovoid.k.val = sp->ovoid_k;
  } else {
    c->type = HARMONIC;
    struct harmonic_t *h = &c->desc.har;
    h->ntheta = sp->theta_steps;
    h->nhar = L;
    for (int k = 0; k < 3; k++) { h->scalefactor[k].state = 'f'; h->scalefactor[k].val = 1.0; }
    h->a = (struct param_t **)tcalloc(sc, L + 1, sizeof(struct param_t *));
    h->b = (struct param_t **)tcalloc(sc, L + 1, sizeof(struct param_t *));
    const double r0 = cbrt(ax[0] * ax[1] * ax[2]);
    for (int l = 0; l <= L; l++) {
      h->a[l] = (struct param_t *)tcalloc(sc, l + 1, sizeof(struct param_t));
      h->b[l] = (struct param_t *)tcalloc(sc, l + 1, sizeof(struct param_t));
      for (int mm = 0; mm <= l; mm++) {
        h->a[l][mm].state = h->b[l][mm].state = 'f';
        h->a[l][mm].val = l == 0 ? r0 : sp->har_sigma * r0 * rng_normal(rg) / (l + 1);
        h->b[l][mm].val = mm == 0 ? 0.0 : sp->har_sigma * r0 * rng_normal(rg) / (l + 1);
      }
    }
  }
}

static void build_model(shp_scene *sc)
{
  const shp_spec *sp = &sc->spec;
  mesh_t m = sp->shape_kind == SHP_SHAPE_ICOSPHERE ? mesh_icosphere(sp->ico_level)
                                                   : mesh_ellipsoid_tess(sp->theta_steps);
  mesh_orient(&m);
  rng_t rg = rng_make(sp->seed_geom);

  struct mod_t *mod = &sc->mod;
  strcpy(mod->name, "synthetic");
  mod->shape.ncomp = 1;
  mod->shape.comp = (struct comp_t *)tcalloc(sc, 1, sizeof(struct comp_t));
  struct comp_t *c = &mod->shape.comp[0];
  c->type = VERTEX;
  for (int k = 0; k < 3; k++) { c->off[k].state = 'c'; c->rot[k].state = 'c'; }
  if (sp->comp_type != 0) {
    build_analytic(sc, c, &m, &rg);
    build_sides(sc, &c->real);
    for (int f = 0; f < m.nf; f++)
      for (int k = 0; k < 3; k++) c->real.v[m.fv[3 * f + k]].naf++;
    for (int i = 0; i < m.nv; i++) {
      c->real.v[i].af = (int *)tcalloc(sc, c->real.v[i].naf, sizeof(int));
      c->real.v[i].naf = 0;
    }
    for (int f = 0; f < m.nf; f++)
      for (int k = 0; k < 3; k++) {
        struct vertex_t *v = &c->real.v[m.fv[3 * f + k]];
        v->af[v->naf++] = f;
      }
    free(m.u);
    free(m.fv);
    build_spin(sc);
    if (c->type == ELLIPSE) {
      add_param(sc, &c->desc.ell.two_a.val, SHP_SIZEPAR);
      add_param(sc, &c->desc.ell.a_over_b.val, SHP_SIZEPAR);
      add_param(sc, &c->desc.ell.b_over_c.val, SHP_SIZEPAR);
    } else if (c->type == OVOID) {
      add_param(sc, &c->desc.ovoid.two_a.val, SHP_SIZEPAR);
      add_param(sc, &c->desc.ovoid.a_over_b.val, SHP_SIZEPAR);
      add_param(sc, &c->desc.ovoid.b_over_c.val, SHP_SIZEPAR);
    } else {
      for (int k = 0; k < 3; k++) add_param(sc, &c->desc.har.scalefactor[k].val, SHP_SIZEPAR);
    }
    add_param(sc, &sc->mod.spin.omega[2].val, SHP_SPINPAR);
    add_param(sc, &sc->mod.spin.angle[0].val, SHP_SPINPAR);
    add_param(sc, &sc->mod.spin.angle[1].val, SHP_SPINPAR);
    if (c->type == OVOID) add_param(sc, &c->desc.ovoid.k.val, SHP_SHAPEPAR);
    if (c->type == HARMONIC && c->desc.har.nhar >= 1) {
      add_param(sc, &c->desc.har.a[0][0].val, SHP_SHAPEPAR);
      add_param(sc, &c->desc.har.a[1][1].val, SHP_SHAPEPAR);
    }
    return;
  }
  struct vertices_t *vt = &c->desc.ver;
  for (int k = 0; k < 3; k++) { vt->scalefactor[k].state = 'f'; vt->scalefactor[k].val = 1.0; }
  vt->nv = m.nv;
  vt->nf = m.nf;
  vt->v = (struct vertex_t *)tcalloc(sc, m.nv, sizeof(struct vertex_t));
  vt->f = (struct facet_t *)tcalloc(sc, m.nf, sizeof(struct facet_t));
  double a2[3];
  for (int k = 0; k < 3; k++) a2[k] = 1.0 / (sp->semi_axes[k] * sp->semi_axes[k]);
  for (int i = 0; i < m.nv; i++) {
    struct vertex_t *v = &vt->v[i];
    double *u = m.u + 3 * i;
    double r = 1.0 / sqrt(a2[0] * u[0] * u[0] + a2[1] * u[1] * u[1] + a2[2] * u[2] * u[2]);
    r *= 1.0 + sp->rdev_sigma * rng_normal(&rg);
    for (int k = 0; k < 3; k++) { v->u[k] = u[k]; v->a[k] = 0.0; }
    v->r.state = 'f';
    v->r.val = r;
    v->act = 1;
    v->v_mirror = i;
  }
  for (int f = 0; f < m.nf; f++) {
    for (int k = 0; k < 3; k++) vt->f[f].v[k] = m.fv[3 * f + k];
    vt->f[f].act = 1;
  }
  /* attached facets, ascending facet order */
  for (int f = 0; f < m.nf; f++)
    for (int k = 0; k < 3; k++) vt->v[m.fv[3 * f + k]].naf++;
  for (int i = 0; i < m.nv; i++) {
    vt->v[i].af = (int *)tcalloc(sc, vt->v[i].naf, sizeof(int));
    vt->v[i].naf = 0;
  }
  for (int f = 0; f < m.nf; f++)
    for (int k = 0; k < 3; k++) {
      struct vertex_t *v = &vt->v[m.fv[3 * f + k]];
      v->af[v->naf++] = f;
    }
  build_sides(sc, vt);
  c->real = *vt; /* setupreal: the VERTEX realization shares the description's arrays */
  free(m.u);
  free(m.fv);

  build_spin(sc);

  add_param(sc, &c->desc.ver.scalefactor[0].val, SHP_SIZEPAR);
  add_param(sc, &c->desc.ver.scalefactor[1].val, SHP_SIZEPAR);
  add_param(sc, &c->desc.ver.scalefactor[2].val, SHP_SIZEPAR);
  add_param(sc, &mod->spin.omega[2].val, SHP_SPINPAR);
  add_param(sc, &mod->spin.angle[0].val, SHP_SPINPAR);
  add_param(sc, &mod->spin.angle[1].val, SHP_SPINPAR);
}

/* epoch of the k-th of n frames */
static double frame_epoch(const shp_spec *sp, rng_t *r, int k, int n)
{
  (void)k; (void)n;
  return sp->t0 + sp->epoch_span_days * rng_uniform(r);
}

static void init_delcor(shp_scene *sc, struct delcor_t *dc, double t0)
{
  dc->t0 = t0;
  dc->n = 1;
  dc->a = (struct param_t *)tcalloc(sc, 2, sizeof(struct param_t));
  dc->a[0].state = 'f';
  dc->a[0].val = sc->spec.delcor_a0;
  dc->a[1].state = 'f';
  dc->a[1].val = sc->spec.delcor_a1;
  dc->delcor0_save = dc->a[0].val;
}

/* Doppler width (bins) the echo needs at resolution dpp: bandwidth plus sinc^2 wings */
static double echo_bins(const shp_spec *sp, double dpp)
{
  double rmax = sp->semi_axes[0] * (1.0 + 3.0 * sp->rdev_sigma);
  double omega = TWOPI / (sp->spin_period_h / 24.0);
  double bw = 2.0 * rmax * omega * KM2HZFACT * sp->Ftx; /* Hz */
  return bw / dpp;
}

static void build_data(shp_scene *sc)
{
  const shp_spec *sp = &sc->spec;
  struct dat_t *dat = &sc->dat;
  rng_t rf = rng_make(sp->seed_frames);
  int n = sp->pos_n;
  strcpy(dat->name, "synthetic");
  alloc_pos(sc, &dat->pos, n, sp->km_per_pixel, 1);

  int nrad_sets = sp->one_set_per_frame ? sp->n_deldop + sp->n_doppler
                                        : (sp->n_deldop > 0) + (sp->n_doppler > 0);
  int nsets = nrad_sets + sp->n_lightcurves;
  dat->nsets = nsets;
  dat->set = (struct set_t *)tcalloc(sc, nsets, sizeof(struct set_t));

  /* photometric laws: one radar law per radar set, one optical law per lightcurve */
  struct photo_t *ph = &sc->mod.photo;
  ph->nradlaws = nrad_sets > 0 ? nrad_sets : 1;
  ph->radtype = (unsigned char *)tcalloc(sc, ph->nradlaws, 1);
  ph->radar = (union radscat_t *)tcalloc(sc, ph->nradlaws, sizeof(union radscat_t));
  for (int l = 0; l < ph->nradlaws; l++) {
    ph->radtype[l] = COSINELAW_DIFF;
    ph->radar[l].RC.R.state = 'f';
    ph->radar[l].RC.R.val = rng_range(&rf, sp->R_min, sp->R_max);
    ph->radar[l].RC.C.state = 'f';
    ph->radar[l].RC.C.val = rng_range(&rf, sp->C_min, sp->C_max);
    ph->radar[l].RC.R_save = ph->radar[l].RC.R.val;
  }
  ph->noptlaws = sp->n_lightcurves > 0 ? sp->n_lightcurves : 1;
  ph->opttype = (unsigned char *)tcalloc(sc, ph->noptlaws, 1);
  ph->optical = (union optscat_t *)tcalloc(sc, ph->noptlaws, sizeof(union optscat_t));
  for (int l = 0; l < ph->noptlaws; l++) {
    ph->opttype[l] = KAASALAINEN;
    struct kaas_t *k = &ph->optical[l].kaas;
    k->R.state = 'f'; k->R.val = 1.0; k->R_save = 1.0;
    k->wt.state = 'c'; k->wt.val = sp->kaas_wt;
    k->A0.state = 'c'; k->A0.val = 0.0;
    k->D.state = 'c'; k->D.val = 0.1;
    k->k.state = 'c'; k->k.val = 0.0;
  }

  double tmid = sp->t0 + 0.5 * sp->epoch_span_days;
  int s = 0;
  double dof_dd = 0, dof_dop = 0, dof_lc = 0;

  /* ---- delay-Doppler ---- */
  int nframes_left = sp->n_deldop;
  while (nframes_left > 0) {
    int nfr = sp->one_set_per_frame ? 1 : nframes_left;
    nframes_left -= nfr;
    struct set_t *set = &dat->set[s];
    set->type = DELAY;
    set->inputnode = s;
    for (int k = 0; k < 3; k++) { set->angleoff[k].state = 'c'; set->omegaoff[k].state = 'c'; }
    struct deldop_t *dd = &set->desc.deldop;
    dd->iradlaw = s;
    dd->Ftx = sp->Ftx;
    init_delcor(sc, &dd->delcor, tmid);
    dd->del_per_pixel = sp->del_per_pixel;
    dd->dop_per_pixel = sp->dop_per_pixel;
    dd->codemethod = sp->codemethod;
    dd->spb = sp->spb;
    dd->stride = sp->stride;
    dd->dopfftlen = sp->dopfftlen;
    dd->dopscale.state = 'c';
    dd->dopscale.val = 1.0;
    dd->dopscale_save = 1.0;
    dd->nviews = sp->radar_nviews > 1 ? sp->radar_nviews : 1;
    dd->v0 = dd->nviews / 2;
    dd->view_interval = sp->view_interval_days;
    dd->nframes = nfr;
    dd->frame = (struct deldopfrm_t *)tcalloc(sc, nfr, sizeof(struct deldopfrm_t));
    double set_dof = 0;
    for (int f = 0; f < nfr; f++) {
      struct deldopfrm_t *fr = &dd->frame[f];
      int ndel = rng_int(&rf, sp->ndel_min, sp->ndel_max);
      int ndop = rng_int(&rf, sp->ndop_min, sp->ndop_max);
      double t = frame_epoch(sp, &rf, f, nfr);
      if (sp->dop_per_pixel <= 0.0) {
        /* choose the Doppler resolution (0.25 Hz steps) so the echo fits */
        double dpp = 0.25;
        while (echo_bins(sp, dpp) + sp->dopfftlen * 0 + 14 > ndop) dpp += 0.25;
        dd->dop_per_pixel = dpp; /* set-level value; frames of a set share it */
      }
      fr->ndel = ndel;
      fr->ndop = ndop;
      fr->t0 = t;
      fr->nlooks = 0;
      fr->weight = 1.0;
      fr->sdev = 1.0;
      fr->cal.state = 'f';
      fr->cal.val = 1.0;
      fr->delcom_vig = 0.5 * (ndel + 1) + rng_range(&rf, -2.0, 2.0);
      fr->dopcom_vig = 0.5 * (ndop + 1) + rng_range(&rf, -2.0, 2.0);
      fr->dopDC_vig = fr->dopcom_vig + rng_range(&rf, -3.0, 3.0);
      fr->delcom = fr->delcom_vig;
      fr->obs = t_matrix(sc, 1, ndel, 1, ndop);
      fr->fit = t_matrix(sc, 1, ndel, 1, ndop);
      fr->oneovervar = t_matrix(sc, 1, ndel, 1, ndop);
      for (int i = 1; i <= ndel; i++)
        for (int j = 1; j <= ndop; j++) fr->oneovervar[i][j] = 1.0 / (fr->sdev * fr->sdev);
      fr->dof = fr->weight * ndel * ndop;
      set_dof += fr->dof;
      snprintf(fr->name, sizeof(fr->name), "dd_%03d_%03d", s, f);
      frame_pos(sc, &fr->pos, 0);
      fr->view = (struct deldopview_t *)tcalloc(sc, dd->nviews, sizeof(struct deldopview_t));
      for (int v = 0; v < dd->nviews; v++) {  /* smeared views around the frame epoch */
        double tv = t + (v - dd->v0) * dd->view_interval, dir[3];
        fr->view[v].t = tv;
        los_dir(sp, sp->epoch_span_days > 0 ? (tv - sp->t0) / sp->epoch_span_days : 0.0, dir);
        frame_from_dir(fr->view[v].oe, dir);
      }
    }
    dd->dof = set_dof;
    set->dof = set_dof;
    dof_dd += set_dof;
    s++;
  }

  /* ---- Doppler ---- */
  nframes_left = sp->n_doppler;
  while (nframes_left > 0) {
    int nfr = sp->one_set_per_frame ? 1 : nframes_left;
    nframes_left -= nfr;
    struct set_t *set = &dat->set[s];
    set->type = DOPPLER;
    set->inputnode = s;
    for (int k = 0; k < 3; k++) { set->angleoff[k].state = 'c'; set->omegaoff[k].state = 'c'; }
    struct doppler_t *dp = &set->desc.doppler;
    dp->iradlaw = s;
    dp->Ftx = sp->Ftx;
    init_delcor(sc, &dp->delcor, tmid);
    dp->dop_per_bin = rng_range(&rf, sp->dop_per_bin_min, sp->dop_per_bin_max);
    dp->dopscale.state = 'c';
    dp->dopscale.val = 1.0;
    dp->dopscale_save = 1.0;
    dp->nviews = sp->radar_nviews > 1 ? sp->radar_nviews : 1;
    dp->v0 = dp->nviews / 2;
    dp->view_interval = sp->view_interval_days;
    dp->nframes = nfr;
    dp->frame = (struct dopfrm_t *)tcalloc(sc, nfr, sizeof(struct dopfrm_t));
    double set_dof = 0;
    for (int f = 0; f < nfr; f++) {
      struct dopfrm_t *fr = &dp->frame[f];
      int ndop = rng_int(&rf, sp->nbin_min, sp->nbin_max);
      double t = frame_epoch(sp, &rf, f, nfr);
      fr->ndop = ndop;
      fr->t0 = t;
      fr->weight = 1.0;
      fr->sdev = 1.0;
      fr->cal.state = 'f';
      fr->cal.val = 1.0;
      fr->dopcom_vig = 0.5 * (ndop + 1) + rng_range(&rf, -2.0, 2.0);
      fr->obs = t_vector(sc, 1, ndop);
      fr->fit = t_vector(sc, 1, ndop);
      fr->oneovervar = t_vector(sc, 1, ndop);
      for (int j = 1; j <= ndop; j++) fr->oneovervar[j] = 1.0;
      fr->dof = fr->weight * ndop;
      set_dof += fr->dof;
      snprintf(fr->name, sizeof(fr->name), "dop_%03d_%03d", s, f);
      frame_pos(sc, &fr->pos, 0);
      fr->view = (struct dopview_t *)tcalloc(sc, dp->nviews, sizeof(struct dopview_t));
      for (int v = 0; v < dp->nviews; v++) {  /* smeared views around the frame epoch */
        double tv = t + (v - dp->v0) * dp->view_interval, dir[3];
        fr->view[v].t = tv;
        los_dir(sp, sp->epoch_span_days > 0 ? (tv - sp->t0) / sp->epoch_span_days : 0.0, dir);
        frame_from_dir(fr->view[v].oe, dir);
      }
    }
    dp->dof = set_dof;
    set->dof = set_dof;
    dof_dop += set_dof;
    s++;
  }

  /* ---- lightcurves ---- */
  for (int l = 0; l < sp->n_lightcurves; l++, s++) {
    struct set_t *set = &dat->set[s];
    set->type = LGHTCRV;
    set->inputnode = s;
    for (int k = 0; k < 3; k++) { set->angleoff[k].state = 'c'; set->omegaoff[k].state = 'c'; }
    struct lghtcrv_t *lc = &set->desc.lghtcrv;
    lc->ioptlaw = l;
    lc->nviews = sp->lc_nviews > 1 ? sp->lc_nviews : 1;
    lc->v0 = lc->nviews / 2;
    lc->view_interval = sp->view_interval_days;
    lc->n = sp->lc_points;
    lc->ncalc = sp->lc_ncalc;
    lc->weight = 1.0;
    lc->cal.state = 'f';
    lc->cal.val = 1.0;
    int nn = lc->n, nc = lc->ncalc;
    lc->t0 = t_vector(sc, 1, nn);
    lc->t = t_matrix(sc, 1, nn, 0, lc->nviews - 1);
    lc->obs = t_vector(sc, 1, nn);
    lc->oneovervar = t_vector(sc, 1, nn);
    lc->fit = t_vector(sc, 1, nn);
    lc->rotphase_calc = t_vector(sc, 1, nc);
    lc->rotphase_obs = t_vector(sc, 1, nn);
    lc->x0 = t_vector(sc, 1, nc);
    lc->x = t_vector(sc, 1, nc);
    lc->y = t_vector(sc, 1, nc);
    lc->y2 = t_vector(sc, 1, nc);
    lc->solar_phase = t_vector(sc, 1, nc);
    lc->solar_azimuth = t_vector(sc, 1, nc);
    lc->rend = (struct crvrend_t *)tcalloc(sc, nc + 1, sizeof(struct crvrend_t));
    /* one night: calc points cover [tstart, tstart+len], observations inside */
    double len = sp->spin_period_h / 24.0 * 1.2;
    double tstart = sp->t0 + (sp->epoch_span_days - len > 0 ? (sp->epoch_span_days - len) * rng_uniform(&rf) : 0.0);
    double phase = rng_range(&rf, sp->phase_min_deg, sp->phase_max_deg) * D2R;
    double dir[3], sdir[3], oe[3][3];
    los_dir(sp, sp->epoch_span_days > 0 ? (tstart - sp->t0) / sp->epoch_span_days : 0.0, dir);
    frame_from_dir(oe, dir);
    /* sun direction: rotate the line of sight by the solar phase angle toward +x */
    for (int k = 0; k < 3; k++) sdir[k] = cos(phase) * oe[2][k] + sin(phase) * oe[0][k];
    for (int i = 1; i <= nc; i++) {
      lc->x[i] = lc->x0[i] = tstart + len * (i - 1) / (nc - 1);
      lc->solar_phase[i] = phase;
      lc->solar_azimuth[i] = 0.0;
      struct crvrend_t *rd = &lc->rend[i];
      memcpy(rd->oe, oe, sizeof(oe));
      frame_from_dir(rd->se, sdir);
      frame_pos(sc, &rd->pos, 1);
    }
    for (int i = 1; i <= nn; i++) {
      double t = tstart + len * (0.02 + 0.96 * (i - 0.5) / nn);
      lc->t0[i] = t;
      for (int v = 0; v < lc->nviews; v++) lc->t[i][v] = t + (v - lc->v0) * lc->view_interval;
      lc->oneovervar[i] = 1.0;
    }
    lc->dof = lc->weight * nn;
    set->dof = lc->dof;
    dof_lc += lc->dof;
    snprintf(lc->name, sizeof(lc->name), "lc_%03d", s);
  }

  dat->dof_deldop = dof_dd;
  dat->dof_doppler = dof_dop;
  dat->dof_lghtcrv = dof_lc;
  dat->dof_poset = 0.0;
  dat->dof = dof_dd + dof_dop + dof_lc;

  /* free parameters beyond the model's: first set's radar law, first set's delay polynomial */
  if (nrad_sets > 0) {
    add_param(sc, &ph->radar[0].RC.R.val, SHP_PHOTOPAR);
    add_param(sc, &ph->radar[0].RC.C.val, SHP_PHOTOPAR);
    struct delcor_t *dc = dat->set[0].type == DELAY ? &dat->set[0].desc.deldop.delcor
                                                    : &dat->set[0].desc.doppler.delcor;
    add_param(sc, &dc->a[0].val, SHP_DELCORPAR);
    add_param(sc, &dc->a[1].val, SHP_DELCORPAR);
  }
  /* last, a vertex deviation: the radius of a vertex a third of the way
     through the list (a shape parameter that moves only its own facets) */
  struct vertices_t *vt = &sc->mod.shape.comp[0].desc.ver;  /* (a union member: VERTEX only) */
  if (sc->mod.shape.comp[0].type == VERTEX && vt->nv > 0) add_param(sc, &vt->v[vt->nv / 3].r.val, SHP_SHAPEPAR);
}

static void build_par(shp_scene *sc)
{
  struct par_t *p = &sc->par;
  const shp_spec *sp = &sc->spec;
  p->action = FIT;
  p->speckle = 0;
  p->pos_smooth = (unsigned char)sp->pos_smooth;
  p->pos_scope = (unsigned char)sp->pos_scope;
  p->int_method = PERIODIC;
  p->bad_objfactor = 2.0;
  /* realize_photo's bounds, as read_par.c sets them by default */
  p->rad_R_max = p->rad_C_max = p->rad_rho_max = HUGENUMBER;
  p->opt_R_max = p->opt_A0_max = p->opt_D_max = p->opt_h_max = p->opt_B0_max = HUGENUMBER;
  p->opt_wt_max = 1.0;
  p->opt_k_min = -HUGENUMBER;
  p->opt_w_max = 1.0;
  p->opt_g_min = -1.0;
  p->opt_g_max = 1.0;
  p->opt_theta_max = PIE / 2;
  p->pos_pixels = 2 * sp->pos_n + 1;
  p->pos_width = sp->km_per_pixel * 2 * sp->pos_n;
  p->dd_gamma = 1.0;
  p->mask_tol = p->pos_width / (p->pos_pixels - 1);
  p->nsinc2 = sp->nsinc2 > 0 ? sp->nsinc2 : 1;
  p->sinc2width = sp->sinc2width > 0 ? sp->sinc2width : 6;
  p->exclude_seen = -1;
  p->mincosine_seen = 0.0;
  p->int_abstol = 1.0e-7;
  p->term_prec = 1.0e-6;
  p->npar_update = 20;
  for (int i = 0; i < MAXSUNMAGS; i++) p->sun_appmag[i] = -26.762;
}

shp_scene *shp_scene_create(const shp_spec *spec)
{
  shp_scene *sc = (shp_scene *)calloc(1, sizeof(shp_scene));
  sc->spec = *spec;
  build_par(sc);
  build_model(sc);
  build_data(sc);
  sc->par.nfpar = sc->nparams;
  sc->par.fpntr = (double **)tcalloc(sc, sc->nparams, sizeof(double *));
  sc->par.fpartype = (int *)tcalloc(sc, sc->nparams, sizeof(int));
  for (int i = 0; i < sc->nparams; i++) { sc->par.fpntr[i] = sc->params[i]; sc->par.fpartype[i] = sc->ptype[i]; }
  return sc;
}

void shp_scene_destroy(shp_scene *sc)
{
  if (!sc) return;
  for (int i = sc->nalloc - 1; i >= 0; i--) free(sc->allocs[i]);
  free(sc->allocs);
  free(sc);
}

struct par_t *shp_scene_par(shp_scene *sc) { return &sc->par; }
struct mod_t *shp_scene_mod(shp_scene *sc) { return &sc->mod; }
struct dat_t *shp_scene_dat(shp_scene *sc) { return &sc->dat; }
int shp_scene_nv(shp_scene *sc) { return sc->mod.shape.comp[0].real.nv; }
int shp_scene_nf(shp_scene *sc) { return sc->mod.shape.comp[0].real.nf; }
double *shp_param_ptr(shp_scene *sc, int i) { return (i >= 0 && i < sc->nparams) ? sc->params[i] : NULL; }
int shp_param_count(shp_scene *sc) { return sc->nparams; }
int shp_param_type(shp_scene *sc, int i) { return (i >= 0 && i < sc->nparams) ? sc->ptype[i] : 0; }

void shp_scene_synthesize_obs(shp_scene *sc, double snr, double lc_rel_sigma, unsigned long long seed)
{
  rng_t rn = rng_make(seed);
  struct dat_t *dat = &sc->dat;
  for (int s = 0; s < dat->nsets; s++) {
    struct set_t *set = &dat->set[s];
    if (set->type == DELAY) {
      for (int f = 0; f < set->desc.deldop.nframes; f++) {
        struct deldopfrm_t *fr = &set->desc.deldop.frame[f];
        double mx = 0;
        for (int i = 1; i <= fr->ndel; i++)
          for (int j = 1; j <= fr->ndop; j++) if (fr->fit[i][j] > mx) mx = fr->fit[i][j];
        double A = mx > 0 ? snr * fr->sdev / mx : 1.0;
        for (int i = 1; i <= fr->ndel; i++)
          for (int j = 1; j <= fr->ndop; j++) fr->obs[i][j] = A * fr->fit[i][j] + fr->sdev * rng_normal(&rn);
        fr->cal.val = A;
      }
    } else if (set->type == DOPPLER) {
      for (int f = 0; f < set->desc.doppler.nframes; f++) {
        struct dopfrm_t *fr = &set->desc.doppler.frame[f];
        double mx = 0;
        for (int j = 1; j <= fr->ndop; j++) if (fr->fit[j] > mx) mx = fr->fit[j];
        double A = mx > 0 ? snr * fr->sdev / mx : 1.0;
        for (int j = 1; j <= fr->ndop; j++) fr->obs[j] = A * fr->fit[j] + fr->sdev * rng_normal(&rn);
        fr->cal.val = A;
      }
    } else if (set->type == LGHTCRV) {
      struct lghtcrv_t *lc = &set->desc.lghtcrv;
      double mean = 0;
      for (int i = 1; i <= lc->n; i++) mean += lc->fit[i];
      mean /= lc->n;
      for (int i = 1; i <= lc->n; i++) {
        double sig = lc_rel_sigma * mean;
        lc->obs[i] = lc->fit[i] + sig * rng_normal(&rn);
        lc->oneovervar[i] = 1.0 / (sig * sig);
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* accessors                                                          */
/* ------------------------------------------------------------------ */
int shp_set_type(struct dat_t *dat, int s) { return dat->set[s].type; }
int shp_dat_nsets(struct dat_t *dat) { return dat->nsets; }

int shp_set_nframes(struct dat_t *dat, int s)
{
  struct set_t *set = &dat->set[s];
  switch (set->type) {
  case DELAY: return set->desc.deldop.nframes;
  case DOPPLER: return set->desc.doppler.nframes;
  case LGHTCRV: return set->desc.lghtcrv.ncalc;
  }
  return 0;
}

void shp_frame_dims(struct dat_t *dat, int s, int f, int out[2])
{
  struct set_t *set = &dat->set[s];
  out[0] = out[1] = 0;
  if (set->type == DELAY) { out[0] = set->desc.deldop.frame[f].ndel; out[1] = set->desc.deldop.frame[f].ndop; }
  else if (set->type == DOPPLER) { out[0] = 1; out[1] = set->desc.doppler.frame[f].ndop; }
  else if (set->type == LGHTCRV) { out[0] = set->desc.lghtcrv.n; out[1] = set->desc.lghtcrv.ncalc; }
}

void shp_frame_image(struct dat_t *dat, int s, int f, int which, double *out)
{
  struct set_t *set = &dat->set[s];
  if (set->type == DELAY) {
    struct deldopfrm_t *fr = &set->desc.deldop.frame[f];
    double **m = which == 0 ? fr->fit : which == 1 ? fr->obs : fr->oneovervar;
    for (int i = 1; i <= fr->ndel; i++)
      for (int j = 1; j <= fr->ndop; j++) *out++ = m[i][j];
  } else if (set->type == DOPPLER) {
    struct dopfrm_t *fr = &set->desc.doppler.frame[f];
    double *m = which == 0 ? fr->fit : which == 1 ? fr->obs : fr->oneovervar;
    for (int j = 1; j <= fr->ndop; j++) *out++ = m[j];
  } else if (set->type == LGHTCRV) {
    struct lghtcrv_t *lc = &set->desc.lghtcrv;
    if (which == 3) { for (int i = 1; i <= lc->ncalc; i++) *out++ = lc->y[i]; return; }
    double *m = which == 0 ? lc->fit : which == 1 ? lc->obs : lc->oneovervar;
    for (int i = 1; i <= lc->n; i++) *out++ = m[i];
  }
}

void shp_frame_set_image(struct dat_t *dat, int s, int f, int which, const double *in)
{
  struct set_t *set = &dat->set[s];
  if (set->type == DELAY) {
    struct deldopfrm_t *fr = &set->desc.deldop.frame[f];
    double **m = which == 0 ? fr->fit : which == 1 ? fr->obs : fr->oneovervar;
    for (int i = 1; i <= fr->ndel; i++)
      for (int j = 1; j <= fr->ndop; j++) m[i][j] = *in++;
  } else if (set->type == DOPPLER) {
    struct dopfrm_t *fr = &set->desc.doppler.frame[f];
    double *m = which == 0 ? fr->fit : which == 1 ? fr->obs : fr->oneovervar;
    for (int j = 1; j <= fr->ndop; j++) m[j] = *in++;
  } else if (set->type == LGHTCRV) {
    struct lghtcrv_t *lc = &set->desc.lghtcrv;
    double *m = which == 0 ? lc->fit : which == 1 ? lc->obs : lc->oneovervar;
    for (int i = 1; i <= lc->n; i++) m[i] = *in++;
  }
}

void shp_frame_scalars(struct dat_t *dat, int s, int f, double *out)
{
  struct set_t *set = &dat->set[s];
  memset(out, 0, sizeof(double) * 18);
  if (set->type == DELAY) {
    struct deldopfrm_t *fr = &set->desc.deldop.frame[f];
    out[0] = fr->chi2; out[1] = fr->cal.val; out[2] = fr->overflow_o2; out[3] = fr->overflow_m2;
    out[4] = fr->badradar_logfactor; out[5] = fr->dof;
    out[6] = fr->idellim[0]; out[7] = fr->idellim[1]; out[8] = fr->idoplim[0]; out[9] = fr->idoplim[1];
    out[10] = fr->dellim[0]; out[11] = fr->dellim[1]; out[12] = fr->doplim[0]; out[13] = fr->doplim[1];
    out[14] = fr->view[set->desc.deldop.v0].km2Hz; out[15] = fr->overflow_xsec;
    out[16] = fr->overflow_dopmean; out[17] = fr->weight;
  } else if (set->type == DOPPLER) {
    struct dopfrm_t *fr = &set->desc.doppler.frame[f];
    out[0] = fr->chi2; out[1] = fr->cal.val; out[2] = fr->overflow_o2; out[3] = fr->overflow_m2;
    out[4] = fr->badradar_logfactor; out[5] = fr->dof;
    out[8] = fr->idoplim[0]; out[9] = fr->idoplim[1];
    out[12] = fr->doplim[0]; out[13] = fr->doplim[1];
    out[14] = fr->view[set->desc.doppler.v0].km2Hz; out[15] = fr->overflow_xsec;
    out[16] = fr->overflow_dopmean; out[17] = fr->weight;
  } else if (set->type == LGHTCRV) {
    struct lghtcrv_t *lc = &set->desc.lghtcrv;
    out[0] = set->chi2; out[1] = lc->cal.val; out[5] = lc->dof; out[17] = lc->weight;
  }
}

struct pos_t *shp_frame_pos(struct dat_t *dat, int s, int f)
{
  struct set_t *set = &dat->set[s];
  if (set->type == DELAY) return &set->desc.deldop.frame[f].pos;
  if (set->type == DOPPLER) return &set->desc.doppler.frame[f].pos;
  if (set->type == LGHTCRV) return &set->desc.lghtcrv.rend[f + 1].pos;
  return NULL;
}

void shp_pos_array(struct pos_t *pos, int which, double *out)
{
  int n = pos->n;
  double **d = NULL;
  int **iv = NULL;
  switch (which) {
  case 0: d = pos->b; break;
  case 1: d = pos->cose; break;
  case 2: d = pos->cosi; break;
  case 3: d = pos->z; break;
  case 4: d = pos->cosill; break;
  case 5: d = pos->zill; break;
  case 10: iv = pos->f; break;
  case 11: iv = pos->body; break;
  case 12: iv = pos->comp; break;
  case 13: iv = pos->fill; break;
  default: return;
  }
  for (int i = -n; i <= n; i++)
    for (int j = -n; j <= n; j++) *out++ = d ? d[i][j] : (double)iv[i][j];
}

/* overwrite POS array `which` (numbering of shp_pos_array) from in[(2n+1)^2] */
void shp_pos_set_array(struct pos_t *pos, int which, const double *in)
{
  int n = pos->n;
  double **d = NULL;
  int **iv = NULL;
  switch (which) {
  case 0: d = pos->b; break;
  case 1: d = pos->cose; break;
  case 2: d = pos->cosi; break;
  case 3: d = pos->z; break;
  case 4: d = pos->cosill; break;
  case 5: d = pos->zill; break;
  case 10: iv = pos->f; break;
  case 11: iv = pos->body; break;
  case 12: iv = pos->comp; break;
  case 13: iv = pos->fill; break;
  default: return;
  }
  for (int i = -n; i <= n; i++)
    for (int j = -n; j <= n; j++) {
      if (d) d[i][j] = *in++;
      else iv[i][j] = (int)*in++;
    }
}

/* set the POS's observer and source frames (row-major 3x3) */
void shp_pos_set_frames(struct pos_t *pos, const double *oe, const double *se)
{
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      pos->oe[r][c] = oe[3 * r + c];
      pos->se[r][c] = se[3 * r + c];
    }
}

void shp_pos_copy(struct pos_t *dst, const struct pos_t *src)
{
  int n = src->n;
  double **dd[] = {dst->b, dst->cose, dst->cosi, dst->z, dst->cosill, dst->zill};
  double **sd[] = {src->b, src->cose, src->cosi, src->z, src->cosill, src->zill};
  int **di[] = {dst->f, dst->body, dst->comp, dst->fill, dst->bodyill, dst->compill};
  int **si[] = {src->f, src->body, src->comp, src->fill, src->bodyill, src->compill};
  for (int a = 0; a < 6; a++) {
    for (int i = -n; i <= n; i++) {
      if (dd[a] && sd[a]) memcpy(&dd[a][i][-n], &sd[a][i][-n], sizeof(double) * (2 * n + 1));
      if (di[a] && si[a]) memcpy(&di[a][i][-n], &si[a][i][-n], sizeof(int) * (2 * n + 1));
    }
  }
  for (int k = 0; k < 2; k++) {
    dst->xlim[k] = src->xlim[k]; dst->ylim[k] = src->ylim[k];
    dst->xlim2[k] = src->xlim2[k]; dst->ylim2[k] = src->ylim2[k];
  }
  memcpy(dst->ae, src->ae, sizeof(src->ae));
  memcpy(dst->oe, src->oe, sizeof(src->oe));
  memcpy(dst->se, src->se, sizeof(src->se));
  dst->bistatic = src->bistatic;
  dst->km_per_pixel = src->km_per_pixel;
  dst->posbnd_logfactor = src->posbnd_logfactor;
}

void shp_pos_scalars(struct pos_t *pos, double out[4])
{
  out[0] = pos->posbnd_logfactor; out[1] = pos->bistatic; out[2] = pos->km_per_pixel; out[3] = pos->n;
}

struct vertices_t *shp_mod_vertices(struct mod_t *mod) { return &mod->shape.comp[0].real; }

void *shp_set_desc(struct dat_t *dat, int s) { return &dat->set[s].desc; }

void shp_pos_limits(struct pos_t *pos, int out[8])
{
  out[0] = pos->xlim[0]; out[1] = pos->xlim[1]; out[2] = pos->ylim[0]; out[3] = pos->ylim[1];
  out[4] = pos->xlim2[0]; out[5] = pos->xlim2[1]; out[6] = pos->ylim2[0]; out[7] = pos->ylim2[1];
}

static void view_ptrs(struct dat_t *dat, int s, int f, int v, double (**ae)[3], double (**oe)[3],
                      double **spin, double **intspin, double **deloff, double **dopoff)
{
  struct set_t *set = &dat->set[s];
  *deloff = *dopoff = NULL;
  if (set->type == DELAY) {
    struct deldopview_t *w = &set->desc.deldop.frame[f].view[v];
    *ae = w->ae; *oe = w->oe; *spin = w->spin; *intspin = w->intspin; *deloff = &w->deloff; *dopoff = &w->dopoff;
  } else if (set->type == DOPPLER) {
    struct dopview_t *w = &set->desc.doppler.frame[f].view[v];
    *ae = w->ae; *oe = w->oe; *spin = w->spin; *intspin = w->intspin; *dopoff = &w->dopoff;
  } else {
    struct crvrend_t *w = &set->desc.lghtcrv.rend[f + 1];
    *ae = w->ae; *oe = w->oe; *spin = w->spin; *intspin = w->intspin;
  }
}

void shp_view_data(struct dat_t *dat, int s, int f, int v, double *out)
{
  double (*ae)[3], (*oe)[3], *spin, *intspin, *deloff, *dopoff;
  view_ptrs(dat, s, f, v, &ae, &oe, &spin, &intspin, &deloff, &dopoff);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) { out[3 * i + j] = ae[i][j]; out[9 + 3 * i + j] = oe[i][j]; }
  for (int k = 0; k < 3; k++) { out[18 + k] = spin[k]; out[21 + k] = intspin[k]; }
  out[24] = deloff ? *deloff : 0.0;
  out[25] = dopoff ? *dopoff : 0.0;
  out[26] = out[27] = out[28] = 0.0;
}

void shp_set_view_data(struct dat_t *dat, int s, int f, int v, const double *in)
{
  double (*ae)[3], (*oe)[3], *spin, *intspin, *deloff, *dopoff;
  view_ptrs(dat, s, f, v, &ae, &oe, &spin, &intspin, &deloff, &dopoff);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) { ae[i][j] = in[3 * i + j]; oe[i][j] = in[9 + 3 * i + j]; }
  for (int k = 0; k < 3; k++) { spin[k] = in[18 + k]; intspin[k] = in[21 + k]; }
  if (deloff) *deloff = in[24];
  if (dopoff) *dopoff = in[25];
}

void shp_geometry(struct mod_t *mod, double *vx, double *vn, double *fn, double *area)
{
  struct vertices_t *rv = &mod->shape.comp[0].real;
  for (int i = 0; i < rv->nv; i++)
    for (int k = 0; k < 3; k++) {
      if (vx) vx[3 * i + k] = rv->v[i].x[k];
      if (vn) vn[3 * i + k] = rv->v[i].n[k];
    }
  for (int f = 0; f < rv->nf; f++) {
    for (int k = 0; k < 3; k++) if (fn) fn[3 * f + k] = rv->f[f].n[k];
    if (area) area[f] = rv->f[f].area;
  }
}

/* angle (which = 0) or spin (which = 1) offset component j of dataset s:
   set its state and value (state 0 leaves the state), read them back */
void shp_set_offset(struct dat_t *dat, int s, int which, int j, int state, double val)
{
  struct param_t *p = which ? &dat->set[s].omegaoff[j] : &dat->set[s].angleoff[j];
  if (state) p->state = (char)state;
  p->val = val;
}
double shp_get_offset(struct dat_t *dat, int s, int which, int j)
{
  return which ? dat->set[s].omegaoff[j].val : dat->set[s].angleoff[j].val;
}

/* what calc_fits does around its per-dataset calls (calc_fits.c:349-361,
   :407-409): reset the flags, log factors and seen flags (start = 1), or
   normalize the summed log factors by the degrees of freedom (start = 0) */
void shp_fits_bracket(struct par_t *par, struct mod_t *mod, struct dat_t *dat, int start)
{
  if (start) {
    par->posbnd = par->badposet = par->badradar = 0;
    par->posbnd_logfactor = par->badposet_logfactor = par->badradar_logfactor = 0.0;
    for (int c = 0; c < mod->shape.ncomp; c++)
      for (int f = 0; f < mod->shape.comp[c].real.nf; f++) mod->shape.comp[c].real.f[f].seen = 0;
  } else {
    par->posbnd_logfactor /= dat->dof;
    par->badposet_logfactor /= dat->dof_poset;
    par->badradar_logfactor /= (dat->dof_deldop + dat->dof_doppler);
  }
}

/* overwrite the realized vertex normals of component 0 (3*nv doubles) */
void shp_set_vertex_normals(struct mod_t *mod, const double *vn)
{
  struct vertices_t *rv = &mod->shape.comp[0].real;
  for (int i = 0; i < rv->nv; i++)
    for (int k = 0; k < 3; k++) rv->v[i].n[k] = vn[3 * i + k];
}

void shp_facet_seen(struct mod_t *mod, unsigned char *out)
{
  struct vertices_t *rv = &mod->shape.comp[0].real;
  for (int f = 0; f < rv->nf; f++) out[f] = rv->f[f].seen;
}

/* frame constants: delcom_vig, dopcom_vig, dopDC_vig, epoch of view 0,
   del_per_pixel, dop_per_pixel (or dop_per_bin), Ftx, codemethod, dopfftlen */
void shp_frame_consts(struct dat_t *dat, int s, int f, double out[9])
{
  struct set_t *set = &dat->set[s];
  memset(out, 0, sizeof(double) * 9);
  if (set->type == DELAY) {
    struct deldop_t *d = &set->desc.deldop;
    struct deldopfrm_t *fr = &d->frame[f];
    out[0] = fr->delcom_vig; out[1] = fr->dopcom_vig; out[2] = fr->dopDC_vig; out[3] = fr->view[0].t;
    out[4] = d->del_per_pixel; out[5] = d->dop_per_pixel; out[6] = d->Ftx; out[7] = d->codemethod;
    out[8] = d->dopfftlen;
  } else if (set->type == DOPPLER) {
    struct doppler_t *d = &set->desc.doppler;
    struct dopfrm_t *fr = &d->frame[f];
    out[1] = fr->dopcom_vig; out[3] = fr->view[0].t; out[5] = d->dop_per_bin; out[6] = d->Ftx;
  } else if (set->type == LGHTCRV) {
    out[3] = set->desc.lghtcrv.x[f + 1];
  }
}

/* par knobs used by tests: 0 nsinc2, 1 sinc2width, 2 pos_smooth, 3 exclude_seen,
   4 mask_tol, 5 speckle, 6 dd_gamma */
void shp_par_set(struct par_t *par, int which, double v)
{
  switch (which) {
  case 0: par->nsinc2 = (int)v; break;
  case 1: par->sinc2width = (int)v; break;
  case 2: par->pos_smooth = (unsigned char)v; break;
  case 3: par->exclude_seen = (int)v; break;
  case 4: par->mask_tol = v; break;
  case 5: par->speckle = (unsigned char)v; break;
  case 6: par->dd_gamma = v; break;
  }
}

/* attached facets of vertex v (up to cap); returns the count */
int shp_vertex_facets(struct mod_t *mod, int v, int *out, int cap)
{
  struct vertex_t *vx = &mod->shape.comp[0].real.v[v];
  for (int q = 0; q < vx->naf && q < cap; q++) out[q] = vx->af[q];
  return vx->naf < cap ? vx->naf : cap;
}

/* cosine-law R and C of the radar law used by set s */
void shp_radar_law(struct mod_t *mod, struct dat_t *dat, int s, double out[2])
{
  struct set_t *set = &dat->set[s];
  int l = set->type == DELAY ? set->desc.deldop.iradlaw : set->desc.doppler.iradlaw;
  out[0] = mod->photo.radar[l].RC.R.val;
  out[1] = mod->photo.radar[l].RC.C.val;
}

/* Replace radar law l by one of the homogeneous laws (types 1-7), keeping
   the addresses of its fittable R and C.  p: cosine law [R, C]; quasispecular
   laws [R, C, cos_cutoff]; hybrids [diffuse R, C, quasispecular R, C,
   cos_cutoff].  Returns 0, or -1 for another type. */
int shp_set_radar_law(struct mod_t *mod, int l, int type, const double *p)
{
  if (l < 0 || l >= mod->photo.nradlaws) return -1;
  union radscat_t *r = &mod->photo.radar[l];
  switch (type) {
  case COSINELAW_DIFF:
    r->RC.R.val = p[0]; r->RC.C.val = p[1]; r->RC.R_save = p[0];
    break;
  case GAUSSIANLAW: case HAGFORSLAW: case COSINELAW_QS:
    r->quasispec.R.val = p[0]; r->quasispec.C.val = p[1]; r->quasispec.cos_cutoff = p[2];
    r->quasispec.R_save = p[0];
    break;
  case GAUSSIAN_COSINE: case HAGFORS_COSINE: case COSINE_COSINE:
    r->hybrid.diff.R.val = p[0]; r->hybrid.diff.C.val = p[1]; r->hybrid.diff.R_save = p[0];
    r->hybrid.qs.R.val = p[2]; r->hybrid.qs.C.val = p[3]; r->hybrid.qs.cos_cutoff = p[4];
    r->hybrid.qs.R_save = p[2];
    break;
  default:
    return -1;
  }
  mod->photo.radtype[l] = (unsigned char)type;
  return 0;
}

/* pointer to the model's photo_t (for radlaw / apply_photo calls) */
struct photo_t *shp_mod_photo(struct mod_t *mod) { return &mod->photo; }

void shp_par_flags(struct par_t *par, double out[6])
{
  out[0] = par->posbnd; out[1] = par->badradar; out[2] = par->baddiam;
  out[3] = par->posbnd_logfactor; out[4] = par->badradar_logfactor; out[5] = par->baddiam_logfactor;
}

/* ---- penalties and the spin quantities they read ----------------------- */

/* par->pen as read_par.c leaves it: n penalties, types and weights 1-based
   (penalties.c:145-1237 reads pen.type[i], pen.weight[i], i = 1..n, and
   writes pen.base[i]) */
int shp_scene_set_penalties(shp_scene *sc, int n, const int *type, const double *weight)
{
  struct pen_t *pen = &sc->par.pen;
  if (n < 0) return -1;
  pen->n = n;
  pen->type = (int *)tcalloc(sc, (size_t)n + 1, sizeof(int));
  pen->weight = (double *)tcalloc(sc, (size_t)n + 1, sizeof(double));
  pen->base = (double *)tcalloc(sc, (size_t)n + 1, sizeof(double));
  for (int i = 1; i <= n; i++) { pen->type[i] = type[i - 1]; pen->weight[i] = weight[i - 1]; }
  return 0;
}

void shp_pen_base(struct par_t *par, double *out)
{
  for (int i = 1; i <= par->pen.n; i++) out[i - 1] = par->pen.base[i];
}

void shp_set_spin_inertia(struct mod_t *mod, const double *in)
{
  for (int k = 0; k < 3; k++) mod->spin.inertia[k].val = in[k];
}

/* n spin impulses at epochs t[k] of (imp[3k], imp[3k+1], imp[3k+2]) rad/day */
int shp_set_impulses(struct mod_t *mod, int n, const double *t, const double *imp)
{
  if (n < 0 || n > MAXIMP) return -1;
  mod->spin.n_impulse = n;
  for (int k = 0; k < n; k++) {
    mod->spin.t_impulse[k] = t[k];
    for (int j = 0; j < 3; j++) { mod->spin.impulse[k][j].val = imp[3 * k + j]; mod->spin.impulse[k][j].state = 'c'; }
  }
  return 0;
}

/* area, volume, com[3], inertia[9] of the model (mode 0) or component c (mode 1 + c) */
void shp_moments(struct mod_t *mod, int mode, double out[14])
{
  const double *area, *volume, *com, *inertia;
  if (mode == 0) {
    area = &mod->shape.area; volume = &mod->shape.volume; com = mod->shape.com; inertia = &mod->shape.inertia[0][0];
  } else {
    struct comp_t *c = &mod->shape.comp[mode - 1];
    area = &c->area; volume = &c->volume; com = c->com; inertia = &c->inertia[0][0];
  }
  out[0] = *area;
  out[1] = *volume;
  for (int k = 0; k < 3; k++) out[2 + k] = com[k];
  for (int k = 0; k < 9; k++) out[5 + k] = inertia[k];
}

/* the harmonic component's coefficients, (l, m) with 0 <= m <= l <= L in
   order, a and b (b[l][0] is 0): which = 0 reads them, 1 writes them;
   returns L, or -1 for another component type */
int shp_harmonic(struct mod_t *mod, int which, double *a, double *b)
{
  struct comp_t *c = &mod->shape.comp[0];
  if (c->type != HARMONIC) return -1;
  const int L = c->desc.har.nhar;
  for (int l = 0, q = 0; l <= L; l++)
    for (int m = 0; m <= l; m++, q++) {
      if (which) {
        c->desc.har.a[l][m].val = a[q];
        if (m > 0) c->desc.har.b[l][m].val = b[q];
      } else {
        a[q] = c->desc.har.a[l][m].val;
        b[q] = m > 0 ? c->desc.har.b[l][m].val : 0.0;
      }
    }
  return L;
}

/* facet corners of the model's first component (3 per facet) */
void shp_facets(struct mod_t *mod, int *out)
{
  struct vertices_t *r = &mod->shape.comp[0].real;
  for (int f = 0; f < r->nf; f++)
    for (int k = 0; k < 3; k++) out[3 * f + k] = r->f[f].v[k];
}

/* ---- joint adjustments (vary_params) ------------------------------------ */

/* par's vary_delcor0, vary_radalb, vary_optalb, vary_dopscale (VARY_NONE 0 ...
   VARY_ALL 3) */
void shp_set_vary(struct par_t *par, int delcor0, int radalb, int optalb, int dopscale)
{
  par->vary_delcor0 = (unsigned char)delcor0;
  par->vary_radalb = (unsigned char)radalb;
  par->vary_optalb = (unsigned char)optalb;
  par->vary_dopscale = (unsigned char)dopscale;
}

/* calibration state and value of frame f of dataset s (the lightcurve's own
   for a lightcurve set, f ignored); which = 1: the Doppler scaling factor's
   state and value of a radar set; which = 2: the state of the delay
   polynomial's constant term; which = 3: the frame's (lightcurve's) weight */
void shp_set_state(struct dat_t *dat, int s, int f, int which, int state, double val)
{
  struct set_t *set = &dat->set[s];
  if (which == 0) {
    struct param_t *cal = set->type == DELAY     ? &set->desc.deldop.frame[f].cal
                          : set->type == DOPPLER ? &set->desc.doppler.frame[f].cal
                                                 : &set->desc.lghtcrv.cal;
    cal->state = (char)state;
    cal->val = val;
  } else if (which == 1) {
    struct param_t *d = set->type == DELAY ? &set->desc.deldop.dopscale : &set->desc.doppler.dopscale;
    d->state = (char)state;
    d->val = val;
    if (set->type == DELAY) set->desc.deldop.dopscale_save = val;
    else set->desc.doppler.dopscale_save = val;
  } else if (which == 2) {
    struct delcor_t *dc = set->type == DELAY ? &set->desc.deldop.delcor : &set->desc.doppler.delcor;
    dc->a[0].state = (char)state;
  } else {
    double *w = set->type == DELAY     ? &set->desc.deldop.frame[f].weight
                : set->type == DOPPLER ? &set->desc.doppler.frame[f].weight
                                       : &set->desc.lghtcrv.weight;
    *w = val;
  }
}

/* the weight sums read_dat.c forms while reading (:421-501, :686-692 and the
   Doppler and lightcurve readers): frames whose delay constant floats,
   absolutely calibrated frames and lightcurves, floating Doppler scales */
void shp_dat_vary_weights(struct dat_t *dat)
{
  dat->sum_deldop_zmax_weights = dat->sum_rad_xsec_weights = 0.0;
  dat->sum_opt_brightness_weights = dat->sum_cos_subradarlat_weights = 0.0;
  for (int s = 0; s < dat->nsets; s++) {
    struct set_t *set = &dat->set[s];
    if (set->type == DELAY) {
      struct deldop_t *dd = &set->desc.deldop;
      dd->sum_deldop_zmax_weights = dd->sum_rad_xsec_weights = dd->sum_cos_subradarlat_weights = 0.0;
      for (int f = 0; f < dd->nframes; f++) {
        if (dd->delcor.a[0].state != 'c') dd->sum_deldop_zmax_weights += dd->frame[f].weight;
        if (dd->frame[f].cal.state == 'c') dd->sum_rad_xsec_weights += dd->frame[f].weight;
        if (dd->dopscale.state != 'c') dd->sum_cos_subradarlat_weights += dd->frame[f].weight;
      }
      dat->sum_deldop_zmax_weights += dd->sum_deldop_zmax_weights;
      dat->sum_rad_xsec_weights += dd->sum_rad_xsec_weights;
      dat->sum_cos_subradarlat_weights += dd->sum_cos_subradarlat_weights;
    } else if (set->type == DOPPLER) {
      struct doppler_t *dp = &set->desc.doppler;
      dp->sum_rad_xsec_weights = dp->sum_cos_subradarlat_weights = 0.0;
      for (int f = 0; f < dp->nframes; f++) {
        if (dp->frame[f].cal.state == 'c') dp->sum_rad_xsec_weights += dp->frame[f].weight;
        if (dp->dopscale.state != 'c') dp->sum_cos_subradarlat_weights += dp->frame[f].weight;
      }
      dat->sum_rad_xsec_weights += dp->sum_rad_xsec_weights;
      dat->sum_cos_subradarlat_weights += dp->sum_cos_subradarlat_weights;
    } else if (set->type == LGHTCRV) {
      struct lghtcrv_t *lc = &set->desc.lghtcrv;
      lc->sum_opt_brightness_weights = lc->cal.state == 'c' ? lc->weight * lc->n : 0.0;
      dat->sum_opt_brightness_weights += lc->sum_opt_brightness_weights;
    }
  }
}

/* rotation (Euler angles, rad) and offset (km) of the model's first component */
void shp_set_comp_transform(struct mod_t *mod, const double *rot, const double *off)
{
  struct comp_t *c = &mod->shape.comp[0];
  for (int j = 0; j < 3; j++) { c->rot[j].val = rot[j]; c->off[j].val = off[j]; }
}

/* vertex deviations r.val of the model's first component */
void shp_set_vertex_r(struct mod_t *mod, const double *r)
{
  struct vertices_t *rv = &mod->shape.comp[0].real;
  for (int i = 0; i < rv->nv; i++) rv->v[i].r.val = r[i];
}

void shp_vertex_dirs(struct mod_t *mod, double *u, double *r)
{
  struct vertices_t *rv = &mod->shape.comp[0].real;
  for (int i = 0; i < rv->nv; i++) {
    for (int k = 0; k < 3; k++) u[3 * i + k] = rv->v[i].u[k];
    r[i] = rv->v[i].r.val;
  }
}

/* side list of the model's first component: per side v0, v1, f0, f1 */
int shp_sides(struct mod_t *mod, int *out, int cap)
{
  struct vertices_t *r = &mod->shape.comp[0].real;
  for (int s = 0; s < r->ns && s < cap; s++) {
    out[4 * s] = r->s[s].v[0]; out[4 * s + 1] = r->s[s].v[1];
    out[4 * s + 2] = r->s[s].f[0]; out[4 * s + 3] = r->s[s].f[1];
  }
  return r->ns;
}

/* ---- structs from other producers (the input-file readers) -------------- */

/* sizes for callers that allocate the ABI structs themselves: 0 par_t,
   1 mod_t, 2 dat_t */
size_t shp_struct_size(int which)
{
  return which == 0 ? sizeof(struct par_t) : which == 1 ? sizeof(struct mod_t) : sizeof(struct dat_t);
}

/* the file name a reader takes from mod_t / dat_t */
void shp_set_name(int which, void *p, const char *name)
{
  char *dst = which == 1 ? ((struct mod_t *)p)->name : ((struct dat_t *)p)->name;
  const size_t cap = which == 1 ? sizeof(((struct mod_t *)0)->name) : sizeof(((struct dat_t *)0)->name);
  snprintf(dst, cap, "%s", name);
}

/* realized vertex and facet counts of component 0 */
int shp_mod_nv(struct mod_t *mod) { return mod->shape.comp[0].real.nv; }
int shp_mod_nf(struct mod_t *mod) { return mod->shape.comp[0].real.nf; }

/* each vertex's attached sides, flattened, with counts: returns the total */
int shp_vertex_sides(struct mod_t *mod, int *count, int *out, int cap)
{
  struct vertices_t *r = &mod->shape.comp[0].real;
  int n = 0;
  for (int v = 0; v < r->nv; v++) {
    count[v] = r->v[v].nas;
    for (int q = 0; q < r->v[v].nas; q++, n++)
      if (n < cap) out[n] = r->v[v].as[q];
  }
  return n;
}

/* facets' side triples */
void shp_facet_sides(struct mod_t *mod, int *out)
{
  struct vertices_t *r = &mod->shape.comp[0].real;
  for (int f = 0; f < r->nf; f++)
    for (int k = 0; k < 3; k++) out[3 * f + k] = r->f[f].s[k];
}

/* a view's epoch and plane-of-sky spin contribution: t, orbspin[3] */
void shp_view_epoch(struct dat_t *dat, int s, int f, int v, double out[4])
{
  struct set_t *set = &dat->set[s];
  if (set->type == DELAY) {
    struct deldopview_t *w = &set->desc.deldop.frame[f].view[v];
    out[0] = w->t; out[1] = w->orbspin[0]; out[2] = w->orbspin[1]; out[3] = w->orbspin[2];
  } else if (set->type == DOPPLER) {
    struct dopview_t *w = &set->desc.doppler.frame[f].view[v];
    out[0] = w->t; out[1] = w->orbspin[0]; out[2] = w->orbspin[1]; out[3] = w->orbspin[2];
  } else {
    struct crvrend_t *w = &set->desc.lghtcrv.rend[f];
    out[0] = set->desc.lghtcrv.x[f]; out[1] = w->orbspin[0]; out[2] = w->orbspin[1]; out[3] = w->orbspin[2];
  }
}

/* The free-parameter slots of a model read from a file, numbered as the
   scene builder numbers its own: 0-2 component 0's size parameters (vertex
   scale factors, or 2a, a/b, b/c of an ellipse or ovoid, or harmonic scale
   factors), 3 the spin rate omega[2], 4-5 the first two Euler angles.
   NULL past the end. */
double *shp_mod_param_ptr(struct mod_t *mod, int i)
{
  struct comp_t *c = &mod->shape.comp[0];
  if (i >= 0 && i < 3) {
    switch (c->type) {
    case VERTEX: return &c->desc.ver.scalefactor[i].val;
    case HARMONIC: return &c->desc.har.scalefactor[i].val;
    case ELLIPSE: return i == 0 ? &c->desc.ell.two_a.val : i == 1 ? &c->desc.ell.a_over_b.val : &c->desc.ell.b_over_c.val;
    default: return i == 0 ? &c->desc.ovoid.two_a.val : i == 1 ? &c->desc.ovoid.a_over_b.val : &c->desc.ovoid.b_over_c.val;
    }
  }
  if (i == 3) return &mod->spin.omega[2].val;
  if (i == 4 || i == 5) return &mod->spin.angle[i - 4].val;
  return NULL;
}

/* A non-principal-axis rotator: spin vector w (rad/day, body frame, held
   constant: state 'c'), principal moments I, and par's int_method (BRUTE or
   PERIODIC); clears the PA flag. */
void shp_set_npa(struct par_t *par, struct mod_t *mod, const double w[3], const double I[3], int method)
{
  for (int k = 0; k < 3; k++) {
    mod->spin.omega[k].val = w[k];
    mod->spin.inertia[k].val = I[k];
    mod->spin.inertia[k].state = 'c';
  }
  mod->spin.omega[0].state = mod->spin.omega[1].state = 'c';
  mod->spin.pa = 0;
  par->int_method = (unsigned char)method;
}

/* spin epoch, Euler angles and spin vector: t0, angle[3], omega[3] */
void shp_spin_state(struct mod_t *mod, double out[7])
{
  out[0] = mod->spin.t0;
  for (int k = 0; k < 3; k++) { out[1 + k] = mod->spin.angle[k].val; out[4 + k] = mod->spin.omega[k].val; }
}

/* a per-facet law's resolved values of facet f (component 0), k of them:
   R (Lambert / Lommel), R C (harmcosine / inhocosine), w h B0 g theta
   (Hapke), R wt A0 D k (Kaasalainen); 0 for the other laws */
int shp_law_local(struct mod_t *mod, int radar, int ilaw, int f, double *out)
{
  const struct param_t *p = NULL;
  int k = 0;
  if (radar) {
    union radscat_t *r = &mod->photo.radar[ilaw];
    if (mod->photo.radtype[ilaw] == HARMCOSINE_DIFF) p = &r->harmcosine.local[0][f].R;
    else if (mod->photo.radtype[ilaw] == INHOCOSINE_DIFF) p = &r->inhocosine.local[0][f].R;
    k = 2;
  } else {
    union optscat_t *o = &mod->photo.optical[ilaw];
    switch (mod->photo.opttype[ilaw]) {
    case HARMLAMBERT: case HARMLOMMEL: p = &o->harmR.local[0][f].R; k = 1; break;
    case INHOLAMBERT: case INHOLOMMEL: p = &o->inhoR.local[0][f].R; k = 1; break;
    case HARMHAPKE: p = &o->harmhapke.local[0][f].w; k = 5; break;
    case INHOHAPKE: p = &o->inhohapke.local[0][f].w; k = 5; break;
    case HARMKAAS: p = &o->harmkaas.local[0][f].R; k = 5; break;
    case INHOKAAS: p = &o->inhokaas.local[0][f].R; k = 5; break;
    }
  }
  if (!p) return 0;
  for (int j = 0; j < k; j++) out[j] = p[j].val;
  return k;
}

/* a tabular law's values rho[0..n-1] (at most cap); returns n */
int shp_law_table(struct mod_t *mod, int ilaw, double *out, int cap)
{
  if (mod->photo.radtype[ilaw] != TABULARLAW) return 0;
  const struct tabular_t *t = &mod->photo.radar[ilaw].tabular;
  for (int i = 0; i < t->n && i < cap; i++) out[i] = t->rho[i].val;
  return t->n;
}

/* facet f of component 0: center x[3], theta, phi */
void shp_facet_geom(struct mod_t *mod, int f, double out[5])
{
  const struct facet_t *fa = &mod->shape.comp[0].real.f[f];
  for (int j = 0; j < 3; j++) out[j] = fa->x[j];
  out[3] = fa->theta;
  out[4] = fa->phi;
}

/* par's bad-photometry flag and log factor */
void shp_badphoto(struct par_t *par, double out[2])
{
  out[0] = par->badphoto;
  out[1] = par->badphoto_logfactor;
}

/* coefficient q of an expansion in the model file's order a00, a10, a11, b11, ... */
static struct param_t *shp_coef(struct harmonic_t *h, int q)
{
  for (int l = 0; l <= h->nhar; l++) {
    if (q == 0) return &h->a[l][0];
    q--;
    for (int m = 1; m <= l; m++) {
      if (q == 0) return &h->a[l][m];
      if (q == 1) return &h->b[l][m];
      q -= 2;
    }
  }
  return NULL;
}

/* address of a scattering law's parameter: a homogeneous law's which-th
   parameter (R C; w h B0 g theta; R wt A0 D k; a tabular law's rho[which]),
   a harmonic law's which-th coefficient of its first expansion, an
   inhomogeneous law's which-th global value (f < 0) or facet f's */
double *shp_law_param_ptr(struct mod_t *mod, int radar, int ilaw, int which, int f)
{
  struct param_t *p = NULL;
  if (radar) {
    union radscat_t *r = &mod->photo.radar[ilaw];
    switch (mod->photo.radtype[ilaw]) {
    case COSINELAW_DIFF: p = &r->RC.R + which; break;
    case TABULARLAW: p = &r->tabular.rho[which]; break;
    case HARMCOSINE_DIFF: p = shp_coef(&r->harmcosine.R, which); break;
    case INHOCOSINE_DIFF: p = (f < 0 ? &r->inhocosine.global.R : &r->inhocosine.local[0][f].R) + which; break;
    }
  } else {
    union optscat_t *o = &mod->photo.optical[ilaw];
    switch (mod->photo.opttype[ilaw]) {
    case GEOMETRICAL: case LAMBERTLAW: case LOMMEL: p = &o->R.R; break;
    case HAPKE: p = &o->hapke.w + which; break;
    case KAASALAINEN: p = &o->kaas.R + which; break;
    case HARMLAMBERT: case HARMLOMMEL: p = shp_coef(&o->harmR.R, which); break;
    case HARMHAPKE: p = shp_coef(&o->harmhapke.w, which); break;
    case HARMKAAS: p = shp_coef(&o->harmkaas.R, which); break;
    case INHOLAMBERT: case INHOLOMMEL: p = f < 0 ? &o->inhoR.global.R : &o->inhoR.local[0][f].R; break;
    case INHOHAPKE: p = (f < 0 ? &o->inhohapke.global.w : &o->inhohapke.local[0][f].w) + which; break;
    case INHOKAAS: p = (f < 0 ? &o->inhokaas.global.R : &o->inhokaas.local[0][f].R) + which; break;
    }
  }
  return p ? &p->val : NULL;
}
