/*
 * io_core.c -- allocation records, whole-file loading, the token grammar,
 * calendar dates, ephemeris interpolation and the side / attached-facet set-up
 * of a realized component, shared by the par, mod and obs readers.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "io_internal.h"

/* ---- errors ------------------------------------------------------------ */
static __thread char io_last_error[512];

const char *shape_io_error(void) { return io_last_error; }

void io_fail(io_ctx *cx, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(cx->err, sizeof cx->err, fmt, ap);
  va_end(ap);
  cx->code = SHAPE_IO_EPARSE;
  longjmp(cx->jb, 1);
}

void io_fail_code(io_ctx *cx, int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(cx->err, sizeof cx->err, fmt, ap);
  va_end(ap);
  cx->code = code;
  longjmp(cx->jb, 1);
}

/* ---- allocation records ------------------------------------------------ */
static void arena_push(io_ctx *cx, void *p)
{
  io_arena *a = &cx->arena;
  if (a->n == a->cap) {
    size_t cap = a->cap ? 2 * a->cap : 64;
    void **q = (void **)realloc(a->p, cap * sizeof(void *));
    if (!q) { free(p); io_fail(cx, "out of host memory"); }
    a->p = q;
    a->cap = cap;
  }
  a->p[a->n++] = p;
}

void *io_calloc(io_ctx *cx, size_t n, size_t sz)
{
  void *p = calloc(n ? n : 1, sz ? sz : 1);
  if (!p) io_fail(cx, "out of host memory (%zu x %zu bytes)", n, sz);
  arena_push(cx, p);
  return p;
}

/* Numerical Recipes offset arrays (nr/nrutil.c): element [nl] is the first */
double *io_vector(io_ctx *cx, int nl, int nh)
{
  double *v = (double *)io_calloc(cx, (size_t)(nh - nl + 1), sizeof(double));
  return v - nl;
}

int *io_ivector(io_ctx *cx, int nl, int nh)
{
  int *v = (int *)io_calloc(cx, (size_t)(nh - nl + 1), sizeof(int));
  return v - nl;
}

/* row pointers into one zeroed block; [nrl][ncl] is the first element */
double **io_matrix(io_ctx *cx, int nrl, int nrh, int ncl, int nch)
{
  const size_t nr = (size_t)(nrh - nrl + 1), nc = (size_t)(nch - ncl + 1);
  double **m = (double **)io_calloc(cx, nr, sizeof(double *));
  double *blk = (double *)io_calloc(cx, nr * nc, sizeof(double));
  for (size_t i = 0; i < nr; i++) m[i] = blk + i * nc - ncl;
  return m - nrl;
}

int **io_imatrix(io_ctx *cx, int nrl, int nrh, int ncl, int nch)
{
  const size_t nr = (size_t)(nrh - nrl + 1), nc = (size_t)(nch - ncl + 1);
  int **m = (int **)io_calloc(cx, nr, sizeof(int *));
  int *blk = (int *)io_calloc(cx, nr * nc, sizeof(int));
  for (size_t i = 0; i < nr; i++) m[i] = blk + i * nc - ncl;
  return m - nrl;
}

/* Allocations of successful calls, by root struct.  A later call under the
   same root adds to that root's list. */
typedef struct root_rec {
  const void *root;
  io_arena arena;
  struct root_rec *next;
} root_rec;

static root_rec *roots;
static pthread_mutex_t roots_mu = PTHREAD_MUTEX_INITIALIZER;

void io_arena_commit(io_ctx *cx, const void *root)
{
  pthread_mutex_lock(&roots_mu);
  root_rec *r = roots;
  while (r && r->root != root) r = r->next;
  if (!r) {
    r = (root_rec *)calloc(1, sizeof *r);
    if (!r) { pthread_mutex_unlock(&roots_mu); io_arena_discard(cx); return; }
    r->root = root;
    r->next = roots;
    roots = r;
  }
  io_arena *a = &r->arena;
  if (a->n + cx->arena.n > a->cap) {
    size_t cap = a->cap ? a->cap : 64;
    while (cap < a->n + cx->arena.n) cap *= 2;
    void **q = (void **)realloc(a->p, cap * sizeof(void *));
    if (q) { a->p = q; a->cap = cap; }
  }
  if (a->n + cx->arena.n <= a->cap) {
    memcpy(a->p + a->n, cx->arena.p, cx->arena.n * sizeof(void *));
    a->n += cx->arena.n;
    free(cx->arena.p);
    memset(&cx->arena, 0, sizeof cx->arena);
  }
  pthread_mutex_unlock(&roots_mu);
}

void io_arena_discard(io_ctx *cx)
{
  for (size_t i = 0; i < cx->arena.n; i++) free(cx->arena.p[i]);
  free(cx->arena.p);
  memset(&cx->arena, 0, sizeof cx->arena);
}

static void free_root(const void *root)
{
  pthread_mutex_lock(&roots_mu);
  root_rec **pp = &roots;
  while (*pp && (*pp)->root != root) pp = &(*pp)->next;
  root_rec *r = *pp;
  if (r) *pp = r->next;
  pthread_mutex_unlock(&roots_mu);
  if (!r) return;
  for (size_t i = 0; i < r->arena.n; i++) free(r->arena.p[i]);
  free(r->arena.p);
  free(r);
}

void shape_io_free_par(struct par_t *par) { free_root(par); }
void shape_io_free_mod(struct mod_t *mod) { free_root(mod); }
void shape_io_free_dat(struct dat_t *dat) { free_root(dat); }

int io_unwind(io_ctx *cx)
{
  io_arena_discard(cx);
  for (int i = 0; i < 16; i++) { free(cx->bufs[i]); cx->bufs[i] = NULL; }
  snprintf(io_last_error, sizeof io_last_error, "%s", cx->err);
  return cx->code ? cx->code : SHAPE_IO_EPARSE;
}

/* ---- files ------------------------------------------------------------- */
static int slurp(FILE *fp, char **buf, size_t *len)
{
  size_t cap = 1 << 16, n = 0;
  char *b = (char *)malloc(cap + 1);
  if (!b) return -1;
  for (;;) {
    if (n == cap) {
      cap *= 2;
      char *q = (char *)realloc(b, cap + 1);
      if (!q) { free(b); return -1; }
      b = q;
    }
    size_t got = fread(b + n, 1, cap - n, fp);
    n += got;
    if (got == 0) break;
  }
  if (ferror(fp)) { free(b); return -1; }
  b[n] = '\0';
  *buf = b;
  *len = n;
  return 0;
}

void io_open(io_ctx *cx, io_file *f, const char *path)
{
  memset(f, 0, sizeof *f);
  snprintf(f->name, sizeof f->name, "%s", path);
  const size_t L = strlen(path);
  FILE *fp;
  int piped = 0;
  if (L > 3 && !strcmp(path + L - 3, ".gz")) {
    /* the reference copies and gunzips to a temporary file (read_dat.c:774-789) */
    char cmd[2 * IO_MAXLEN + 32];
    for (const char *c = path; *c; c++)
      if (*c == '\'') io_fail(cx, "can't open compressed file %s (quote in name)", path);
    snprintf(cmd, sizeof cmd, "gzip -dc '%s'", path);
    fp = popen(cmd, "r");
    piped = 1;
  } else {
    fp = fopen(path, "rb");
  }
  if (!fp) io_fail_code(cx, SHAPE_IO_EFILE, "can't open file %s: %s", path, strerror(errno));
  int rc = slurp(fp, &f->buf, &f->len);
  int st = piped ? pclose(fp) : fclose(fp);
  if (rc) io_fail_code(cx, SHAPE_IO_EFILE, "can't read file %s", path);
  int slot = 0;
  while (slot < 16 && cx->bufs[slot]) slot++;
  if (slot == 16) { free(f->buf); f->buf = NULL; io_fail(cx, "too many open files in one read"); }
  cx->bufs[slot] = f->buf;
  if (piped && st != 0) io_fail_code(cx, SHAPE_IO_EFILE, "can't decompress file %s", path);
}

void io_close(io_ctx *cx, io_file *f)
{
  io_temp_free(cx, f->buf);
  f->buf = NULL;
}

void *io_temp_alloc(io_ctx *cx, size_t bytes)
{
  int slot = 0;
  while (slot < 16 && cx->bufs[slot]) slot++;
  if (slot == 16) io_fail(cx, "too many temporaries in one read");
  char *p = (char *)calloc(bytes ? bytes : 1, 1);
  if (!p) io_fail(cx, "out of host memory (%zu bytes)", bytes);
  cx->bufs[slot] = p;
  return p;
}

void io_temp_free(io_ctx *cx, void *p)
{
  if (!p) return;
  for (int i = 0; i < 16; i++)
    if (cx->bufs[i] == p) cx->bufs[i] = NULL;
  free(p);
}

/* ---- tokens ------------------------------------------------------------ */
static int fgetc_m(io_file *f) { return f->pos < f->len ? (unsigned char)f->buf[f->pos++] : EOF; }

/* skip whitespace and {comments}; returns the first character of a word or EOF */
static int skip_blank(io_file *f, int *unclosed)
{
  int c;
  for (;;) {
    c = fgetc_m(f);
    if (c == EOF) return EOF;
    if (c == '{') {
      while ((c = fgetc_m(f)) != '}')
        if (c == EOF) { *unclosed = 1; return EOF; }
      continue;
    }
    if (!isspace(c)) return c;
  }
}

int io_word(io_file *f, char *w, size_t cap)
{
  int unclosed = 0;
  int c = skip_blank(f, &unclosed);
  size_t n = 0;
  if (c == EOF) { w[0] = '\0'; return 0; }
  do {
    if (n + 1 < cap) w[n++] = (char)c;
    c = fgetc_m(f);
    if (c == '{') { f->pos--; break; } /* the comment belongs to the next scan */
  } while (c != EOF && !isspace(c));
  w[n] = '\0';
  return 1;
}

int io_str_to_int(io_ctx *cx, const char *s)
{
  char *end;
  errno = 0;
  long v = strtol(s, &end, 10);
  if (*s == '\0' || *end != '\0') io_fail(cx, "expected a base-10 integer, got '%s'", s);
  if (errno == ERANGE || v < INT32_MIN || v > INT32_MAX) io_fail(cx, "integer '%s' out of range", s);
  return (int)v;
}

double io_str_to_double(io_ctx *cx, const char *s)
{
  char *end;
  double v = strtod(s, &end);
  if (*s == '\0' || *end != '\0') io_fail(cx, "expected a floating-point number, got '%s'", s);
  return v;
}

int io_int(io_ctx *cx, io_file *f)
{
  char w[IO_MAXLEN + 1];
  if (!io_word(f, w, sizeof w)) io_fail(cx, "%s: end of file where an integer was expected", f->name);
  return io_str_to_int(cx, w);
}

double io_double(io_ctx *cx, io_file *f)
{
  char w[IO_MAXLEN + 1];
  if (!io_word(f, w, sizeof w)) io_fail(cx, "%s: end of file where a number was expected", f->name);
  return io_str_to_double(cx, w);
}

int io_param(io_ctx *cx, io_file *f, struct param_t *p)
{
  char w[IO_MAXLEN + 1];
  if (!io_word(f, w, sizeof w)) return -1;
  if (!strcmp(w, "c")) p->state = 'c';
  else if (!strcmp(w, "f")) p->state = 'f';
  else if (!strcmp(w, "=")) p->state = '=';
  else io_fail(cx, "%s: expected a parameter state c, f or =, got '%s'", f->name, w);
  p->val = io_double(cx, f);
  return p->state == 'f';
}

double io_caldate(io_ctx *cx, io_file *f)
{
  int c[6];
  for (int i = 0; i < 6; i++) c[i] = io_int(cx, f);
  return io_cal2jd(c[0], c[1], c[2], c[3], c[4], c[5]);
}

void io_restofline(io_file *f, char *line, size_t cap)
{
  size_t n = 0;
  int c;
  while ((c = fgetc_m(f)) != EOF && c != '\n')
    if (n + 1 < cap) line[n++] = (char)c;
  line[n] = '\0';
  /* comments on the line, each opened and closed on it (read_par.c:1884) */
  char *o;
  while ((o = strchr(line, '{')) != NULL) {
    char *e = strchr(o, '}');
    if (!e) break;
    memmove(o, e + 1, strlen(e + 1) + 1);
  }
}

int io_count_words(io_file *f)
{
  const size_t save = f->pos;
  char w[IO_MAXLEN + 1];
  int n = 0;
  while (io_word(f, w, sizeof w)) n++;
  f->pos = save;
  return n;
}

int io_no_more_words(io_file *f)
{
  const size_t save = f->pos;
  char w[8];
  int more = io_word(f, w, sizeof w);
  f->pos = save;
  return !more;
}

/* ---- calendar ---------------------------------------------------------- */
/* Julian day number of a calendar date at noon: Gregorian from 1582 October
   15, Julian before (the calendar reform), by the standard integer form of
   the day count (Fliegel & Van Flandern 1968 for the Gregorian branch). */
static long jdn(int y, int m, int d)
{
  if (y < 0) y++; /* no year zero: 1 BC is year -1 */
  const long a = (14 - m) / 12;
  const long yy = y + 4800 - a, mm = m + 12 * a - 3;
  const int gregorian = (y > 1582) || (y == 1582 && (m > 10 || (m == 10 && d >= 15)));
  if (gregorian)
    return d + (153 * mm + 2) / 5 + 365 * yy + yy / 4 - yy / 100 + yy / 400 - 32045;
  return d + (153 * mm + 2) / 5 + 365 * yy + yy / 4 - 32083;
}

/* jdcal.c:10-14: the day number at noon, less half a day, plus the time */
double io_cal2jd(int yy, int mo, int dd, int hh, int mi, int ss)
{
  double jd = (double)jdn(yy, mo, dd) - 0.5;
  jd += ((hh * 60 + mi) * 60 + ss) / 86400.0;
  return jd;
}

double shape_io_cal2jd(int yy, int mo, int dd, int hh, int mi, int ss)
{
  return io_cal2jd(yy, mo, dd, hh, mi, ss);
}

/* ---- vectors ----------------------------------------------------------- */
static void v_unit(double u[3])
{
  const double n = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  if (n != 0.0)
    for (int i = 0; i < 3; i++) u[i] /= n;
}

static void v_cross(double z[3], const double x[3], const double y[3])
{
  const double a = x[1] * y[2] - x[2] * y[1];
  const double b = x[2] * y[0] - x[0] * y[2];
  const double c = x[0] * y[1] - x[1] * y[0];
  z[0] = a; z[1] = b; z[2] = c;
}

/* Equatorial -> ecliptic (J2000): a rotation about x by the obliquity
   epsilon_0 = 23 deg 26' 21.406" (IAU 2006); its cosine and sine to 16
   digits, the values of astro/eq2ec.c:18-19. */
static void eq_to_ecl(double u[3])
{
  const double ce = 0.9174821430652418, se = 0.3977769691126060;
  const double y = u[1] * ce + u[2] * se;
  u[2] = u[2] * ce - u[1] * se;
  u[1] = y;
}

/* z-x-z Euler matrix, the convention of util/euler.c */
static void euler_matrix(double m[3][3], double phi, double theta, double psi)
{
  const double ct = cos(theta), st = sin(theta), cp = cos(phi), sp = sin(phi);
  const double cs = cos(psi), ss = sin(psi);
  m[0][0] = cs * cp - ct * sp * ss;
  m[0][1] = cs * sp + ct * cp * ss;
  m[0][2] = ss * st;
  m[1][0] = -ss * cp - ct * sp * cs;
  m[1][1] = -ss * sp + ct * cp * cs;
  m[1][2] = cs * st;
  m[2][0] = st * sp;
  m[2][1] = -st * cp;
  m[2][2] = ct;
}

/* The interval of an ephemeris holding t and the linear interpolation of
   RA, Dec and distance there, RA rate taken across the 0/360 line
   (ephem2mat.c:51-95).  Fails outside the ephemeris. */
typedef struct { double ra, dec, dist, rarate, decrate; } eph_at;

static eph_at ephem_at(io_ctx *cx, const struct ephem_t *e, double t, const char *who)
{
  if (e->n < 2 || !e->pnt) io_fail(cx, "ephem2mat: the %s ephemeris needs at least two points", who);
  if (e->pnt[0].t > t)
    io_fail(cx, "ephem2mat: epoch %.5f before the %s ephemeris (starts %.5f)", t, who, e->pnt[0].t);
  int i = 0;
  while (e->pnt[i + 1].t < t) {
    ++i;
    if (i == e->n - 1)
      io_fail(cx, "ephem2mat: epoch %.5f after the %s ephemeris (ends %.5f)", t, who, e->pnt[i].t);
  }
  const double span = e->pnt[i + 1].t - e->pnt[i].t, dt = t - e->pnt[i].t;
  const double twopi = 2 * PIE;
  const double ra0 = e->pnt[i].ra, ra1 = e->pnt[i + 1].ra, dra = ra1 - ra0;
  eph_at r;
  if (dra > PIE) r.rarate = (ra1 - (ra0 + twopi)) / span;
  else if (dra < -PIE) r.rarate = ((ra1 + twopi) - ra0) / span;
  else r.rarate = dra / span;
  r.ra = e->pnt[i].ra + r.rarate * dt;
  r.ra -= twopi * floor(r.ra / twopi);
  r.decrate = (e->pnt[i + 1].dec - e->pnt[i].dec) / span;
  r.dec = e->pnt[i].dec + r.decrate * dt;
  r.dist = e->pnt[i].dist + (e->pnt[i + 1].dist - e->pnt[i].dist) / span * dt;
  return r;
}

double io_ephem2mat(io_ctx *cx, const struct ephem_t *ast, const struct ephem_t *sun, double t,
                    double oe[3][3], double se[3][3], double s[3], double *solar_phase,
                    double *solar_azimuth, int bistatic)
{
  const eph_at a = ephem_at(cx, ast, t, "asteroid");
  const double cd = cos(a.dec), sd = sin(a.dec), cr = cos(a.ra), sr = sin(a.ra);
  /* z: target -> observer; x: celestial pole x z (west); y = z x x (north) */
  double z[3] = {-cd * cr, -cd * sr, -sd}, x[3] = {cd * sr, -cd * cr, 0.0}, y[3];
  double ee[3];
  for (int k = 0; k < 3; k++) ee[k] = z[k] * a.dist;
  v_unit(x);
  v_cross(y, z, x);
  eq_to_ecl(x);
  eq_to_ecl(y);
  eq_to_ecl(z);
  for (int k = 0; k < 3; k++) { oe[0][k] = x[k]; oe[1][k] = y[k]; oe[2][k] = z[k]; }
  /* plane-of-sky spin: d(z)/dt x z, the rate taken in equatorial axes */
  double zdot[3] = {a.decrate * sd * cr + a.rarate * cd * sr,
                    a.decrate * sd * sr - a.rarate * cd * cr,
                    -a.decrate * cd};
  eq_to_ecl(zdot);
  v_cross(s, zdot, oe[2]);
  if (!bistatic) {
    *solar_phase = 0.0;
    *solar_azimuth = 0.0;
    return a.dist;
  }
  const eph_at b = ephem_at(cx, sun, t, "solar");
  double ss[3] = {cos(b.dec) * cos(b.ra) * b.dist + ee[0],
                  cos(b.dec) * sin(b.ra) * b.dist + ee[1],
                  sin(b.dec) * b.dist + ee[2]};
  eq_to_ecl(ss);
  v_unit(ss);
  euler_matrix(se, atan2(ss[1], ss[0]) + PIE / 2, acos(ss[2]), 0.0);
  double so[3];
  for (int i = 0; i < 3; i++) so[i] = oe[i][0] * ss[0] + oe[i][1] * ss[1] + oe[i][2] * ss[2];
  v_unit(so);
  *solar_phase = acos(so[2]);
  if (so[0] == 0.0 && so[1] == 0.0) {
    *solar_azimuth = 0.0;
  } else {
    *solar_azimuth = atan2(so[1], so[0]) - PIE / 2;
    if (*solar_azimuth < 0.0) *solar_azimuth += 2 * PIE;
  }
  return a.dist;
}

double shape_io_ephem2mat(const struct ephem_t *ast, const struct ephem_t *sun, double t,
                          double oe[3][3], double se[3][3], double s[3], double *solar_phase,
                          double *solar_azimuth, int bistatic)
{
  io_ctx cx;
  memset(&cx, 0, sizeof cx);
  if (setjmp(cx.jb)) {
    io_unwind(&cx);
    return -1.0;
  }
  return io_ephem2mat(&cx, ast, sun, t, oe, se, s, solar_phase, solar_azimuth, bistatic);
}

/* ---- sides and attached facets of a realized component ----------------- */
/* Sides are numbered by first appearance over the facets in order, each
   facet's corners taken sorted (a < b < c) as the pairs (a,b), (a,c), (b,c);
   a side's f[0] is the lower-numbered facet; each facet's s[] ascends; each
   vertex lists its sides and its facets in ascending order
   (read_mod.c:2386-2596).  A hash of the vertex pair replaces the
   reference's scan of the side list, which is quadratic in the facets. */
static void *setup_alloc(io_ctx *cx, size_t n, size_t sz)
{
  if (cx) return io_calloc(cx, n, sz);
  void *p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "setupsides: out of host memory\n"); exit(2); }
  return p;
}

static void setup_fail(io_ctx *cx, const char *msg, int f, const int v[3])
{
  if (cx) io_fail(cx, "facet %d (vertices %d %d %d): %s", f, v[0], v[1], v[2], msg);
  fprintf(stderr, "Facet %d (vertices %d %d %d) %s\nERROR: routine setupsides\n", f, v[0], v[1], v[2], msg);
  exit(2);
}

void io_setup_sides(io_ctx *cx, struct vertices_t *vt)
{
  const int nf = vt->nf, nv = vt->nv;
  vt->ns = (3 * nf) / 2;
  vt->s = (struct side_t *)setup_alloc(cx, (size_t)vt->ns, sizeof(struct side_t));
  size_t cap = 16;
  while (cap < 4 * (size_t)(vt->ns + 1)) cap <<= 1;
  long long *key = (long long *)malloc(cap * sizeof(long long));
  int *sid = (int *)malloc(cap * sizeof(int));
  if (!key || !sid) { free(key); free(sid); setup_fail(cx, "out of host memory", 0, (int[3]){0, 0, 0}); }
  for (size_t i = 0; i < cap; i++) sid[i] = -1;
  int ns = 0;
  for (int s = 0; s < vt->ns; s++) vt->s[s].f[0] = vt->s[s].f[1] = -1;
  for (int f = 0; f < nf; f++) {
    int v[3] = {vt->f[f].v[0], vt->f[f].v[1], vt->f[f].v[2]};
    for (int k = 0; k < 3; k++)
      if (v[k] < 0 || v[k] >= nv) { free(key); free(sid); setup_fail(cx, "has out-of-bounds vertex numbers", f, v); }
    if (v[0] == v[1] || v[1] == v[2] || v[2] == v[0]) {
      free(key); free(sid);
      setup_fail(cx, "does not have three different corners", f, v);
    }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2 - i; j++)
        if (v[j] > v[j + 1]) { int t = v[j]; v[j] = v[j + 1]; v[j + 1] = t; }
    const int pa[3] = {v[0], v[0], v[1]}, pb[3] = {v[1], v[2], v[2]};
    int got[3];
    for (int k = 0; k < 3; k++) {
      const long long kk = (long long)pa[k] * nv + pb[k];
      size_t h = (size_t)(((unsigned long long)kk * 0x9E3779B97F4A7C15ULL) >> 17) & (cap - 1);
      while (sid[h] >= 0 && key[h] != kk) h = (h + 1) & (cap - 1);
      if (sid[h] < 0) {
        if (ns >= vt->ns) {
          free(key); free(sid);
          if (cx) io_fail(cx, "more than 3F/2 = %d sides: the component is not a closed surface", vt->ns);
          fprintf(stderr, "Should find %d sides for component with %d vertices, found more\n", vt->ns, nv);
          exit(2);
        }
        key[h] = kk;
        sid[h] = ns;
        vt->s[ns].v[0] = pa[k];
        vt->s[ns].v[1] = pb[k];
        ns++;
      }
      const int s = sid[h];
      if (vt->s[s].f[0] < 0) vt->s[s].f[0] = f;
      else vt->s[s].f[1] = f;
      got[k] = s;
    }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2 - i; j++)
        if (got[j] > got[j + 1]) { int t = got[j]; got[j] = got[j + 1]; got[j + 1] = t; }
    for (int k = 0; k < 3; k++) vt->f[f].s[k] = got[k];
  }
  free(key);
  free(sid);
  if (ns != vt->ns) {
    if (cx) io_fail(cx, "found %d sides where a closed surface of %d facets has %d", ns, nf, vt->ns);
    fprintf(stderr, "Should find %d sides for component with %d vertices, found %d instead\n", vt->ns, nv, ns);
    exit(2);
  }
}

void io_setup_vertices(io_ctx *cx, struct vertices_t *vt)
{
  for (int v = 0; v < vt->nv; v++) vt->v[v].naf = vt->v[v].nas = 0;
  for (int s = 0; s < vt->ns; s++) { vt->v[vt->s[s].v[0]].nas++; vt->v[vt->s[s].v[1]].nas++; }
  for (int f = 0; f < vt->nf; f++)
    for (int k = 0; k < 3; k++) vt->v[vt->f[f].v[k]].naf++;
  for (int v = 0; v < vt->nv; v++) {
    vt->v[v].as = (int *)setup_alloc(cx, (size_t)vt->v[v].nas, sizeof(int));
    vt->v[v].af = (int *)setup_alloc(cx, (size_t)vt->v[v].naf, sizeof(int));
    vt->v[v].nas = vt->v[v].naf = 0;
  }
  for (int s = 0; s < vt->ns; s++)
    for (int k = 0; k < 2; k++) {
      struct vertex_t *x = &vt->v[vt->s[s].v[k]];
      x->as[x->nas++] = s;
    }
  for (int f = 0; f < vt->nf; f++)
    for (int k = 0; k < 3; k++) {
      struct vertex_t *x = &vt->v[vt->f[f].v[k]];
      x->af[x->naf++] = f;
    }
}

void setupsides(struct vertices_t *vt) { io_setup_sides(NULL, vt); }
void setupvertices(struct vertices_t *vt) { io_setup_vertices(NULL, vt); }
