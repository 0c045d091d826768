/*
 * io_dat.c -- the obs file and the data files it names (read_dat.c:409-2432):
 * delay-Doppler, Doppler, lightcurve and heliocurve datasets, their
 * ephemerides, observing geometry per view, vignetting by pixel-weighting
 * masks, the plane-of-sky frames (pos_scope) and the delay-correction and
 * Doppler-scale bookkeeping read_dat does before returning.
 */
#define _POSIX_C_SOURCE 200809L
#include <ctype.h>
#include <math.h>
#include <stdint.h>
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
  int set;
} datrd;

static void rp(datrd *r, struct param_t *p, const char *what)
{
  const int k = io_param(r->cx, r->f, p);
  if (k < 0) io_fail(r->cx, "read_dat: %s: end of file where %s was expected", r->f->name, what);
  r->nfpar += k;
}

static void read_ephem(datrd *r, struct ephem_t *e)
{
  e->n = io_int(r->cx, r->f);
  if (e->n < 1) io_fail(r->cx, "read_dat: set %d: an ephemeris needs at least one point", r->set);
  e->pnt = (struct ephpnt_t *)io_calloc(r->cx, (size_t)e->n, sizeof(struct ephpnt_t));
  for (int i = 0; i < e->n; i++) {
    e->pnt[i].t = io_caldate(r->cx, r->f);
    e->pnt[i].ra = io_double(r->cx, r->f) * D2R;
    e->pnt[i].dec = io_double(r->cx, r->f) * D2R;
    e->pnt[i].dist = io_double(r->cx, r->f);
  }
}

static void read_delcor(datrd *r, struct delcor_t *d)
{
  d->t0 = io_caldate(r->cx, r->f);
  d->n = io_int(r->cx, r->f);
  if (d->n < 0 || d->n >= MAXDELCORPAR) io_fail(r->cx, "read_dat: set %d: delay polynomial degree %d", r->set, d->n);
  d->a = (struct param_t *)io_calloc(r->cx, (size_t)d->n + 1, sizeof(struct param_t));
  d->equals_set = io_ivector(r->cx, 0, d->n);
  for (int i = 0; i <= d->n; i++) {
    rp(r, &d->a[i], "a delay polynomial coefficient");
    d->equals_set[i] = -1;
  }
  d->delcor0_save = d->a[0].val;
}

/* views per frame, their spacing (s) and whether the frame epoch is the
   centre or the first view */
static void read_smearing(datrd *r, int *nviews, double *interval, unsigned char *mode, int *v0)
{
  char w[IO_MAXLEN + 1];
  *nviews = io_int(r->cx, r->f);
  if (*nviews < 1) io_fail(r->cx, "read_dat: set %d: at least one view per frame", r->set);
  *interval = io_double(r->cx, r->f) / 86400;
  io_word(r->f, w, sizeof w);
  if (!strcmp(w, "center")) { *mode = SMEAR_CENTER; *v0 = *nviews / 2; }
  else if (!strcmp(w, "first")) { *mode = SMEAR_FIRST; *v0 = 0; }
  else io_fail(r->cx, "read_dat: set %d: unknown smearing mode '%s'", r->set, w);
}

static void read_name(datrd *r, char *dst, size_t cap, const char *what)
{
  char w[IO_MAXLEN + 1];
  if (!io_word(r->f, w, sizeof w)) io_fail(r->cx, "read_dat: set %d: end of file where %s was expected", r->set, what);
  if (strlen(w) >= cap) io_fail(r->cx, "read_dat: set %d: %s '%s' longer than %zu characters", r->set, what, w, cap - 1);
  strcpy(dst, w);
}

/* ---- pixel-weighting masks ------------------------------------------------ */
/* mask_SS_FF.txt (FFF when a set has more than 100 frames) in maskdir, or in
   the set's data directory (read_dat.c:642-652) */
static double *read_mask(datrd *r, const char *dir, int nframes, int frame, int nvals)
{
  char path[2 * IO_MAXLEN + 32];
  const char *base = r->par->maskdir[0] ? r->par->maskdir : dir;
  if (nframes > 100) snprintf(path, sizeof path, "%s/mask_%02d_%03d.txt", base, r->set, frame);
  else snprintf(path, sizeof path, "%s/mask_%02d_%02d.txt", base, r->set, frame);
  io_file m;
  io_open(r->cx, &m, path);
  const int n = io_count_words(&m);
  if (n != nvals) io_fail(r->cx, "read_dat: expected %d mask weights in %s, found %d", nvals, path, n);
  double *w = (double *)io_temp_alloc(r->cx, (size_t)nvals * sizeof(double));
  for (int i = 0; i < nvals; i++) w[i] = io_double(r->cx, &m);
  io_close(r->cx, &m);
  return w;
}

/* ---- data files ------------------------------------------------------------ */
enum { DF_ASCII, DF_BINARY, DF_RDF };

static float swap_f32(float x)
{
  uint32_t u;
  memcpy(&u, &x, 4);
  u = (u >> 24) | ((u >> 8) & 0xff00u) | ((u << 8) & 0xff0000u) | (u << 24);
  memcpy(&x, &u, 4);
  return x;
}

/* The type of a data file from its first 128 bytes (read_dat.c:797-820):
   any non-ASCII byte means raw binary; otherwise "SIMPLE" means FITS and
   "type" (any case) an RDF header. */
static int data_kind(datrd *r, io_file *d)
{
  char head[129];
  const size_t n = d->len < 128 ? d->len : 128;
  int binary = 0;
  for (size_t i = 0; i < n; i++)
    if ((unsigned char)d->buf[i] > 127) binary = 1;
  memcpy(head, d->buf, n);
  head[n] = '\0';
  if (binary) return DF_BINARY;
  if (strstr(head, "SIMPLE"))
    io_fail_code(r->cx, SHAPE_IO_EUNSUP, "read_dat: %s is a FITS file (FITS data are not read by this library)", d->name);
  for (size_t i = 0; i < n; i++) head[i] = (char)tolower((unsigned char)head[i]);
  return strstr(head, "type") ? DF_RDF : DF_ASCII;
}

static int host_little_endian(void)
{
  const int one = 1;
  return *(const char *)&one == 1;
}

/* nvals values of an image or spectrum, in file order, each rounded to FP32
   as the reference stores it (read_dat.c:2074-2256) */
static void read_values(datrd *r, io_file *d, float *out, size_t nvals)
{
  const int kind = data_kind(r, d);
  if (kind == DF_ASCII) {
    for (size_t i = 0; i < nvals; i++) out[i] = (float)io_double(r->cx, d);
    return;
  }
  size_t pos = 0;
  if (kind == DF_RDF) {
    /* header lines (up to 199 characters at a time) until one begins with '.' */
    for (;;) {
      if (pos >= d->len) io_fail(r->cx, "read_dat: can't read RDF data file %s", d->name);
      const size_t start = pos;
      while (pos < d->len && pos - start < 199 && d->buf[pos] != '\n') pos++;
      if (pos < d->len && d->buf[pos] == '\n' && pos - start < 199) pos++;
      if (d->buf[start] == '.') break;
    }
  }
  if (d->len - pos < 4 * nvals) io_fail(r->cx, "read_dat: error reading binary data file %s", d->name);
  const int swap = (host_little_endian() && r->par->endian == BIG_ENDIAN_DATA) ||
                   (!host_little_endian() && r->par->endian == LITTLE_ENDIAN_DATA);
  for (size_t i = 0; i < nvals; i++) {
    float v;
    memcpy(&v, d->buf + pos + 4 * i, 4);
    out[i] = swap ? swap_f32(v) : v;
  }
}

static void load_values(datrd *r, const char *path, float *out, size_t nvals)
{
  io_file d;
  io_open(r->cx, &d, path);
  read_values(r, &d, out, nvals);
  io_close(r->cx, &d);
}

/* ---- observing geometry of each view ---------------------------------------- */
/* view epoch t0 + (k - v0) dt; monostatic ephemeris geometry there, then at
   the light-time corrected epoch (read_dat.c:862-898) */
static void radar_view(datrd *r, const struct ephem_t *e, double tk, double *t, double oe[3][3], double orbspin[3])
{
  double se[3][3], ph, az;
  *t = tk;
  const double dist = io_ephem2mat(r->cx, e, e, *t, oe, se, orbspin, &ph, &az, 0);
  if (r->par->perform_ltc) {
    *t -= DAYSPERAU * dist;
    io_ephem2mat(r->cx, e, e, *t, oe, se, orbspin, &ph, &az, 0);
  }
}

static void frame_impulse_init(int *n_integrate, double *t_integrate, double (*impulse)[3])
{
  *n_integrate = -999;
  for (int n = 0; n < MAXIMP + 2; n++) {
    t_integrate[n] = -HUGENUMBER;
    for (int j = 0; j < 3; j++) impulse[n][j] = 0.0;
  }
}

/* ---- delay-Doppler (read_dat.c:528-919) -------------------------------------- */
static void read_deldop(datrd *r, struct deldop_t *dd, double *chi2_variance)
{
  io_ctx *cx = r->cx;
  io_file *f = r->f;
  char w[IO_MAXLEN + 1], path[2 * IO_MAXLEN + 2];
  dd->dof = 0.0;
  *chi2_variance = 0.0;
  dd->sum_deldop_zmax_weights = dd->sum_rad_xsec_weights = dd->sum_cos_subradarlat_weights = 0.0;
  dd->iradlaw = io_int(cx, f);
  if (dd->iradlaw < 0 || dd->iradlaw >= r->mod->photo.nradlaws)
    io_fail(cx, "read_dat: set %d: must have 0 <= radar scattering law <= %d", r->set, r->mod->photo.nradlaws - 1);
  read_ephem(r, &dd->astephem);
  dd->Ftx = io_double(cx, f);
  dd->ndel = io_int(cx, f);
  dd->del_per_pixel = io_double(cx, f);
  dd->spb = io_int(cx, f);
  dd->stride = io_int(cx, f);
  if (dd->ndel < 1 || dd->spb < 1 || dd->stride < 1 || dd->spb % dd->stride != 0)
    io_fail(cx, "read_dat: set %d: bad delay dimensions (stride must divide spb evenly)", r->set);
  io_word(f, w, sizeof w);
  if (!strcmp(w, "short")) dd->codemethod = SHORT;
  else if (!strcmp(w, "long1") || !strcmp(w, "long_orig")) dd->codemethod = LONG_ORIG;
  else if (!strcmp(w, "long2") || !strcmp(w, "long_mod")) dd->codemethod = LONG_MOD;
  else io_fail(cx, "read_dat: set %d: unknown code method '%s'", r->set, w);
  dd->ndop = io_int(cx, f);
  dd->dop_per_pixel = io_double(cx, f);
  dd->dopcom = io_double(cx, f);
  dd->dopDC = io_double(cx, f);
  dd->dopfftlen = io_int(cx, f);
  if (dd->ndop < 1) io_fail(cx, "read_dat: set %d: bad Doppler dimension", r->set);
  read_delcor(r, &dd->delcor);
  rp(r, &dd->dopscale, "the Doppler scaling factor");
  dd->dopscale_save = dd->dopscale.val;
  read_smearing(r, &dd->nviews, &dd->view_interval, &dd->smearing_mode, &dd->v0);
  read_name(r, dd->dir, sizeof dd->dir, "the data directory");
  dd->nframes = io_int(cx, f);
  if (dd->nframes < 1) io_fail(cx, "read_dat: set %d: no frames", r->set);
  dd->frame = (struct deldopfrm_t *)io_calloc(cx, (size_t)dd->nframes, sizeof(struct deldopfrm_t));
  const size_t nimg = (size_t)dd->ndel * dd->ndop;
  float *img = (float *)io_temp_alloc(cx, nimg * sizeof(float));
  for (int i = 0; i < dd->nframes; i++) {
    struct deldopfrm_t *fr = &dd->frame[i];
    fr->view = (struct deldopview_t *)io_calloc(cx, (size_t)dd->nviews, sizeof(struct deldopview_t));
    read_name(r, fr->name, sizeof fr->name, "a data file name");
    snprintf(path, sizeof path, "%s/%s", dd->dir, fr->name);
    fr->t0 = io_caldate(cx, f);
    fr->sdev = io_double(cx, f);
    io_param(cx, f, &fr->cal);
    if (fr->cal.val <= 0.0 && fr->cal.state == 'c')
      io_fail(cx, "read_dat: set %d: frame %d has constant calfact <= 0.0", r->set, i);
    fr->nlooks = io_double(cx, f);
    const double lookfact = fr->nlooks > 0 ? 1 / sqrt(fr->nlooks) : 0.0;
    fr->delcom = io_double(cx, f);
    fr->weight = io_double(cx, f);
    if (dd->delcor.a[0].state != 'c') dd->sum_deldop_zmax_weights += fr->weight;
    if (fr->cal.state == 'c') dd->sum_rad_xsec_weights += fr->weight;
    if (dd->dopscale.state != 'c') dd->sum_cos_subradarlat_weights += fr->weight;
    fr->pixels_weighted = io_int(cx, f);
    /* the frame's vignetted window: the mask's positive-weight bounding box */
    int idel[2] = {1, dd->ndel}, idop[2] = {1, dd->ndop};
    double dof = 0.0, *mask = NULL;
    if (fr->pixels_weighted) {
      mask = read_mask(r, dd->dir, dd->nframes, i, (int)nimg);
      idel[0] = dd->ndel + 1; idel[1] = 0; idop[0] = dd->ndop + 1; idop[1] = 0;
      for (int j = 1; j <= dd->ndel; j++)
        for (int k = 1; k <= dd->ndop; k++)
          if (mask[(size_t)(j - 1) * dd->ndop + (k - 1)] > 0.0) {
            dof += fr->weight;
            *chi2_variance += 2 * fr->weight * fr->weight;
            if (j < idel[0]) idel[0] = j;
            if (j > idel[1]) idel[1] = j;
            if (k < idop[0]) idop[0] = k;
            if (k > idop[1]) idop[1] = k;
          }
      if (idel[1] < idel[0]) io_fail(cx, "read_dat: set %d frame %d: the mask keeps no pixel", r->set, i);
    } else {
      dof = fr->weight * dd->ndel * dd->ndop;
      *chi2_variance += 2 * fr->weight * fr->weight * dd->ndel * dd->ndop;
    }
    fr->dof = dof;
    dd->dof += dof;
    fr->ndel = idel[1] - idel[0] + 1;
    fr->ndop = idop[1] - idop[0] + 1;
    fr->idelvig[0] = idel[0];
    fr->idelvig[1] = idel[1];
    fr->idopvig[0] = idop[0];
    fr->idopvig[1] = idop[1];
    fr->delcom_vig = fr->delcom - idel[0] + 1;
    fr->dopcom_vig = dd->dopcom - idop[0] + 1;
    fr->dopDC_vig = dd->dopDC - idop[0] + 1;
    fr->obs = io_matrix(cx, 1, fr->ndel, 1, fr->ndop);
    fr->fit = io_matrix(cx, 1, fr->ndel, 1, fr->ndop);
    fr->oneovervar = io_matrix(cx, 1, fr->ndel, 1, fr->ndop);
    load_values(r, path, img, nimg);
    const double var = pow(fr->sdev, 2.0);
    for (int j = 1; j <= fr->ndel; j++)
      for (int k = 1; k <= fr->ndop; k++) {
        const size_t src = (size_t)(j + idel[0] - 2) * dd->ndop + (k + idop[0] - 2);
        double o = img[src] * fr->sdev;
        if (r->par->dd_gamma != 1.0 && o > 0.0) o = pow(o, 1 / r->par->dd_gamma);
        fr->obs[j][k] = o;
        fr->oneovervar[j][k] = r->par->speckle ? 1 / (var + pow(lookfact * o, 2.0)) : 1 / var;
        if (fr->pixels_weighted) fr->oneovervar[j][k] *= mask[src];
      }
    io_temp_free(cx, mask);
    for (int k = 0; k < dd->nviews; k++)
      radar_view(r, &dd->astephem, fr->t0 + (k - dd->v0) * dd->view_interval, &fr->view[k].t,
                 fr->view[k].oe, fr->view[k].orbspin);
    frame_impulse_init(&fr->n_integrate, fr->t_integrate, fr->impulse);
  }
  io_temp_free(cx, img);
}

/* ---- Doppler (read_dat.c:921-1237) -------------------------------------------- */
static void read_doppler(datrd *r, struct doppler_t *dp, double *chi2_variance)
{
  io_ctx *cx = r->cx;
  io_file *f = r->f;
  char path[2 * IO_MAXLEN + 2];
  dp->dof = 0.0;
  *chi2_variance = 0.0;
  dp->sum_rad_xsec_weights = dp->sum_cos_subradarlat_weights = 0.0;
  dp->iradlaw = io_int(cx, f);
  if (dp->iradlaw < 0 || dp->iradlaw >= r->mod->photo.nradlaws)
    io_fail(cx, "read_dat: set %d: must have 0 <= radar scattering law <= %d", r->set, r->mod->photo.nradlaws - 1);
  read_ephem(r, &dp->astephem);
  dp->Ftx = io_double(cx, f);
  dp->ndop = io_int(cx, f);
  dp->dop_per_bin = io_double(cx, f);
  dp->dopcom = io_double(cx, f);
  if (dp->ndop < 1) io_fail(cx, "read_dat: set %d: bad Doppler dimension", r->set);
  read_delcor(r, &dp->delcor);
  if (dp->delcor.a[0].state == 'f')
    io_fail(cx, "read_dat: set %d: can't use \"f\" state for a Doppler dataset's 0th-order delay coefficient", r->set);
  rp(r, &dp->dopscale, "the Doppler scaling factor");
  dp->dopscale_save = dp->dopscale.val;
  read_smearing(r, &dp->nviews, &dp->view_interval, &dp->smearing_mode, &dp->v0);
  read_name(r, dp->dir, sizeof dp->dir, "the data directory");
  dp->nframes = io_int(cx, f);
  if (dp->nframes < 1) io_fail(cx, "read_dat: set %d: no frames", r->set);
  dp->frame = (struct dopfrm_t *)io_calloc(cx, (size_t)dp->nframes, sizeof(struct dopfrm_t));
  float *spec = (float *)io_temp_alloc(cx, (size_t)dp->ndop * sizeof(float));
  for (int i = 0; i < dp->nframes; i++) {
    struct dopfrm_t *fr = &dp->frame[i];
    fr->view = (struct dopview_t *)io_calloc(cx, (size_t)dp->nviews, sizeof(struct dopview_t));
    read_name(r, fr->name, sizeof fr->name, "a data file name");
    snprintf(path, sizeof path, "%s/%s", dp->dir, fr->name);
    fr->t0 = io_caldate(cx, f);
    fr->sdev = io_double(cx, f);
    io_param(cx, f, &fr->cal);
    if (fr->cal.val <= 0.0 && fr->cal.state == 'c')
      io_fail(cx, "read_dat: set %d: frame %d has constant calfact <= 0.0", r->set, i);
    fr->nlooks = io_double(cx, f);
    const double lookfact = fr->nlooks > 0 ? 1 / sqrt(fr->nlooks) : 0.0;
    fr->weight = io_double(cx, f);
    if (fr->cal.state == 'c') dp->sum_rad_xsec_weights += fr->weight;
    if (dp->dopscale.state != 'c') dp->sum_cos_subradarlat_weights += fr->weight;
    fr->pixels_weighted = io_int(cx, f);
    int idop[2] = {1, dp->ndop};
    double dof = 0.0, *mask = NULL;
    if (fr->pixels_weighted) {
      mask = read_mask(r, dp->dir, dp->nframes, i, dp->ndop);
      idop[0] = dp->ndop + 1; idop[1] = 0;
      for (int k = 1; k <= dp->ndop; k++)
        if (mask[k - 1] > 0.0) {
          dof += fr->weight;
          *chi2_variance += 2 * fr->weight * fr->weight;
          if (k < idop[0]) idop[0] = k;
          if (k > idop[1]) idop[1] = k;
        }
      if (idop[1] < idop[0]) io_fail(cx, "read_dat: set %d frame %d: the mask keeps no bin", r->set, i);
    } else {
      dof = fr->weight * dp->ndop;
      *chi2_variance += 2 * fr->weight * fr->weight * dp->ndop;
    }
    fr->dof = dof;
    dp->dof += dof;
    fr->ndop = idop[1] - idop[0] + 1;
    fr->idopvig[0] = idop[0];
    fr->idopvig[1] = idop[1];
    fr->dopcom_vig = dp->dopcom - idop[0] + 1;
    fr->obs = io_vector(cx, 1, fr->ndop);
    fr->fit = io_vector(cx, 1, fr->ndop);
    fr->oneovervar = io_vector(cx, 1, fr->ndop);
    load_values(r, path, spec, (size_t)dp->ndop);
    const double var = pow(fr->sdev, 2.0);
    for (int k = 1; k <= fr->ndop; k++) {
      const int src = k + idop[0] - 2;
      const double o = spec[src] * fr->sdev;
      fr->obs[k] = o;
      fr->oneovervar[k] = r->par->speckle ? 1 / (var + pow(lookfact * o, 2.0)) : 1 / var;
      if (fr->pixels_weighted) fr->oneovervar[k] *= mask[src];
    }
    io_temp_free(cx, mask);
    for (int k = 0; k < dp->nviews; k++)
      radar_view(r, &dp->astephem, fr->t0 + (k - dp->v0) * dp->view_interval, &fr->view[k].t,
                 fr->view[k].oe, fr->view[k].orbspin);
    frame_impulse_init(&fr->n_integrate, fr->t_integrate, fr->impulse);
  }
  io_temp_free(cx, spec);
}

/* ---- lightcurves (read_dat.c:1667-2072) ---------------------------------------- */
static int cmp_double(const void *a, const void *b)
{
  const double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

/* calculation epochs x0 (and light-time corrected x) with their geometry */
static void lc_calc_points(datrd *r, struct lghtcrv_t *lc, const double *epochs)
{
  io_ctx *cx = r->cx;
  const int n = lc->ncalc;
  lc->x0 = io_vector(cx, 1, n);
  lc->x = io_vector(cx, 1, n);
  lc->y = io_vector(cx, 1, n);
  lc->y2 = io_vector(cx, 1, n);
  lc->rotphase_calc = io_vector(cx, 1, n);
  lc->solar_phase = io_vector(cx, 1, n);
  lc->solar_azimuth = io_vector(cx, 1, n);
  lc->rend = (struct crvrend_t *)io_calloc(cx, (size_t)n + 1, sizeof(struct crvrend_t));
  for (int i = 1; i <= n; i++) lc->x[i] = lc->x0[i] = epochs[i - 1];
  for (int i = 1; i <= n; i++) {
    struct crvrend_t *rd = &lc->rend[i];
    const double dist = io_ephem2mat(cx, &lc->astephem, &lc->solephem, lc->x0[i], rd->oe, rd->se, rd->orbspin,
                                     &lc->solar_phase[i], &lc->solar_azimuth[i], 1);
    if (r->par->perform_ltc) {
      lc->x[i] = lc->x0[i] - DAYSPERAU * dist;
      io_ephem2mat(cx, &lc->astephem, &lc->solephem, lc->x[i], rd->oe, rd->se, rd->orbspin, &lc->solar_phase[i],
                   &lc->solar_azimuth[i], 1);
    }
  }
}

static void read_lghtcrv(datrd *r, struct lghtcrv_t *lc, int helio, double *chi2_variance)
{
  io_ctx *cx = r->cx;
  io_file *f = r->f;
  lc->ioptlaw = io_int(cx, f);
  if (lc->ioptlaw < 0 || lc->ioptlaw >= r->mod->photo.noptlaws)
    io_fail(cx, "read_dat: set %d: must have 0 <= optical scattering law <= %d", r->set, r->mod->photo.noptlaws - 1);
  if (!helio) {
    lc->astephem_sun.n = lc->obsephem_sun.n = -1;
    read_ephem(r, &lc->astephem);
    read_ephem(r, &lc->solephem);
  } else {
    /* target and observer seen from the Sun -> target and Sun seen by the
       observer (read_dat.c:1736-1803) */
    read_ephem(r, &lc->astephem_sun);
    read_ephem(r, &lc->obsephem_sun);
    const int n = lc->astephem_sun.n;
    if (lc->obsephem_sun.n != n)
      io_fail(cx, "read_dat: set %d: the two heliocentric ephemerides differ in length", r->set);
    lc->astephem.n = lc->solephem.n = n;
    lc->astephem.pnt = (struct ephpnt_t *)io_calloc(cx, (size_t)n, sizeof(struct ephpnt_t));
    lc->solephem.pnt = (struct ephpnt_t *)io_calloc(cx, (size_t)n, sizeof(struct ephpnt_t));
    for (int i = 0; i < n; i++) {
      const struct ephpnt_t *a = &lc->astephem_sun.pnt[i], *o = &lc->obsephem_sun.pnt[i];
      if (o->t != a->t) io_fail(cx, "read_dat: set %d: heliocentric ephemeris epochs differ", r->set);
      const double x = a->dist * cos(a->dec) * cos(a->ra) - o->dist * cos(o->dec) * cos(o->ra);
      const double y = a->dist * cos(a->dec) * sin(a->ra) - o->dist * cos(o->dec) * sin(o->ra);
      const double z = a->dist * sin(a->dec) - o->dist * sin(o->dec);
      struct ephpnt_t *e = &lc->astephem.pnt[i];
      e->t = a->t;
      e->dist = sqrt(x * x + y * y + z * z);
      e->ra = atan2(y, x);
      e->dec = asin(z / e->dist);
      struct ephpnt_t *s = &lc->solephem.pnt[i];
      s->t = o->t;
      s->ra = o->ra - PIE;
      if (s->ra < 0.0) s->ra += 2 * PIE;
      s->dec = -o->dec;
      s->dist = o->dist;
    }
  }
  /* calculation epochs: an explicit list (> 0), an even grid (0), or the
     observation epochs (< 0) */
  const int np = lc->ncalc_obsfile = io_int(cx, f);
  if (np != 0) lc->jdstart = lc->jdstop = lc->jdinterval = -HUGENUMBER;
  if (np >= 0) {
    double *ep;
    if (np > 0) {
      lc->ncalc = np;
      ep = (double *)io_calloc(cx, (size_t)np, sizeof(double));
      for (int i = 0; i < np; i++) ep[i] = io_double(cx, f);
      for (int i = 1; i < np; i++)
        if (ep[i] <= ep[i - 1]) io_fail(cx, "read_dat: set %d: calculation epochs must increase", r->set);
    } else {
      const double t1 = lc->jdstart = io_double(cx, f);
      const double t2 = lc->jdstop = io_double(cx, f);
      const double dt = lc->jdinterval = io_double(cx, f);
      if (t2 < t1 || dt <= 0) io_fail(cx, "read_dat: set %d: bad calculation epoch grid", r->set);
      lc->ncalc = ((int)floor((t2 - t1) / dt)) + 1;
      ep = (double *)io_calloc(cx, (size_t)lc->ncalc, sizeof(double));
      for (int i = 0; i < lc->ncalc; i++) ep[i] = t1 + i * dt;
    }
    lc_calc_points(r, lc, ep);
  }
  read_smearing(r, &lc->nviews, &lc->view_interval, &lc->smearing_mode, &lc->v0);
  lc->n = io_int(cx, f);
  if (lc->n < 1) io_fail(cx, "read_dat: set %d: a lightcurve needs at least one point", r->set);
  read_name(r, lc->name, sizeof lc->name, "a lightcurve file name");
  io_param(cx, f, &lc->cal);
  lc->weight = io_double(cx, f);
  lc->dof = lc->weight * lc->n;
  *chi2_variance = 2 * lc->weight * lc->weight * lc->n;
  lc->sum_opt_brightness_weights = lc->cal.state == 'c' ? lc->weight * lc->n : 0.0;
  lc->t0 = io_vector(cx, 1, lc->n);
  lc->t = io_matrix(cx, 1, lc->n, 0, lc->nviews - 1);
  lc->obs = io_vector(cx, 1, lc->n);
  lc->fit = io_vector(cx, 1, lc->n);
  lc->oneovervar = io_vector(cx, 1, lc->n);
  lc->rotphase_obs = io_vector(cx, 1, lc->n);
  /* each point: epoch (JD), magnitude, its rms error -> intensity relative to
     the Sun's and its inverse variance (read_dat.c:1933-1972) */
  if (lc->ioptlaw >= MAXSUNMAGS) io_fail(cx, "read_dat: set %d: optical law %d has no sun_appmag entry", r->set, lc->ioptlaw);
  const double sunmag = r->par->sun_appmag[lc->ioptlaw];
  io_file d;
  io_open(cx, &d, lc->name);
  for (int i = 1; i <= lc->n; i++) {
    const double epoch = io_double(cx, &d), mag = io_double(cx, &d), magerr = io_double(cx, &d);
    const double intens = exp(-0.4 * LN10 * (mag - sunmag));
    const double err = (0.4 * LN10 * magerr) * intens;
    lc->t0[i] = epoch;
    lc->obs[i] = intens;
    lc->oneovervar[i] = 1.0 / (err * err);
    for (int k = 0; k < lc->nviews; k++) {
      lc->t[i][k] = lc->t0[i] + (k - lc->v0) * lc->view_interval;
      if (r->par->perform_ltc) {
        double oe[3][3], se[3][3], spin[3], ph, az;
        const double dist =
            io_ephem2mat(cx, &lc->astephem, &lc->solephem, lc->t[i][k], oe, se, spin, &ph, &az, 1);
        lc->t[i][k] -= DAYSPERAU * dist;
      }
    }
  }
  if (!io_no_more_words(&d))
    io_fail(cx, "read_dat: set %d: %s holds more than %d points", r->set, lc->name, lc->n);
  io_close(cx, &d);
  if (np < 0) {
    /* the sorted distinct observation epochs */
    double *ep = (double *)io_calloc(cx, (size_t)lc->n, sizeof(double));
    for (int i = 0; i < lc->n; i++) ep[i] = lc->t0[i + 1];
    qsort(ep, (size_t)lc->n, sizeof(double), cmp_double);
    int u = 1;
    for (int i = 1; i < lc->n; i++)
      if (ep[i] > ep[u - 1]) ep[u++] = ep[i];
    lc->ncalc = u;
    lc_calc_points(r, lc, ep);
  }
  for (int i = 1; i <= lc->ncalc; i++)
    frame_impulse_init(&lc->rend[i].n_integrate, lc->rend[i].t_integrate, lc->rend[i].impulse);
}

/* ---- plane-of-sky frames (read_dat.c:1239-1390) ---------------------------------- */
static void pos_alloc(io_ctx *cx, struct pos_t *pos, int n, double kmpp, int bistatic)
{
  pos->n = n;
  pos->km_per_pixel = kmpp;
  pos->b = io_matrix(cx, -n, n, -n, n);
  pos->cosi = io_matrix(cx, -n, n, -n, n);
  pos->cose = io_matrix(cx, -n, n, -n, n);
  pos->z = io_matrix(cx, -n, n, -n, n);
  pos->body = io_imatrix(cx, -n, n, -n, n);
  pos->comp = io_imatrix(cx, -n, n, -n, n);
  pos->f = io_imatrix(cx, -n, n, -n, n);
  pos->cosill = io_matrix(cx, -n, n, -n, n);
  pos->zill = io_matrix(cx, -n, n, -n, n);
  pos->bodyill = io_imatrix(cx, -n, n, -n, n);
  pos->compill = io_imatrix(cx, -n, n, -n, n);
  pos->fill = io_imatrix(cx, -n, n, -n, n);
  pos->bistatic = bistatic;
}

/* GLOBAL: one frame shared by every radar frame and lightcurve point (each
   holds a copy of the struct); LOCAL: a frame of its own for each */
static void set_up_pos(io_ctx *cx, struct par_t *par, struct dat_t *dat)
{
  if (par->pos_pixels < 1) io_fail(cx, "read_dat: pos_pixels must be set (it is %d)", par->pos_pixels);
  const int n = (par->pos_pixels - 1) / 2;
  const double kmpp = par->pos_width / (2.0 * n);
  pos_alloc(cx, &dat->pos, n, kmpp, 1);
  if (par->pos_scope != GLOBAL && par->pos_scope != LOCAL) io_fail(cx, "read_dat: undefined pos_scope");
  const int global = par->pos_scope == GLOBAL;
  for (int s = 0; s < dat->nsets; s++) {
    struct set_t *st = &dat->set[s];
    if (st->type == DELAY) {
      for (int f = 0; f < st->desc.deldop.nframes; f++) {
        struct pos_t *p = &st->desc.deldop.frame[f].pos;
        if (global) *p = dat->pos; else pos_alloc(cx, p, n, kmpp, 0);
      }
    } else if (st->type == DOPPLER) {
      for (int f = 0; f < st->desc.doppler.nframes; f++) {
        struct pos_t *p = &st->desc.doppler.frame[f].pos;
        if (global) *p = dat->pos; else pos_alloc(cx, p, n, kmpp, 0);
      }
    } else if (st->type == LGHTCRV) {
      for (int i = 1; i <= st->desc.lghtcrv.ncalc; i++) {
        struct pos_t *p = &st->desc.lghtcrv.rend[i].pos;
        if (global) *p = dat->pos; else pos_alloc(cx, p, n, kmpp, 1);
      }
    }
  }
}

/* ---- delay correction and Doppler scale bookkeeping ------------------------------ */
/* initialize_delcor (read_dat.c:2374-2432): for each coefficient, the set an
   '=' coefficient copies; an '=' needs an earlier polynomial of the same
   degree and epoch.  Sets whose polynomial has fewer coefficients are skipped
   for the higher ones (the reference reads past their arrays). */
static struct delcor_t *set_delcor(struct set_t *st)
{
  if (st->type == DELAY) return &st->desc.deldop.delcor;
  if (st->type == DOPPLER) return &st->desc.doppler.delcor;
  return NULL;
}

static void initialize_delcor(io_ctx *cx, struct dat_t *dat)
{
  int nmax = -1;
  for (int s = 0; s < dat->nsets; s++) {
    const struct delcor_t *d = set_delcor(&dat->set[s]);
    if (d && d->n > nmax) nmax = d->n;
  }
  for (int i = 0; i <= nmax; i++) {
    int s_src = -1, n_src = -1;
    double t0_src = -1.0;
    for (int s = 0; s < dat->nsets; s++) {
      struct delcor_t *d = set_delcor(&dat->set[s]);
      if (!d || d->n < i) continue;
      if (d->a[i].state == '=') {
        if (s_src < 0) io_fail(cx, "read_dat: can't use \"=\" state for the first delay polynomial");
        if (d->n != n_src) io_fail(cx, "read_dat: delay polynomials must have the same degree if state = \"=\"");
        if (fabs(d->t0 - t0_src) > HALFSECOND) io_fail(cx, "read_dat: delay polynomials must have the same t0 if state = \"=\"");
      } else {
        s_src = s;
        n_src = d->n;
        t0_src = d->t0;
      }
      d->equals_set[i] = s_src;
    }
  }
}

/* realize_delcor(dat, 0, 0) (realize_delcor.c:50-160): '=' polynomials take
   the last explicit one's epoch and coefficients, then each view's delay
   and Doppler offsets in image bins (deldopoffs.c, dopoffs.c) */
static void delcor_offsets(io_ctx *cx, struct dat_t *dat)
{
  const struct delcor_t *src = NULL;
  for (int s = 0; s < dat->nsets; s++) {
    struct set_t *st = &dat->set[s];
    struct delcor_t *d = set_delcor(st);
    if (!d) continue;
    if (d->a[0].state != '=') {
      src = d;
      if (d->a[0].state == 'f') d->delcor0_save = d->a[0].val;
    } else {
      if (!src) io_fail(cx, "read_dat: can't use \"=\" state for the first delay polynomial");
      if (d->n != src->n) io_fail(cx, "read_dat: delay polynomials must have the same degree if state = \"=\"");
      if (fabs(d->t0 - src->t0) > HALFSECOND) io_fail(cx, "read_dat: delay polynomials must have the same t0 if state = \"=\"");
      d->t0 = src->t0;
      for (int i = 0; i <= d->n; i++) d->a[i].val = src->a[i].val;
    }
    if (st->type == DELAY) {
      struct deldop_t *dd = &st->desc.deldop;
      for (int f = 0; f < dd->nframes; f++)
        for (int k = 0; k < dd->nviews; k++) {
          double x = 1.0, dop = 0.0, del = d->a[0].val;
          const double arg = dd->frame[f].view[k].t - d->t0;
          for (int n = 1; n <= d->n; n++) {
            dop += n * d->a[n].val * x;
            del += d->a[n].val * (x *= arg);
          }
          dd->frame[f].view[k].deloff = del / dd->del_per_pixel;
          dd->frame[f].view[k].dopoff = -dop * dd->Ftx / (dd->dop_per_pixel * 86400.0);
        }
    } else {
      struct doppler_t *dp = &st->desc.doppler;
      for (int f = 0; f < dp->nframes; f++)
        for (int k = 0; k < dp->nviews; k++) {
          double x = 1.0, dop = 0.0;
          const double arg = dp->frame[f].view[k].t - d->t0;
          for (int n = 1; n <= d->n; n++) {
            dop += n * d->a[n].val * x;
            x *= arg;
          }
          dp->frame[f].view[k].dopoff = -dop * dp->Ftx / (dp->dop_per_bin * 86400.0);
        }
    }
  }
}

/* realize_dopscale(par, dat, 1, 0) (realize_dopscale.c:47-140): '=' factors
   copy the last explicit one; explicit factors outside
   (dopscale_min, dopscale_max) set the baddopscale flag and penalty */
static void dopscale_check(struct par_t *par, double v)
{
  if (v <= par->dopscale_min) { par->baddopscale = 1; par->baddopscale_logfactor += log(1 + par->dopscale_min - v); }
  if (v >= par->dopscale_max) { par->baddopscale = 1; par->baddopscale_logfactor += log(1 + v - par->dopscale_max); }
}

static void realize_dopscale(io_ctx *cx, struct par_t *par, struct dat_t *dat)
{
  par->baddopscale = 0;
  par->baddopscale_logfactor = 0.0;
  const struct param_t *src = NULL;
  for (int s = 0; s < dat->nsets; s++) {
    struct set_t *st = &dat->set[s];
    struct param_t *p;
    double *save;
    if (st->type == DELAY) { p = &st->desc.deldop.dopscale; save = &st->desc.deldop.dopscale_save; }
    else if (st->type == DOPPLER) { p = &st->desc.doppler.dopscale; save = &st->desc.doppler.dopscale_save; }
    else continue;
    if (p->state != '=') {
      src = p;
      if (p->state == 'f') *save = p->val;
      dopscale_check(par, p->val);
    } else {
      if (!src) io_fail(cx, "read_dat: can't use \"=\" state for the first (delay-)Doppler dataset's Doppler scale");
      p->val = src->val;
    }
  }
}

/* ---- the obs file ------------------------------------------------------------- */
static int dat_body(io_ctx *cx, struct par_t *par, struct mod_t *mod, struct dat_t *dat)
{
  char name[sizeof dat->name], w[IO_MAXLEN + 1];
  memcpy(name, dat->name, sizeof name);
  name[sizeof name - 1] = '\0';
  memset(dat, 0, sizeof *dat);
  memcpy(dat->name, name, sizeof name);
  io_file f;
  io_open(cx, &f, name);
  datrd r = {cx, &f, par, mod, 0, 0};
  dat->nsets = io_int(cx, &f);
  if (dat->nsets < 1) io_fail(cx, "read_dat: %s: no datasets", name);
  dat->set = (struct set_t *)io_calloc(cx, (size_t)dat->nsets, sizeof(struct set_t));
  for (int s = 0; s < dat->nsets; s++) {
    struct set_t *st = &dat->set[s];
    double chi2_variance = 0.0;
    r.set = s;
    st->inputnode = io_int(cx, &f);
    for (int i = 0; i < 3; i++) {
      rp(&r, &st->angleoff[i], "an angle offset");
      st->angleoff[i].val += 360.0 * floor((180.0 - st->angleoff[i].val) / 360.0);
      st->angleoff[i].val *= D2R;
    }
    for (int i = 0; i < 3; i++) {
      rp(&r, &st->omegaoff[i], "a spin offset");
      st->omegaoff[i].val *= D2R;
    }
    io_word(&f, w, sizeof w);
    if (!strcmp(w, "doppler")) {
      st->type = DOPPLER;
      read_doppler(&r, &st->desc.doppler, &chi2_variance);
      dat->dof_doppler += st->desc.doppler.dof;
      dat->sum_rad_xsec_weights += st->desc.doppler.sum_rad_xsec_weights;
      dat->sum_cos_subradarlat_weights += st->desc.doppler.sum_cos_subradarlat_weights;
    } else if (!strcmp(w, "delay-doppler")) {
      st->type = DELAY;
      read_deldop(&r, &st->desc.deldop, &chi2_variance);
      dat->dof_deldop += st->desc.deldop.dof;
      dat->sum_deldop_zmax_weights += st->desc.deldop.sum_deldop_zmax_weights;
      dat->sum_rad_xsec_weights += st->desc.deldop.sum_rad_xsec_weights;
      dat->sum_cos_subradarlat_weights += st->desc.deldop.sum_cos_subradarlat_weights;
    } else if (!strcmp(w, "lightcurve") || !strcmp(w, "heliocurve")) {
      st->type = LGHTCRV;
      read_lghtcrv(&r, &st->desc.lghtcrv, w[0] == 'h', &chi2_variance);
      dat->dof_lghtcrv += st->desc.lghtcrv.dof;
      dat->sum_opt_brightness_weights += st->desc.lghtcrv.sum_opt_brightness_weights;
    } else if (!strcmp(w, "plane-of-sky")) {
      io_fail_code(cx, SHAPE_IO_EUNSUP, "read_dat: set %d: plane-of-sky datasets are not read by this library", s);
    } else {
      io_fail(cx, "read_dat: can't handle type '%s' for dataset %d", w, s);
    }
    dat->chi2_variance += chi2_variance;
  }
  dat->dof = dat->dof_doppler + dat->dof_deldop + dat->dof_poset + dat->dof_lghtcrv;
  io_close(cx, &f);
  initialize_delcor(cx, dat);
  delcor_offsets(cx, dat);
  realize_dopscale(cx, par, dat);
  set_up_pos(cx, par, dat);
  return r.nfpar;
}

int shape_read_dat(struct par_t *par, struct mod_t *mod, struct dat_t *dat)
{
  io_ctx cx;
  memset(&cx, 0, sizeof cx);
  if (setjmp(cx.jb)) return io_unwind(&cx);
  shape_io_free_dat(dat);
  const int n = dat_body(&cx, par, mod, dat);
  io_arena_commit(&cx, dat);
  return n;
}

int read_dat(struct par_t *par, struct mod_t *mod, struct dat_t *dat)
{
  const int n = shape_read_dat(par, mod, dat);
  if (n < 0) {
    printf("%s\n", shape_io_error());
    fprintf(stderr, "ERROR: %s\n", shape_io_error());
    exit(2);
  }
  return n;
}
