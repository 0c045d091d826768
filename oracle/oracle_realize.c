/*
 * oracle_realize.c -- model realization, spin state and COM offsets.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdio.h>
#include <string.h>

#include "oracle_internal.h"

#define O_SMALLRATIO 0.01 /* realize_mod.c:138 */
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

/* realize_mod.c:153-177 -> realize_coordinates :183, VERTEX branch :444-467,
   coordinates :475-479, component rotation/offset :484-512, facets :532-546,
   vertex normals :552-563, check_surface for one component (:580-593: every
   vertex, facet and side active), compute_moments :757 (oracle_pen.c). */
void oracle_realize_mod(struct par_t *par, struct mod_t *mod)
{
  par->baddiam = 0;
  par->baddiam_logfactor = 0.0;
  for (int c = 0; c < mod->shape.ncomp; c++) {
    struct comp_t *comp = &mod->shape.comp[c];
    struct vertices_t *real = &comp->real;
    int nv = real->nv, nf = real->nf;
    if (comp->type != VERTEX) {
      oracle_set_error("oracle_realize_mod: only VERTEX components are restated");
      return;
    }
    for (int i = 0; i < nv; i++)
      if (real->v[i].r.state == '=') real->v[i].r.val = real->v[real->v[i].v_mirror].r.val;
    for (int j = 0; j <= 2; j++) {
      if (j > 0 && comp->desc.ver.scalefactor[j].state == '=')
        comp->desc.ver.scalefactor[j].val = comp->desc.ver.scalefactor[j - 1].val;
      real->scalefactor[j].val = comp->desc.ver.scalefactor[j].val;
      if (real->scalefactor[j].val <= O_SMALLRATIO) {
        par->baddiam = 1;
        par->baddiam_logfactor += log(1 + O_SMALLRATIO - real->scalefactor[j].val);
        real->scalefactor[j].val = O_SMALLRATIO / (1 + O_SMALLRATIO - real->scalefactor[j].val);
      }
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

/* inteuler.c:93-113: principal-axis rotator kinematics */
static int o_inteuler_pa(const struct spin_t *spin, const double t[], double imp[][3], int n,
                         double w[3], double m[3][3])
{
  if (!spin->pa) {
    oracle_set_error("oracle_realize_spin: NPA rotators are not restated");
    return 0;
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
static void o_spin_view(struct mod_t *mod, struct set_t *set, double t, double t_int[],
                        double imp[][3], int *n_int, double intspin[3], double ae[3][3],
                        const double orbspin[3], double spin[3])
{
  o_realize_impulse(&mod->spin, t, t_int, imp, n_int);
  o_inteuler_pa(&mod->spin, t_int, imp, *n_int, intspin, ae);
  for (int j = 0; j <= 2; j++) intspin[j] += set->omegaoff[j].val;
  o_cotrans(intspin, ae, intspin, -1);
  for (int j = 0; j <= 2; j++) spin[j] = orbspin[j] + intspin[j];
}

/* realize_spin.c:41-280 (angle/omega offsets realized with '=' states, :283-342) */
void oracle_realize_spin(struct par_t *par, struct mod_t *mod, struct dat_t *dat)
{
  (void)par;
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
          o_spin_view(mod, set, v->t, fr->t_integrate, fr->impulse, &fr->n_integrate, v->intspin,
                      v->ae, v->orbspin, v->spin);
        }
    } else if (set->type == DOPPLER) {
      struct doppler_t *dp = &set->desc.doppler;
      for (int f = 0; f < dp->nframes; f++)
        for (int k = 0; k < dp->nviews; k++) {
          struct dopfrm_t *fr = &dp->frame[f];
          struct dopview_t *v = &fr->view[k];
          o_spin_view(mod, set, v->t, fr->t_integrate, fr->impulse, &fr->n_integrate, v->intspin,
                      v->ae, v->orbspin, v->spin);
        }
    } else if (set->type == LGHTCRV) {
      struct lghtcrv_t *lc = &set->desc.lghtcrv;
      for (int i = 1; i <= lc->ncalc; i++) {
        struct crvrend_t *r = &lc->rend[i];
        o_spin_view(mod, set, lc->x[i], r->t_integrate, r->impulse, &r->n_integrate, r->intspin,
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
