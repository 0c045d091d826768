/*
 * oracle_internal.h -- small 3-vector / 3x3 helpers used by the oracle.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Each helper evaluates in the same operation order as the reference's
 * src/util routine it stands for (cotrans.c, mmmul.c, mtrnsps.c, dot.c,
 * cross.c, normalize.c, euler.c, iround.c), so that the oracle rounds the way
 * the reference does.
 */
#ifndef SHAPE_ORACLE_INTERNAL_H
#define SHAPE_ORACLE_INTERNAL_H

#include <math.h>

#include "oracle.h"

#define O_MIN(a, b) (((a) < (b)) ? (a) : (b))
#define O_MAX(a, b) (((a) > (b)) ? (a) : (b))

void oracle_set_error(const char *msg);

/* y = a x (dir = 1) or y = a^T x (dir = -1); y may alias x (util/cotrans.c) */
static inline void o_cotrans(double y[3], double a[3][3], const double x[3], int dir)
{
  double t[3];
  for (int i = 0; i < 3; i++) {
    t[i] = 0.0;
    for (int j = 0; j < 3; j++) t[i] += (dir == 1) ? a[i][j] * x[j] : a[j][i] * x[j];
  }
  y[0] = t[0]; y[1] = t[1]; y[2] = t[2];
}

/* x = y z (util/mmmul.c) */
static inline void o_mmmul(double x[3][3], double y[3][3], double z[3][3])
{
  double t[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      t[i][j] = 0.0;
      for (int k = 0; k < 3; k++) t[i][j] += y[i][k] * z[k][j];
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) x[i][j] = t[i][j];
}

/* a = b^T (util/mtrnsps.c) */
static inline void o_mtrnsps(double a[3][3], double b[3][3])
{
  double t[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t[i][j] = b[j][i];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = t[i][j];
}

static inline double o_dot(const double x[3], const double y[3])
{
  return x[0] * y[0] + x[1] * y[1] + x[2] * y[2];
}

/* z = x cross y, returns |z| (util/cross.c) */
static inline double o_cross(double z[3], const double x[3], const double y[3])
{
  double t0 = x[1] * y[2] - x[2] * y[1];
  double t1 = x[2] * y[0] - x[0] * y[2];
  double t2 = x[0] * y[1] - x[1] * y[0];
  z[0] = t0; z[1] = t1; z[2] = t2;
  return sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
}

/* normalize in place, return the original length (util/normalize.c) */
static inline double o_normalize(double u[3])
{
  double norm = 0.0;
  for (int i = 0; i < 3; i++) norm += u[i] * u[i];
  norm = sqrt(norm);
  if (norm != 0.0)
    for (int i = 0; i < 3; i++) u[i] /= norm;
  return norm;
}

/* Euler angles to rotation matrix (util/euler.c) */
static inline void o_euler2mat(double m[3][3], double phi, double theta, double psi)
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

/* round half away from zero (util/iround.c) */
static inline int o_iround(double x) { return x < 0.0 ? (int)(x - 0.5) : (int)(x + 0.5); }

/* sinc^2 and triangle responses (macros/func.h:6-8) */
static inline double o_sinc2(double x)
{
  double t = fabs(3.141592653589793 * x);
  if (t > 1.0e-4) { t = sin(t) / t; return t * t; }
  return 1.0;
}
static inline double o_tri(double x) { double t = 1 - fabs(x); return t > 0.0 ? t : 0.0; }
static inline double o_tri2(double x) { double t = 1 - fabs(x); return t > 0.0 ? t * t : 0.0; }

/* per-thread accumulators for the par_t bad-model log factors (the serial
   path adds straight into par; the threaded baseline gives each thread its own) */
struct o_flags {
  int posbnd, badradar;
  double posbnd_logfactor, badradar_logfactor;
};

void oracle_calc_deldop(struct par_t *par, struct mod_t *mod, struct deldop_t *deldop, int s,
                        struct o_flags *fl);
void oracle_calc_doppler(struct par_t *par, struct mod_t *mod, struct doppler_t *doppler, int s,
                         struct o_flags *fl);
void oracle_calc_lghtcrv(struct par_t *par, struct mod_t *mod, struct lghtcrv_t *lghtcrv, int s,
                         struct o_flags *fl);

#endif
