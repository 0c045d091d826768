/*
 * oracle_npa.c -- the spin state of a non-principal-axis (NPA) rotator:
 * Euler's equations for torque-free rotation integrated from the spin epoch
 * to a view's epoch, spin impulses added between intervals
 * (inteuler.c:114-392).  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * State: the body-frame spin w (rad/day) and the ecliptic-to-body matrix m,
 *   dw0/dt = (I1 - I2)/I0 w1 w2, dw1/dt = (I2 - I0)/I1 w2 w0,
 *   dw2/dt = (I0 - I1)/I2 w0 w1,
 *   d(column j of m)/dt = (column j) x w   (inteuler.c:118-151).
 *
 * Integrator (not the reference's adaptive odeint / bsstep): fixed
 * Bulirsch-Stoer macro steps -- Gragg's modified midpoint rule with
 * n = 2, 4, ..., 12 substeps, polynomially extrapolated in h^2 to h = 0
 * (Aitken-Neville) -- with the macro step count set by the angle swept,
 * ceil(|w| sqrt(Imax/Imin) |dt| / 0.25).  Its error is ~1e-13 on the state;
 * the reference integrates to par->int_abstol (1e-7 by default), so the two
 * agree to that tolerance.  The device integrates with the same steps.
 *
 * Two methods, as par->int_method selects (inteuler.c:95-112):
 *   BRUTE    -- integrate the whole interval (intnpa1, inteuler.c:156-205);
 *   PERIODIC -- integrate one period T of w in the body frame, read the
 *               rotation psiT about the angular momentum L it leaves, integrate
 *               only the remainder dt = t - t0 - N T, and compose N psiT about
 *               L (intnpa2, inteuler.c:208-391).  T = 4 K(k) sqrt(I1 I2 I3 /
 *               ((I3 - I2)(M^2 - 2E I1))) with the moments sorted (largest and
 *               smallest exchanged for a long-axis mode), k^2 as Landau &
 *               Lifshitz sec. 37 give it; K by the arithmetic-geometric mean
 *               to full precision (the reference calls Numerical Recipes'
 *               cel, an iteration stopped at a fixed tolerance).
 */
#include <math.h>

#include "oracle_internal.h"

#define NPA_LEVELS 6 /* modified-midpoint substeps 2, 4, ..., 12 */
#define NPA_MAX_STEPS 200000 /* longest span (macro steps) integrated */

static void npa_derivs(const double I[3], const double y[12], double d[12])
{
  const double w0 = y[0], w1 = y[1], w2 = y[2];
  d[0] = ((I[1] - I[2]) / I[0]) * w1 * w2;
  d[1] = ((I[2] - I[0]) / I[1]) * w2 * w0;
  d[2] = ((I[0] - I[1]) / I[2]) * w0 * w1;
  for (int j = 0; j < 3; j++) {
    const double *x = y + 3 + 3 * j;
    double *dx = d + 3 + 3 * j;
    dx[0] = x[1] * w2 - x[2] * w1;
    dx[1] = x[2] * w0 - x[0] * w2;
    dx[2] = x[0] * w1 - x[1] * w0;
  }
}

/* Gragg's modified midpoint rule over H with n substeps */
static void npa_midpoint(const double I[3], const double y[12], double H, int n, double out[12])
{
  const double h = H / n;
  double z0[12], z1[12], d[12];
  npa_derivs(I, y, d);
  for (int i = 0; i < 12; i++) { z0[i] = y[i]; z1[i] = y[i] + h * d[i]; }
  for (int k = 1; k < n; k++) {
    npa_derivs(I, z1, d);
    for (int i = 0; i < 12; i++) {
      const double z2 = z0[i] + 2.0 * h * d[i];
      z0[i] = z1[i];
      z1[i] = z2;
    }
  }
  npa_derivs(I, z1, d);
  for (int i = 0; i < 12; i++) out[i] = 0.5 * (z1[i] + z0[i] + h * d[i]);
}

/* one macro step: the midpoint results for n = 2..12, extrapolated */
static void npa_bs_step(const double I[3], double y[12], double H)
{
  double T[NPA_LEVELS][12];
  for (int k = 0; k < NPA_LEVELS; k++) {
    npa_midpoint(I, y, H, 2 * (k + 1), T[k]);
    /* Neville: T[k][.] becomes the order-k extrapolant, T[k-1] the previous row */
    for (int j = k - 1; j >= 0; j--) {
      const double r = (double)(k + 1) / (j + 1);
      const double den = r * r - 1.0;
      for (int i = 0; i < 12; i++) T[j][i] = T[j + 1][i] + (T[j + 1][i] - T[j][i]) / den;
    }
  }
  for (int i = 0; i < 12; i++) y[i] = T[0][i];
}

int oracle_npa_steps(const double I[3], const double w[3], double dt)
{
  const double imax = O_MAX(I[0], O_MAX(I[1], I[2])), imin = O_MIN(I[0], O_MIN(I[1], I[2]));
  const double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double a = wn * sqrt(imax / imin) * fabs(dt) / 0.25;
  const int n = (int)ceil(a);
  return n < 1 ? 1 : n;
}

/* intnpa1: evolve (w, m) over a span of time.  Spans shorter than 1e-6 day
   are skipped (inteuler.c:165-168); 0 for a span longer than NPA_MAX_STEPS
   macro steps. */
static int npa_span(const double I[3], double span, double w[3], double m[3][3])
{
  if (fabs(span) < 1.0e-6) return 1;
  {
    const double imax = O_MAX(I[0], O_MAX(I[1], I[2])), imin = O_MIN(I[0], O_MIN(I[1], I[2]));
    const double need = ceil(sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) * sqrt(imax / imin) * fabs(span) / 0.25);
    if (!(need <= NPA_MAX_STEPS)) return 0;
  }
  double y[12];
  for (int i = 0; i < 3; i++) y[i] = w[i];
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) y[3 + 3 * j + i] = m[i][j];
  const int n = oracle_npa_steps(I, w, span);
  const double H = span / n;
  for (int s = 0; s < n; s++) npa_bs_step(I, y, H);
  for (int i = 0; i < 3; i++) w[i] = y[i];
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) m[i][j] = y[3 + 3 * j + i];
  return 1;
}

void oracle_npa_integrate(const double I[3], double t0, double t, double w[3], double m[3][3])
{
  (void)npa_span(I, t - t0, w, m);
}

/* complete elliptic integral of the first kind K(k) from the complementary
   modulus kc = sqrt(1 - k^2): pi / (2 AGM(1, kc)) */
double oracle_ellip_k(double kc)
{
  double a = 1.0, b = kc;
  for (int i = 0; i < 64 && fabs(a - b) > 1e-16 * a; i++) {
    const double an = 0.5 * (a + b);
    b = sqrt(a * b);
    a = an;
  }
  return PIE / (a + b);
}

/* the body-frame period of w (inteuler.c:238-262); <= 0 where it is not
   finite (a rotation on the separatrix) */
double oracle_npa_period(const double inertia[3], const double w[3])
{
  double I[4], twoE = 0.0, M2 = 0.0;
  for (int i = 1; i <= 3; i++) {
    I[i] = inertia[i - 1];
    const double tmp = I[i] * w[i - 1] * w[i - 1];
    twoE += tmp;
    M2 += I[i] * tmp;
  }
  for (int j = 2; j <= 3; j++) { /* ascending */
    const double a = I[j];
    int i = j - 1;
    while (i > 0 && I[i] > a) { I[i + 1] = I[i]; i--; }
    I[i + 1] = a;
  }
  if (M2 < twoE * I[2]) { const double t = I[1]; I[1] = I[3]; I[3] = t; }
  const double k2 = ((I[2] - I[1]) * (twoE * I[3] - M2)) / ((I[3] - I[2]) * (M2 - twoE * I[1]));
  if (!(k2 < 1.0) || !(k2 >= 0.0)) return -1.0;
  const double K = oracle_ellip_k(sqrt(1 - k2));
  return 4 * K * sqrt((I[1] * I[2] * I[3]) / ((I[3] - I[2]) * (M2 - twoE * I[1])));
}

/* intnpa2 (inteuler.c:208-391): one period integrated, its rotation about L
   repeated N times, the remainder integrated.  The reference integrates from
   t0 to t0 + T and to t0 + dt, epochs near JD 2.5e6 that round T and dt by up
   to 2e-10 day (5e-9 rad at 25 rad/day, times N for the period); here the
   spans T and dt are integrated as they are. */
static int npa_periodic(const double inertia[3], double t0, double t, double w[3], double m[3][3])
{
  double w0[3], m0[3][3], L[3];
  for (int i = 0; i < 3; i++) {
    w0[i] = w[i];
    L[i] = inertia[i] * w[i];
    for (int j = 0; j < 3; j++) m0[i][j] = m[i][j];
  }
  const double T = oracle_npa_period(inertia, w);
  if (!(T > 0.0) || !(fabs((t - t0) / T) < 1.0e9)) return 0;
  o_cotrans(L, m0, L, -1);
  o_normalize(L);
  const double Ltheta = acos(L[2]), Lphi = atan2(L[1], L[0]) + PIE / 2;
  const int N = o_iround((t - t0) / T);
  const double dt = t - t0 - N * T;
  double wT[3] = {w0[0], w0[1], w0[2]}, mT[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) mT[i][j] = m0[i][j];
  if (!npa_span(inertia, T, wT, mT)) return 0;
  /* a unit vector normal to L, and the angle it turns about L over T */
  int jmin = 0;
  double amin = 1.0;
  for (int i = 0; i < 3; i++)
    if (fabs(L[i]) < amin) { amin = fabs(L[i]); jmin = i; }
  double u[3] = {0.0, 0.0, 0.0}, v[3];
  u[jmin] = 1.0;
  o_cross(u, L, u);
  o_normalize(u);
  o_cotrans(v, m0, u, 1);
  o_cotrans(v, mT, v, -1);
  const double c = o_dot(u, v);
  o_cross(u, u, v);
  const double s = o_dot(L, u);
  const double psiT = atan2(s, c);
  if (!npa_span(inertia, dt, w, m)) return 0;
  double m1[3][3], m2[3][3];
  o_euler2mat(m1, Lphi, Ltheta, N * psiT);
  o_euler2mat(m2, 0.0, -Ltheta, -Lphi);
  o_mmmul(mT, m2, m1);
  o_mmmul(m, m, mT);
  return 1;
}

/* inteuler.c:114-152: from the spin state at t[0] through the impulse epochs
   to t[n-1]; returns 0 where the periodic method has no finite period or a
   span is longer than NPA_MAX_STEPS macro steps */
int oracle_inteuler_npa(const struct spin_t *spin, const double t[], double imp[][3], int n, int method,
                        double w[3], double m[3][3])
{
  double inertia[3];
  for (int i = 0; i < 3; i++) {
    inertia[i] = spin->inertia[i].val;
    w[i] = spin->omega[i].val;
  }
  o_euler2mat(m, spin->angle[0].val, spin->angle[1].val, spin->angle[2].val);
  for (int k = 0; k < n - 1; k++) {
    if (method == BRUTE) { if (!npa_span(inertia, t[k + 1] - t[k], w, m)) return 0; }
    else if (!npa_periodic(inertia, t[k], t[k + 1], w, m)) return 0;
    for (int i = 0; i < 3; i++) w[i] += imp[k + 1][i];
  }
  return 1;
}
