// npa_dev.h -- the spin state of a non-principal-axis rotator on the device
// (inteuler.c:114-392): Euler's equations for torque-free rotation, one view
// per thread, integrated from the spin epoch through the impulse epochs to the
// view's epoch.
//
// State y[12]: the body-frame spin w (rad/day) and the columns of the
// ecliptic-to-body matrix m (y[3+3j+i] = m[i][j]); dw/dt by Euler's equations,
// d(column)/dt = column x w.  Integrator: fixed Bulirsch-Stoer macro steps
// (modified midpoint with 2, 4, ..., 12 substeps, Aitken-Neville extrapolation
// in h^2), ceil(|w| sqrt(Imax/Imin) |span| / 0.25) macro steps per span.
// PERIODIC integrates one body-frame period T = 4 K(k) sqrt(...) and composes
// its rotation about the angular momentum N times, then integrates the
// remainder; BRUTE integrates the whole span.  Included by k_fp64.hip (no FMA
// contraction), so it rounds as the host restatement does.
#pragma once

namespace shp {

constexpr int NPA_LEVELS = 6;
// longest integration one span may take (macro steps); past it the state is
// refused (NaN), so a thread's loop always ends
constexpr int NPA_MAX_STEPS = 200000;

__device__ inline void npa_derivs(const double I[3], const double y[12], double d[12]) {
  const double w0 = y[0], w1 = y[1], w2 = y[2];
  d[0] = ((I[1] - I[2]) / I[0]) * w1 * w2;
  d[1] = ((I[2] - I[0]) / I[1]) * w2 * w0;
  d[2] = ((I[0] - I[1]) / I[2]) * w0 * w1;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const double x0 = y[3 + 3 * j], x1 = y[4 + 3 * j], x2 = y[5 + 3 * j];
    d[3 + 3 * j] = x1 * w2 - x2 * w1;
    d[4 + 3 * j] = x2 * w0 - x0 * w2;
    d[5 + 3 * j] = x0 * w1 - x1 * w0;
  }
}

__device__ __noinline__ void npa_midpoint(const double I[3], const double y[12], double H, int n, double out[12]) {
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

__device__ __noinline__ void npa_bs_step(const double I[3], double y[12], double H) {
  double T[NPA_LEVELS][12];
  for (int k = 0; k < NPA_LEVELS; k++) {
    npa_midpoint(I, y, H, 2 * (k + 1), T[k]);
    for (int j = k - 1; j >= 0; j--) {
      const double r = (double)(k + 1) / (j + 1);
      const double den = r * r - 1.0;
      for (int i = 0; i < 12; i++) T[j][i] = T[j + 1][i] + (T[j + 1][i] - T[j][i]) / den;
    }
  }
  for (int i = 0; i < 12; i++) y[i] = T[0][i];
}

__device__ inline int npa_steps(const double I[3], const double w[3], double span) {
  const double imax = fmax(I[0], fmax(I[1], I[2])), imin = fmin(I[0], fmin(I[1], I[2]));
  const double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const int n = (int)ceil(wn * sqrt(imax / imin) * fabs(span) / 0.25);
  return n < 1 ? 1 : n;
}

// evolve (w, m) over a span of time; m row-major [3*i+j]; spans under 1e-6
// day are skipped (inteuler.c:165-168); false past NPA_MAX_STEPS
__device__ __noinline__ bool npa_span(const double I[3], double span, double w[3], double m[9]) {
  if (fabs(span) < 1.0e-6) return true;
  const double need = ceil(sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) *
                           sqrt(fmax(I[0], fmax(I[1], I[2])) / fmin(I[0], fmin(I[1], I[2]))) * fabs(span) / 0.25);
  if (!(need <= NPA_MAX_STEPS)) return false;
  double y[12];
  for (int i = 0; i < 3; i++) y[i] = w[i];
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) y[3 + 3 * j + i] = m[3 * i + j];
  const int n = npa_steps(I, w, span);
  const double H = span / n;
  for (int s = 0; s < n; s++) npa_bs_step(I, y, H);
  for (int i = 0; i < 3; i++) w[i] = y[i];
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) m[3 * i + j] = y[3 + 3 * j + i];
  return true;
}

// K(k) from kc = sqrt(1 - k^2) by the arithmetic-geometric mean
__device__ inline double npa_ellip_k(double kc) {
  double a = 1.0, b = kc;
  for (int i = 0; i < 64 && fabs(a - b) > 1e-16 * a; i++) {
    const double an = 0.5 * (a + b);
    b = sqrt(a * b);
    a = an;
  }
  return 3.141592653589793 / (a + b);
}

// the body-frame period of w (Landau & Lifshitz sec. 37); <= 0 if not finite
__device__ inline double npa_period(const double inertia[3], const double w[3]) {
  double I1 = inertia[0], I2 = inertia[1], I3 = inertia[2], twoE = 0.0, M2 = 0.0;
  for (int i = 0; i < 3; i++) {
    const double tmp = inertia[i] * w[i] * w[i];
    twoE += tmp;
    M2 += inertia[i] * tmp;
  }
  // ascending
  if (I2 < I1) { const double t = I1; I1 = I2; I2 = t; }
  if (I3 < I2) {
    const double t = I3; I3 = I2; I2 = t;
    if (I2 < I1) { const double u = I1; I1 = I2; I2 = u; }
  }
  if (M2 < twoE * I2) { const double t = I1; I1 = I3; I3 = t; }
  const double k2 = ((I2 - I1) * (twoE * I3 - M2)) / ((I3 - I2) * (M2 - twoE * I1));
  if (!(k2 < 1.0) || !(k2 >= 0.0)) return -1.0;
  const double K = npa_ellip_k(sqrt(1 - k2));
  return 4 * K * sqrt((I1 * I2 * I3) / ((I3 - I2) * (M2 - twoE * I1)));
}

__device__ inline void npa_euler9(double m[9], double phi, double theta, double psi) {
  const double ct = cos(theta), st = sin(theta), cp = cos(phi), sp = sin(phi), cs = cos(psi), ss = sin(psi);
  m[0] = cs * cp - ct * sp * ss;
  m[3] = -ss * cp - ct * sp * cs;
  m[6] = st * sp;
  m[1] = cs * sp + ct * cp * ss;
  m[4] = -ss * sp + ct * cp * cs;
  m[7] = -st * cp;
  m[2] = ss * st;
  m[5] = cs * st;
  m[8] = ct;
}

__device__ inline void npa_mmmul(double x[9], const double y[9], const double z[9]) {
  double t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += y[3 * i + k] * z[3 * k + j];
      t[3 * i + j] = s;
    }
  for (int q = 0; q < 9; q++) x[q] = t[q];
}

__device__ inline void npa_unit(double u[3]) {
  double n = 0.0;
  for (int i = 0; i < 3; i++) n += u[i] * u[i];
  n = sqrt(n);
  if (n != 0.0)
    for (int i = 0; i < 3; i++) u[i] /= n;
}

__device__ inline void npa_cross(double z[3], const double x[3], const double y[3]) {
  const double a = x[1] * y[2] - x[2] * y[1], b = x[2] * y[0] - x[0] * y[2], c = x[0] * y[1] - x[1] * y[0];
  z[0] = a; z[1] = b; z[2] = c;
}

// one interval by the periodic method (intnpa2); false if T is not finite
__device__ __noinline__ bool npa_periodic(const double I[3], double span, double w[3], double m[9]) {
  double w0[3], m0[9], L[3];
  for (int i = 0; i < 3; i++) { w0[i] = w[i]; L[i] = I[i] * w[i]; }
  for (int q = 0; q < 9; q++) m0[q] = m[q];
  const double T = npa_period(I, w);
  if (!(T > 0.0) || !(fabs(span / T) < 1.0e9)) return false;
  double Le[3];  // L in ecliptic coordinates: m0^T L
  for (int i = 0; i < 3; i++) {
    double t = 0.0;
    for (int j = 0; j < 3; j++) t += m0[3 * j + i] * L[j];
    Le[i] = t;
  }
  npa_unit(Le);
  const double Ltheta = acos(Le[2]), Lphi = atan2(Le[1], Le[0]) + 3.141592653589793 / 2;
  const double x = span / T;
  const int N = x < 0.0 ? (int)(x - 0.5) : (int)(x + 0.5);
  const double dt = span - N * T;
  double wT[3] = {w0[0], w0[1], w0[2]}, mT[9];
  for (int q = 0; q < 9; q++) mT[q] = m0[q];
  if (!npa_span(I, T, wT, mT)) return false;
  int jmin = 0;
  double amin = 1.0;
  for (int i = 0; i < 3; i++)
    if (fabs(Le[i]) < amin) { amin = fabs(Le[i]); jmin = i; }
  double u[3] = {0.0, 0.0, 0.0}, v[3], b[3];
  u[jmin] = 1.0;
  npa_cross(u, Le, u);
  npa_unit(u);
  for (int i = 0; i < 3; i++) {  // ecliptic -> body at the start
    double t = 0.0;
    for (int j = 0; j < 3; j++) t += m0[3 * i + j] * u[j];
    b[i] = t;
  }
  for (int i = 0; i < 3; i++) {  // body -> ecliptic one period later
    double t = 0.0;
    for (int j = 0; j < 3; j++) t += mT[3 * j + i] * b[j];
    v[i] = t;
  }
  const double c = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
  npa_cross(u, u, v);
  const double s = Le[0] * u[0] + Le[1] * u[1] + Le[2] * u[2];
  const double psiT = atan2(s, c);
  if (!npa_span(I, dt, w, m)) return false;
  double m1[9], m2[9];
  npa_euler9(m1, Lphi, Ltheta, N * psiT);
  npa_euler9(m2, 0.0, -Ltheta, -Lphi);
  npa_mmmul(mT, m2, m1);
  npa_mmmul(m, m, mT);
  return true;
}

// From the spin state at tint[0] through the impulse epochs to tint[n-1]:
// angle (the Euler angles with the set's offsets), omega, moments; imp3[k]
// added to w after interval k-1.  method: 0 periodic, -1 brute.  Returns
// false where the periodic method has no finite period or a span would take
// more than NPA_MAX_STEPS macro steps.
__device__ __noinline__ bool npa_state(const double ang[3], const double omega[3], const double inertia[3],
                                 const double *tint, const double (*imp3)[3], int n, int method,
                                 double w[3], double m[9]) {
  for (int i = 0; i < 3; i++) w[i] = omega[i];
  npa_euler9(m, ang[0], ang[1], ang[2]);
  for (int k = 0; k < n - 1; k++) {
    const double span = tint[k + 1] - tint[k];
    if (method < 0) { if (!npa_span(inertia, span, w, m)) return false; }
    else if (!npa_periodic(inertia, span, w, m)) return false;
    for (int i = 0; i < 3; i++) w[i] += imp3[k + 1][i];
  }
  return true;
}

}  // namespace shp
