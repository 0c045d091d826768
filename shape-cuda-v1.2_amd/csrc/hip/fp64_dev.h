// fp64_dev.h -- FP64 projection, coverage test and shading of one facet at one
// pixel, in the reference's operation order (posvis.c:142-275, src/util
// cotrans / normalize).  Every function body disables FP contraction itself,
// so translation units built with contraction on (the mixed kernels) still
// reproduce the CPU's z, facet ownership and cos(scattering) bit for bit.
#pragma once

#include <climits>

#include "common.h"
#include "engine.h"

namespace shp {

__device__ inline void cotrans3(double y[3], const double a[9], const double x[3]) {
#pragma clang fp contract(off)
  double t0 = 0.0, t1 = 0.0, t2 = 0.0;
  t0 += a[0] * x[0]; t0 += a[1] * x[1]; t0 += a[2] * x[2];
  t1 += a[3] * x[0]; t1 += a[4] * x[1]; t1 += a[5] * x[2];
  t2 += a[6] * x[0]; t2 += a[7] * x[1]; t2 += a[8] * x[2];
  y[0] = t0; y[1] = t1; y[2] = t2;
}
__device__ inline double normalize3(double u[3]) {
#pragma clang fp contract(off)
  double nn = 0.0;
  nn += u[0] * u[0]; nn += u[1] * u[1]; nn += u[2] * u[2];
  nn = sqrt(nn);
  if (nn != 0.0) { u[0] /= nn; u[1] /= nn; u[2] /= nn; }
  return nn;
}
__device__ inline double dmin(double a, double b) { return a < b ? a : b; }
__device__ inline double dmax(double a, double b) { return a > b ? a : b; }

struct FacetProj {
  double v0[3], v1[3], v2[3], nrm[3], den;
  int i1, i2, j1, j2;
  double imin_d, imax_d, jmin_d, jmax_d;
  bool front, inside, outbnd;
};

__device__ inline void project_facet(const ViewRec &V, const GeomSlot &G, int f, int src, FacetProj &P) {
#pragma clang fp contract(off)
  const double *oa = src ? V.sa : V.oa;
  double nb[3] = {G.fn[3 * f], G.fn[3 * f + 1], G.fn[3 * f + 2]};
  cotrans3(P.nrm, oa, nb);
  P.front = P.nrm[2] > 0.0;
  if (!P.front) return;
  const int *fv = G.fv + 3 * f;
  cotrans3(P.v0, oa, G.x + 3 * fv[0]);
  cotrans3(P.v1, oa, G.x + 3 * fv[1]);
  cotrans3(P.v2, oa, G.x + 3 * fv[2]);
  for (int k = 0; k < 3; k++) {  // orbit offset (posvis.c:166-170; zero for single bodies)
    P.v0[k] += V.orb[k];
    P.v1[k] += V.orb[k];
    P.v2[k] += V.orb[k];
  }
  double kmpp = V.kmpp;
  int n = V.n;
  P.imin_d = floor(dmin(P.v0[0], dmin(P.v1[0], P.v2[0])) / kmpp - 1.0e-10 + 0.5);
  P.imax_d = floor(dmax(P.v0[0], dmax(P.v1[0], P.v2[0])) / kmpp + 1.0e-10 + 0.5);
  P.jmin_d = floor(dmin(P.v0[1], dmin(P.v1[1], P.v2[1])) / kmpp - 1.0e-10 + 0.5);
  P.jmax_d = floor(dmax(P.v0[1], dmax(P.v1[1], P.v2[1])) / kmpp + 1.0e-10 + 0.5);
  int imin = P.imin_d < -2147483648.0 ? INT_MIN : (int)P.imin_d;
  int imax = P.imax_d > 2147483647.0 ? INT_MAX : (int)P.imax_d;
  int jmin = P.jmin_d < -2147483648.0 ? INT_MIN : (int)P.jmin_d;
  int jmax = P.jmax_d > 2147483647.0 ? INT_MAX : (int)P.jmax_d;
  P.outbnd = (imin < -n || imax > n || jmin < -n || jmax > n);
  P.i1 = max(imin, -n); P.i2 = min(imax, n);
  P.j1 = max(jmin, -n); P.j2 = min(jmax, n);
  P.inside = !(P.i1 > n || P.i2 < -n || P.j1 > n || P.j2 < -n);
  P.den = 1 / ((P.v1[0] - P.v0[0]) * (P.v2[1] - P.v1[1]) - (P.v2[0] - P.v1[0]) * (P.v1[1] - P.v0[1]));
}

// project_facet without the bounding box (its four divisions): for the
// per-pixel resolve and replay, which test given pixels only.  The facing,
// vertices and den are the same expressions, so pixel_hit takes the same
// bits.
__device__ inline void project_facet_px(const ViewRec &V, const GeomSlot &G, int f, int src, FacetProj &P) {
#pragma clang fp contract(off)
  const double *oa = src ? V.sa : V.oa;
  double nb[3] = {G.fn[3 * f], G.fn[3 * f + 1], G.fn[3 * f + 2]};
  cotrans3(P.nrm, oa, nb);
  P.front = P.nrm[2] > 0.0;
  if (!P.front) return;
  const int *fv = G.fv + 3 * f;
  cotrans3(P.v0, oa, G.x + 3 * fv[0]);
  cotrans3(P.v1, oa, G.x + 3 * fv[1]);
  cotrans3(P.v2, oa, G.x + 3 * fv[2]);
  for (int k = 0; k < 3; k++) {
    P.v0[k] += V.orb[k];
    P.v1[k] += V.orb[k];
    P.v2[k] += V.orb[k];
  }
  P.den = 1 / ((P.v1[0] - P.v0[0]) * (P.v2[1] - P.v1[1]) - (P.v2[0] - P.v1[0]) * (P.v1[1] - P.v0[1]));
}

// barycentric test and depth at pixel (i, j); same expressions as posvis.c:217-228
__device__ inline bool pixel_hit(const FacetProj &P, double kmpp, int i, int j, double &s, double &t, double &z) {
#pragma clang fp contract(off)
  double x0 = i * kmpp, x1 = j * kmpp;
  s = ((x0 - P.v0[0]) * (P.v2[1] - P.v1[1]) - (P.v2[0] - P.v1[0]) * (x1 - P.v0[1])) * P.den;
  if (!((s >= -1.0e-10) && (s <= 1.0 + 1.0e-10))) return false;
  t = ((P.v1[0] - P.v0[0]) * (x1 - P.v0[1]) - (x0 - P.v0[0]) * (P.v1[1] - P.v0[1])) * P.den;
  if (!((t >= -1.0e-10) && (t <= s + 1.0e-10))) return false;
  z = P.v0[2] + s * (P.v1[2] - P.v0[2]) + t * (P.v2[2] - P.v1[2]);
  return true;
}

// shading of the winning facet at one pixel (posvis.c:243-275)
__device__ inline void shade64(const ViewRec &V, const GeomSlot &G, int f, int src, int smooth, double s, double t,
                               double &nz, double &cosb) {
#pragma clang fp contract(off)
  double nrm[3];
  const double *oa = src ? V.sa : V.oa;
  if (smooth) {
    const int *fv = G.fv + 3 * f;
    const double *a = G.vn + 3 * fv[0], *b = G.vn + 3 * fv[1], *c = G.vn + 3 * fv[2];
    for (int k = 0; k < 3; k++) nrm[k] = a[k] + s * (b[k] - a[k]) + t * (c[k] - b[k]);
    double tmp[3];
    cotrans3(tmp, oa, nrm);
    nrm[0] = tmp[0]; nrm[1] = tmp[1]; nrm[2] = tmp[2];
    normalize3(nrm);
  } else {
    double nb[3] = {G.fn[3 * f], G.fn[3 * f + 1], G.fn[3 * f + 2]};
    cotrans3(nrm, oa, nb);
  }
  nz = nrm[2];
  cosb = 0.0;
  if (!src && V.bistatic) cosb = nrm[0] * V.usrc[0] + nrm[1] * V.usrc[1] + nrm[2] * V.usrc[2];
}

// radar scattering laws in FP64 (radlaw.c:51-117), in the reference's order
// of operations; pl: [type, R, C, cos_cutoff, qsR, qsC, qs_cutoff]
__device__ inline double radlaw64(const double *pl, int type, double c) {
#pragma clang fp contract(off)
  double d = 0.0;
  switch (type) {
  case COSINELAW_DIFF: d = pl[1] * (pl[2] + 1) * pow(c, 2 * pl[2] - 1); break;
  case GAUSSIANLAW:
    if (c >= pl[3]) { double t2 = 1 / (c * c) - 1; d = pl[1] * pl[2] * exp(-pl[2] * t2) / (c * c * c * c * c); }
    break;
  case HAGFORSLAW:
    if (c >= pl[3]) { double s2 = 1 - c * c, a = c * c * c * c + pl[2] * s2; d = 0.5 * pl[1] * pl[2] * pow(a, -1.5) / c; }
    break;
  case COSINELAW_QS:
    if (c >= pl[3]) d = pl[1] * (pl[2] + 1) * pow(c, 2 * pl[2] - 1);
    break;
  case GAUSSIAN_COSINE:
    d = pl[1] * (pl[2] + 1) * pow(c, 2 * pl[2] - 1);
    if (c >= pl[6]) { double t2 = 1 / (c * c) - 1; d += pl[4] * pl[5] * exp(-pl[5] * t2) / (c * c * c * c * c); }
    break;
  case HAGFORS_COSINE:
    d = pl[1] * (pl[2] + 1) * pow(c, 2 * pl[2] - 1);
    if (c >= pl[6]) { double s2 = 1 - c * c, a = c * c * c * c + pl[5] * s2; d += 0.5 * pl[4] * pl[5] * pow(a, -1.5) / c; }
    break;
  case COSINE_COSINE:
    d = pl[1] * (pl[2] + 1) * pow(c, 2 * pl[2] - 1);
    if (c >= pl[6]) d += pl[4] * (pl[5] + 1) * pow(c, 2 * pl[5] - 1);
    break;
  }
  return d;
}

// radlaw with the laws whose values are per facet or tabulated
// (radlaw.c:56-66, :118-135): loc = the evaluation's resolved values (a
// tabular law's rho[0..n-1]; harmcosine / inhocosine: R, C per facet), f the
// pixel's facet.  The table index is clamped to n - 2: at cos(inc) below
// ~1e-16 the reference would read rho[n].
__device__ inline double radlaw64x(const double *pl, int type, double c, const double *loc, int f) {
#pragma clang fp contract(off)
  if (type == TABULARLAW) {
    if (f < 0) return 0.0;
    const int n = (int)pl[2];
    const double inc = acos(c), angres = (3.141592653589793 / 2) / (n - 1), ratio = inc / angres;
    int i = (int)floor(ratio);
    if (i > n - 2) i = n - 2;
    const double r1 = loc[i], r2 = loc[i + 1];
    return (r1 + (ratio - i) * (r2 - r1)) / c;
  }
  if (type == HARMCOSINE_DIFF || type == INHOCOSINE_DIFF) {
    if (f < 0) return 0.0;
    const double R = loc[2 * f], C = loc[2 * f + 1];
    return R * (C + 1) * pow(c, 2 * C - 1);
  }
  return radlaw64(pl, type, c);
}

// Hapke's bidirectional reflectance (astro/hapke.c:41-176; Hapke 1984 with
// the Henyey-Greenstein phase function, the opposition surge, the
// macroscopic roughness correction and Hapke 2002's H-function
// approximation); 0 for parameters out of range, as the reference gives
__device__ inline double hapke64(double cosi, double cose, double phase, double w, double h, double B0, double g,
                                 double theta) {
#pragma clang fp contract(off)
  const double pi = 4 * atan(1.0);
  if (w <= 0.0 || w >= 1.0 || h <= 0.0 || B0 < 0.0 || fabs(g) >= 1.0 || theta < 0.0 || theta >= pi / 2 ||
      phase < 0.0 || phase > pi)
    return 0.0;
  const double mu0 = cosi, mu = cose;
  const double cosphase = (g != 0.0 || theta != 0.0) ? cos(phase) : -999.99;
  const double s1w = sqrt(1 - w), r0 = (1 - s1w) / (1 + s1w);
  const double P = (g == 0.0) ? 1.0 : (1 - g * g) / pow(1 + 2 * g * cosphase + g * g, 1.5);
  const double B = (B0 == 0.0) ? 0.0 : B0 / (1 + tan(phase / 2.0) / h);
  double mu0p = mu0, mup = mu, S = 1.0;
  if (theta != 0.0) {
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
    double E1e = 0.0, E2e = 0.0, E1i = 0.0, E2i = 0.0;
    if (e >= pi / 2) { E1e = 1.0; E2e = 1.0; }
    else if (e > 0.0) { const double a = ct / tan(e); E1e = exp(-(2 / pi) * a); E2e = exp(-(1 / pi) * a * a); }
    if (i >= pi / 2) { E1i = 1.0; E2i = 1.0; }
    else if (i > 0.0) { const double a = ct / tan(i); E1i = exp(-(2 / pi) * a); E2i = exp(-(1 / pi) * a * a); }
    const double mu0p0 = C * (mu0 + sini * tt * E2i / (2 - E1i));
    const double mup0 = C * (mu + sine * tt * E2e / (2 - E1e));
    if (i < e) {
      const double den = 2 - E1e - (psi / pi) * E1i;
      mu0p = C * (mu0 + sini * tt * (cospsi * E2e + s2p2 * E2i) / den);
      mup = C * (mu + sine * tt * (E2e - s2p2 * E2i) / den);
      S = ((mup * mu0) / (mup0 * mu0p0)) * C / (1 - f * (1 - C * mu0 / mu0p0));
    } else {
      const double den = 2 - E1i - (psi / pi) * E1e;
      mu0p = C * (mu0 + sini * tt * (E2i - s2p2 * E2e) / den);
      mup = C * (mu + sine * tt * (cospsi * E2i + s2p2 * E2e) / den);
      S = ((mup * mu0) / (mup0 * mu0p0)) * C / (1 - f * (1 - C * mu / mup0));
    }
  }
  const double H0 = mu0p > 1e-40 ? 1 / (1 - w * mu0p * (r0 + 0.5 * (1 - 2 * r0 * mu0p) * log(1 + 1 / mu0p))) : 1.0;
  const double H1 = mup > 1e-40 ? 1 / (1 - w * mup * (r0 + 0.5 * (1 - 2 * r0 * mup) * log(1 + 1 / mup))) : 1.0;
  const double r = ((w * mu0p) / (4 * pi * (mu0p + mup))) * ((1 + B) * P + H0 * H1 - 1.0) * S;
  return r > 0.0 ? r : 0.0;
}

// one lit pixel's optical brightness (apply_photo.c:78-309), in the
// reference's order of operations: pl the law's parameter slots, loc the
// evaluation's per-facet values of a harmonic / inhomogeneous law (k per
// facet), f the pixel's facet (< 0: none, no brightness for a per-facet law)
__device__ inline double optlaw64(const double *pl, int type, double inf, double ci, double ce, double phase,
                                  const double *loc, int k, int f) {
#pragma clang fp contract(off)
  const double pi = 3.141592653589793;
  double q[5] = {pl[1], pl[2], pl[3], pl[4], pl[5]};
  if (type >= HARMLAMBERT && type != HAPKE && type != KAASALAINEN) {  // per-facet laws
    if (f < 0) return 0.0;
    for (int j = 0; j < k; j++) q[j] = loc[(long long)k * f + j];
  }
  switch (type) {
  case LAMBERTLAW: case HARMLAMBERT: case INHOLAMBERT: return inf * (q[0] / pi) * ci;
  case LOMMEL: case HARMLOMMEL: case INHOLOMMEL: return inf * (q[0] / (4 * pi)) * ci / (ci + ce);
  case GEOMETRICAL: return inf * q[0];
  case HAPKE: case HARMHAPKE: case INHOHAPKE: return inf * hapke64(ci, ce, phase, q[0], q[1], q[2], q[3], q[4]);
  default: {  // KAASALAINEN, HARMKAAS, INHOKAAS
    const double pf = q[2] * exp(-phase / q[3]) + q[4] * phase + 1;
    const double sls = (1 - q[1]) * pf * q[0] / (4 * pi);
    const double sl = q[1] * pf * q[0] / pi;
    return inf * ci * (sls / (ci + ce) + sl);
  }
  }
}

}  // namespace shp
