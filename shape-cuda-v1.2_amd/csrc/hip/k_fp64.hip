// k_fp64.hip -- FP64 kernels: per-view setup, realization, and the plain
// FP64 forward model (exact z-buffer, radar synthesis, photometry, chi-square).
//
// This translation unit is compiled with -ffp-contract=off: every expression
// below is evaluated in the same order and with the same roundings as the
// reference's CPU code (no fused multiply-adds), so the FP64 path reproduces
// the CPU's z values, facet ownership and cos(scattering) bit for bit; only
// transcendental functions (sin, cos, pow, exp) may differ in the last ulp.
#include "common.h"
#include "engine.h"
#include "kernels.h"
#include "fp64_dev.h"
#include "npa_dev.h"

namespace shp {

// ---------------------------------------------------------------------------
// small FP64 helpers (same operation order as src/util/*.c)
// ---------------------------------------------------------------------------
__device__ inline void cotrans3T(double y[3], const double a[9], const double x[3]) {
  double t0 = 0.0, t1 = 0.0, t2 = 0.0;
  t0 += a[0] * x[0]; t0 += a[3] * x[1]; t0 += a[6] * x[2];
  t1 += a[1] * x[0]; t1 += a[4] * x[1]; t1 += a[7] * x[2];
  t2 += a[2] * x[0]; t2 += a[5] * x[1]; t2 += a[8] * x[2];
  y[0] = t0; y[1] = t1; y[2] = t2;
}
// oa = b * a^T  (mtrnsps + mmmul of posvis.c:101,108,122)
__device__ inline void mul_bat(double oa[9], const double b[9], const double a[9]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double t = 0.0;
      for (int k = 0; k < 3; k++) t += b[3 * i + k] * a[3 * j + k];
      oa[3 * i + j] = t;
    }
}
__device__ inline void euler2mat9(double m[9], double phi, double theta, double psi) {
  double ct = cos(theta), st = sin(theta), cp = cos(phi), sp = sin(phi), cs = cos(psi), ss = sin(psi);
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
__device__ inline double sinc2d(double x) {
  double t = fabs(3.141592653589793 * x);
  if (t > 1.0e-4) { t = sin(t) / t; return t * t; }
  return 1.0;
}
__device__ inline double trid(double x) { double t = 1 - fabs(x); return t > 0.0 ? t : 0.0; }
__device__ inline double tri2d(double x) { double t = 1 - fabs(x); return t > 0.0 ? t * t : 0.0; }


// wavefront reductions (wave64)
__device__ inline int wave_min_i(int v) { for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64)); return v; }
__device__ inline int wave_max_i(int v) { for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64)); return v; }
__device__ inline double wave_min_d(double v) { for (int o = 32; o > 0; o >>= 1) v = dmin(v, __shfl_xor(v, o, 64)); return v; }
__device__ inline double wave_max_d(double v) { for (int o = 32; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o, 64)); return v; }
__device__ inline bool lane0() { return (threadIdx.x & 63) == 0; }

// ---------------------------------------------------------------------------
// per-view setup: spin state, COM offsets, projection and Doppler constants
// (realize_spin.c:41-280, inteuler.c:93-113, deldopoffs.c, dopoffs.c,
//  posvis.c:96-135, pos2deldop.c:224-295, pos2doppler.c:179-232)
// ---------------------------------------------------------------------------
__global__ void k_setup_views(ViewRec *views, int nviews, const double *params) {
  int iv = blockIdx.x * blockDim.x + threadIdx.x;
  if (iv >= nviews) return;
  ViewRec &V = views[iv];
  const double *pb = params + V.pb;
  if (!V.host_spin) {
    const double *ps = pb + V.pb_set;
    double ang[3];
    for (int j = 0; j < 3; j++) ang[j] = pb[PB_ANGLE + j] + ps[PB_SET_ANGLEOFF + j];
    // impulse list from t0 to t (realize_spin.c:350-398)
    double t0 = pb[PB_T0], t = V.t;
    double tint[MAXIMP + 2], imp3[MAXIMP + 2][3];
    int k = 0, nimp = (int)pb[PB_NIMP];
    tint[0] = t0; imp3[0][0] = imp3[0][1] = imp3[0][2] = 0.0;
    if (t >= t0) {
      for (int q = 0; q < nimp; q++) {
        double tq = pb[PB_TIMP + q];
        if (tq > t0 && tq <= t) {
          k++; tint[k] = tq;
          for (int j = 0; j < 3; j++) imp3[k][j] = pb[PB_IMP + 3 * q + j];
        }
      }
      if (tint[k] < t) { k++; tint[k] = t; imp3[k][0] = imp3[k][1] = imp3[k][2] = 0.0; }
    } else {
      for (int q = nimp - 1; q >= 0; q--) {
        double tq = pb[PB_TIMP + q];
        if (tq < t0 && tq >= t) {
          k++; tint[k] = tq;
          for (int j = 0; j < 3; j++) imp3[k][j] = -pb[PB_IMP + 3 * q + j];
        }
      }
      if (tint[k] > t) { k++; tint[k] = t; imp3[k][0] = imp3[k][1] = imp3[k][2] = 0.0; }
    }
    int n = k + 1;
    double w[3];
    if (pb[PB_PA] > 0.5) {
      // principal-axis kinematics (inteuler.c:93-113)
      double dt = tint[n - 1] - tint[0];
      double psi = ang[2] + pb[PB_OMEGA + 2] * dt + 0.5 * pb[PB_OMEGADOT + 2] * dt * dt +
                   pb[PB_LIBAMP] * sin(pb[PB_LIBFREQ] * dt + pb[PB_LIBPHASE]);
      for (int i = 0; i <= 1; i++) w[i] = pb[PB_OMEGA + i] + pb[PB_OMEGADOT + i] * dt;
      w[2] = pb[PB_OMEGA + 2] + pb[PB_OMEGADOT + 2] * dt +
             pb[PB_LIBAMP] * pb[PB_LIBFREQ] * cos(pb[PB_LIBFREQ] * dt + pb[PB_LIBPHASE]);
      for (int q = 0; q < n; q++) { psi += imp3[q][2] * (tint[n - 1] - tint[q]); w[2] += imp3[q][2]; }
      euler2mat9(V.ae, ang[0], ang[1], psi);
    } else {
      // non-principal-axis rotation: Euler's equations (inteuler.c:114-392)
      double om[3] = {pb[PB_OMEGA], pb[PB_OMEGA + 1], pb[PB_OMEGA + 2]};
      double in[3] = {pb[PB_INERTIA], pb[PB_INERTIA + 1], pb[PB_INERTIA + 2]};
      if (!npa_state(ang, om, in, tint, imp3, n, pb[PB_PA] < -0.5 ? -1 : 0, w, V.ae)) {
        for (int q = 0; q < 9; q++) V.ae[q] = __longlong_as_double(0x7ff8000000000000LL);  // NaN: no finite period
      }
    }
    for (int j = 0; j < 3; j++) w[j] += ps[PB_SET_OMEGAOFF + j];
    double is[3];
    cotrans3T(is, V.ae, w);
    for (int j = 0; j < 3; j++) { V.intspin[j] = is[j]; V.spin[j] = V.orbspin[j] + is[j]; }
    // COM offsets from the delay-correction polynomial
    if (V.type == VT_DELAY || V.type == VT_DOPPLER) {
      const double *pdc = pb + V.pb_dc;
      int nd = (int)pdc[PB_SET_DCN];
      double x = 1.0, dop = 0.0, del = pdc[PB_SET_DCA];
      double arg = V.t - pdc[PB_SET_DCT0];
      for (int q = 1; q <= nd; q++) {
        dop += q * pdc[PB_SET_DCA + q] * x;
        if (V.type == VT_DELAY) del += pdc[PB_SET_DCA + q] * (x *= arg);
        else x *= arg;
      }
      V.deloff = (V.type == VT_DELAY) ? del / V.del_per_pixel : 0.0;
      V.dopoff = -dop * V.Ftx / (V.dop_per_pixel * 86400.0);
    }
  }
  mul_bat(V.oa, V.oe, V.ae);
  if (V.bistatic) {
    mul_bat(V.sa, V.se, V.ae);
    double u[3] = {0.0, 0.0, 1.0};
    cotrans3T(u, V.se, u);
    cotrans3(V.usrc, V.oe, u);
  }
  // radar constants
  if (V.type == VT_DELAY || V.type == VT_DOPPLER) {
    const double *ps = pb + V.pb_set;
    double w[3];
    cotrans3(w, V.oe, V.spin);
    double dopfact = ps[PB_SET_DOPSCALE] * 7.72139109254982e-5 * V.kmpp * V.Ftx / V.dop_per_pixel;
    V.ax = -w[1] * dopfact;
    V.ay = w[0] * dopfact;
    V.km2Hz = sqrt(V.ax * V.ax + V.ay * V.ay) * V.dop_per_pixel / V.kmpp;
    if (w[0] != 0.0 || w[1] != 0.0)
      V.dop_extra = V.km2Hz * 0.5 * V.kmpp * sqrt(w[0] * w[0] + w[1] * w[1]) / dmax(fabs(w[0]), fabs(w[1]));
    else
      V.dop_extra = 0.0;
    int nsinc2 = V.nsinc2, width = V.sinc2width;
    V.xincr = V.ax / nsinc2;
    V.yincr = V.ay / nsinc2;
    V.dopdiff_bl = -(nsinc2 - 1) * (V.xincr + V.yincr) / 2;
    V.dopdiff_max = (nsinc2 - 1) * (fabs(V.xincr) + fabs(V.yincr)) / 2;
    V.badradar_sinc = 0;
    V.badradar_sinc_log = 0.0;
    if (2 * V.dopdiff_max + width + 1 > 100) {
      V.badradar_sinc = 1;
      V.badradar_sinc_log = log((2 * V.dopdiff_max + width + 1) / 100);
    }
    if (V.codemethod != 1) {
      V.const1 = (3 * V.spb - 1) / (2.0 * V.stride);
      V.const2 = (V.spb - 1) / (2.0 * V.stride);
    } else {
      V.const1 = (double)(V.spb / V.stride);
      V.const2 = 0.0;
    }
    V.delfact = -6.67128190396304 / V.del_per_pixel;
    V.delshift = V.delcom_vig + V.deloff;
    V.dopshift = V.dopcom_vig + V.dopoff + V.orb_dop;
    V.idop0 = 1000 - (int)floor(V.dopshift + 0.5);
  }
  // FP32 constants of the mixed method (anchored bin coordinates)
  for (int q = 0; q < 9; q++) { V.oa32[q] = (float)V.oa[q]; V.sa32[q] = (float)V.sa[q]; }
  for (int q = 0; q < 3; q++) V.usrc32[q] = (float)V.usrc[q];
  V.kmpp32 = (float)V.kmpp;
  V.inv_kmpp32 = (float)(1.0 / V.kmpp);
  // FP32 projected coordinates of the window carry ~3 ulps of n + 1 pixels;
  // times 2 (s and t), in pixels per km of facet edge
  V.shade_err32 = (float)(2.0 * 3.0 * 5.9604645e-8 * (V.n + 1) / V.kmpp);
  V.limb32 = 0.0f;
  if ((V.type == VT_DELAY || V.type == VT_DOPPLER) && V.law_type == COSINELAW_DIFF &&
      params[V.pb + V.pb_law + 2] < 0.5)
    V.limb32 = 0.05f;  // the trial's own C
  if (V.type == VT_DELAY || V.type == VT_DOPPLER) {
    double d0 = floor(V.delshift), p0 = floor(V.dopshift);
    V.D0 = (int)d0; V.P0 = (int)p0;
    V.dfr = (float)(V.delshift - d0);
    V.pfr = (float)(V.dopshift - p0);
    V.delfact32 = (float)V.delfact;
    V.ax32 = (float)V.ax; V.ay32 = (float)V.ay;
    V.const1_32 = (float)V.const1;
    V.dcrel32 = (float)(p0 - V.dopDC_vig);
  }
  // reset accumulators
  if (!V.keep_rect) {
    V.xlim[0] = V.ylim[0] = V.xlim2[0] = V.ylim2[0] = V.n;
    V.xlim[1] = V.ylim[1] = V.xlim2[1] = V.ylim2[1] = -V.n;
  }
  V.outbnd = V.outbnd2 = V.any_overflow = V.nstale = V.nclaimed = 0;
  for (int q = 0; q < 4; q++) V.ov[q] = V.ov2[q] = ord64(1.0e20);
  V.idellim[0] = V.ndel + 999999; V.idellim[1] = -999999;
  V.idoplim[0] = V.ndop + 999999; V.idoplim[1] = -999999;
  V.dellim[0] = V.dellim[1] = V.doplim[0] = V.doplim[1] = ord64(1.0e20);
  V.ysum = 0.0;
}

// projection matrices of a caller's POS (posvis.c:96-135): oa = oe * ae^T,
// sa = se * ae^T and the source direction in observer coordinates
__global__ void k_pos_setup(ViewRec *views) {
  ViewRec &V = views[blockIdx.x];
  if (threadIdx.x != 0) return;
  mul_bat(V.oa, V.oe, V.ae);
  if (V.bistatic) {
    mul_bat(V.sa, V.se, V.ae);
    double u[3] = {0.0, 0.0, 1.0};
    cotrans3T(u, V.se, u);
    cotrans3(V.usrc, V.oe, u);
  }
}

// seed the z-buffer from a POS that already holds claims: a pixel with an
// owner starts at its depth, so only nearer facets claim it (posvis.c:236)
__global__ void k_pos_seed(ViewRec *views, PosDev pd, int src) {
  const ViewRec &V = views[blockIdx.y];
  int w = 2 * V.n + 1;
  int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= w * w) return;
  long long p = V.pos_off + pix;
  int fac = (src ? pd.fill : pd.f)[p];
  double z = (src ? pd.zill : pd.z)[p];
  pd.zkey[p] = fac >= 0 ? (unsigned long long)ord64(z) : 0ULL;
  pd.fbuf[p] = 0x7fffffff;
  pd.stale_idx[p] = -1;
}

// ---------------------------------------------------------------------------
// realization (realize_mod.c:444-563)
// ---------------------------------------------------------------------------
// Real roots of coeff[3] x^3 + coeff[2] x^2 + coeff[1] x + coeff[0]
// (realize_mod.c:852-941): the quadratic or linear case below a small cubic
// coefficient, otherwise the trigonometric (three roots) or Cardano (one
// root) solution of Numerical Recipes in C, 2nd ed., section 5.6, with
// r^2 - q^3 written out in the coefficients.  Returns the root count, or -1
// where the reference bails out.
__device__ int cubic_roots(const double coeff[4], double root[3]) {
  root[0] = root[1] = root[2] = -1.0e20;
  if (fabs(coeff[3]) < 1.0e-5) {
    const double a = coeff[2], b = coeff[1], c = coeff[0];
    if (fabs(a) < 1.0e-10) {
      if (fabs(b) < 1.0e-10) return -1;
      root[0] = -c / b;
      return 1;
    }
    const double disc = b * b - 4 * a * c;
    if (disc < 0.0) return -1;
    if (fabs(b) < 1.0e-10) {
      root[0] = sqrt(disc) / (2 * a);
      root[1] = -root[0];
    } else {
      const int bsign = (b < 0.0) ? -1 : 1;
      const double q = -0.5 * (b + bsign * sqrt(disc));
      root[0] = q / a;
      root[1] = c / q;
    }
    return 2;
  }
  const double a = coeff[2] / coeff[3], b = coeff[1] / coeff[3], c = coeff[0] / coeff[3];
  const double q = (a * a - 3 * b) / 9;
  const double r = (2 * a * a * a - 9 * a * b + 27 * c) / 54;
  const double d = (4 * a * a * a * c - a * a * b * b - 18 * a * b * c + 27 * c * c + 4 * b * b * b) / 108;
  if (d >= 0.0) {
    const double rsign = (r < 0.0) ? -1 : 1;
    const double s = -rsign * pow(fabs(r) + sqrt(d), 1.0 / 3);
    const double t = (fabs(s) >= 1.0e-10) ? q / s : 0.0;
    root[0] = s + t - a / 3;
    return 1;
  }
  const double qs = sqrt(q), theta = acos(r / (q * qs));
  root[0] = -2 * qs * cos(theta / 3) - a / 3;
  root[1] = -2 * qs * cos((theta + 2 * 3.141592653589793) / 3) - a / 3;
  root[2] = -2 * qs * cos((theta - 2 * 3.141592653589793) / 3) - a / 3;
  return 3;
}

constexpr double HAIRWIDTH_D = 1.0e-7, SMALLRATIO_D = 0.01;

// the harmonic radius of vertex i before the HAIRWIDTH guard
__device__ inline double harmonic_r(const GeomSlot &G, const double *coef, int i) {
  const int nb = (G.nhar + 1) * (G.nhar + 1);
  const double *h = G.hbasis + (size_t)i * nb;
  double r = 0.0;
  for (int l = 0, q = 0; l <= G.nhar; l++) {
    r += coef[q] * h[q];
    q++;
    for (int m = 1; m <= l; m++, q += 2) r += coef[q] * h[q] + coef[q + 1] * h[q + 1];
  }
  return r;
}

// Vertex coordinates x = s (u r + a), turned and shifted (realize_mod.c:444-512)
// with r by component type (:217-442): VERTEX the slot's deviation; ELLIPSE
// and OVOID from their parameters (kept in bounds with the baddiam flag,
// scale factors 1); HARMONIC the coefficients times the vertex's factors.
// Every thread derives the component's constants; thread 0 of the slot
// writes the per-component baddiam flag and log factor (the harmonic
// per-vertex guard adds to them in k_realize_harmonic_bad).  An ovoid vertex
// with no admissible root (the reference bails out) gets a NaN position.
__global__ void k_realize_vertices(GeomSlot *slots, int nslot, const double *params, int *slot_flags,
                                   double *slot_logf) {
  int s = blockIdx.y;
  GeomSlot &G = slots[s];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double *pc = params + G.comp_pb;
  const int type = G.comp_type;
  double sf[3] = {1.0, 1.0, 1.0};
  int bad = 0;
  double blog = 0.0;
  double ea[3] = {0.0, 0.0, 0.0};                        // ELLIPSE: inverse squared radii
  double o_ar = 0.0, o_ab = 0.0, o_bc = 0.0, o_k = 0.0, o_x0 = 0.0;  // OVOID
  if (type == ELLIPSE) {
    double diam = pc[0], ratio;
    if (diam > HAIRWIDTH_D) ea[0] = 2.0 / diam;
    else { ea[0] = (2.0 / HAIRWIDTH_D) * (1 + HAIRWIDTH_D - diam); bad = 1; blog += log(1 + HAIRWIDTH_D - diam); }
    diam = (2.0 / ea[0]);
    ratio = pc[1];
    if (ratio > SMALLRATIO_D) ea[1] = 2.0 / (diam / ratio);
    else {
      ea[1] = (2.0 / (diam / SMALLRATIO_D)) / (1 + SMALLRATIO_D - ratio);
      bad = 1; blog += log(1 + SMALLRATIO_D - ratio);
    }
    diam = (2.0 / ea[1]);
    ratio = pc[2];
    if (ratio > SMALLRATIO_D) ea[2] = 2.0 / (diam / ratio);
    else {
      ea[2] = (2.0 / (diam / SMALLRATIO_D)) / (1 + SMALLRATIO_D - ratio);
      bad = 1; blog += log(1 + SMALLRATIO_D - ratio);
    }
    for (int j = 0; j < 3; j++) ea[j] *= ea[j];
  } else if (type == OVOID) {
    o_ar = pc[0] / 2;
    if (o_ar <= HAIRWIDTH_D / 2) {
      bad = 1; blog += log(1 + HAIRWIDTH_D - 2 * o_ar);
      o_ar = (HAIRWIDTH_D / 2) / (1 + HAIRWIDTH_D - 2 * o_ar);
    }
    o_ab = pc[1];
    if (o_ab <= SMALLRATIO_D) { bad = 1; blog += log(1 + SMALLRATIO_D - o_ab); o_ab = SMALLRATIO_D / (1 + SMALLRATIO_D - o_ab); }
    o_bc = pc[2];
    if (o_bc <= SMALLRATIO_D) { bad = 1; blog += log(1 + SMALLRATIO_D - o_bc); o_bc = SMALLRATIO_D / (1 + SMALLRATIO_D - o_bc); }
    o_k = pc[PB_COMP_OVOID_K];
    if (fabs(o_k) > 1 - 1.0e-10) {
      bad = 1; blog += log(fabs(o_k) + 1.0e-10);
      o_k = o_k > 0.0 ? 1 - 1.0e-10 * (1 - 1.0e-10) / o_k : -1 - 1.0e-10 * (1 - 1.0e-10) / o_k;
    }
    if (fabs(o_k) > 0.01) {
      const double x0term = 3 * (1 - o_k * o_k) * log((1 + o_k) / (1 - o_k));
      const double numer = 2.0 * o_k * (3 - 2 * o_k * o_k) - x0term;
      const double denom = 2.0 * o_k * (3 - o_k * o_k) - x0term;
      o_x0 = (o_ar / o_k) * (numer / denom);
    } else {
      o_x0 = 0.4 * o_k * o_ar;
    }
  } else {
    double raw[3];
    for (int j = 0; j < 3; j++) {
      raw[j] = (j > 0 && ((G.sf_tie >> j) & 1)) ? raw[j - 1] : pc[j];
      sf[j] = raw[j];
      if (sf[j] <= SMALLRATIO_D) {
        bad = 1;
        blog += log(1 + SMALLRATIO_D - sf[j]);
        sf[j] = SMALLRATIO_D / (1 + SMALLRATIO_D - sf[j]);
      }
    }
  }
  if (i == 0) { slot_flags[s] = bad; slot_logf[s] = blog; }
  if (i >= G.nv) return;
  const double *u = G.u + 3 * i;
  double r;
  if (type == ELLIPSE) {
    double den = 0.0;
    for (int j = 0; j < 3; j++) den += ea[j] * (u[j] * u[j]);
    r = 1 / sqrt(den);
    G.r[i] = r;
  } else if (type == OVOID) {
    const double ac = o_ab * o_bc;
    const double h = o_ab * o_ab * u[1] * u[1] + ac * ac * u[2] * u[2];
    const double alpha0 = o_x0 / o_ar, ux = u[0];
    double coeff[4], root[3], good;
    coeff[3] = (h - ux * ux) * o_k * ux;
    coeff[2] = (1 + 3 * o_k * alpha0) * ux * ux + h * (1 - o_k * alpha0);
    coeff[1] = (o_k - (2 + 3 * o_k * alpha0) * alpha0) * ux;
    coeff[0] = -(1 - alpha0 * alpha0) * (1 + o_k * alpha0);
    if (fabs(o_k) <= 1.0e-6) {
      good = 1 / sqrt(ux * ux + h);
      good -= (coeff[3] * good * good * good + coeff[1] * good) /
              (3 * coeff[3] * good * good + 2 * coeff[2] * good + coeff[1]);
    } else {
      const int nr = cubic_roots(coeff, root);
      good = -1.0e20;
      for (int q = 0; q < nr; q++)
        if (root[q] >= 0.0 && fabs(root[q] * ux - alpha0) - 1 < 1.0e-6) good = fmax(good, root[q]);
    }
    r = good < 0.0 ? __builtin_nan("") : good * o_ar;
    G.r[i] = r;
  } else if (type == HARMONIC) {
    r = harmonic_r(G, pc + PB_COMP_HAR, i);
    if (!(r > HAIRWIDTH_D / 2)) r = (HAIRWIDTH_D / 2) / (1 + HAIRWIDTH_D - 2 * r);
    G.r[i] = r;
  } else {
    r = G.r[i];
  }
  double x[3];
  for (int j = 0; j < 3; j++) x[j] = sf[j] * (u[j] * r + G.a[3 * i + j]);
  const double *rot = pc + 3, *off = pc + 6;
  if (!(rot[0] == 0 && rot[1] == 0 && rot[2] == 0)) {
    double m[9], y[3];
    euler2mat9(m, rot[0], rot[1], rot[2]);
    for (int j = 0; j < 3; j++) {
      y[j] = 0.0;
      for (int k = 0; k < 3; k++) y[j] += m[3 * j + k] * x[k];
    }
    for (int j = 0; j < 3; j++) x[j] = y[j];
  }
  if (!(off[0] == 0.0 && off[1] == 0.0 && off[2] == 0.0))
    for (int j = 0; j < 3; j++) x[j] += off[j];
  for (int j = 0; j < 3; j++) G.x[3 * i + j] = x[j];
  G.x32[i] = make_float4((float)x[0], (float)x[1], (float)x[2], 0.0f);
}

// The harmonic component's per-vertex guard (realize_mod.c:410-415): for
// every vertex whose radius is not above HAIRWIDTH / 2, log(1 + HAIRWIDTH -
// 2 r) / (L+1)^2 joins the slot's baddiam log factor.  One workgroup per
// slot, a fixed tree: the sum repeats from replay to replay.
__global__ __launch_bounds__(256) void k_realize_harmonic_bad(GeomSlot *slots, const double *params, int *slot_flags,
                                                              double *slot_logf) {
  const int s = blockIdx.x;
  const GeomSlot &G = slots[s];
  if (G.comp_type != HARMONIC) return;
  __shared__ double sh[256];
  __shared__ int any[256];
  const double *coef = params + G.comp_pb + PB_COMP_HAR;
  const double norm = (double)((G.nhar + 1) * (G.nhar + 1));
  double acc = 0.0;
  int a = 0;
  for (int i = threadIdx.x; i < G.nv; i += 256) {
    const double r = harmonic_r(G, coef, i);
    if (!(r > HAIRWIDTH_D / 2)) { acc += log(1 + HAIRWIDTH_D - 2 * r) / norm; a = 1; }
  }
  sh[threadIdx.x] = acc;
  any[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { sh[threadIdx.x] += sh[threadIdx.x + w]; any[threadIdx.x] |= any[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0 && any[0]) {
    slot_flags[s] = 1;
    slot_logf[s] = sh[0] + slot_logf[s];
  }
}

__global__ void k_realize_facets(GeomSlot *slots, int nslot) {
  GeomSlot &G = slots[blockIdx.y];
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= G.nf) return;
  const int *fv = G.fv + 3 * f;
  const double *p0 = G.x + 3 * fv[0], *p1 = G.x + 3 * fv[1], *p2 = G.x + 3 * fv[2];
  double a[3], b[3], n[3];
  for (int i = 0; i < 3; i++) { a[i] = p1[i] - p0[i]; b[i] = p2[i] - p1[i]; }
  n[0] = a[1] * b[2] - a[2] * b[1];
  n[1] = a[2] * b[0] - a[0] * b[2];
  n[2] = a[0] * b[1] - a[1] * b[0];
  double area = 0.5 * sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  normalize3(n);
  for (int i = 0; i < 3; i++) {
    G.fn[3 * f + i] = n[i];
    G.fx[3 * f + i] = (p0[i] + p1[i] + p2[i]) / 3;
  }
  G.area[f] = area;
  G.fn32[f] = make_float4((float)n[0], (float)n[1], (float)n[2], 0.0f);
}

__global__ void k_realize_vnormals(GeomSlot *slots, int nslot) {
  GeomSlot &G = slots[blockIdx.y];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.nv) return;
  double n[3] = {0.0, 0.0, 0.0};
  for (int q = G.af_off[i]; q < G.af_off[i + 1]; q++) {
    int f = G.af_idx[q];
    for (int k = 0; k < 3; k++) n[k] += G.fn[3 * f + k];
  }
  normalize3(n);
  for (int k = 0; k < 3; k++) G.vn[3 * i + k] = n[k];
  G.vn32[i] = make_float4((float)n[0], (float)n[1], (float)n[2], 0.0f);
}

// ---------------------------------------------------------------------------
// FP64 rasterization (posvis.c:142-295) in three passes over the facets:
//   1: z-buffer  -- atomicMax of the ordered depth of every covered pixel
//   2: ownership -- lowest facet index whose depth equals the maximum
//   3: history   -- facets covering pixels whose winner has a back-facing
//                   smoothed normal (their cos(scattering) is inherited from
//                   an earlier claimant; replayed in k_replay64)
// ---------------------------------------------------------------------------
__global__ void k_raster64(ViewRec *views, const GeomSlot *slots, PosDev pd, int src, int pass) {
  ViewRec &V = views[blockIdx.y];
  if (V.fp64_skip) return;
  if (pass == 3 && !V.nstale) return;  // no replay record in this view: no candidates to collect
  const GeomSlot &G = slots[V.slot];
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  int n = V.n, w = 2 * n + 1;
  long long base = V.pos_off;
  // pass-1 accumulators: in-frame rectangle and overall extent
  int xl0 = n, xl1 = -n, yl0 = n, yl1 = -n, ob = 0;
  double o0 = 1.0e20, o1 = -1.0e20, o2 = 1.0e20, o3 = -1.0e20;
  if (f < G.nf) {
    FacetProj P;
    project_facet(V, G, f, src, P);
    if (P.front) {
      ob = P.outbnd;
      if (!P.inside) {
        // posvis.c:194-199: facet wholly outside -> only the rectangles grow
        o0 = P.imin_d; o1 = P.imax_d; o2 = P.jmin_d; o3 = P.jmax_d;
        int imin = P.imin_d < -2147483648.0 ? INT_MIN : (int)P.imin_d;
        int imax = P.imax_d > 2147483647.0 ? INT_MAX : (int)P.imax_d;
        int jmin = P.jmin_d < -2147483648.0 ? INT_MIN : (int)P.jmin_d;
        int jmax = P.jmax_d > 2147483647.0 ? INT_MAX : (int)P.jmax_d;
        xl0 = max(imin, -n); xl1 = min(imax, n); yl0 = max(jmin, -n); yl1 = min(jmax, n);
        if (xl0 > n) xl0 = n;       // MAX(imin,-n) only lowers xlim[0] when below it
        if (xl1 < -n) xl1 = -n;
        if (yl0 > n) yl0 = n;
        if (yl1 < -n) yl1 = -n;
      } else {
        for (int i = P.i1; i <= P.i2; i++)
          for (int j = P.j1; j <= P.j2; j++) {
            double s, t, z;
            if (!pixel_hit(P, V.kmpp, i, j, s, t, z)) continue;
            long long p = base + (long long)(i + n) * w + (j + n);
            if (pass == 1) {
              atomicMax(&pd.zkey[p], (unsigned long long)ord64(z));
              if (!V.rect_by_claim) {
                xl0 = min(xl0, i); xl1 = max(xl1, i); yl0 = min(yl0, j); yl1 = max(yl1, j);
              }
              o0 = dmin(o0, (double)i); o1 = dmax(o1, (double)i);
              o2 = dmin(o2, (double)j); o3 = dmax(o3, (double)j);
            } else if (pass == 2) {
              if ((unsigned long long)ord64(z) == pd.zkey[p]) atomicMin(&pd.fbuf[p], f);
            } else {
              int si = pd.stale_idx[p];
              if (si >= 0) {
                int q = atomicAdd(&pd.stale[si].ncand, 1);
                if (q < STALE_CAND) pd.stale[si].cand[q] = f;
              }
            }
          }
      }
    }
  }
  if (pass != 1) return;
  xl0 = wave_min_i(xl0); xl1 = wave_max_i(xl1); yl0 = wave_min_i(yl0); yl1 = wave_max_i(yl1);
  ob = wave_max_i(ob);
  o0 = wave_min_d(o0); o1 = wave_max_d(o1); o2 = wave_min_d(o2); o3 = wave_max_d(o3);
  if (lane0()) {
    int *xl = src ? V.xlim2 : V.xlim, *yl = src ? V.ylim2 : V.ylim;
    unsigned long long *ov = src ? V.ov2 : V.ov;
    if (xl0 < n) atomicMin(&xl[0], xl0);
    if (xl1 > -n) atomicMax(&xl[1], xl1);
    if (yl0 < n) atomicMin(&yl[0], yl0);
    if (yl1 > -n) atomicMax(&yl[1], yl1);
    if (ob) atomicOr(src ? &V.outbnd2 : &V.outbnd, 1);
    if (o0 < 1.0e20) atomic_min_d(&ov[0], o0);
    if (o1 > -1.0e20) atomic_min_d(&ov[1], -o1);
    if (o2 < 1.0e20) atomic_min_d(&ov[2], o2);
    if (o3 > -1.0e20) atomic_min_d(&ov[3], -o3);
  }
}

// Resolve the z-buffer into POS arrays.  merge: pixels already owned before
// this call keep their owner on depth ties (the `fac < 0` rule of
// posvis.c:236); blank: pixels without a new claimant are written as blank sky.
__global__ void k_resolve64(ViewRec *views, const GeomSlot *slots, PosDev pd, int src, int merge, int blank,
                            int body, int comp) {
  ViewRec &V = views[blockIdx.y];
  if (V.fp64_skip) return;
  int n = V.n, w = 2 * n + 1;
  int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= w * w) return;
  int i = pix / w - n, j = pix % w - n;
  const int *xl = src ? V.xlim2 : V.xlim, *yl = src ? V.ylim2 : V.ylim;
  if (!merge && (i < xl[0] || i > xl[1] || j < yl[0] || j > yl[1])) return;
  long long p = V.pos_off + pix;
  unsigned long long k = pd.zkey[p];
  int fw = pd.fbuf[p];
  pd.zkey[p] = 0ULL;
  pd.fbuf[p] = 0x7fffffff;
  double *zz = src ? pd.zill : pd.z;
  double *cosa = src ? pd.cosill : pd.cose;
  int *fac = src ? pd.fill : pd.f;
  if (fw == 0x7fffffff) {
    if (blank) {
      zz[p] = -1.0e20; cosa[p] = 0.0; fac[p] = -1;
      if (!src && pd.cosi) pd.cosi[p] = 0.0;
    }
    return;
  }
  int oldf = fac[p];
  double oldz = zz[p];
  if (merge && oldf >= 0 && k == (unsigned long long)ord64(oldz)) return;  // earlier owner keeps the tie
  const GeomSlot &G = slots[V.slot];
  FacetProj P;
  project_facet(V, G, fw, src, P);
  double s, t, z;
  pixel_hit(P, V.kmpp, i, j, s, t, z);
  int smooth = src ? V.smooth_src : V.smooth;
  double nz, cosb;
  shade64(V, G, fw, src, smooth, s, t, nz, cosb);
  zz[p] = z;
  fac[p] = fw;
  if (V.rect_by_claim) {  // posvis.c:322-364 for a pixel this call claims
    int *xl = src ? V.xlim2 : V.xlim, *yl = src ? V.ylim2 : V.ylim;
    if (i < xl[0]) atomicMin(&xl[0], i);
    if (i > xl[1]) atomicMax(&xl[1], i);
    if (j < yl[0]) atomicMin(&yl[0], j);
    if (j > yl[1]) atomicMax(&yl[1], j);
  }
  if (pd.body) {
    (src ? pd.bodyill : pd.body)[p] = body;
    (src ? pd.compill : pd.comp)[p] = comp;
  }
  if (nz > 0.0) {
    cosa[p] = nz;
    if (!src && V.bistatic) {
      pd.cosi[p] = cosb;
      if (cosb <= 0.0) cosa[p] = 0.0;
    }
  } else {
    // cos(scattering) is inherited from an earlier claimant: record for replay
    V.nstale = 1;  // (pass 3 collects candidates only in views with records)
    int si = atomicAdd(pd.nstale, 1);
    if (si < pd.stale_cap) {
      StaleRec &S = pd.stale[si];
      S.view = blockIdx.y; S.pix = pix;
      S.oldf = merge ? oldf : -1;          // without merging the pixel started blank
      S.oldz = merge ? oldz : -1.0e20;
      S.oldcose = merge ? cosa[p] : 0.0;
      S.oldcosi = (!src && V.bistatic) ? (merge ? pd.cosi[p] : 0.0) : 0.0;
      S.ncand = 0;
      S.src = src;
      pd.stale_idx[p] = si;
      if (!merge) { cosa[p] = 0.0; if (!src && V.bistatic) pd.cosi[p] = 0.0; }
    } else {
      atomicOr(pd.error, 1);
    }
  }
}

// Replay, in facet order, the claims made on one pixel whose final owner has a
// back-facing smoothed normal, reproducing the CPU's inherited values.
__device__ void replay64_one(ViewRec *views, const GeomSlot *slots, const PosDev &pd, int si) {
  StaleRec &S = pd.stale[si];
  const ViewRec &V = views[S.view];
  const GeomSlot &G = slots[V.slot];
  int n = V.n, w = 2 * n + 1;
  int i = S.pix / w - n, j = S.pix % w - n;
  long long p = V.pos_off + S.pix;
  int nc = min(S.ncand, STALE_CAND);
  if (S.ncand > STALE_CAND) atomicOr(pd.error, 2);
  int c[STALE_CAND];
  for (int q = 0; q < nc; q++) c[q] = S.cand[q];
  for (int a = 1; a < nc; a++) {  // insertion sort: facet order
    int v = c[a], b = a - 1;
    while (b >= 0 && c[b] > v) { c[b + 1] = c[b]; b--; }
    c[b + 1] = v;
  }
  double zz = S.oldz, cosa = S.oldcose, cosb = S.oldcosi;
  int fac = S.oldf;
  int src = S.src, smooth = src ? V.smooth_src : V.smooth;
  for (int q = 0; q < nc; q++) {
    FacetProj P;
    project_facet(V, G, c[q], src, P);
    double s, t, z;
    if (!pixel_hit(P, V.kmpp, i, j, s, t, z)) continue;
    if (!((z > zz) || (fac < 0))) continue;
    zz = z;
    double nz, cb;
    shade64(V, G, c[q], src, smooth, s, t, nz, cb);
    if (nz > 0.0) {
      cosa = nz;
      if (!src && V.bistatic) { cosb = cb; if (cb <= 0.0) cosa = 0.0; }
    }
    fac = c[q];
  }
  (src ? pd.cosill : pd.cose)[p] = cosa;
  if (!src && V.bistatic) pd.cosi[p] = cosb;
  pd.stale_idx[p] = -1;
}

// grid-stride over the records (the grid is sized for the cap, not the count)
__global__ void k_replay64(ViewRec *views, const GeomSlot *slots, PosDev pd) {
  int ns = min(*pd.nstale, pd.stale_cap);
  for (int si = blockIdx.x * blockDim.x + threadIdx.x; si < ns; si += gridDim.x * blockDim.x)
    replay64_one(views, slots, pd, si);
}

// ---------------------------------------------------------------------------
// shadowing (posmask.c:58-174); pixels outside the sun-view rectangle are
// blank sky by construction
// ---------------------------------------------------------------------------
__device__ inline double zill_at(const ViewRec &V, const PosDev &pd, int i, int j) {
  if (i < V.xlim2[0] || i > V.xlim2[1] || j < V.ylim2[0] || j > V.ylim2[1]) return -1.0e20;
  int w = 2 * V.n + 1;
  return pd.zill[V.pos_off + (long long)(i + V.n) * w + (j + V.n)];
}
__device__ inline int ill_at(const ViewRec &V, const int *a, int i, int j) {
  if (i < V.xlim2[0] || i > V.xlim2[1] || j < V.ylim2[0] || j > V.ylim2[1]) return -1;
  int w = 2 * V.n + 1;
  return a[V.pos_off + (long long)(i + V.n) * w + (j + V.n)];
}
__device__ inline int fill_at(const ViewRec &V, const PosDev &pd, int i, int j) { return ill_at(V, pd.fill, i, j); }

__global__ void k_posmask64(ViewRec *views, PosDev pd, double tol) {
  const ViewRec &V = views[blockIdx.y];
  if (V.fp64_skip) return;
  int n = V.n, w = 2 * n + 1;
  int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= w * w) return;
  int i = pix / w - n, j = pix % w - n;
  if (i < V.xlim[0] || i > V.xlim[1] || j < V.ylim[0] || j > V.ylim[1]) return;
  long long p = V.pos_off + pix;
  if (pd.cose[p] == 0.0) return;
  double so[9];
  mul_bat(so, V.se, V.oe);  // so = se * oe^T: observer -> source
  double kmpp = V.kmpp, ppk = 1 / kmpp;
  double x[3] = {i * kmpp, j * kmpp, pd.z[p]}, y[3];
  cotrans3(y, so, x);
  double i0 = y[0] * ppk, j0 = y[1] * ppk;
  int im = i0 < 0.0 ? (int)(i0 - 0.5) : (int)(i0 + 0.5);
  int jm = j0 < 0.0 ? (int)(j0 - 0.5) : (int)(j0 + 0.5);
  double big = 0.99 * 1.0e20;
  if (!(fabs(i0) < n && fabs(j0) < n)) return;
  if (!(zill_at(V, pd, im, jm) > -big)) return;
  bool other = pd.f[p] != fill_at(V, pd, im, jm);
  if (pd.comp && pd.compill) other = other || pd.comp[p] != ill_at(V, pd.compill, im, jm) ||
                                    pd.body[p] != ill_at(V, pd.bodyill, im, jm);
  if (!other) return;
  int i1 = (int)floor(i0), j1 = (int)floor(j0);
  double zill;
  double z11 = zill_at(V, pd, i1, j1), z21 = zill_at(V, pd, i1 + 1, j1), z12 = zill_at(V, pd, i1, j1 + 1),
         z22 = zill_at(V, pd, i1 + 1, j1 + 1);
  if (z11 > -big && z21 > -big && z12 > -big && z22 > -big) {
    double t = i0 - i1, u = j0 - j1;
    zill = (1 - t) * (1 - u) * z11 + t * (1 - u) * z21 + t * u * z22 + (1 - t) * u * z12;
  } else {
    double zm = zill_at(V, pd, im, jm);
    zill = zm;
    int is = (i0 >= im) ? 1 : -1;
    int i2 = im + is;
    if (abs(i2) <= n && zill_at(V, pd, i2, jm) > -big) {
      zill += fabs(i0 - im) * (zill_at(V, pd, i2, jm) - zm);
    } else {
      i2 = im - is;
      if (abs(i2) <= n && zill_at(V, pd, i2, jm) > -big) zill -= fabs(i0 - im) * (zill_at(V, pd, i2, jm) - zm);
    }
    int js = (j0 >= jm) ? 1 : -1;
    int j2 = jm + js;
    if (abs(j2) <= n && zill_at(V, pd, im, j2) > -big) {
      zill += fabs(j0 - jm) * (zill_at(V, pd, im, j2) - zm);
    } else {
      j2 = jm - js;
      if (abs(j2) <= n && zill_at(V, pd, im, j2) > -big) zill -= fabs(j0 - jm) * (zill_at(V, pd, im, j2) - zm);
    }
  }
  if (zill - y[2] > tol) pd.cose[p] = 0.0;
}

// ---------------------------------------------------------------------------
// optical brightness (apply_photo.c:78-309, every optical law) summed per
// lightcurve point
// ---------------------------------------------------------------------------
__global__ void k_photo64(ViewRec *views, PosDev pd, const double *params) {
  ViewRec &V = views[blockIdx.y];
  if (V.fp64_skip) return;
  int n = V.n, w = 2 * n + 1;
  int pix = blockIdx.x * blockDim.x + threadIdx.x;
  double b = 0.0;
  if (pix < w * w) {
    int i = pix / w - n, j = pix % w - n;
    if (i >= V.xlim[0] && i <= V.xlim[1] && j >= V.ylim[0] && j <= V.ylim[1]) {
      long long p = V.pos_off + pix;
      double ce = pd.cose[p], ci = pd.cosi[p];
      if (ce > 0.0 && ci > 0.0 && (pd.body ? pd.body[p] == V.body : true)) {
        const double *pl = params + V.pb + V.pb_law;
        b = optlaw64(pl, V.law_type, V.intensityfactor, ci, ce, V.phase, V.lawloc, V.lawk, V.lawloc ? pd.f[p] : -1);
        if (pd.b) pd.b[p] = b;
      }
    }
  }
  // block reduction, then one FP64 atomic per block
  __shared__ double red[256];
  red[threadIdx.x] = b;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] != 0.0) atomicAdd(&V.ysum, red[0]);
}

// ---------------------------------------------------------------------------
// delay-Doppler and Doppler synthesis (pos2deldop.c:305-535,
// pos2doppler.c:244-439) in FP64: one workgroup per SYN_TILE x SYN_TILE tile
// of a view's POS, one pixel per thread.  The tile's in-frame footprints are
// accumulated in an FP64 LDS window bounded exactly by the tile's own delay
// rows and Doppler columns, then flushed once to the fit image.
// ---------------------------------------------------------------------------
constexpr int DELMAX = 4, DOPMAX = 16;  // register footprint; wider ones take the direct path
constexpr int WIN64 = 2048;             // LDS window (doubles)

// block-wide minimum of four ints (one barrier pair); sh holds 4 per wave
__device__ inline void blk_min4_64(int &a, int &b, int &c, int &d, int (*sh)[4]) {
  a = wave_min_i(a); b = wave_min_i(b); c = wave_min_i(c); d = wave_min_i(d);
  int wv = threadIdx.x >> 6;
  if (lane0()) { sh[wv][0] = a; sh[wv][1] = b; sh[wv][2] = c; sh[wv][3] = d; }
  __syncthreads();
  a = sh[0][0]; b = sh[0][1]; c = sh[0][2]; d = sh[0][3];
  for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
    a = min(a, sh[q][0]); b = min(b, sh[q][1]); c = min(c, sh[q][2]); d = min(d, sh[q][3]);
  }
  __syncthreads();
}

// the view's delay-Doppler limits, once per workgroup; limits only move
// outward, so updates a read of the current value shows useless are skipped
__device__ inline void limits_block64(ViewRec &V, bool dd, int il0, int il1, int jl0, int jl1, double dl0, double dl1,
                                      double pl0, double pl1, int ovf, int (*shi)[8], double (*shd)[4]) {
  il0 = wave_min_i(il0); il1 = wave_max_i(il1); jl0 = wave_min_i(jl0); jl1 = wave_max_i(jl1);
  dl0 = wave_min_d(dl0); dl1 = wave_max_d(dl1); pl0 = wave_min_d(pl0); pl1 = wave_max_d(pl1);
  ovf = wave_max_i(ovf);
  int wv = threadIdx.x >> 6;
  if (lane0()) {
    shi[wv][0] = il0; shi[wv][1] = il1; shi[wv][2] = jl0; shi[wv][3] = jl1; shi[wv][4] = ovf;
    shd[wv][0] = dl0; shd[wv][1] = dl1; shd[wv][2] = pl0; shd[wv][3] = pl1;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
    il0 = min(il0, shi[q][0]); il1 = max(il1, shi[q][1]); jl0 = min(jl0, shi[q][2]); jl1 = max(jl1, shi[q][3]);
    ovf |= shi[q][4];
    dl0 = dmin(dl0, shd[q][0]); dl1 = dmax(dl1, shd[q][1]); pl0 = dmin(pl0, shd[q][2]); pl1 = dmax(pl1, shd[q][3]);
  }
  if (jl0 == 0x7fffffff) return;
  if (dd) {
    if (il0 < V.idellim[0]) atomicMin(&V.idellim[0], il0);
    if (il1 > V.idellim[1]) atomicMax(&V.idellim[1], il1);
    unsigned long long k0 = ord64(dl0), k1 = ord64(-dl1);
    if (k0 < V.dellim[0]) atomicMin(&V.dellim[0], k0);
    if (k1 < V.dellim[1]) atomicMin(&V.dellim[1], k1);
  }
  if (jl0 < V.idoplim[0]) atomicMin(&V.idoplim[0], jl0);
  if (jl1 > V.idoplim[1]) atomicMax(&V.idoplim[1], jl1);
  unsigned long long k2 = ord64(pl0), k3 = ord64(-pl1);
  if (k2 < V.doplim[0]) atomicMin(&V.doplim[0], k2);
  if (k3 < V.doplim[1]) atomicMin(&V.doplim[1], k3);
  if (ovf && !V.any_overflow) atomicOr(&V.any_overflow, 1);
}

__device__ inline double sinc2_mean64(const ViewRec &V, double dopPOS, int idop) {
  int ns = V.nsinc2;
  if (ns == 1) return sinc2d(dopPOS - idop);
  if (ns == 2) {
    double bl = dopPOS + V.dopdiff_bl - idop;
    return (sinc2d(bl) + sinc2d(bl + V.xincr) + sinc2d(bl + V.yincr) + sinc2d(bl + V.xincr + V.yincr)) / 4;
  }
  double al = dopPOS + V.dopdiff_bl - idop, sum = 0.0;
  for (int a = 0; a < ns; a++) {
    double g = al;
    for (int b = 0; b < ns; b++) { sum += sinc2d(g); g += V.xincr; }
    al += V.yincr;
  }
  return sum / (ns * ns);
}

__device__ inline double delay_weight64(const ViewRec &V, double delPOS, int idel) {
  if (V.codemethod != 1) {
    if (V.spb == 1) return tri2d(delPOS - idel);
    int sos = V.spb / V.stride;
    if (V.spb == 2) {
      double arg = (delPOS - (idel - V.const2)) / sos;
      double d = trid(arg) + trid(arg - 0.5);
      return d * (d / 4);
    }
    double d = 0.0, osp = 1.0 / V.spb;
    int m_min = max((int)floor((delPOS - idel - V.const2) * V.stride), 0);
    int m_max = min((int)ceil((delPOS - idel + V.const1) * V.stride), V.spb) - 1;
    double arg = (delPOS - (idel - V.const2)) / sos - m_min * osp;
    for (int m = m_min; m <= m_max; m++) { d += trid(arg); arg -= osp; }
    return d * (d / (V.spb * V.spb));
  }
  return tri2d((delPOS - idel) / (V.spb / V.stride));
}

// one radar scattering-law value (radlaw.c:44-148)
__global__ void k_radlaw_one(const double *pl, int type, double c, const double *loc, int f, double *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = radlaw64x(pl, type, c, loc, f);
}

__global__ __launch_bounds__(256) void k_deldop64(ViewRec *views, FrameRec *frames, PosDev pd, double *fit,
                                                  const double *params) {
  ViewRec &V = views[blockIdx.y];
  if (V.type != VT_DELAY) return;
  const int n = V.n, w = 2 * n + 1, tx = (w + SYN_TILE - 1) / SYN_TILE;
  const int ti = blockIdx.x / tx, tj = blockIdx.x % tx;
  if (ti >= tx) return;
  const int i0 = ti * SYN_TILE - n, j0 = tj * SYN_TILE - n;
  if (i0 > V.xlim[1] || i0 + SYN_TILE - 1 < V.xlim[0] || j0 > V.ylim[1] || j0 + SYN_TILE - 1 < V.ylim[0]) return;
  __shared__ double win[WIN64];
  __shared__ int shw[4][4];
  __shared__ int shi[4][8];
  __shared__ double shd[4][4];
  const int x = i0 + (int)threadIdx.x / SYN_TILE, y = j0 + (int)threadIdx.x % SYN_TILE;
  const long long p = V.pos_off + (long long)(x + n) * w + (y + n);
  const int ndel = V.ndel, ndop = V.ndop, width = V.sinc2width;
  bool valid = x <= n && y <= n && x >= V.xlim[0] && x <= V.xlim[1] && y >= V.ylim[0] && y <= V.ylim[1] &&
               pd.cose[p] > 0.0 && (pd.body ? pd.body[p] == V.body : true) && pd.z[p] != -1e20;
  double delPOS = 0.0, dopPOS = 0.0;
  int idel_min = 0, idel_max = -1, idop_min = 0, idop_max = -1;
  if (valid) {
    delPOS = pd.z[p] * V.delfact + V.delshift;
    idel_min = (int)floor(delPOS - V.const1) + 1;
    idel_max = (int)ceil(delPOS + V.const1) - 1;
    dopPOS = V.ax * (x - V.orb_x) + V.ay * (y - V.orb_y) + V.dopshift;
    idop_min = (int)floor(dopPOS - V.dopdiff_max + 1 - width / 2.0);
    idop_max = (int)floor(dopPOS + V.dopdiff_max + width / 2.0);
  }
  const bool in_bounds = valid && (idel_min >= 1) && (idel_max <= ndel) && (idop_min >= 1) && (idop_max <= ndop);
  const int nd = idel_max - idel_min + 1, np = idop_max - idop_min + 1;
  const bool small = in_bounds && nd <= DELMAX && np <= DOPMAX;
  // window over the tile's in-frame register footprints
  int lo_d = small ? idel_min : 0x7fffffff, nhi_d = small ? -idel_max : 0x7fffffff;
  int lo_p = small ? idop_min : 0x7fffffff, nhi_p = small ? -idop_max : 0x7fffffff;
  blk_min4_64(lo_d, nhi_d, lo_p, nhi_p, shw);
  const int wd = -nhi_d - lo_d + 1, wp = -nhi_p - lo_p + 1;
  const bool use_lds = lo_d != 0x7fffffff && (long long)wd * wp <= WIN64;
  if (use_lds) {
    for (int q = threadIdx.x; q < wd * wp; q += blockDim.x) win[q] = 0.0;
    __syncthreads();
  }
  double *img = fit + frames[V.frame_id].off;
  if (small) {
    double codefactor = (V.codemethod == 0) ? sinc2d((dopPOS - V.dopDC_vig) / V.dopfftlen) : 1.0;
    double dc[DELMAX], pc[DOPMAX];
#pragma unroll
    for (int a = 0; a < DELMAX; a++) dc[a] = a < nd ? delay_weight64(V, delPOS, idel_min + a) : 0.0;
#pragma unroll
    for (int b = 0; b < DOPMAX; b++) pc[b] = b < np ? sinc2_mean64(V, dopPOS, idop_min + b) : 0.0;
    double sumw = 0.0;
#pragma unroll
    for (int a = 0; a < DELMAX; a++)
#pragma unroll
      for (int b = 0; b < DOPMAX; b++)
        if (a < nd && b < np) sumw += dc[a] * pc[b];
    const double *pl = params + V.pb + V.pb_law;
    double amp = radlaw64x(pl, V.law_type, pd.cose[p], V.lawloc, V.lawloc ? pd.f[p] : -1) * V.kmpp * V.kmpp * codefactor / sumw;
#pragma unroll
    for (int a = 0; a < DELMAX; a++) {
      if (a < nd) {
        double ad = amp * dc[a];
        double *row = use_lds ? win + (long long)(idel_min + a - lo_d) * wp + (idop_min - lo_p)
                              : img + (long long)(idel_min + a - 1) * ndop + (idop_min - 1);
#pragma unroll
        for (int b = 0; b < DOPMAX; b++)
          if (b < np) atomicAdd(&row[b], ad * pc[b]);
      }
    }
  } else if (in_bounds) {
    // wide footprints: recompute weights instead of holding them in registers
    double codefactor = (V.codemethod == 0) ? sinc2d((dopPOS - V.dopDC_vig) / V.dopfftlen) : 1.0;
    double sumw = 0.0;
    for (int a = idel_min; a <= idel_max; a++) {
      double d = delay_weight64(V, delPOS, a);
      for (int b = idop_min; b <= idop_max; b++) sumw += d * sinc2_mean64(V, dopPOS, b);
    }
    const double *pl = params + V.pb + V.pb_law;
    double amp = radlaw64x(pl, V.law_type, pd.cose[p], V.lawloc, V.lawloc ? pd.f[p] : -1) * V.kmpp * V.kmpp * codefactor / sumw;
    for (int a = idel_min; a <= idel_max; a++) {
      double d = delay_weight64(V, delPOS, a);
      for (int b = idop_min; b <= idop_max; b++)
        atomicAdd(&img[(long long)(a - 1) * ndop + (b - 1)], amp * d * sinc2_mean64(V, dopPOS, b));
    }
  }
  if (use_lds) {
    __syncthreads();
    for (int q = threadIdx.x; q < wd * wp; q += blockDim.x) {
      double v = win[q];
      if (v != 0.0) atomicAdd(&img[(long long)(lo_d + q / wp - 1) * ndop + (lo_p + q % wp - 1)], v);
    }
  }
  limits_block64(V, true, valid ? idel_min : 0x7fffffff, valid ? idel_max : -0x7fffffff,
                 valid ? idop_min : 0x7fffffff, valid ? idop_max : -0x7fffffff, valid ? delPOS : 1.0e20,
                 valid ? delPOS : -1.0e20, valid ? dopPOS : 1.0e20, valid ? dopPOS : -1.0e20, 0, shi, shd);
}

// Doppler synthesis with the overflow spectrum (pos2doppler.c:244-439)
__global__ __launch_bounds__(256) void k_doppler64(ViewRec *views, FrameRec *frames, PosDev pd, double *fit,
                                                   double *overflow, const double *params) {
  ViewRec &V = views[blockIdx.y];
  if (V.type != VT_DOPPLER) return;
  const int n = V.n, w = 2 * n + 1, tx = (w + SYN_TILE - 1) / SYN_TILE;
  const int ti = blockIdx.x / tx, tj = blockIdx.x % tx;
  if (ti >= tx) return;
  const int i0 = ti * SYN_TILE - n, j0 = tj * SYN_TILE - n;
  if (i0 > V.xlim[1] || i0 + SYN_TILE - 1 < V.xlim[0] || j0 > V.ylim[1] || j0 + SYN_TILE - 1 < V.ylim[0]) return;
  __shared__ double win[WIN64];
  __shared__ int shw[4][4];
  __shared__ int shi[4][8];
  __shared__ double shd[4][4];
  const int x = i0 + (int)threadIdx.x / SYN_TILE, y = j0 + (int)threadIdx.x % SYN_TILE;
  const long long p = V.pos_off + (long long)(x + n) * w + (y + n);
  const int ndop = V.ndop, width = V.sinc2width;
  bool valid = x <= n && y <= n && x >= V.xlim[0] && x <= V.xlim[1] && y >= V.ylim[0] && y <= V.ylim[1] &&
               pd.cose[p] > 0.0 && (pd.body ? pd.body[p] == V.body : true);
  double dopPOS = 0.0;
  int idop_min = 0, idop_max = -1;
  if (valid) {
    dopPOS = V.ax * (x - V.orb_x) + V.ay * (y - V.orb_y) + V.dopshift;
    idop_min = (int)floor(dopPOS - V.dopdiff_max + 1 - width / 2.0);
    idop_max = (int)floor(dopPOS + V.dopdiff_max + width / 2.0);
  }
  const bool in_bounds = valid && (idop_min >= 1) && (idop_max <= ndop);
  const int np = idop_max - idop_min + 1;
  const bool small = in_bounds && np <= DOPMAX;
  int lo_p = small ? idop_min : 0x7fffffff, nhi_p = small ? -idop_max : 0x7fffffff, u0 = 0, u1 = 0;
  blk_min4_64(lo_p, nhi_p, u0, u1, shw);
  const int wp = -nhi_p - lo_p + 1;
  const bool use_lds = lo_p != 0x7fffffff && wp <= WIN64;
  if (use_lds) {
    for (int q = threadIdx.x; q < wp; q += blockDim.x) win[q] = 0.0;
    __syncthreads();
  }
  double *img = fit + frames[V.frame_id].off;
  if (valid) {
    // weights in bin order, as the CPU sums them; held in registers for the
    // first DOPMAX bins, recomputed (same value) beyond
    double sumw = 0.0;
    double pc[DOPMAX];
#pragma unroll
    for (int b = 0; b < DOPMAX; b++) {
      double v = b < np ? sinc2_mean64(V, dopPOS, idop_min + b) : 0.0;
      pc[b] = v;
      if (b < np) sumw += v;
    }
    for (int b = DOPMAX; b < np; b++) sumw += sinc2_mean64(V, dopPOS, idop_min + b);
    const double *pl = params + V.pb + V.pb_law;
    double amp = radlaw64x(pl, V.law_type, pd.cose[p], V.lawloc, V.lawloc ? pd.f[p] : -1) * V.kmpp * V.kmpp / sumw;
    if (small) {
      double *row = use_lds ? win + (idop_min - lo_p) : img + (idop_min - 1);
#pragma unroll
      for (int b = 0; b < DOPMAX; b++)
        if (b < np) atomicAdd(&row[b], amp * pc[b]);
    } else if (in_bounds) {
      for (int b = 0; b < np; b++) atomicAdd(&img[idop_min + b - 1], amp * sinc2_mean64(V, dopPOS, idop_min + b));
    } else {
      int idop0 = V.idop0;
      int i1 = max(idop_min, -idop0), i2 = min(idop_max, -idop0 + 2000 - 1);
      double *ov = overflow + (long long)blockIdx.y * 2000;
      for (int b = i1; b <= i2; b++) atomicAdd(&ov[b + idop0], amp * sinc2_mean64(V, dopPOS, b));
    }
  }
  if (use_lds) {
    __syncthreads();
    for (int q = threadIdx.x; q < wp; q += blockDim.x) {
      double v = win[q];
      if (v != 0.0) atomicAdd(&img[lo_p + q - 1], v);
    }
  }
  limits_block64(V, false, 0, 0, valid ? idop_min : 0x7fffffff, valid ? idop_max : -0x7fffffff, 0.0, 0.0,
                 valid ? dopPOS : 1.0e20, valid ? dopPOS : -1.0e20, valid && !in_bounds ? 1 : 0, shi, shd);
}

// ---------------------------------------------------------------------------
// per-view finishing: POS-bound log factor (posvis.c:301-307), delay-Doppler
// limits in usec/Hz (pos2deldop.c:543-548), Doppler overflow terms and the
// beyond-overflow check (pos2doppler.c:447-499)
// ---------------------------------------------------------------------------
__global__ void k_finish_views(ViewRec *views, int nviews, double *overflow, const FrameRec *frames,
                               double *view_out) {
  int iv = blockIdx.x * blockDim.x + threadIdx.x;
  if (iv >= nviews) return;
  ViewRec &V = views[iv];
  double *o = view_out + (long long)iv * VIEW_OUT;
  for (int q = 0; q < VIEW_OUT; q++) o[q] = 0.0;
  int n = V.n;
  for (int src = 0; src < (V.bistatic ? 2 : 1); src++) {
    int ob = src ? V.outbnd2 : V.outbnd;
    const unsigned long long *ov = src ? V.ov2 : V.ov;
    double lf = 0.0;
    if (ob) {
      double imin = unord64(ov[0]), imax = -unord64(ov[1]), jmin = unord64(ov[2]), jmax = -unord64(ov[3]);
      double xf = (dmax(imax, (double)n) - dmin(imin, (double)-n) + 1) / (2 * n + 1);
      double yf = (dmax(jmax, (double)n) - dmin(jmin, (double)-n) + 1) / (2 * n + 1);
      lf = log(xf * yf);
    }
    o[OUT_OUTBND + src] = ob;
    o[OUT_POSBND_LOG + src] = lf;
  }
  if (V.type == VT_DELAY || V.type == VT_DOPPLER) {
    double dl0 = unord64(V.dellim[0]), dl1 = -unord64(V.dellim[1]);
    double pl0 = unord64(V.doplim[0]), pl1 = -unord64(V.doplim[1]);
    if (V.type == VT_DELAY) {
      o[OUT_DELLIM] = (dl0 - V.delshift) * V.del_per_pixel;
      o[OUT_DELLIM + 1] = (dl1 - V.delshift) * V.del_per_pixel;
    }
    o[OUT_DOPLIM] = (pl0 - V.dopshift) * V.dop_per_pixel - V.dop_extra;
    o[OUT_DOPLIM + 1] = (pl1 - V.dopshift) * V.dop_per_pixel + V.dop_extra;
    int bad = V.badradar_sinc;
    double blog = V.badradar_sinc_log;
    double o2 = 0.0, m2 = 0.0, xs = 0.0, dm = 0.0;
    if (V.type == VT_DOPPLER && V.any_overflow) {
      double *ovb = overflow + (long long)iv * 2000;
      int idop0 = V.idop0;
      int j1 = max(V.idoplim[0] + idop0, 0), j2 = min(V.idoplim[1] + idop0, 1999);
      const FrameRec &F = frames[V.frame_id];
      double sd2 = F.sdev_sq, var = sd2;
      for (int j = j1; j <= j2; j++) {
        double fo = ovb[j];
        if (fo != 0.0) {
          if (F.speckle) var = sd2 + F.lookfact * fo * fo;
          o2 += 1.0;
          m2 += fo * fo / var;
          xs += fo;
          dm += (j - idop0) * fo;
        }
      }
      if (xs != 0.0) dm /= xs;
      if ((V.idoplim[0] + idop0) < 0 || (V.idoplim[1] + idop0) >= 2000) {
        bad = 1;
        double df = (max(V.idoplim[1] + idop0, 2000) - min(V.idoplim[0] + idop0, 0)) / (1.0 * 2000);
        blog += log(df);
      }
      for (int j = 0; j < 2000; j++) ovb[j] = 0.0;
    }
    o[OUT_OVF] = o2; o[OUT_OVF + 1] = m2; o[OUT_OVF + 2] = xs; o[OUT_OVF + 3] = dm;
    o[OUT_BADRADAR] = bad;
    o[OUT_BADRADAR_LOG] = blog;
    o[OUT_IDELLIM] = V.idellim[0]; o[OUT_IDELLIM + 1] = V.idellim[1];
    o[OUT_IDOPLIM] = V.idoplim[0]; o[OUT_IDOPLIM + 1] = V.idoplim[1];
    o[OUT_KM2HZ] = V.km2Hz;
  }
  if (V.type == VT_LGHTCRV) o[OUT_YSUM] = (V.ysum == 0.0) ? 1.0e-40 : V.ysum;
}

// ---------------------------------------------------------------------------
// lightcurve interpolation: natural cubic spline of y(x) over the calculated
// points, evaluated at each observation epoch and view (calc_fits.c:1923-1933)
// ---------------------------------------------------------------------------
__global__ void k_spline(FrameRec *frames, int nframes, const double *view_out, const ViewRec *views,
                         const double *lc_x, const double *lc_t, double *lc_y, double *lc_y2, double *fit) {
  int fi = blockIdx.x;
  if (fi >= nframes) return;
  const FrameRec &F = frames[fi];
  if (F.type != VT_LGHTCRV) return;
  int nc = F.lc_ncalc;
  const double *x = lc_x + F.lc_x_off;     // 1-based
  double *y = lc_y + F.lc_y_off, *y2 = lc_y2 + F.lc_y2_off;
  if (threadIdx.x == 0) {
    for (int i = 1; i <= nc; i++) y[i] = view_out[(long long)(F.first_view + i - 1) * VIEW_OUT + OUT_YSUM];
    // nr/spline.c with natural end conditions; u reuses y2's scratch half
    double *u = y2 + nc + 1;
    y2[1] = u[1] = 0.0;
    for (int i = 2; i <= nc - 1; i++) {
      double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
      double p = sig * y2[i - 1] + 2.0;
      y2[i] = (sig - 1.0) / p;
      u[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
      u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    double qn = 0.0, un = 0.0;
    y2[nc] = (un - qn * u[nc - 1]) / (qn * y2[nc - 1] + 1.0);
    for (int k = nc - 1; k >= 1; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
  }
  __syncthreads();
  const double *t = lc_t + F.lc_t_off;     // t[(i-1)*nviews + v]
  double *out = fit + F.off;               // 0-based fit[i-1]
  for (int i = 1 + threadIdx.x; i <= F.lc_n; i += blockDim.x) {
    double acc = 0.0;
    for (int v = 0; v < F.lc_nviews; v++) {
      double xv = t[(long long)(i - 1) * F.lc_nviews + v];
      int klo = 1, khi = nc;
      while (khi - klo > 1) {
        int k = (khi + klo) >> 1;
        if (x[k] > xv) khi = k; else klo = k;
      }
      double h = x[khi] - x[klo];
      double a = (x[khi] - xv) / h, b = (xv - x[klo]) / h;
      acc += a * y[klo] + b * y[khi] + ((a * a * a - a) * y2[klo] + (b * b * b - b) * y2[khi]) * (h * h) / 6.0;
    }
    out[i - 1] = acc / F.lc_nviews;
  }
}

// ---------------------------------------------------------------------------
// chi-square per frame (chi2.c:418-461, :752-785, :1137-1170): FP64 partial
// sums with Neumaier compensation and a fixed-order block reduction
// ---------------------------------------------------------------------------
__device__ inline void neumaier(double &s, double &c, double x) {
  double t = s + x;
  if (fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s;
  s = t;
}

__global__ void k_chi2_frames64(FrameRec *frames, int nframes, double *fit, const double *obs, const double *oov,
                                const double *view_out, int clear_fit, int ovf_from_host, int lc_only) {
  int fi = blockIdx.x;
  if (fi >= nframes) return;
  FrameRec &F = frames[fi];
  if (lc_only && F.type != VT_LGHTCRV) return;  // mixed method: radar frames have their own kernel
  long long N = (F.type == VT_DELAY) ? (long long)F.ndel * F.ndop : (F.type == VT_DOPPLER ? F.ndop : F.lc_n);
  double *ft = fit + F.off;
  const double *ob = obs + F.obs_off, *w = oov + F.obs_off;
  double inv_nv = 1.0 / F.nviews;
  double m2s = 0, m2c = 0, oms = 0, omc = 0;
  for (long long q = threadIdx.x; q < N; q += blockDim.x) {
    double fv = ft[q];
    // (the buffer keeps the sum over a smeared frame's views: readers divide)
    if (F.nviews > 1 && F.type != VT_LGHTCRV) fv = fv * inv_nv;
    if (clear_fit) ft[q] = 0.0;
    neumaier(m2s, m2c, fv * fv * w[q]);
    neumaier(oms, omc, fv * ob[q] * w[q]);
  }
  __shared__ double r0[256], r1[256], r2[256], r3[256];
  r0[threadIdx.x] = m2s; r1[threadIdx.x] = m2c; r2[threadIdx.x] = oms; r3[threadIdx.x] = omc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, ac = 0, b = 0, bc = 0;
    for (int q = 0; q < (int)blockDim.x; q++) {
      neumaier(a, ac, r0[q]); ac += r1[q];
      neumaier(b, bc, r2[q]); bc += r3[q];
    }
    double o2 = F.o2, m2 = a + ac, om = b + bc;
    if (ovf_from_host) {
      // chi2() on host-resident fits: overflow terms come from the frame (chi2.c:430-431)
      o2 = F.overflow_o2 + F.o2;
      m2 = F.overflow_m2 + m2;
    } else if (F.type == VT_DOPPLER) {
      // overflow terms, averaged over the views of the frame (calc_fits.c:939-960)
      double ovo2 = 0, ovm2 = 0, ovxs = 0, ovdm = 0, blog = 0;
      int bad = 0;
      for (int v = 0; v < F.nviews; v++) {
        const double *o = view_out + (long long)(F.first_view + v) * VIEW_OUT;
        ovo2 += o[OUT_OVF]; ovm2 += o[OUT_OVF + 1]; ovxs += o[OUT_OVF + 2]; ovdm += o[OUT_OVF + 3];
        if (o[OUT_BADRADAR] != 0.0) { bad = 1; blog += o[OUT_BADRADAR_LOG]; }
      }
      if (F.nviews > 1) { ovo2 /= F.nviews; ovm2 /= F.nviews; ovxs /= F.nviews; ovdm /= F.nviews; }
      F.overflow_o2 = ovo2; F.overflow_m2 = ovm2; F.overflow_xsec = ovxs; F.overflow_dopmean = ovdm;
      o2 = ovo2 + F.o2;
      m2 = ovm2 + m2;
      F.badradar = bad; F.badradar_logfactor = blog;
    } else if (F.type == VT_DELAY) {
      int bad = 0;
      double blog = 0;
      for (int v = 0; v < F.nviews; v++) {
        const double *o = view_out + (long long)(F.first_view + v) * VIEW_OUT;
        if (o[OUT_BADRADAR] != 0.0) { bad = 1; blog += o[OUT_BADRADAR_LOG]; }
      }
      F.overflow_o2 = F.overflow_m2 = F.overflow_xsec = F.overflow_dopmean = 0.0;
      F.badradar = bad; F.badradar_logfactor = blog;
    }
    double cal = F.cal_fixed;
    if (F.cal_floats) cal = (om > 0.0) ? om / m2 : 1.0e-40;
    F.cal = cal;
    F.m2 = m2;
    F.om = om;
    F.chi2 = F.weight * (o2 - 2 * cal * om + cal * cal * m2);
  }
}

// objective per evaluation (bestfit.c:789-866; the penalty sum from k_pen_eval), in
// fixed frame order; log factors follow calc_fits.c:407-409, :511-512,
// :547-550, :1654-1677, :1947
__global__ __launch_bounds__(256) void k_objective(EvalRec *evals, int nevals, const FrameRec *frames,
                                                   const ViewRec *views, const double *view_out, const int *slot_flags,
                                                   const double *slot_logf, const int *eval_slot,
                                                   const double *pen_total, const int *job_flag,
                                                   const double *job_log, int photo_nblk) {
  // one workgroup per evaluation: the threads fetch 256 frames' terms at a
  // time into LDS, one thread adds them in frame order (the CPU's order)
  const int e = blockIdx.x;
  if (e >= nevals) return;
  EvalRec &E = evals[e];
  __shared__ double s_chi[256], s_pb[256], s_br[256];
  __shared__ int s_fl[256];
  double chi2 = 0.0, pb = 0.0, br = 0.0;
  int posbnd = 0, badradar = 0;
  for (int q0 = 0; q0 < E.nframes; q0 += 256) {
    const int q = q0 + threadIdx.x;
    if (q < E.nframes) {
      const FrameRec &F = frames[E.first_frame + q];
      double fpb = 0.0, fbr = 0.0;
      int fl = 0;
      if (F.type == VT_LGHTCRV) {
        double lf = 0.0;
        for (int i = 0; i < F.lc_ncalc; i++) {
          const double *o = view_out + (long long)(F.first_view + i) * VIEW_OUT;
          if (o[OUT_OUTBND] != 0.0) { fl |= 1; lf += 0.5 * o[OUT_POSBND_LOG]; }
          if (o[OUT_OUTBND + 1] != 0.0) { fl |= 1; lf += 0.5 * o[OUT_POSBND_LOG + 1]; }
        }
        fpb = F.dof * (lf / F.lc_ncalc);
      } else {
        for (int v = 0; v < F.nviews; v++) {
          const ViewRec &V = views[F.first_view + v];
          const double *o = view_out + (long long)(F.first_view + v) * VIEW_OUT;
          if (V.is_v0 && o[OUT_OUTBND] != 0.0) { fl |= 1; fpb += F.dof * o[OUT_POSBND_LOG]; }
          if (o[OUT_BADRADAR] != 0.0) { fl |= 2; fbr += F.dof * o[OUT_BADRADAR_LOG] / F.nviews; }
        }
      }
      s_chi[threadIdx.x] = F.chi2; s_pb[threadIdx.x] = fpb; s_br[threadIdx.x] = fbr; s_fl[threadIdx.x] = fl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 0; k < min(256, E.nframes - q0); k++) {
        chi2 += s_chi[k];
        if (s_fl[k] & 1) { posbnd = 1; pb += s_pb[k]; }
        if (s_fl[k] & 2) { badradar = 1; br += s_br[k]; }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  E.chi2 = chi2;
  E.posbnd = posbnd;
  E.badradar = badradar;
  E.posbnd_logfactor = pb / E.dof;
  E.badradar_logfactor = br / (E.dof_deldop + E.dof_doppler);
  int sl = eval_slot[e];
  E.baddiam = slot_flags[sl];
  E.baddiam_logfactor = slot_logf[sl];
  E.badphoto = 0;
  E.badphoto_logfactor = 0.0;
  for (int q = 0; q < E.photo_njobs * photo_nblk; q++) {  // the jobs in law order, their blocks in order
    E.badphoto |= job_flag[E.photo_job0 * photo_nblk + q];
    E.badphoto_logfactor += job_log[E.photo_job0 * photo_nblk + q];
  }
  double err = chi2 / E.dof;
  if (pen_total) err += pen_total[e];  // bestfit.c:807-811
  if (E.baddiam) err *= E.bad_objfactor * exp(E.baddiam_logfactor);
  if (E.badphoto) err *= E.bad_objfactor * exp(E.badphoto_logfactor);
  if (E.posbnd) err *= E.bad_objfactor * exp(E.posbnd_logfactor);
  if (E.badradar) err *= E.bad_objfactor * exp(E.badradar_logfactor);
  E.objective = err;
}

}  // namespace shp

namespace shp {

// ---------------------------------------------------------------------------
// realize_photo (realize_photo.c:122-969) as the objective calls it (albedo
// factors 1): grid (facet blocks, jobs), one job per scattering law per
// evaluation; each workgroup leaves its flag and log factor at
// job * gridDim.x + block
// ---------------------------------------------------------------------------

// checkphotopar (realize_photo.c:965-1004): the log factor a value adds;
// mode bit 0 makes the lower bound exclusive (modes 1, 3), bit 1 the upper
__device__ inline double photo_check(double v, double lo, double hi, int mode, int &bad) {
  double add = 0.0;
  if ((mode == 0 || mode == 2) ? v < lo : v <= lo) { bad = 1; add += log(1 + lo - v); }
  if ((mode == 0 || mode == 1) ? v > hi : v >= hi) { bad = 1; add += log(1 + v - hi); }
  return add;
}

// associated Legendre function P_l^m(x) with the Condon-Shortley phase, by
// the three-term recurrence in l from P_m^m
__device__ double photo_plm(int l, int m, double x) {
  double pmm = 1.0;
  if (m > 0) {
    const double s = sqrt((1.0 - x) * (1.0 + x));
    double fact = 1.0;
    for (int i = 1; i <= m; i++) { pmm *= -fact * s; fact += 2.0; }
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

// sqrt((2l+1) (l-m)! / (l+m)!)
__device__ inline double photo_nlm(int l, int m) {
  double r = 1.0;
  for (int k = l - m + 1; k <= l + m; k++) r /= k;
  return sqrt((2 * l + 1) * r);
}

// an expansion's value at (cos theta, phi); c in the model file's order
// a00, a10, a11, b11, a20, a21, b21, ...
__device__ double photo_harm(const double *c, int L, double ct, double phi) {
  double v = 0.0;
  int q = 0;
  for (int l = 0; l <= L; l++) {
    v += c[q++] * (photo_nlm(l, 0) * photo_plm(l, 0, ct));
    for (int m = 1; m <= l; m++) {
      const double p = photo_nlm(l, m) * photo_plm(l, m, ct);
      const double a = c[q], b = c[q + 1];
      q += 2;
      v += a * (cos(m * phi) * p) + b * (sin(m * phi) * p);
    }
  }
  return v;
}

__global__ __launch_bounds__(256) void k_realize_photo(const PhotoJob *jobs, const double *params,
                                                       const GeomSlot *slots, double *lawloc, int *job_flag,
                                                       double *job_log) {
  const PhotoJob &J = jobs[blockIdx.y];
  const double *d = params + J.data;
  double *out = lawloc + J.loc;
  double lg = 0.0;
  int bad = 0;
  const int k = J.k;
  const bool first = blockIdx.x == 0;  // the global values and tables: the job's first workgroup
  if (J.kind == PHOTO_HOMOG) {
    if (first && threadIdx.x == 0)
      for (int q = 0; q < J.nchk; q++) lg += photo_check(params[J.chk[q]], J.lo[q], J.hi[q], J.mode[q], bad);
  } else if (J.kind == PHOTO_TABLE) {
    if (first && threadIdx.x == 0)
      for (int i = 0; i < J.n; i++) {
        const double v = (i > 0 && d[J.n + i] != 0.0) ? out[i - 1] : d[i];
        out[i] = v;
        lg += photo_check(v, J.lo[0], J.hi[0], J.mode[0], bad);
      }
  } else {
    if (J.kind == PHOTO_INHO && first && threadIdx.x == 0)
      for (int j = 0; j < k; j++) lg += photo_check(d[j], J.lo[j], J.hi[j], J.mode[j], bad);
    const GeomSlot &G = slots[J.slot];
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < J.n; f += gridDim.x * blockDim.x) {
      double v[5];
      if (J.kind == PHOTO_HARM) {
        const double *x = G.fx + 3 * (long long)f;
        const double theta = atan2(sqrt(x[0] * x[0] + x[1] * x[1]), x[2]), phi = atan2(x[1], x[0]);
        const double ct = cos(theta);
        for (int j = 0; j < k; j++) v[j] = photo_harm(d + J.coef[j], J.nhar[j], ct, phi);
      } else {
        const double *loc = d + k + (long long)k * f, *tie = d + k + (long long)k * J.n + (long long)k * f;
        for (int j = 0; j < k; j++) v[j] = tie[j] != 0.0 ? d[j] : loc[j];
      }
      for (int j = 0; j < k; j++) {
        out[(long long)k * f + j] = v[j];
        lg += photo_check(v[j], J.lo[j], J.hi[j], J.mode[j], bad);
      }
    }
  }
  __shared__ double s_lg[256];
  __shared__ int s_bad[256];
  s_lg[threadIdx.x] = lg;
  s_bad[threadIdx.x] = bad;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s_lg[threadIdx.x] += s_lg[threadIdx.x + o];
      s_bad[threadIdx.x] |= s_bad[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int q = blockIdx.y * gridDim.x + blockIdx.x;
    job_flag[q] = s_bad[0];
    job_log[q] = s_lg[0];
  }
}

// facets seen from Earth (calc_fits.c:522-532, :903-913, :1699-1710)
__global__ void k_mark_seen(ViewRec *views, GeomSlot *slots, PosDev pd, double mincos) {
  const ViewRec &V = views[blockIdx.y];
  if (V.fp64_skip) return;
  if (V.exclude_seen || !V.is_v0) return;
  int n = V.n, w = 2 * n + 1;
  int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= w * w) return;
  int i = pix / w - n, j = pix % w - n;
  if (i < V.xlim[0] || i > V.xlim[1] || j < V.ylim[0] || j > V.ylim[1]) return;
  long long p = V.pos_off + pix;
  int f = pd.f[p];
  if (f < 0 || !(pd.cose[p] > mincos)) return;
  if (V.type == VT_LGHTCRV && !(pd.cosi[p] > mincos)) return;
  slots[V.slot].seen[f] = 1;
}

__global__ void k_fill_i32(int *p, long long n, int v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_f64(double *p, long long n, double v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u8(unsigned char *p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0;
}
__global__ void k_zero_counters(int *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

}  // namespace shp

namespace shp {

// o2 = sum obs^2/variance per frame (chi2.c:438), computed once per scene
__global__ void k_frame_o2(FrameRec *frames, int nframes, const double *obs, const double *oov) {
  int fi = blockIdx.x;
  if (fi >= nframes) return;
  FrameRec &F = frames[fi];
  long long N = (F.type == VT_DELAY) ? (long long)F.ndel * F.ndop : (F.type == VT_DOPPLER ? F.ndop : F.lc_n);
  const double *ob = obs + F.obs_off, *w = oov + F.obs_off;
  double s = 0, c = 0;
  for (long long q = threadIdx.x; q < N; q += blockDim.x) neumaier(s, c, ob[q] * ob[q] * w[q]);
  __shared__ double r0[256], r1[256];
  r0[threadIdx.x] = s; r1[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, ac = 0;
    for (int q = 0; q < (int)blockDim.x; q++) { neumaier(a, ac, r0[q]); ac += r1[q]; }
    F.o2 = a + ac;
  }
}

// chi2 totals per set and for the data set, in set order (chi2.c:270-300)
__global__ void k_chi2_sets(const FrameRec *frames, int nframes, double *set_chi2, int nsets, double *total) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int s = 0; s < nsets; s++) set_chi2[s] = 0.0;
  for (int q = 0; q < nframes; q++) set_chi2[frames[q].set] += frames[q].chi2;
  double t = 0.0;
  for (int s = 0; s < nsets; s++) t += set_chi2[s];
  *total = t;
}

}  // namespace shp
