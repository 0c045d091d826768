// k_pen.hip -- the model's moments (realize_mod.c:757-840, facmom.c) and the
// objective's penalties (penalties.c:120-1240) on the device, FP64.
//
// Per geometry slot, in four launches (the last two only when a penalty
// reads vertices one by one):
//   k_moments_part   each workgroup reduces a contiguous range of facets,
//                    sides and vertices: facet area, volume, first moment and
//                    inertia tensor contributions; the nonsmooth and
//                    concavity terms of each side; the vertex extents;
//   k_moments_fin    adds the workgroups' partials in workgroup order,
//                    divides out the COM, and diagonalizes the inertia tensor
//                    (cyclic Jacobi, one thread) when a penalty needs the
//                    principal axes;
//   k_pen_vert_part  per vertex, with the effective radius and principal
//                    axes known: rdev / maxrdev deviations, the distance to
//                    the DEEVE (maxellipdev), bifurcation zone sums;
//   k_pen_vert_fin   the same ordered reduction of those partials.
// k_pen_eval then forms each evaluation's penalties from its slot's terms
// and its parameter block (spin inertia, impulses, dataset angle / spin
// offsets), and par->pen.base's rule for negative weights.
//
// Every reduction has a fixed shape (workgroup ranges, LDS tree, then
// workgroups in order), so results repeat bit for bit from replay to replay;
// they differ from the CPU's facet-order sums by rounding only.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace shp {

namespace {

constexpr int PT = 256;
constexpr int NBIF = 20, NZONES = NBIF - 1;  // penalties.c:116-117
constexpr double PI_D = 3.141592653589793, R2D_D = 57.29577951308232;
constexpr double TINYPEN_D = 1.0e-10, SMALLVAL_D = 1.0e-10, HUGE_D = 1.0e20;

__device__ inline double dmin(double a, double b) { return a < b ? a : b; }
__device__ inline double dmax(double a, double b) { return a > b ? a : b; }

// op 0: sum, 1: min, 2: max of one value per thread, over the workgroup in a
// fixed tree; every thread gets the result
__device__ double wg_reduce(double v, double *sh, int op) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = PT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double a = sh[threadIdx.x], b = sh[threadIdx.x + w];
      sh[threadIdx.x] = op == 0 ? a + b : op == 1 ? dmin(a, b) : dmax(a, b);
    }
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// One facet's volume between it and a coordinate plane, that volume's first
// moment and its inertia tensor (uniform density), taking each axis in turn
// as "z" by the cyclic relabelling (x, y, z) -> (z, x, y).  out: dv, dvr[3],
// dI00, dI11, dI22, dI01, dI02, dI12.
__device__ void facet_moments(const double *p0, const double *p1, const double *p2, const double *fn,
                              double out[10]) {
  double v0[3] = {p0[0], p0[1], p0[2]}, v1[3] = {p1[0], p1[1], p1[2]}, v2[3] = {p2[0], p2[1], p2[2]};
  double n[3] = {fn[0], fn[1], fn[2]};
  double dvr[3], d00 = 0.0, d11 = 0.0, d22 = 0.0, d01 = 0.0, d02 = 0.0, d12 = 0.0, dv = 0.0;
  for (int i = 0; i < 3; i++) {
    const int sign = n[2] < 0.0 ? -1 : 1;
    const double jac = fabs((v1[0] - v0[0]) * (v2[1] - v1[1]) - (v2[0] - v1[0]) * (v1[1] - v0[1]));
    const double zs = v0[2] + v1[2] + v2[2];
    if (i == 0) dv = sign * jac * zs / 6.0;
    const double r = sign * jac * (v0[2] * zs + v1[2] * (v1[2] + v2[2]) + v2[2] * v2[2]) / 24.0;
    const double z2 = v0[2] * v0[2] + v1[2] * v1[2] + v2[2] * v2[2];
    const double s = -(sign * jac / 120.0) *
                     ((v0[1] + v1[1] + v2[1]) * z2 + 2.0 * (v0[1] * v0[2] + v1[1] * v1[2] + v2[1] * v2[2]) * zs +
                      (v0[1] * v1[2] * v2[2] + v1[1] * v0[2] * v2[2] + v2[1] * v0[2] * v1[2]));
    const double t = (sign * jac / 60.0) * (z2 * zs + v0[2] * v1[2] * v2[2]);
    if (i == 0) { dvr[2] = r; d12 = s; d00 = t; d11 = t; }
    else if (i == 1) { dvr[1] = r; d01 = s; d00 += t; d22 = t; }
    else { dvr[0] = r; d02 = s; d11 += t; d22 += t; }
    double tmp;
    tmp = v0[2]; v0[2] = v0[1]; v0[1] = v0[0]; v0[0] = tmp;
    tmp = v1[2]; v1[2] = v1[1]; v1[1] = v1[0]; v1[0] = tmp;
    tmp = v2[2]; v2[2] = v2[1]; v2[1] = v2[0]; v2[0] = tmp;
    tmp = n[2]; n[2] = n[1]; n[1] = n[0]; n[0] = tmp;
  }
  out[0] = dv; out[1] = dvr[0]; out[2] = dvr[1]; out[3] = dvr[2];
  out[4] = d00; out[5] = d11; out[6] = d22; out[7] = d01; out[8] = d02; out[9] = d12;
}

__device__ inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Eigenvalues d and eigenvectors (columns of v) of a symmetric 3x3 matrix by
// cyclic Jacobi rotations (Numerical Recipes in C, 2nd ed., 11.1): sweeps over
// the upper triangle in row order; a threshold of 0.2 * sum|off-diagonal| / 9
// in the first three sweeps; after the fourth, an element negligible beside
// both diagonal entries is set to zero; the diagonal is refreshed from the
// accumulated updates after each sweep.  Returns 0, or 1 after 50 sweeps.
__device__ int jacobi3(const double ain[9], double d[3], double v[9]) {
  double a[3][3], b[3], z[3];
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++) { a[p][q] = ain[3 * p + q]; v[3 * p + q] = p == q ? 1.0 : 0.0; }
  for (int p = 0; p < 3; p++) { b[p] = d[p] = a[p][p]; z[p] = 0.0; }
  for (int sweep = 1; sweep <= 50; sweep++) {
    const double sm = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
    if (sm == 0.0) return 0;
    const double tresh = sweep < 4 ? 0.2 * sm / 9 : 0.0;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        const double g = 100.0 * fabs(a[p][q]);
        if (sweep > 4 && fabs(d[p]) + g == fabs(d[p]) && fabs(d[q]) + g == fabs(d[q])) {
          a[p][q] = 0.0;
        } else if (fabs(a[p][q]) > tresh) {
          double h = d[q] - d[p], t;
          if (fabs(h) + g == fabs(h)) {
            t = a[p][q] / h;
          } else {
            const double theta = 0.5 * h / a[p][q];
            t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
            if (theta < 0.0) t = -t;
          }
          const double c = 1.0 / sqrt(1 + t * t), s = t * c, tau = s / (1.0 + c);
          h = t * a[p][q];
          z[p] -= h; z[q] += h; d[p] -= h; d[q] += h;
          a[p][q] = 0.0;
          for (int j = 0; j < 3; j++) {
            double *x, *y;
            if (j < p) { x = &a[j][p]; y = &a[j][q]; }
            else if (j > p && j < q) { x = &a[p][j]; y = &a[j][q]; }
            else if (j > q) { x = &a[p][j]; y = &a[q][j]; }
            else continue;
            const double gg = *x, hh = *y;
            *x = gg - s * (hh + gg * tau);
            *y = hh + s * (gg - hh * tau);
          }
          for (int j = 0; j < 3; j++) {
            const double gg = v[3 * j + p], hh = v[3 * j + q];
            v[3 * j + p] = gg - s * (hh + gg * tau);
            v[3 * j + q] = hh + s * (gg - hh * tau);
          }
        }
      }
    for (int p = 0; p < 3; p++) { b[p] += z[p]; d[p] = b[p]; z[p] = 0.0; }
  }
  return 1;
}

__device__ inline double reff_of(double volume) { return pow(3 * volume / (4 * PI_D), 1.0 / 3.0); }

__device__ inline int cdiv_d(int a, int b) { return (a + b - 1) / b; }

}  // namespace

int pen_supported(int type) {
  return (type >= NONSMOOTH && type <= MAXELLIPDEV) || type == BIFURCATION || type == IMPULSE;
}

// the law-reading penalties, evaluated by the batch (k_pen_laws) only
int pen_supported_laws(int type) { return (type >= OPTALBDEL && type <= RAD_C_VAR) || type == NONCOSINE; }

int pen_needs(int type) {
  switch (type) {
  case NONSMOOTH: case CONCAVITY: case OPTALBDEL: case THETADEL: case RADALBDEL: case RAD_C_DEL: return PEN_NEED_SIDES;
  case RDEV: case MAXRDEV: return PEN_NEED_VERTS;
  case MAXELLIPDEV: case BIFURCATION: return PEN_NEED_VERTS | PEN_NEED_PA;
  case INERTIADEV_UNI: case PA3TILT: case NONPA_UNI: case FLATTENING: return PEN_NEED_PA;
  default: return 0;
  }
}

// grid (nb, slots): workgroup b reduces facets, sides and vertices of its
// contiguous range; part[(slot * nb + b) * MP1_N + q]:
//   0 area, 1 dv, 2-4 dvr, 5-10 dI (00, 11, 22, 01, 02, 12), 11 nonsmooth,
//   12 concavity, 13-15 min x, 16-18 max x
__global__ __launch_bounds__(256) void k_moments_part(const GeomSlot *slots, double *part, int nb) {
  const GeomSlot &G = slots[blockIdx.y];
  __shared__ double sh[PT];
  double acc[13];
  for (int q = 0; q < 13; q++) acc[q] = 0.0;
  double lo[3] = {HUGE_D, HUGE_D, HUGE_D}, hi[3] = {-HUGE_D, -HUGE_D, -HUGE_D};
  const int b = blockIdx.x;
  {
    const int ch = cdiv_d(G.nf, nb), f0 = b * ch, f1 = min(G.nf, f0 + ch);
    for (int f = f0 + (int)threadIdx.x; f < f1; f += PT) {
      const int *fv = G.fv + 3 * f;
      double m[10];
      facet_moments(G.x + 3 * fv[0], G.x + 3 * fv[1], G.x + 3 * fv[2], G.fn + 3 * f, m);
      acc[0] += G.area[f];
      for (int q = 0; q < 10; q++) acc[1 + q] += m[q];
    }
  }
  if (G.sides) {
    const int ch = cdiv_d(G.ns, nb), s0 = b * ch, s1 = min(G.ns, s0 + ch);
    for (int s = s0 + (int)threadIdx.x; s < s1; s += PT) {
      const int4 S = G.sides[s];  // v0, v1, f0, f1
      const double *n1 = G.fn + 3 * S.z, *n2 = G.fn + 3 * S.w;
      const double c = dot3(n1, n2), x = 1 - c;
      acc[11] += x * x * x * x;
      // far corner of the first facet; concave when it lies in front of the second facet's plane
      const int *fv = G.fv + 3 * S.z;
      int v3 = 0;
      for (int j = 0; j < 3; j++)
        if (fv[j] != S.x && fv[j] != S.y) v3 = fv[j];
      double disp[3];
      for (int j = 0; j < 3; j++) disp[j] = G.x[3 * v3 + j] - G.x[3 * S.x + j];
      if (dot3(disp, n2) > 0.0) acc[12] += x * x;
    }
  }
  {
    const int ch = cdiv_d(G.nv, nb), v0 = b * ch, v1 = min(G.nv, v0 + ch);
    for (int v = v0 + (int)threadIdx.x; v < v1; v += PT)
      for (int j = 0; j < 3; j++) {
        lo[j] = dmin(lo[j], G.x[3 * v + j]);
        hi[j] = dmax(hi[j], G.x[3 * v + j]);
      }
  }
  double *out = part + ((size_t)blockIdx.y * nb + b) * MP1_N;
  for (int q = 0; q < 13; q++) {
    const double r = wg_reduce(acc[q], sh, 0);
    if (threadIdx.x == 0) out[q] = r;
  }
  for (int j = 0; j < 3; j++) {
    const double a = wg_reduce(lo[j], sh, 1), c = wg_reduce(hi[j], sh, 2);
    if (threadIdx.x == 0) { out[13 + j] = a; out[16 + j] = c; }
  }
}

// grid (slots), 64 threads.  given (drop-in penalties: one slot) holds the
// caller's area, volume, com[3], inertia[9] and realized scale factors[3];
// otherwise the moments are the partials' sums and the scale factors the
// parameter block's, made safe as realize_mod.c:458-465 does.
__global__ void k_moments_fin(const GeomSlot *slots, const double *params, const double *part, int nb, double *mom,
                              const double *given, int need_pa) {
  const GeomSlot &G = slots[blockIdx.x];
  __shared__ double r[MP1_N];
  const int q = threadIdx.x;
  if (q < MP1_N) {
    const double *p = part + (size_t)blockIdx.x * nb * MP1_N + q;
    double a = q < 13 ? 0.0 : q < 16 ? HUGE_D : -HUGE_D;
    for (int b = 0; b < nb; b++) {
      const double x = p[(size_t)b * MP1_N];
      a = q < 13 ? a + x : q < 16 ? dmin(a, x) : dmax(a, x);
    }
    r[q] = a;
  }
  __syncthreads();
  if (q != 0) return;
  double *m = mom + (size_t)blockIdx.x * MOM_N;
  if (given) {
    m[MOM_AREA] = given[0];
    m[MOM_VOL] = given[1];
    for (int j = 0; j < 3; j++) m[MOM_COM + j] = given[2 + j];
    for (int k = 0; k < 9; k++) m[MOM_I + k] = given[5 + k];
    for (int j = 0; j < 3; j++) m[MOM_SF + j] = given[14 + j];
  } else {
    const double vol = r[1];
    m[MOM_AREA] = r[0];
    m[MOM_VOL] = vol;
    for (int j = 0; j < 3; j++) m[MOM_COM + j] = r[2 + j] / vol;
    const double d00 = r[5], d11 = r[6], d22 = r[7], d01 = r[8], d02 = r[9], d12 = r[10];
    const double I[9] = {d00, d01, d02, d01, d11, d12, d02, d12, d22};
    for (int k = 0; k < 9; k++) m[MOM_I + k] = I[k];
    const double *pc = params + G.comp_pb;
    for (int j = 0; j < 3; j++) {
      double sf = pc[j];
      if (sf <= 0.01) sf = 0.01 / (1 + 0.01 - sf);
      m[MOM_SF + j] = sf;
    }
  }
  m[MOM_NONSMOOTH] = r[11] / G.ns;
  m[MOM_CONCAV] = r[12] / G.ns;
  for (int j = 0; j < 3; j++) { m[MOM_LO + j] = r[13 + j]; m[MOM_HI + j] = r[16 + j]; }
  m[MOM_JACOBI_FAIL] = 0.0;
  if (need_pa) m[MOM_JACOBI_FAIL] = jacobi3(m + MOM_I, m + MOM_PM, m + MOM_AP);
}

// grid (nb, slots): per vertex, with the slot's effective radius, scale
// factors and principal axes; part[(slot * nb + b) * MP2_N + q]:
//   0 sum of squared deviations (rdev), 1 their maximum (maxrdev),
//   2 maximum squared deviation from the DEEVE (maxellipdev),
//   3 + k  bifurcation zone k's weighted sum of squared axis distance,
//   3 + NZONES + k  its weight
__global__ __launch_bounds__(256) void k_pen_vert_part(const GeomSlot *slots, const double *mom, double *part, int nb,
                                                       int need) {
  const GeomSlot &G = slots[blockIdx.y];
  const double *m = mom + (size_t)blockIdx.y * MOM_N;
  __shared__ double sh[PT];
  const double vol = m[MOM_VOL], reff = reff_of(vol);
  const double *sf = m + MOM_SF, *pm = m + MOM_PM, *ap = m + MOM_AP;
  const bool want_pa = (need & PEN_NEED_PA) != 0;
  // DEEVE radii (penalties.c:781-796)
  double deeve[3] = {1.0, 1.0, 1.0};
  int jmax = 0, j1 = 1, j2 = 2;
  double lo = 0.0, incr = 1.0;
  if (want_pa) {
    double rad[3];
    rad[0] = sqrt((5 / vol) * (-pm[0] + pm[1] + pm[2]) / 2);
    rad[1] = sqrt((5 / vol) * (pm[0] - pm[1] + pm[2]) / 2);
    rad[2] = sqrt((5 / vol) * (pm[0] + pm[1] - pm[2]) / 2);
    const double scale = pow(3 * vol / (4 * PI_D * rad[0] * rad[1] * rad[2]), 1.0 / 3.0);
    for (int j = 0; j < 3; j++) deeve[j] = scale * rad[j];
    // the axis of the smallest principal moment (penalties.c:1116-1121)
    if (pm[0] <= pm[1] && pm[0] <= pm[2]) jmax = 0;
    else if (pm[1] <= pm[0] && pm[1] <= pm[2]) jmax = 1;
    else jmax = 2;
    j1 = (jmax + 1) % 3;
    j2 = (jmax + 2) % 3;
    lo = m[MOM_LO + jmax];
    incr = (m[MOM_HI + jmax] - lo + SMALLVAL_D) / NBIF;
  }
  double rsum = 0.0, rmax = 0.0, emax = 0.0;
  const bool vertex_comp = G.comp_type != ELLIPSE && G.comp_type != OVOID && G.comp_type != HARMONIC;
  double zs[NZONES], zc[NZONES];
#pragma unroll
  for (int k = 0; k < NZONES; k++) zs[k] = zc[k] = 0.0;
  const int ch = cdiv_d(G.nv, nb), v0 = blockIdx.x * ch, v1 = min(G.nv, v0 + ch);
  for (int v = v0 + (int)threadIdx.x; v < v1; v += PT) {
    const double *x = G.x + 3 * v;
    double scale = 0.0;
    for (int j = 0; j < 3; j++) {
      const double y = G.u[3 * v + j] * sf[j];
      scale += y * y;
    }
    scale = sqrt(scale);
    if (vertex_comp) {  // rdev / maxrdev visit VERTEX components only (penalties.c:701, :741)
      const double dx = scale * G.r[v] / reff, d2 = dx * dx;
      rsum += d2;
      rmax = dmax(rmax, d2);
    }
    if (want_pa) {
      // body -> principal-axis frame: w = ap^T x
      double w[3];
      for (int i = 0; i < 3; i++) w[i] = ap[0 + i] * x[0] + ap[3 + i] * x[1] + ap[6 + i] * x[2];
      const double rr = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
      const double theta = atan2(sqrt(w[0] * w[0] + w[1] * w[1]), w[2]), phi = atan2(w[1], w[0]);
      const double xt = sin(theta) * cos(phi) / deeve[0], yt = sin(theta) * sin(phi) / deeve[1],
                   zt = cos(theta) / deeve[2];
      const double rd = 1 / sqrt(xt * xt + yt * yt + zt * zt);
      double e = (rr - rd) / reff;
      e *= e;
      emax = dmax(emax, e);
      // bifurcation zones (penalties.c:1146-1178)
      const double rho2 = x[j1] * x[j1] + x[j2] * x[j2];
      const double xz = (x[jmax] - lo) / incr;
      const int k = (int)floor(xz);
      const double y = xz - k;
      int k1, k2;
      double w1, w2;
      if (k == 0) { k1 = k2 = 0; w1 = 0.0; w2 = 1.0; }
      else if (k == NZONES) { k1 = k2 = NZONES - 1; w1 = 1.0; w2 = 0.0; }
      else { k1 = k - 1; k2 = k; w1 = 1 - pow(y, 6); w2 = 1 - pow(1 - y, 6); }
#pragma unroll
      for (int z = 0; z < NZONES; z++) {
        if (z == k1) { zs[z] += w1 * rho2; zc[z] += w1; }
        if (z == k2) { zs[z] += w2 * rho2; zc[z] += w2; }
      }
    }
  }
  double *out = part + ((size_t)blockIdx.y * nb + blockIdx.x) * MP2_N;
  double r;
  r = wg_reduce(rsum, sh, 0); if (threadIdx.x == 0) out[0] = r;
  r = wg_reduce(rmax, sh, 2); if (threadIdx.x == 0) out[1] = r;
  if (!want_pa) return;  // (uniform over the workgroup)
  r = wg_reduce(emax, sh, 2); if (threadIdx.x == 0) out[2] = r;
#pragma unroll
  for (int z = 0; z < NZONES; z++) {
    r = wg_reduce(zs[z], sh, 0); if (threadIdx.x == 0) out[3 + z] = r;
    r = wg_reduce(zc[z], sh, 0); if (threadIdx.x == 0) out[3 + NZONES + z] = r;
  }
}

// grid (slots), 64 threads
__global__ void k_pen_vert_fin(const GeomSlot *slots, const double *part, int nb, double *mom, int need) {
  const GeomSlot &G = slots[blockIdx.x];
  __shared__ double r[MP2_N];
  const int q = threadIdx.x;
  const bool want_pa = (need & PEN_NEED_PA) != 0;
  const int nq = want_pa ? MP2_N : 2;
  if (q < nq) {
    const double *p = part + (size_t)blockIdx.x * nb * MP2_N + q;
    const bool mx = q == 1 || q == 2;
    double a = 0.0;
    for (int b = 0; b < nb; b++) {
      const double x = p[(size_t)b * MP2_N];
      a = mx ? dmax(a, x) : a + x;
    }
    r[q] = a;
  }
  __syncthreads();
  if (q != 0) return;
  double *m = mom + (size_t)blockIdx.x * MOM_N;
  m[MOM_RDEV] = r[0] / G.nv;
  m[MOM_MAXRDEV] = r[1];
  if (!want_pa) return;
  m[MOM_MAXELL] = r[2];
  double mean[NZONES];
  for (int k = 0; k < NZONES; k++) mean[k] = r[3 + k] / r[3 + NZONES + k];
  double deepest = HUGE_D;
  for (int k = 1; k < NZONES - 1; k++)
    for (int k1 = 0; k1 < k; k1++)
      if (mean[k1] > mean[k])
        for (int k2 = k + 1; k2 < NZONES; k2++)
          if (mean[k2] > mean[k]) deepest = dmin(deepest, mean[k] / ((mean[k1] + mean[k2]) / 2));
  m[MOM_BIF] = deepest < 1.0 ? (1 / deepest - 1) * (1 / deepest - 1) : 0.0;
}

// grid (evaluations), 64 threads (thread 0 works): each penalty of the list
// from the evaluation's slot terms and parameter block (penalties.c:145-1237)
__global__ void k_pen_eval(const EvalRec *evals, int nevals, const int *eval_slot, const double *params, PenDev P) {
  const int e = blockIdx.x;
  if (e >= nevals || threadIdx.x != 0) return;
  const EvalRec &E = evals[e];
  const double *m = P.mom + (size_t)eval_slot[e] * MOM_N;
  const double *pb = params + E.pb;
  const double *I = m + MOM_I, *pm = m + MOM_PM, *ap = m + MOM_AP, *si = pb + PB_INERTIA;
  double sum = 0.0;
  for (int i = 0; i < P.n; i++) {
    double pen = 0.0;
    switch (P.type[i]) {
    case NONSMOOTH: pen = m[MOM_NONSMOOTH]; break;
    case CONCAVITY: pen = m[MOM_CONCAV]; break;
    case RDEV: pen = m[MOM_RDEV]; break;
    case MAXRDEV: pen = m[MOM_MAXRDEV]; break;
    case MAXELLIPDEV: pen = m[MOM_MAXELL]; break;
    case BIFURCATION: pen = m[MOM_BIF]; break;
    case VOLUME: pen = m[MOM_VOL]; break;
    case COMDEV: {
      const double reff = reff_of(m[MOM_VOL]);
      for (int k = 0; k < 3; k++) pen += m[MOM_COM + k] * m[MOM_COM + k];
      pen /= (reff * reff);
      break;
    }
    case INERTIADEV:
    case INERTIADEV_UNI: {
      const double *pv = P.type[i] == INERTIADEV ? si : pm;
      double a = 0.0, b = 0.0, adotb = 0.0;
      for (int j = 0; j < 3; j++) {
        a += pv[j] * pv[j];
        adotb += pv[j] * I[4 * j];
        for (int k = 0; k < 3; k++) b += I[3 * j + k] * I[3 * j + k];
      }
      pen = 1 - adotb / sqrt(a * b);
      break;
    }
    case PA3TILT: pen = dmax(0.0, 1 - ap[8] * ap[8]); break;
    case NONPA:
    case NONPA_UNI: {
      const double *pv = P.type[i] == NONPA ? si : pm;
      pen = dmax((pv[0] / pv[2]) - 1, (pv[1] / pv[2]) - 1) + 0.01;
      if (pen < 0.0) pen = 0.0;
      break;
    }
    case FLATTENING: {
      const double boc = sqrt(dmin((-pm[0] + pm[1] + pm[2]), (pm[0] - pm[1] + pm[2])) / (pm[0] + pm[1] - pm[2]));
      if (boc > 1.0) {
        const double x = boc - 1;
        pen = x * x * x * x;
      }
      break;
    }
    case EULEROFFS: {
      int ntot = 0;
      for (int k = 0; k < E.pen_nsets; k++) {
        const double *ps = pb + E.pen_set_off + k * PB_SET_STRIDE;
        for (int j = 0; j < 3; j++) {
          pen += ps[PB_SET_ANGLEOFF + j] * ps[PB_SET_ANGLEOFF + j];
          pen += ps[PB_SET_OMEGAOFF + j] * ps[PB_SET_OMEGAOFF + j];
          ntot += 2;
        }
      }
      pen *= R2D_D * R2D_D / ntot;
      break;
    }
    case IMPULSE: {
      const int nimp = (int)pb[PB_NIMP];
      for (int k = 0; k < nimp; k++)
        for (int j = 0; j < 3; j++) pen += pb[PB_IMP + 3 * k + j] * pb[PB_IMP + 3 * k + j];
      pen *= R2D_D * R2D_D / (3 * nimp);
      break;
    }
    case OPTALBDEL: case OPTALBVAR: case THETAVAR: case THETADEL: case RADALBDEL: case RADALBVAR: case RAD_C_DEL:
    case RAD_C_VAR: case NONCOSINE:
      pen = P.law[(size_t)e * PEN_LAW_N + pen_law_slot(P.type[i])];
      break;
    default: break;  // rejected when the penalty list is packed
    }
    const double base = P.weight[i] >= 0.0 ? pen : 1.0 / dmax(pen, TINYPEN_D);
    P.base[(size_t)e * P.n + i] = base;
    sum += fabs(P.weight[i]) * base;
  }
  P.total[e] = sum;
}

// ---------------------------------------------------------------------------
// The law-reading penalties (penalties.c:150-588) from the evaluation's
// resolved per-facet values (k_realize_photo): one workgroup per evaluation.
// kind 0 the optical albedo (R, or Hapke's w), 1 Hapke's theta, 2 the radar
// R, 3 the radar C.
// ---------------------------------------------------------------------------
__device__ inline int law_value_index(const PhotoJob &J, int kind) {
  if (J.kind != PHOTO_HARM && J.kind != PHOTO_INHO) return -1;
  if (kind >= 2) return J.radar ? kind - 2 : -1;
  if (J.radar) return -1;
  const bool hapke = J.type == HARMHAPKE || J.type == INHOHAPKE;
  if (kind == 1) return hapke ? 4 : -1;
  return 0;  // R or w
}

__device__ inline void blk_sum2(double &a, double &b, double *sh) {
  sh[threadIdx.x] = a;
  sh[256 + threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sh[threadIdx.x] += sh[threadIdx.x + o];
      sh[256 + threadIdx.x] += sh[256 + threadIdx.x + o];
    }
    __syncthreads();
  }
  a = sh[0];
  b = sh[256];
  __syncthreads();
}

__device__ inline double logrho(double rho) {
  return rho > 1.0e-20 ? log10(rho) : log10(1.0e-20) + (rho / 1.0e-20 - 1) / LN10;
}

// need: bit pen_law_slot(type) per requested type; noncos: the table angles'
// log10 cos (n - 1 of them) then their mean
__global__ __launch_bounds__(256) void k_pen_laws(const EvalRec *evals, int nevals, const int *eval_slot,
                                                  const GeomSlot *slots, const PhotoJob *jobs, const double *lawloc,
                                                  const double *noncos, int need, double *out) {
  const int e = blockIdx.x;
  if (e >= nevals) return;
  const EvalRec &E = evals[e];
  const GeomSlot &G = slots[eval_slot[e]];
  __shared__ double sh[512];
  __shared__ int shn[256];
  for (int q = 0; q < 8; q++) {
    if (!(need >> q & 1)) continue;
    const int type = q + 1, kind = (type == OPTALBDEL || type == OPTALBVAR) ? 0
                                   : (type == THETADEL || type == THETAVAR) ? 1
                                   : (type == RADALBDEL || type == RADALBVAR) ? 2 : 3;
    const bool del = type == OPTALBDEL || type == THETADEL || type == RADALBDEL || type == RAD_C_DEL;
    double a = 0.0, b = 0.0;
    int ntot = 0;
    for (int jq = 0; jq < E.photo_njobs; jq++) {
      const PhotoJob &J = jobs[E.photo_job0 + jq];
      const int j = law_value_index(J, kind);
      if (j < 0) continue;
      const double *loc = lawloc + J.loc;
      const int k = J.k;
      if (del) {
        for (int s = threadIdx.x; s < G.ns; s += blockDim.x) {
          const int4 sd = G.sides[s];
          const double *x1 = G.x + 3 * (long long)sd.x, *x2 = G.x + 3 * (long long)sd.y;
          const double len = sqrt((x1[0] - x2[0]) * (x1[0] - x2[0]) + (x1[1] - x2[1]) * (x1[1] - x2[1]) +
                                  (x1[2] - x2[2]) * (x1[2] - x2[2]));
          const double x = loc[(long long)k * sd.z + j], y = loc[(long long)k * sd.w + j];
          a += len * fabs(x + y);
          b += len * fabs(x - y);
        }
      } else {
        // the reference's optalbvar does not count harmonic-Hapke facets
        const bool count = !(type == OPTALBVAR && J.type == HARMHAPKE);
        for (int f = threadIdx.x; f < J.n; f += blockDim.x) {
          const double x = loc[(long long)k * f + j];
          a += x;
          b += x * x;
          if (count) ntot++;
        }
      }
    }
    blk_sum2(a, b, sh);
    shn[threadIdx.x] = ntot;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) shn[threadIdx.x] += shn[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      double pen;
      if (del) pen = b / a;
      else {
        pen = shn[0] * (b / (a * a)) - 1.0;
        if (pen < 0.0) pen = 0.0;
      }
      out[(size_t)e * PEN_LAW_N + q] = pen;
    }
    __syncthreads();
  }
  if ((need >> 8 & 1) && threadIdx.x == 0) {
    int ntot = 0;
    double resid2sum = 0.0;
    for (int jq = 0; jq < E.photo_njobs; jq++) {
      const PhotoJob &J = jobs[E.photo_job0 + jq];
      if (!(J.radar && J.kind == PHOTO_TABLE)) continue;
      const double *rho = lawloc + J.loc, xmean = noncos[J.n - 1];
      const int n = J.n;
      ntot += n;
      double ymean = 0.0;
      for (int i = 0; i < n - 1; i++) ymean += logrho(rho[i]);
      ymean /= (n - 1);
      double x2 = 0.0, y2 = 0.0, xy = 0.0;
      for (int i = 0; i < n - 1; i++) {
        const double x = noncos[i] - xmean, y = logrho(rho[i]) - ymean;
        x2 += x * x;
        y2 += y * y;
        xy += x * y;
      }
      const double slope = xy / x2, intercept = ymean - slope * xmean;
      resid2sum += y2 - xy * xy / x2;
      if (rho[n - 1] > rho[n - 2]) {
        const double y = logrho(rho[n - 1]) - (slope * noncos[n - 2] + intercept);
        resid2sum += y * y;
      }
    }
    out[(size_t)e * PEN_LAW_N + 8] = dmax(resid2sum / ntot, 0.0);
  }
}

}  // namespace shp
