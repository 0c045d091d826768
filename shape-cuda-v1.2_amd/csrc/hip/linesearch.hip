// linesearch.hip -- the fit's one-parameter line search, k trial values per
// round (host code).
//
// bestfit.c:446-468 minimizes objective(x) over the hot parameter with
// mnbrak (golden-ratio bracketing from x0 and x0 + step) and brent_abs
// (Brent's method), one objective evaluation after another: ~15-20
// sequential evaluations per parameter.  Here every round evaluates k trial
// values at once (one batched replay on the device), so a parameter takes a
// few rounds instead:
//   * bracketing: the first round spreads x0 + step * {0, +1, -1, +g, -g,
//     +g^2, ...} (g the golden ratio) on both sides; while the lowest point
//     is an end point, the next round continues past it at distances
//     span * g, g^2, ... (span: the gap to its inner neighbour);
//   * refining: with the lowest point x inside (a, b), its neighbours, a
//     round evaluates the vertex u of the parabola through the three, u -+
//     tol1, and an even grid over (a, b) for the other points.  A smooth
//     minimum is pinned by u -+ tol1 one round after bracketing.
// It stops with brent_abs's test: |x - (a + b) / 2| <= 2 tol1 - (b - a) / 2,
// tol1 = tol |x| + max(abstol, 1e-10).
#include <algorithm>
#include <cmath>
#include <vector>

#include "shape_abi.h"
#include "shape_hip.h"

namespace {

constexpr double kGold = 1.618034;

struct Pt {
  double x, f;
};

// evaluated points, sorted by x; a failed or non-finite objective counts as +inf
void insert_point(std::vector<Pt> &pts, double x, double f) {
  if (!std::isfinite(f)) f = HUGE_VAL;
  auto it = std::lower_bound(pts.begin(), pts.end(), x, [](const Pt &p, double v) { return p.x < v; });
  if (it != pts.end() && it->x == x) {
    it->f = std::min(it->f, f);
    return;
  }
  pts.insert(it, Pt{x, f});
}

int lowest(const std::vector<Pt> &pts) {
  int m = 0;
  for (int q = 1; q < (int)pts.size(); q++)
    if (pts[q].f < pts[m].f) m = q;
  return m;
}

double nearest_gap(const std::vector<Pt> &pts, double x) {
  auto it = std::lower_bound(pts.begin(), pts.end(), x, [](const Pt &p, double v) { return p.x < v; });
  double g = HUGE_VAL;
  if (it != pts.end()) g = std::fabs(it->x - x);
  if (it != pts.begin()) g = std::min(g, std::fabs(x - (it - 1)->x));
  return g;
}

}  // namespace

extern "C" int shape_line_search(shape_eval_fn eval, void *ctx, int k, double x0, double step, double tol,
                                 double abstol, int max_rounds, double *xmin, double *fmin, int *rounds) {
  if (!eval || !xmin || !fmin || k < 3 || k > SHAPE_SEARCH_MAX_K || !std::isfinite(x0) || !std::isfinite(step) ||
      step == 0.0 || !(tol >= 0.0) || max_rounds < 1)
    return -1;
  const double atol = std::max(abstol, 1e-10);
  std::vector<double> xs(k), fs(k);
  std::vector<Pt> pts;
  int r = 0;
  // evaluate xs[0, nnew) (the rest of the k slots repeat xs[0]; their results are dropped)
  auto run = [&](int nnew) -> int {
    for (int q = nnew; q < k; q++) xs[q] = xs[0];
    if (eval(ctx, k, xs.data(), fs.data())) return -2;
    for (int q = 0; q < nnew; q++) insert_point(pts, xs[q], fs[q]);
    r++;
    return 0;
  };
  for (int q = 0; q < k; q++) {
    const double mag = q == 0 ? 0.0 : std::pow(kGold, (q - 1) / 2);
    xs[q] = x0 + step * ((q & 1) ? mag : -mag);
  }
  if (run(k)) return -2;
  int status = 1;
  std::vector<double> cand;
  for (;;) {
    const int m = lowest(pts), n = (int)pts.size();
    if (m == 0 || m == n - 1) {  // lowest at an end: continue past it
      if (r >= max_rounds) break;
      const int dir = m == 0 ? -1 : 1;
      const double edge = pts[m].x, span = std::fabs(edge - pts[m - dir].x);
      double d = span;
      for (int q = 0; q < k; q++) {
        d *= kGold;
        xs[q] = edge + dir * d;
      }
      if (run(k)) return -2;
      continue;
    }
    const double a = pts[m - 1].x, b = pts[m + 1].x, x = pts[m].x;
    const double tol1 = tol * std::fabs(x) + atol, xm = 0.5 * (a + b);
    if (std::fabs(x - xm) <= 2.0 * tol1 - 0.5 * (b - a)) {
      status = 0;
      break;
    }
    if (r >= max_rounds) break;
    cand.clear();
    // vertex of the parabola through (a, fa), (x, fx), (b, fb)
    const double fa = pts[m - 1].f, fx = pts[m].f, fb = pts[m + 1].f;
    const double p = (x - a) * (fx - fb), q = (x - b) * (fx - fa);
    const double den = 2.0 * (p - q);
    if (den != 0.0 && std::isfinite(fa) && std::isfinite(fb)) {
      const double u = x - ((x - a) * p - (x - b) * q) / den;
      if (u > a && u < b) {
        cand.push_back(u);
        cand.push_back(u - tol1);
        cand.push_back(u + tol1);
      }
    }
    const int ng = k - (int)cand.size();
    for (int i = 1; i <= ng; i++) cand.push_back(a + (b - a) * i / (ng + 1));
    // new points inside (a, b), at least tol1 / 2 from the evaluated ones and each other
    int nnew = 0;
    for (double c : cand) {
      if (!(c > a && c < b) || nearest_gap(pts, c) < 0.5 * tol1) continue;
      bool close = false;
      for (int s = 0; s < nnew; s++) close |= std::fabs(xs[s] - c) < 0.5 * tol1;
      if (!close) xs[nnew++] = c;
    }
    if (nnew == 0) {  // nothing left to resolve at this tolerance
      status = 0;
      break;
    }
    if (run(nnew)) return -2;
  }
  const int m = lowest(pts);
  *xmin = pts[m].x;
  *fmin = pts[m].f;
  if (rounds) *rounds = r;
  return status;
}
