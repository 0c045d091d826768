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

// One search's state between rounds: propose() writes the next round's k
// trial values (the first nnew new, the rest repeating the first), accept()
// takes their objective values.  A search that has stopped proposes its
// best point k times (its results are dropped), so several searches can
// share every round's replay.
struct Search {
  int k = 0, max_rounds = 0, r = 0, status = 1;
  bool done = false, started = false;
  double x0 = 0, step = 0, tol = 0, atol = 0;
  std::vector<Pt> pts;
  std::vector<double> cand;
  int nnew = 0;

  int init(int k_, double x0_, double step_, double tol_, double abstol, int max_rounds_) {
    if (k_ < 3 || k_ > SHAPE_SEARCH_MAX_K || !std::isfinite(x0_) || !std::isfinite(step_) || step_ == 0.0 ||
        !(tol_ >= 0.0) || max_rounds_ < 1)
      return -1;
    k = k_; x0 = x0_; step = step_; tol = tol_; atol = std::max(abstol, 1e-10); max_rounds = max_rounds_;
    return 0;
  }
  double best_x() const { return pts.empty() ? x0 : pts[lowest(pts)].x; }
  double best_f() const { return pts.empty() ? HUGE_VAL : pts[lowest(pts)].f; }
  void fill(double *xs, int n) {
    nnew = n;
    for (int q = n; q < k; q++) xs[q] = xs[0];
  }
  // the next round's points; false when the search has stopped
  bool propose(double *xs) {
    if (done) {
      for (int q = 0; q < k; q++) xs[q] = best_x();
      nnew = 0;
      return false;
    }
    if (!started) {
      started = true;
      for (int q = 0; q < k; q++) {
        const double mag = q == 0 ? 0.0 : std::pow(kGold, (q - 1) / 2);
        xs[q] = x0 + step * ((q & 1) ? mag : -mag);
      }
      fill(xs, k);
      return true;
    }
    const int m = lowest(pts), n = (int)pts.size();
    if (m == 0 || m == n - 1) {  // lowest at an end: continue past it
      if (r >= max_rounds) return stop(1, xs);
      const int dir = m == 0 ? -1 : 1;
      const double edge = pts[m].x, span = std::fabs(edge - pts[m - dir].x);
      double d = span;
      for (int q = 0; q < k; q++) {
        d *= kGold;
        xs[q] = edge + dir * d;
      }
      fill(xs, k);
      return true;
    }
    const double a = pts[m - 1].x, b = pts[m + 1].x, x = pts[m].x;
    const double tol1 = tol * std::fabs(x) + atol, xm = 0.5 * (a + b);
    if (std::fabs(x - xm) <= 2.0 * tol1 - 0.5 * (b - a)) return stop(0, xs);
    if (r >= max_rounds) return stop(1, xs);
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
    int nn = 0;
    for (double c : cand) {
      if (!(c > a && c < b) || nearest_gap(pts, c) < 0.5 * tol1) continue;
      bool close = false;
      for (int s2 = 0; s2 < nn; s2++) close |= std::fabs(xs[s2] - c) < 0.5 * tol1;
      if (!close) xs[nn++] = c;
    }
    if (nn == 0) return stop(0, xs);  // nothing left to resolve at this tolerance
    fill(xs, nn);
    return true;
  }
  bool stop(int st, double *xs) {
    status = st;
    done = true;
    return propose(xs);
  }
  void accept(const double *xs, const double *fs) {
    for (int q = 0; q < nnew; q++) insert_point(pts, xs[q], fs[q]);
    if (nnew > 0) r++;
  }
};

}  // namespace

extern "C" int shape_line_search(shape_eval_fn eval, void *ctx, int k, double x0, double step, double tol,
                                 double abstol, int max_rounds, double *xmin, double *fmin, int *rounds) {
  if (!eval || !xmin || !fmin) return -1;
  Search S;
  if (S.init(k, x0, step, tol, abstol, max_rounds)) return -1;
  std::vector<double> xs(k), fs(k);
  while (S.propose(xs.data())) {
    if (eval(ctx, k, xs.data(), fs.data())) return -2;
    S.accept(xs.data(), fs.data());
  }
  *xmin = S.best_x();
  *fmin = S.best_f();
  if (rounds) *rounds = S.r;
  return S.status;
}

// nsearch independent searches, k trial values each, sharing every round:
// eval receives nsearch * k values (search s at [s k, (s + 1) k)).  A search
// that has stopped repeats its best point until the others stop.  rounds:
// the rounds (calls of eval) taken.  Returns 0 (every search converged), 1
// (some reached max_rounds), -1 bad arguments, -2 an evaluation failed.
extern "C" int shape_line_searches(shape_eval_fn eval, void *ctx, int nsearch, int k, const double *x0,
                                   const double *step, double tol, double abstol, int max_rounds, double *xmin,
                                   double *fmin, int *status, int *rounds) {
  if (!eval || !x0 || !step || !xmin || !fmin || nsearch < 1) return -1;
  std::vector<Search> S(nsearch);
  for (int s = 0; s < nsearch; s++)
    if (S[s].init(k, x0[s], step[s], tol, abstol, max_rounds)) return -1;
  std::vector<double> xs((size_t)nsearch * k), fs((size_t)nsearch * k);
  int r = 0;
  for (;;) {
    bool any = false;
    for (int s = 0; s < nsearch; s++) any |= S[s].propose(xs.data() + (size_t)s * k);
    if (!any) break;
    if (eval(ctx, nsearch * k, xs.data(), fs.data())) return -2;
    r++;
    for (int s = 0; s < nsearch; s++) S[s].accept(xs.data() + (size_t)s * k, fs.data() + (size_t)s * k);
  }
  int worst = 0;
  for (int s = 0; s < nsearch; s++) {
    xmin[s] = S[s].best_x();
    fmin[s] = S[s].best_f();
    if (status) status[s] = S[s].status;
    worst = std::max(worst, S[s].status);
  }
  if (rounds) *rounds = r;
  return worst;
}
