/* oracle_search.c -- the fit's one-parameter line search, restated for the
 * CPU (test infrastructure only: tests/ load it as the checker of the
 * batched line search).
 *
 * bestfit.c:446-468 brackets a minimum of objective(x) for the hot parameter
 * with mnbrak from (x0, x0 + step) and then homes in with brent_abs, Brent's
 * method with the fractional tolerance tol and an absolute tolerance abstol
 * (floored at 1e-10).  Both follow the published Numerical Recipes in C
 * routines (Press et al. 1992, sections 10.1-10.2): golden-ratio expansion
 * with parabolic extrapolation limited to GLIMIT times the last interval,
 * then parabolic steps guarded by golden-section steps.
 */
#include <math.h>

#include "oracle.h"
#include "oracle_internal.h"

#define O_GOLD 1.618034
#define O_GLIMIT 100.0
#define O_TINY 1.0e-20
#define O_CGOLD 0.3819660
#define O_ZEPS 1.0e-10
#define O_ITMAX 100

static double o_sign(double a, double b) { return b > 0.0 ? fabs(a) : -fabs(a); }

/* nr/mnbrak.c:10-73: on return b lies between a and c, f(b) < f(a), f(b) <= f(c) */
void oracle_mnbrak(double *a, double *b, double *c, double *fa, double *fb, double *fc, oracle_fn f)
{
  *fa = f(*a);
  *fb = f(*b);
  if (*fb > *fa) {  /* go downhill from a to b */
    double t = *a; *a = *b; *b = t;
    t = *fa; *fa = *fb; *fb = t;
  }
  *c = *b + O_GOLD * (*b - *a);
  *fc = f(*c);
  while (*fb > *fc) {
    /* parabola through (a, b, c): its vertex u, and how far to look */
    double r = (*b - *a) * (*fb - *fc);
    double q = (*b - *c) * (*fb - *fa);
    double den = 2.0 * o_sign(fmax(fabs(q - r), O_TINY), q - r);
    double u = *b - ((*b - *c) * q - (*b - *a) * r) / den;
    double ulim = *b + O_GLIMIT * (*c - *b);
    double fu;
    if ((*b - u) * (u - *c) > 0.0) {          /* vertex between b and c */
      fu = f(u);
      if (fu < *fc) {                         /* minimum between b and c */
        *a = *b; *b = u; *fa = *fb; *fb = fu;
        return;
      }
      if (fu > *fb) {                         /* minimum between a and u */
        *c = u; *fc = fu;
        return;
      }
      u = *c + O_GOLD * (*c - *b);            /* the vertex did not help */
      fu = f(u);
    } else if ((*c - u) * (u - ulim) > 0.0) { /* vertex between c and the limit */
      fu = f(u);
      if (fu < *fc) {
        *b = *c; *c = u; u = *c + O_GOLD * (*c - *b);
        *fb = *fc; *fc = fu; fu = f(u);
      }
    } else if ((u - ulim) * (ulim - *c) >= 0.0) { /* vertex beyond the limit: clip */
      u = ulim;
      fu = f(u);
    } else {                                  /* vertex behind: golden step */
      u = *c + O_GOLD * (*c - *b);
      fu = f(u);
    }
    *a = *b; *b = *c; *c = u;
    *fa = *fb; *fb = *fc; *fc = fu;
  }
}

/* nr/brent_abs.c:13-86: Brent's method on the bracket (ax, bx, cx) until the
 * best point is within tol*|x| + max(abstol, 1e-10) of the bracket's middle;
 * returns f there and sets *xmin (the last point if ITMAX steps pass) */
double oracle_brent_abs(double ax, double bx, double cx, oracle_fn f, double tol, double abstol, double *xmin)
{
  const double atol = abstol > O_ZEPS ? abstol : O_ZEPS;
  double a = ax < cx ? ax : cx, b = ax > cx ? ax : cx;
  double x = bx, w = bx, v = bx;            /* best, second best, previous w */
  double fx = f(x), fw = fx, fv = fx;
  double d = 0.0, e = 0.0;                  /* this step, the step before last */
  for (int it = 0; it < O_ITMAX; it++) {
    const double xm = 0.5 * (a + b);
    const double tol1 = tol * fabs(x) + atol, tol2 = 2.0 * tol1;
    if (fabs(x - xm) <= tol2 - 0.5 * (b - a)) {
      *xmin = x;
      return fx;
    }
    int golden = 1;
    if (fabs(e) > tol1) {                   /* try a parabolic step through x, w, v */
      double r = (x - w) * (fx - fv);
      double q = (x - v) * (fx - fw);
      double p = (x - v) * q - (x - w) * r;
      q = 2.0 * (q - r);
      if (q > 0.0) p = -p;
      q = fabs(q);
      const double eprev = e;
      e = d;
      if (!(fabs(p) >= fabs(0.5 * q * eprev) || p <= q * (a - x) || p >= q * (b - x))) {
        d = p / q;
        const double u = x + d;
        if (u - a < tol2 || b - u < tol2) d = o_sign(tol1, xm - x);
        golden = 0;
      }
    }
    if (golden) {                           /* golden section into the larger part */
      e = x >= xm ? a - x : b - x;
      d = O_CGOLD * e;
    }
    const double u = fabs(d) >= tol1 ? x + d : x + o_sign(tol1, d);
    const double fu = f(u);
    if (fu <= fx) {
      if (u >= x) a = x; else b = x;
      v = w; w = x; x = u;
      fv = fw; fw = fx; fx = fu;
    } else {
      if (u < x) a = u; else b = u;
      if (fu <= fw || w == x) {
        v = w; w = u;
        fv = fw; fw = fu;
      } else if (fu <= fv || v == x || v == w) {
        v = u;
        fv = fu;
      }
    }
  }
  oracle_set_error("oracle_brent_abs: too many iterations");
  *xmin = x;
  return fx;
}

/* bestfit.c:446-468 for one parameter: bracket from (x0, x0 + step), then
 * Brent; returns the minimum's objective and sets *xmin */
double oracle_line_search(oracle_fn f, double x0, double step, double tol, double abstol, double *xmin)
{
  double a = x0, b = x0 + step, c, fa, fb, fc;
  oracle_mnbrak(&a, &b, &c, &fa, &fb, &fc, f);
  return oracle_brent_abs(a, b, c, f, tol, abstol, xmin);
}
