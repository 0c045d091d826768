"""The fit's one-parameter line search (CPU only).

The oracle's mnbrak + brent_abs (bestfit.c:446-468) is pinned against
scipy.optimize's bracket and brent (an independent implementation of the
same published Numerical Recipes methods) and against exact minimizers; the
batched search (shape_line_search, the library's host code, driven here
through a Python objective) is checked against the exact minimizers and the
oracle, on the tolerance brent_abs stops on: tol1 = tol |x| + abstol.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import optimize

import helpers as H

sh = H.sh

# (objective, exact minimizer, x0, step)
FUNCS = [
    (lambda x: (x - 1.3) ** 2 + 0.5, 1.3, 1.0, 0.05),
    (lambda x: math.cosh(x - 0.7) + 0.1 * x, 0.7 + math.asinh(-0.1), 0.0, 0.1),
    (lambda x: x ** 4 - 3.0 * x + 2.0, (0.75) ** (1.0 / 3.0), 2.0, -0.1),
    (lambda x: -math.exp(-(x - 20.0) ** 2 / 50.0), 20.0, 0.0, 0.5),   # far from x0: many brackets
    (lambda x: 1e3 * (x - 5e-3) ** 2, 5e-3, 0.0, 1e-3),
]


def _tol1(x, tol, abstol):
    return tol * abs(x) + max(abstol, 1e-10)


@pytest.fixture(scope="module")
def orc():
    return H.load_oracle()


def _counted(f):
    n = [0]

    def g(x):
        n[0] += 1
        return f(x)
    return g, n


@pytest.mark.parametrize("case", range(len(FUNCS)))
def test_oracle_mnbrak_matches_scipy_bracket(orc, case):
    f, _, x0, step = FUNCS[case]
    vals = [C.c_double(v) for v in (x0, x0 + step, 0.0, 0.0, 0.0, 0.0)]
    orc.oracle_mnbrak(*[C.byref(v) for v in vals], H.ORACLE_FN(f))
    a, b, c, fa, fb, fc = [v.value for v in vals]
    xa, xb, xc, ga, gb, gc, _ = optimize.bracket(f, xa=x0, xb=x0 + step, grow_limit=100.0)
    assert (a, b, c) == pytest.approx((xa, xb, xc), rel=1e-12, abs=1e-15)
    assert fb < fa and fb <= fc and (b - a) * (c - b) > 0


@pytest.mark.parametrize("case", range(len(FUNCS)))
def test_oracle_line_search_finds_the_minimum(orc, case):
    f, xstar, x0, step = FUNCS[case]
    tol, abstol = 1e-6, 1e-10
    xmin = C.c_double()
    g, n = _counted(f)
    fmin = orc.oracle_line_search(H.ORACLE_FN(g), x0, step, tol, abstol, C.byref(xmin))
    assert orc.oracle_last_error() in (None, b"")
    t1 = _tol1(xstar, tol, abstol)
    assert abs(xmin.value - xstar) <= 2 * t1
    assert fmin == f(xmin.value)
    # scipy's Brent on the oracle's bracket lands within the same tolerance
    a, b = C.c_double(x0), C.c_double(x0 + step)
    c, fa, fb, fc = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    orc.oracle_mnbrak(C.byref(a), C.byref(b), C.byref(c), C.byref(fa), C.byref(fb), C.byref(fc), H.ORACLE_FN(f))
    xs = optimize.brent(f, brack=(a.value, b.value, c.value), tol=tol)
    assert abs(xmin.value - xs) <= 4 * t1
    print("case", case, "oracle evaluations", n[0])


@pytest.mark.parametrize("k", [3, 8])
@pytest.mark.parametrize("case", range(len(FUNCS)))
def test_batched_search_finds_the_minimum(orc, case, k):
    f, xstar, x0, step = FUNCS[case]
    tol, abstol = 1e-6, 1e-10
    rc, xmin, fmin, rounds, seen = sh.line_search(f, k, x0, step, tol, abstol)
    assert rc == 0
    t1 = _tol1(xstar, tol, abstol)
    assert abs(xmin - xstar) <= 2 * t1
    assert fmin == f(xmin)
    assert rounds == len(seen) and all(len(r) == k for r in seen)
    xo = C.c_double()
    orc.oracle_line_search(H.ORACLE_FN(f), x0, step, tol, abstol, C.byref(xo))
    assert abs(xmin - xo.value) <= 4 * t1
    print("case", case, "k", k, "rounds", rounds)


def test_batched_search_rounds_on_a_parabola():
    """Bracketed in the first round, pinned by the parabola's vertex and
    vertex -+ tol1 in the second, converged without a third."""
    rc, xmin, _, rounds, seen = sh.line_search(lambda x: (x - 0.137) ** 2, 8, 0.1, 0.05, 1e-8, 1e-12)
    assert rc == 0 and rounds == 2
    assert xmin == pytest.approx(0.137, abs=2e-9)
    assert seen[0][:3] == [0.1, 0.15000000000000002, 0.05]


def test_batched_search_limits_and_errors():
    f = lambda x: (x - 3.0) ** 2
    rc, xmin, _, rounds, _ = sh.line_search(f, 8, 0.0, 0.01, 1e-6, 1e-10, max_rounds=1)
    assert rc == 1 and rounds == 1 and xmin > 0.0       # still climbing out of the first round
    # a region where the objective fails counts as +inf
    g = lambda x: float("nan") if x < 0.5 else (x - 0.8) ** 2
    rc, xmin, _, _, _ = sh.line_search(g, 8, 0.6, 0.3, 1e-6, 1e-10)
    assert rc == 0 and abs(xmin - 0.8) <= 2 * _tol1(0.8, 1e-6, 1e-10)
    lib = sh.load_hip()
    x, fv, r = C.c_double(), C.c_double(), C.c_int()
    cb = sh.SEARCH_FN(lambda ctx, k, xs, out: 0)
    assert lib.shape_line_search(cb, None, 2, 0.0, 1.0, 1e-6, 0.0, 10, C.byref(x), C.byref(fv), C.byref(r)) == -1
    assert lib.shape_line_search(cb, None, 8, 0.0, 0.0, 1e-6, 0.0, 10, C.byref(x), C.byref(fv), C.byref(r)) == -1
    bad = sh.SEARCH_FN(lambda ctx, k, xs, out: 1)
    assert lib.shape_line_search(bad, None, 8, 0.0, 1.0, 1e-6, 0.0, 10, C.byref(x), C.byref(fv), C.byref(r)) == -2


def _run_searches(k, tol, abstol, max_rounds=40):
    """shape_line_searches over every FUNCS objective at once (search s gets
    x values [s k, (s + 1) k) of each round)."""
    lib = sh.load_hip()
    ns = len(FUNCS)
    calls = [0]

    def cb(ctx, kk, x, out):
        assert kk == ns * k
        calls[0] += 1
        for q in range(kk):
            out[q] = FUNCS[q // k][0](x[q])
        return 0
    fn = sh.SEARCH_FN(cb)
    dp = C.POINTER(C.c_double)
    x0 = np.array([f[2] for f in FUNCS])
    st = np.array([f[3] for f in FUNCS])
    xm, fm, ss = np.zeros(ns), np.zeros(ns), np.zeros(ns, dtype=np.int32)
    r = C.c_int()
    rc = lib.shape_line_searches(fn, None, ns, k, x0.ctypes.data_as(dp), st.ctypes.data_as(dp), tol, abstol,
                                 max_rounds, xm.ctypes.data_as(dp), fm.ctypes.data_as(dp),
                                 ss.ctypes.data_as(C.POINTER(C.c_int)), C.byref(r))
    return rc, xm, fm, ss, r.value, calls[0]


@pytest.mark.parametrize("k", [4, 8])
def test_searches_sharing_rounds_equal_single_searches(k):
    """Searches sharing every round (shape_line_searches) reach exactly the
    minimum each reaches alone (shape_line_search): a search that stops early
    only repeats its best point; the shared run takes as many rounds as the
    slowest search."""
    tol, abstol = 1e-6, 1e-10
    rc, xm, fm, ss, rounds, calls = _run_searches(k, tol, abstol)
    assert rc == 0 and list(ss) == [0] * len(FUNCS) and calls == rounds
    most = 0
    for s, (f, xstar, x0, step) in enumerate(FUNCS):
        rc1, x1, f1, r1, _ = sh.line_search(f, k, x0, step, tol, abstol)
        assert rc1 == 0
        assert xm[s] == x1 and fm[s] == f1, (s, xm[s], x1)
        assert abs(xm[s] - xstar) <= 2 * _tol1(xstar, tol, abstol)
        most = max(most, r1)
    assert rounds == most


def test_searches_report_max_rounds():
    rc, xm, fm, ss, rounds, calls = _run_searches(4, 1e-12, 0.0, max_rounds=2)
    assert rc == 1 and rounds == 2 and 1 in list(ss)
