"""Batched searches on the GPU against the oracle's sequential mnbrak +
brent_abs (bestfit.c:392-470).

* Config 5's batched fit: 32 models, 8 trial values of the first scale
  factor per model per round, 32 searches sharing every replay
  (shape_batch_line_searches): each model's minimum against the oracle's
  search on that model.
* Config 1's pass over 10 parameters (3 scale factors, spin rate, pole,
  radar law C, delay polynomial a0, a1 and a vertex deviation;
  shape_batch_fit_pass), from a start 2% off the truth: every parameter
  after the pass against the oracle's bestfit pass, and the objective
  there.  (The radar law's R is left out: with the calibration factor
  floating, chi-square does not depend on it, and any value is a minimum.
  The vertex deviation moves a vertex on the far side of config 1's
  views: chi-square is flat in it up to a radius of ~0.835, and where the
  oracle's objective takes the same value at both minima, both are minima.)

Each step of a pass is checked against the oracle's search of the same
parameter from the same state (the batched pass's parameters before it).
FP64 method: minima agree to the searches' stopping tolerance (brent_abs's
test, tol1 = tol |x| + abstol; a batched round stops within 2 tol1,
brent_abs within 2 tol1, so 4 tol1 apart at most); objectives 1e-10.  Mixed
method: its objective is the oracle's to the 1e-6 contract -- near the
truth, where every frame's chi-square is at its noise floor, errors come
close to that -- which moves a shallow minimum by more than tol1, so its
minima are checked in the oracle's own objective: no more than twice the
contract (2e-6) above the oracle's minimum.
"""
import ctypes as C

import numpy as np
import pytest

import helpers as H

sh = H.sh
pytestmark = pytest.mark.gpu


def _oracle_search(orc, sc, i, x0, step, tol, abstol):
    def f(x):
        return H.oracle_objective(orc, sc, {i: x})
    xo = C.c_double()
    fo = orc.oracle_line_search(H.ORACLE_FN(f), x0, step, tol, abstol, C.byref(xo))
    return xo.value, fo


@pytest.mark.parametrize("method,tol_f", [(sh.METHOD_MIXED, 1e-6), (sh.METHOD_FP64, 1e-10)])
def test_many_model_searches_in_one_replay(method, tol_f):
    orc = H.load_oracle()
    nm, k = (32, 8) if method == sh.METHOD_MIXED else (8, 8)
    scenes = [H.truth_scene(5, seed_geom=5000 + m, seed_frames=5500 + m) for m in range(nm)]
    p0 = [sc.get_param(0) for sc in scenes]
    B = sh.Batch(scenes, [(m, [(0, p0[m])]) for m in range(nm) for _ in range(k)], method=method,
                 flags=sh.BATCH_GRAPH)
    tol, abstol = 1e-4, 1e-9
    x0 = [1.02 * p for p in p0]
    step = [0.01 * p for p in p0]
    rc, xb, fb, st, rounds = B.line_searches(x0, step, tol, abstol)
    assert rc == 0 and all(s == 0 for s in st)
    print("method", method, "models", nm, "rounds", rounds)
    worst = 0.0
    for m in range(nm):
        xo, fo = _oracle_search(orc, scenes[m], 0, x0[m], step[m], tol, abstol)
        t1 = tol * abs(xo) + abstol
        worst = max(worst, abs(xb[m] - xo) / t1)
        fo_at = H.oracle_objective(orc, scenes[m], {0: xb[m]})
        assert abs(fb[m] - fo_at) <= tol_f * abs(fo_at), (m, fb[m], fo_at)
        if method == sh.METHOD_FP64:
            assert abs(xb[m] - xo) <= 4 * t1, (m, xb[m], xo)
        else:
            assert fo_at <= fo + 2e-6 * abs(fo), (m, fo_at, fo)
    print("worst distance to the oracle's minimum, in tol1:", worst)


PASS_PARAMS = [0, 1, 2, 3, 4, 5, 7, 8, 9, 10]
PASS_STEPS = [0.01, 0.01, 0.01, 0.01 * 25.13, 0.01, 0.01, 0.05, 0.1, 0.01, 0.01]


@pytest.mark.parametrize("method,tol_f", [(sh.METHOD_FP64, 1e-10), (sh.METHOD_MIXED, 1e-6)])
def test_fit_pass_matches_oracle_bestfit_pass(method, tol_f):
    orc = H.load_oracle()
    sc = H.truth_scene(1)
    truth = [sc.get_param(i) for i in PASS_PARAMS]
    start = [v * 1.02 if v != 0.0 else 0.05 for v in truth]
    tol = [1e-4] * len(PASS_PARAMS)
    abstol = [1e-7] * len(PASS_PARAMS)
    k = 8
    B = sh.Batch([sc], [(0, [(i, start[q]) for q, i in enumerate(PASS_PARAMS)]) for _ in range(k)],
                 method=method, flags=sh.BATCH_GRAPH)
    rc, xb, fb, rounds = B.fit_pass(start, PASS_STEPS, tol, abstol)
    assert rc == 0
    # each step against the oracle's mnbrak + brent_abs of that parameter from
    # the batched pass's state before it
    P = len(PASS_PARAMS)
    calls = [0]
    worst = 0.0
    for q, i in enumerate(PASS_PARAMS):
        state = {PASS_PARAMS[r]: (xb[r] if r < q else start[r]) for r in range(P)}

        def f(x, state=state, i=i):
            calls[0] += 1
            return H.oracle_objective(orc, sc, {**state, i: x})
        xv = C.c_double()
        fo = orc.oracle_line_search(H.ORACLE_FN(f), start[q], PASS_STEPS[q], tol[q], abstol[q], C.byref(xv))
        t1 = tol[q] * abs(xv.value) + abstol[q]
        fb_q = f(xb[q])
        flat = abs(fb_q - fo) <= 1e-13 * abs(fo)
        if not flat:
            worst = max(worst, abs(xb[q] - xv.value) / t1)
        if method == sh.METHOD_FP64:
            assert flat or abs(xb[q] - xv.value) <= 4 * t1, (q, xb[q], xv.value, (xb[q] - xv.value) / t1)
        else:
            assert fb_q <= fo + 2e-6 * abs(fo), (q, fb_q, fo)
    fo_at_b = H.oracle_objective(orc, sc, {i: xb[q] for q, i in enumerate(PASS_PARAMS)})
    f_start = H.oracle_objective(orc, sc, {i: start[q] for q, i in enumerate(PASS_PARAMS)})
    print("method", method, "replays", rounds, "oracle evaluations", calls[0], "objective", f_start, "->", fb,
          "worst step distance in tol1 (parameters not flat)", worst)
    assert abs(fb - fo_at_b) <= tol_f * abs(fo_at_b), (fb, fo_at_b)
    assert fo_at_b < f_start
    assert rounds < calls[0]
