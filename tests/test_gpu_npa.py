"""Non-principal-axis rotators on the device against the oracle.

The setup kernel integrates Euler's equations per view (npa_dev.h) with the
oracle's fixed Bulirsch-Stoer steps (oracle_npa.c), so the view matrices
agree to the libm-vs-device ulps of the trigonometric calls (1e-12 here),
for both of par's integration methods, with and without a spin impulse.  The
FP64 drop-in's chi-square follows to 1e-12 and the mixed batch's objectives,
for trial values of the spin rate and pole, to 1e-6.
"""
import numpy as np
import pytest

import helpers as H
import test_oracle_npa as N

pytestmark = pytest.mark.gpu
sh = H.sh


def _pair(method, impulses=None, **over):
    a, b = H.scene_pair(1, n_deldop=3, epoch_span_days=20.0, **over)
    for sc in (a, b):
        sc.set_npa((2.0, 0.3, 25.0), N.I3, method)
        if impulses:
            t, imp = impulses
            sc.lib.shp_set_impulses(sc.mod, len(t), np.ascontiguousarray(t, dtype=float).ctypes.data_as(H.DP),
                                    np.ascontiguousarray(imp, dtype=float).ctypes.data_as(H.DP))
    return a, b


@pytest.mark.parametrize("method,impulse", [("periodic", False), ("brute", False), ("periodic", True)])
def test_realize_spin_npa_matches_oracle(method, impulse):
    hip, orc = H.load_hip(), H.load_oracle()
    t0 = H.sh.spec_for(1).t0
    imps = ([t0 + 7.0], [[0.5, -0.4, 0.3]]) if impulse else None
    a, b = _pair(method, imps)
    orc.oracle_realize_spin(a.par, a.mod, a.dat)
    assert not orc.oracle_last_error()
    hip.shape_hip_clear_error()
    hip.realize_spin(b.par, b.mod, b.dat)
    assert hip.shape_hip_last_error() == 0
    worst = 0.0
    for s, f in a.all_frames():
        va, vb = a.view(s, max(f, 0)), b.view(s, max(f, 0))
        worst = max(worst, np.abs(va[:9] - vb[:9]).max(), np.abs(va[18:24] - vb[18:24]).max() / 25.0)
    print(method, "impulse" if impulse else "", "worst view difference", worst)
    assert worst < 1e-12, worst


def test_dropin_chain_npa_matches_oracle():
    hip, orc = H.load_hip(), H.load_oracle()
    a, b = _pair("periodic")
    ca = H.oracle_forward(orc, a)
    hip.shape_hip_clear_error()
    hip.realize_mod(b.par, b.mod)
    hip.realize_spin(b.par, b.mod, b.dat)
    hip.realize_delcor(b.dat, 0.0, 0)
    hip.calc_fits(b.par, b.mod, b.dat)
    cb = hip.chi2(b.par, b.dat, 0)
    assert hip.shape_hip_last_error() == 0
    print("NPA chain chi2", ca, cb)
    assert abs(cb - ca) <= 1e-11 * abs(ca), (ca, cb)


@pytest.mark.parametrize("method", [sh.METHOD_FP64, sh.METHOD_MIXED])
def test_batch_npa_matches_oracle(method):
    orc = H.load_oracle()
    sc, _ = _pair("periodic")
    trials = [[], [(3, sc.get_param(3) * 1.0003)], [(4, sc.get_param(4) + 0.01), (5, sc.get_param(5) - 0.02)]]
    want = []
    for t in trials:
        saved = {i: sc.get_param(i) for i, _ in t}
        for i, v in t:
            sc.set_param(i, v)
        H.oracle_forward(orc, sc)
        want.append(orc.oracle_objective_tail(sc.par, sc.dat))
        for i, v in saved.items():
            sc.set_param(i, v)
    B = sh.Batch([sc], [(0, t) for t in trials], method=method)
    obj, _ = B.run()
    print("NPA batch", method, want, list(obj))
    tol = 1e-11 if method == sh.METHOD_FP64 else 1e-6
    for w, g in zip(want, obj):
        assert abs(g - w) <= tol * abs(w), (w, g)
    B.close()
