"""The device path on a job read from SHAPE's own input files.

A parameter file, a model file (the 2,292-facet config-1 vertex model) and an
obs file with two delay-Doppler frames (one vignetted by a pixel-weighting
mask), a Doppler set of two smeared views and a lightcurve, with data in raw
binary, RDF and gzip'd ASCII, are read by libshape_io.so.  Then:
* the FP64 drop-in (realize_mod, realize_spin, realize_delcor, calc_fits,
  chi2) must match the oracle run on a second read of the same files to the
  tolerances of the other drop-in tests (chi2 and fits to 1e-12 relative);
* the batched mixed method must match the oracle's objective (chi-square
  over degrees of freedom plus the file's nonsmooth and volume penalties) for
  trial values of the size and spin parameters to 1e-6 relative.
"""
import numpy as np
import pytest

import helpers as H
import test_io as TI
import test_oracle_pen as P

pytestmark = pytest.mark.gpu
sh = H.sh


def test_dropin_on_read_job(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    files = TI.write_job(".")[0]
    hip, orc = H.load_hip(), H.load_oracle()
    a, b = sh.Job(*files), sh.Job(*files)
    for job in (a, b):
        job.set_param(0, 1.02)
    ca = H.oracle_forward(orc, a)
    hip.shape_hip_clear_error()
    hip.realize_mod(b.par, b.mod)
    hip.realize_spin(b.par, b.mod, b.dat)
    hip.realize_delcor(b.dat, 0.0, 0)
    hip.shape_hip_set_strict(1)
    hip.calc_fits(b.par, b.mod, b.dat)
    hip.shape_hip_set_strict(0)
    cb = hip.chi2(b.par, b.dat, 0)
    assert hip.shape_hip_last_error() == 0
    print("read job: oracle chi2", ca, "device chi2", cb)
    assert abs(cb - ca) <= 1e-12 * abs(ca), (ca, cb)
    worst = 0.0
    for s, f in a.all_frames():
        fa, fb = a.image(s, max(f, 0), 0), b.image(s, max(f, 0), 0)
        assert np.max(np.abs(fa)) > 0, (s, f)
        worst = max(worst, H.rel_err(fb, fa))
    assert worst < 1e-12, worst
    a.close()
    b.close()


def test_mixed_batch_on_read_job(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    files = TI.write_job(".")[0]
    orc = P._oracle()
    ref = sh.Job(*files)
    trials = [[(0, 1.0)], [(0, 1.03)], [(3, ref.get_param(3) * 1.0005)], [(4, ref.get_param(4) + 0.01)]]
    want = []
    for t in trials:
        saved = {i: ref.get_param(i) for i, _ in t}
        for i, v in t:
            ref.set_param(i, v)
        H.oracle_forward(orc, ref)
        want.append(orc.oracle_objective(ref.par, ref.mod, ref.dat))
        for i, v in saved.items():
            ref.set_param(i, v)
    job = sh.Job(*files)
    batch = sh.Batch([job], [(0, t) for t in trials], method=sh.METHOD_MIXED)
    obj, chi = batch.run()
    print("read job objectives: oracle", want, "mixed", list(obj))
    for w, g in zip(want, obj):
        assert abs(g - w) <= 1e-6 * abs(w), (w, g)
    batch.close()
    job.close()
    ref.close()
