"""Launch settings (shape_batch_tune / set_tune) change how the mixed
method's work is spread over the device, never its results: the persistent
tile kernel, its per-XCD work counters over two replays in a row (they reset
themselves), other replay and collection grids, all against the per-tile
grid on the same batch: objective and chi-square to FP64 atomic-ordering
rounding (1e-12), per-frame chi-squares, replayed-pixel counts and facet
maps identical.  The tuner returns valid settings, keeps results, and a
second batch of the same shape takes the cached choice."""
import numpy as np
import pytest

import helpers as H

sh = H.sh
pytestmark = pytest.mark.gpu


def _scene(config, **over):
    """Scene whose observations are the FP64 drop-in's fits plus noise (the
    comparisons here are between launch settings, not against the oracle)."""
    hip = H.load_hip()
    sc = sh.Scene(sh.spec_for(config, **over))
    hip.shape_hip_clear_error()
    hip.realize_mod(sc.par, sc.mod)
    hip.realize_spin(sc.par, sc.mod, sc.dat)
    hip.realize_delcor(sc.dat, 0.0, 0)
    hip.calc_fits(sc.par, sc.mod, sc.dat)
    assert hip.shape_hip_last_error() == 0
    sc.synthesize_obs(snr=20.0)
    return sc


def _batch(config, flags=0, **over):
    sc = _scene(config, **over)
    trials = [[], [(0, sc.get_param(0) * 1.01)], [(3, sc.get_param(3) * 1.0005)]]
    B = sh.Batch([sc], [(0, t) for t in trials], method=sh.METHOD_MIXED, flags=flags | sh.BATCH_FACET_MAP)
    return sc, B


def _state(sc, B):
    obj, chi = B.run()
    n = sc.spec.pos_n
    maps = [B.facet_map(e, q, n) for e in range(B.nevals) for q, (s, f) in enumerate(sc.all_frames())
            if f >= 0 and sc.set_type(s) != sh.LGHTCRV]
    frames = [B.frame_results(e)[0] for e in range(B.nevals)]
    return obj, chi, frames, maps, B.counters()


def _same(a, b):
    obj_a, chi_a, fr_a, maps_a, cnt_a = a
    obj_b, chi_b, fr_b, maps_b, cnt_b = b
    assert np.allclose(obj_a, obj_b, rtol=1e-12, atol=0) and np.allclose(chi_a, chi_b, rtol=1e-12, atol=0)
    for x, y in zip(fr_a, fr_b):
        assert np.allclose(x, y, rtol=1e-11, atol=0)
    assert cnt_a[0] == cnt_b[0] and cnt_a[1] == 0 and cnt_b[1] == 0
    assert all(np.array_equal(x, y) for x, y in zip(maps_a, maps_b))


@pytest.mark.parametrize("config,over", [(1, {}), (5, {}), (4, {"n_deldop": 8})])
def test_persistent_tile_kernel_matches_grid(config, over):
    sc, B = _batch(config, **over)
    assert B.get_tune()[0] == 0
    ref = _state(sc, B)
    for grid in (8, 512, 1536):
        B.set_tune([grid, 2048, 2048])
        _same(ref, _state(sc, B))
        _same(ref, _state(sc, B))  # the work counters were left at zero
    B.set_tune([0, 256, 256])
    _same(ref, _state(sc, B))
    B.set_tune([0, 8192, 4096])
    _same(ref, _state(sc, B))


def test_set_tune_rejects_bad_grids():
    sc, B = _batch(1)
    with pytest.raises(RuntimeError):
        B.set_tune([12, 2048, 2048])  # not a multiple of 8
    with pytest.raises(RuntimeError):
        B.set_tune([0, 0, 2048])


def test_tuner_keeps_results_and_caches(tmp_path):
    sc, B = _batch(5, flags=sh.BATCH_GRAPH)
    ref = _state(sc, B)
    settings, ms = B.tune(replays=3, force=True)
    assert ms > 0 and settings[0] % 8 == 0 and settings[1] >= 1 and settings[2] >= 1
    assert B.get_tune() == settings
    _same(ref, _state(sc, B))
    sc2, B2 = _batch(5, flags=sh.BATCH_GRAPH)
    s2, ms2 = B2.tune(replays=3)
    assert s2 == settings and ms2 == ms  # taken from the cache, not timed again
    _same(ref, _state(sc2, B2))
