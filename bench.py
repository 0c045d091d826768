#!/usr/bin/env python3
"""Chi-square evaluations/sec of SHAPE's model-evaluation path on one MI355X.

Workload (default, the large-model headline): a 163,842-vertex / 327,680-facet
icosphere model (0.9 x 0.72 x 0.54 km, 0.2% radial noise), POS 1023 x 1023 at
2 m/px, 256 delay-Doppler frames of 128-512 rows x 64-256 columns, one view
per frame, cosine radar law.  Synthetic: the scene comes from the package's
scene builder with fixed seeds; the "observed" images are the plain FP64 GPU
drop-in's fits plus Gaussian noise.

One step is one replay of a batch of `--evals` objective evaluations (trial
values of the first size parameter, as in a line search): vertex realization,
rasterization, delay-Doppler synthesis, chi-square and the objective of every
evaluation, on the device.  The timed region launches K replays, each followed
by a wait on its stream, so every step is a complete, synchronous evaluation
round as the fitting loop uses it.

Printed on one JSON line: the mixed method's evaluations/sec (`value`), the
plain FP64 drop-in and the device-resident FP64 batch beside it, the mixed
method's chi-square error against the FP64 method on the same evaluations,
the roofline of the dominant kernel (algorithmic bytes / event-timed launch
duration on the batch's stream), and the oracle (CPU restatement of the
reference's FP64 path) timed on a bounded sample on the host cores.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "shape-cuda-v1.2_amd"))

import shape_hip as sh  # noqa: E402

HBM_PEAK_GBS = 8000.0
NFRAMES_FULL = 256
KSTEPS = 5


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--evals", type=int, default=4, help="objective evaluations per replay")
    ap.add_argument("--frames", type=int, default=NFRAMES_FULL, help="delay-Doppler frames (256: full workload)")
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--kernel-steps", type=int, default=5, help="eager replays timed per kernel")
    ap.add_argument("--cpu-frames", type=int, default=NFRAMES_FULL, help="frames in the CPU-baseline sample")
    ap.add_argument("--cpu-evals", type=int, default=20, help="timed oracle evaluations on all threads")
    ap.add_argument("--cpu1-frames", type=int, default=16, help="frames in the 1-thread CPU sample")
    ap.add_argument("--cpu1-evals", type=int, default=4, help="timed 1-thread oracle evaluations")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device (calibrates HBM byte counters), then exit")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch settings (no shape_batch_tune before the timed replays)")
    ap.add_argument("--configs", type=str, default=None,
                    help="comma-separated workload configs (1-5) to report one JSON line each for: "
                         "mixed and FP64 rates, latency, accuracy, facet-map mismatch, CPU baselines")
    ap.add_argument("--mixed-only", action="store_true",
                    help="time only the mixed replays (for profiler runs)")
    ap.add_argument("--line-search", type=int, default=0, metavar="K",
                    help="time the batched line search (K trial values per replay, mixed method) over the "
                         "first scale factor of the large-model workload, and replays of K evaluations")
    return ap.parse_args()


# ---------------------------------------------------------------------------
# scene and observations
# ---------------------------------------------------------------------------
def make_scene(frames):
    return dropin_obs(sh.Scene(sh.spec_for(4, n_deldop=frames)))


def dropin_obs(sc):
    """Observations = the plain FP64 drop-in's fits (the noiseless truth) + noise."""
    hip = sh.load_hip()
    # fits of the plain FP64 drop-in are the noiseless "truth"
    hip.shape_hip_clear_error()
    hip.realize_mod(sc.par, sc.mod)
    hip.realize_spin(sc.par, sc.mod, sc.dat)
    hip.realize_delcor(sc.dat, 0.0, 0)
    hip.calc_fits(sc.par, sc.mod, sc.dat)
    err = hip.shape_hip_last_error()
    if err:
        raise RuntimeError("FP64 drop-in failed while synthesizing observations (%d)" % err)
    sc.synthesize_obs(snr=20.0)
    return sc


def trial_evals(sc, n):
    """n trial values of parameter 0 (first scale factor) around its value."""
    p0 = sc.get_param(0)
    vals = [p0 * (1.0 + 0.01 * (k - (n - 1) / 2.0)) for k in range(n)]
    return [(0, [(0, v)]) for v in vals]


# ---------------------------------------------------------------------------
# timing
# ---------------------------------------------------------------------------
def time_batch(B, steps, warmup):
    for _ in range(warmup):
        B.launch()
        B.wait(want=False)
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        B.launch()
        B.wait(want=False)
        lat.append(B.last_ms())
    t1 = time.perf_counter()
    return (t1 - t0), np.array(lat)


def tune_batch(B, a):
    """The launch settings the replays run with: shape_batch_tune's choice
    (timed before, outside the measured region) unless --no-tune."""
    if not a.no_tune:
        log("tuning launch settings")
        B.tune(replays=3)
    t = B.get_tune()
    return {"tile_grid": t[0], "replay_grid": t[1], "collect_grid": t[2], "tuned": not a.no_tune}


def time_dropin(sc, steps):
    hip = sh.load_hip()
    hip.shape_hip_clear_error()
    chis = []
    t0 = time.perf_counter()
    for _ in range(steps):
        hip.realize_mod(sc.par, sc.mod)
        hip.realize_spin(sc.par, sc.mod, sc.dat)
        hip.realize_delcor(sc.dat, 0.0, 0)
        hip.calc_fits(sc.par, sc.mod, sc.dat)
        chis.append(hip.chi2(sc.par, sc.dat, 0))
    t1 = time.perf_counter()
    if hip.shape_hip_last_error():
        raise RuntimeError("FP64 drop-in failed (%d)" % hip.shape_hip_last_error())
    return (t1 - t0) / steps, chis[-1]


# Algorithmic HBM bytes per launch of the mixed kernels (one launch covers every
# evaluation of the batch), on the path's byte contract: per frame view, 16 B
# per pixel of the model's POS bounding box (an 8-byte (depth, facet) key
# written and read once) and 8 B per image bin (FP32 obs*w and w); per
# evaluation, 28 B per vertex read and 24 B written by the realization and
# 12 B per facet of indices.  fp: footprint of the replay (bbox_pix and
# image_cells summed over the views; nv and nf over the geometry slots).
def kernel_bytes(name, fp, claimed):
    if name in ("k_tile32", "k_shade_radar32"):
        # the POS raster and synthesis: the key traffic of the contract (the
        # binned kernel keeps its keys in LDS, so it moves far less) plus the
        # realized vertices (positions and normals) and facet indices it reads
        return 16 * fp["bbox_pix"] + 24 * fp["nv"] + 12 * fp["nf"]
    if name == "k_chi2_mixed":
        return 8 * fp["image_cells"]              # FP32 obs*w and w per bin
    return None


def path_bytes(fp, nevals, nv, nf):
    """The contract's bytes for one replay of nevals evaluations (B_eval summed)."""
    return nevals * (52 * nv + 12 * nf) + 16 * fp["bbox_pix"] + 8 * fp["image_cells"]


def image_error(sc, evals, nframes, stride=1):
    """Largest |mixed - FP64| / peak(FP64) over every stride-th frame's model
    image, for the first evaluation (the per-pixel image contract: <= 1e-4)."""
    Bm = sh.Batch([sc], evals, method=sh.METHOD_MIXED, flags=sh.BATCH_KEEP_FIT)
    Bf = sh.Batch([sc], evals, method=sh.METHOD_FP64, flags=sh.BATCH_KEEP_FIT)
    Bm.run()
    Bf.run()
    worst = 0.0
    frames = sc.all_frames()
    for q in range(0, min(nframes, len(frames)), stride):
        s_, f_ = frames[q]
        shape = sc.image(s_, max(f_, 0), 0).shape
        im, imf = Bm.fit(0, q, shape), Bf.fit(0, q, shape)
        worst = max(worst, float(np.max(np.abs(im - imf)) / max(float(np.max(np.abs(imf))), 1e-300)))
    Bm.close()
    Bf.close()
    return worst


def kernel_roofline(sc, nev, tuned):
    B = sh.Batch([sc], trial_evals(sc, nev), method=sh.METHOD_MIXED, flags=0)
    B.run()
    if tuned:
        B.tune(replays=3)  # the timed batch's settings (same shape: the tuner's cached choice)
    B.time_kernels(True)
    for _ in range(KSTEPS):
        B.launch()
        B.wait(want=False)
    stats = B.kernel_stats()
    fp = B.footprint()
    claimed = fp["claimed"]
    B.close()
    rows = {}
    for name, (ms, n) in stats.items():
        avg = ms / max(n, 1)
        by = kernel_bytes(name, fp, claimed)
        rows[name] = {"avg_ms": round(avg, 4), "launches": n,
                      "GBps": None if by is None or avg <= 0 else round(by / (avg * 1e-3) / 1e9, 1),
                      "bytes": by}
    dom = max(stats.items(), key=lambda kv: kv[1][0])[0]
    return dom, rows, fp, claimed


# ---------------------------------------------------------------------------
# CPU baseline: the oracle (the FP64 CPU restatement of the reference's
# path), timed on the host's cores beside the GPU, as the HostMultiThreading
# variant partitions the work (frames dealt to threads, one POS per frame)
# ---------------------------------------------------------------------------
def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(nframes, nevals, nframes1, nevals1):
    """Oracle evaluations/s of the large-model workload: nevals timed
    evaluations (after one warm-up) of the first nframes frames on every
    thread of the process's CPU share (at most 16: a one-GPU box's share),
    and nevals1 single-thread evaluations of the first nframes1 frames; both
    scaled to one 256-frame evaluation."""
    path = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    if not os.path.exists(path):
        return None
    orc = C.CDLL(path)
    vp = C.c_void_p
    for name, res, args in [("oracle_realize_mod", None, [vp, vp]), ("oracle_realize_spin", None, [vp, vp, vp]),
                            ("oracle_realize_delcor", None, [vp, C.c_double, C.c_int]),
                            ("oracle_calc_fits_threads", None, [vp, vp, vp, C.c_int]),
                            ("oracle_chi2", C.c_double, [vp, vp, C.c_int])]:
        fn = getattr(orc, name)
        fn.restype, fn.argtypes = res, args
    orc.oracle_last_error.restype = C.c_char_p
    share = len(os.sched_getaffinity(0))
    threads = max(1, min(16, share))

    def rate(frames, evals, nt):
        # one POS per frame (LOCAL scope): the threads render frames concurrently
        sc = sh.Scene(sh.spec_for(4, n_deldop=frames, pos_scope=sh.LOCAL))

        def forward():
            orc.oracle_realize_mod(sc.par, sc.mod)
            orc.oracle_realize_spin(sc.par, sc.mod, sc.dat)
            orc.oracle_realize_delcor(sc.dat, 0.0, 0)
            orc.oracle_calc_fits_threads(sc.par, sc.mod, sc.dat, nt)
            return orc.oracle_chi2(sc.par, sc.dat, 0)

        forward()  # warm-up, and the truth fits the observations come from
        if orc.oracle_last_error():
            raise RuntimeError("oracle: %s" % orc.oracle_last_error().decode())
        sc.synthesize_obs(snr=20.0)
        t0 = time.perf_counter()
        for _ in range(evals):
            forward()
        dt = time.perf_counter() - t0
        sc.close()
        return evals * frames / (NFRAMES_FULL * dt), dt

    v, dt = rate(nframes, nevals, threads)
    v1, dt1 = rate(nframes1, nevals1, 1)
    return {"value": round(v, 4), "unit": "chi-square evals/s", "cores": threads, "kind": "port",
            "sample": "oracle (FP64 CPU restatement of the reference's path) realize+calc_fits+chi2 of the "
                      "large-model workload: %d timed evaluations of the first %d of its %d frames after one "
                      "warm-up, frames dealt to %d threads, %.1f s wall; scaled to one %d-frame evaluation"
                      % (nevals, nframes, NFRAMES_FULL, threads, dt, NFRAMES_FULL),
            "one_thread": {"value": round(v1, 5), "unit": "chi-square evals/s", "cores": 1,
                           "sample": "%d evaluations of the first %d frames on one thread, %.1f s wall; scaled "
                                     "to one %d-frame evaluation" % (nevals1, nframes1, dt1, NFRAMES_FULL)},
            "nproc": os.cpu_count(), "cpu_share": share, "cpu_model": cpu_model()}


def copy_probe():
    """One plain 1 GiB device-to-device copy: a kernel of known traffic, read
    by the same FETCH_SIZE / WRITE_SIZE counters as the evaluation kernels."""
    import torch
    n = 1 << 30
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n}), flush=True)


# VALU issue peak of the MI355X: 256 CUs x 4 SIMDs, one wave64 VALU
# instruction per SIMD every 4 cycles (a 16-lane SIMD takes 4 passes) at
# 2.4 GHz: 614.4e9 wave-instructions/s.
VALU_PEAK = 256 * 4 * 2.4e9 / 4


def measured_valu(kernel, workload):
    """VALU wave-instructions per launch of `kernel` on `workload` (the
    SQ_INSTS_VALU counter summarized in profiles/sq_counters.json); None when
    no summary for this kernel and workload exists."""
    path = os.path.join(ROOT, "profiles", "sq_counters.json")
    try:
        rec = json.load(open(path))
    except (OSError, ValueError):
        return None
    k = rec.get("kernels", {}).get(kernel)
    if not k or rec.get("workload") != workload:
        return None
    return k.get("SQ_INSTS_VALU")


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload`, from the counter runs
    summarized in profiles/hbm_traffic.json (FETCH_SIZE + WRITE_SIZE, scaled by
    the copy probe); None when no summary for this kernel and workload exists."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        rec = json.load(open(path))
    except (OSError, ValueError):
        return None
    k = rec.get("kernels", {}).get(kernel)
    if not k or rec.get("workload") != workload:
        return None
    return k.get("bytes_per_launch")


# ---------------------------------------------------------------------------
# per-config report (--configs): every workload of the measurement contract
# ---------------------------------------------------------------------------
def load_oracle():
    path = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    if not os.path.exists(path):
        return None
    orc = C.CDLL(path)
    vp = C.c_void_p
    for name, res, args in [("oracle_realize_mod", None, [vp, vp]), ("oracle_realize_spin", None, [vp, vp, vp]),
                            ("oracle_realize_delcor", None, [vp, C.c_double, C.c_int]),
                            ("oracle_calc_fits_threads", None, [vp, vp, vp, C.c_int]),
                            ("oracle_chi2", C.c_double, [vp, vp, C.c_int])]:
        fn = getattr(orc, name)
        fn.restype, fn.argtypes = res, args
    return orc


def cpu_rate(orc, config, threads, max_frames):
    """Oracle evaluations/s on `threads` host threads: realize + calc_fits +
    chi2 of one model of `config`, on at most max_frames radar frames (scaled
    to the config's full frame count; lightcurves always whole)."""
    spec = sh.spec_for(config)
    full = spec.n_deldop + spec.n_doppler
    over = {"pos_scope": sh.LOCAL}
    if full > max_frames:
        nd = min(spec.n_deldop, max_frames)
        over.update(n_deldop=nd, n_doppler=min(spec.n_doppler, max_frames - nd))
    sc = sh.Scene(sh.spec_for(config, **over))
    frames = sc.spec.n_deldop + sc.spec.n_doppler

    def forward():
        orc.oracle_realize_mod(sc.par, sc.mod)
        orc.oracle_realize_spin(sc.par, sc.mod, sc.dat)
        orc.oracle_realize_delcor(sc.dat, 0.0, 0)
        orc.oracle_calc_fits_threads(sc.par, sc.mod, sc.dat, threads)
        return orc.oracle_chi2(sc.par, sc.dat, 0)

    forward()
    sc.synthesize_obs(snr=20.0)
    t0 = time.perf_counter()
    forward()
    dt = time.perf_counter() - t0
    sc.close()
    return frames / max(full, 1) / dt, frames


def config_scenes(config):
    if config == 5:  # 32 models, per-model seeds; 8 trial values of the first size parameter each
        return [dropin_obs(sh.Scene(sh.spec_for(5, seed_geom=5000 + m, seed_frames=5500 + m))) for m in range(32)]
    return [dropin_obs(sh.Scene(sh.spec_for(config)))]


def config_evals(scenes, per_model):
    ev = []
    for m, sc in enumerate(scenes):
        ev += [(m, t) for _, t in trial_evals(sc, per_model)]
    return ev


def facet_map_mismatch(sc, ev0):
    """Owner-facet mismatch of the mixed method against FP64 over the radar
    frames of one evaluation: (mismatched, owned, mismatched off facet edges)."""
    flags = sh.BATCH_KEEP_FIT | sh.BATCH_FACET_MAP
    Bm = sh.Batch([sc], [ev0], method=sh.METHOD_MIXED, flags=flags)
    Bf = sh.Batch([sc], [ev0], method=sh.METHOD_FP64, flags=flags)
    Bm.run()
    Bf.run()
    n = sc.spec.pos_n
    mism = owned = off = 0
    img = 0.0
    for q, (s_, f_) in enumerate(sc.all_frames()):
        if f_ < 0 or sc.set_type(s_) == sh.LGHTCRV:
            continue
        fm, ff = Bm.facet_map(0, q, n), Bf.facet_map(0, q, n)
        d = fm != ff
        edge = np.zeros(ff.shape, dtype=bool)
        edge[1:, :] |= ff[1:, :] != ff[:-1, :]
        edge[:-1, :] |= ff[:-1, :] != ff[1:, :]
        edge[:, 1:] |= ff[:, 1:] != ff[:, :-1]
        edge[:, :-1] |= ff[:, :-1] != ff[:, 1:]
        mism += int(np.count_nonzero(d))
        owned += int(np.count_nonzero(ff >= 0))
        off += int(np.count_nonzero(d & ~edge))
        shape = sc.image(s_, f_, 0).shape
        im, imf = Bm.fit(0, q, shape), Bf.fit(0, q, shape)
        img = max(img, float(np.max(np.abs(im - imf)) / max(float(np.max(np.abs(imf))), 1e-300)))
    Bm.close()
    Bf.close()
    return mism, owned, off, img


def config_report(config, a, orc):
    log("config %d: scenes" % config)
    scenes = config_scenes(config)
    per_model = 8 if config == 5 else a.evals
    ev = config_evals(scenes, per_model)
    ne = len(ev)
    steps = a.steps if config == 4 else max(a.steps, 100)
    Bm = sh.Batch(scenes, ev, method=sh.METHOD_MIXED, flags=sh.BATCH_GRAPH)
    Bm.run()
    tune = tune_batch(Bm, a)
    t_m, lat = time_batch(Bm, steps, a.warmup)
    obj_m, chi_m = Bm.run()
    fp = Bm.footprint()
    Bm.close()
    fsteps = max(3, min(20, steps // 5))
    Bf = sh.Batch(scenes, ev, method=sh.METHOD_FP64, flags=sh.BATCH_GRAPH)
    obj_f, chi_f = Bf.run()
    t_f, _ = time_batch(Bf, fsteps, 1)
    Bf.close()
    t_d, _ = time_dropin(scenes[0], 3)
    mism, owned, off, img = facet_map_mismatch(scenes[0], ev[0])
    nv = sum(sc.nv for sc in scenes) // len(scenes)
    nf = sum(sc.nf for sc in scenes) // len(scenes)
    ms = 1e3 * t_m / steps
    pb = path_bytes(fp, ne, nv, nf)
    out = {"config": config, "launch": tune, "models": len(scenes), "evals_per_replay": ne, "vertices": nv, "facets": nf,
           "frames_per_model": len(scenes[0].all_frames()),
           "mixed_evals_per_s": round(ne * steps / t_m, 2), "replay_ms_p50": round(float(np.percentile(lat, 50)), 4),
           "replay_ms_p99": round(float(np.percentile(lat, 99)), 4), "replays": steps,
           "fp64_batch_evals_per_s": round(ne * fsteps / t_f, 2), "fp64_dropin_evals_per_s": round(1.0 / t_d, 2),
           "chi2_rel_err_vs_fp64": float(np.max(np.abs(chi_m - chi_f) / np.abs(chi_f))),
           "image_rel_err_vs_fp64": img, "facet_map_mismatch": mism / max(owned, 1),
           "facet_map_mismatch_off_edges": off, "owned_pixels": owned,
           "path_roofline_frac": round(pb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if orc is not None and not a.no_cpu:
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        r1, n1 = cpu_rate(orc, config, 1, 16)
        rt, nt = cpu_rate(orc, config, threads, 256)
        out["cpu_1thread_evals_per_s"] = round(r1, 4)
        out["cpu_threads"] = threads
        out["cpu_threads_evals_per_s"] = round(rt, 4)
        out["cpu_sample_frames"] = [n1, nt]
    for sc in scenes:
        sc.close()
    return out


def line_search_report(a):
    """Wall time of shape_batch_line_search (bestfit.c:446-468's search for
    one parameter) on the large-model workload: k evaluations of the first
    scale factor per replay, from 1.02 x its value, step 0.01, tol 1e-4."""
    k = a.line_search
    sc = make_scene(a.frames)
    p0 = sc.get_param(0)
    B = sh.Batch([sc], [(0, [(0, p0)]) for _ in range(k)], method=sh.METHOD_MIXED, flags=sh.BATCH_GRAPH)
    wall, lat = time_batch(B, a.steps, a.warmup)
    out = {"metric": "batched line search wall time", "unit": "ms", "k": k, "frames": a.frames,
           "replay_ms_p50": round(float(np.median(lat)), 3), "replays_timed": a.steps}
    searches = []
    for _ in range(3):
        t0 = time.perf_counter()
        rc, xmin, fmin, rounds = B.line_search(1.02 * p0, 0.01 * p0, 1e-4, 1e-9)
        searches.append((time.perf_counter() - t0) * 1e3)
    out.update({"status": rc, "xmin": xmin, "fmin": fmin, "rounds": rounds,
                "search_ms": round(float(np.median(searches)), 2), "start": 1.02 * p0, "truth": p0})
    return out


def main():
    global KSTEPS
    a = parse()
    if a.copy_probe:
        copy_probe()
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build runs the evaluation path on a single MI355X "
              "(--gpus %d requested). Run with --gpus 1." % a.gpus, file=sys.stderr)
        sys.exit(2)
    KSTEPS = a.kernel_steps
    hip = sh.load_hip()
    if hip.shape_hip_device_count() < 1:
        raise RuntimeError("no HIP device")
    if a.line_search:
        print(json.dumps(line_search_report(a)), flush=True)
        return
    if a.configs:
        orc = load_oracle()
        for c in [int(x) for x in a.configs.split(",")]:
            print(json.dumps(config_report(c, a, orc)), flush=True)
        return
    log("building scene: %d frames" % a.frames)
    sc = make_scene(a.frames)
    evals = trial_evals(sc, a.evals)

    log("mixed batch: %d evaluations per replay" % a.evals)
    Bm = sh.Batch([sc], evals, method=sh.METHOD_MIXED, flags=sh.BATCH_GRAPH)
    obj_m, chi_m = Bm.run()
    tune = tune_batch(Bm, a)
    t_m, lat_m = time_batch(Bm, a.steps, a.warmup)
    obj_m, chi_m = Bm.run()
    cnt = Bm.counters()
    fpm = Bm.footprint()
    Bm.close()
    evals_per_s = a.evals * a.steps / t_m
    ms_step = 1e3 * t_m / a.steps
    log("mixed: %.2f evals/s, %.3f ms/step" % (evals_per_s, ms_step))
    out = {"metric": "chi-square evals/sec (batched mixed vs plain FP64) at <=1e-6 rel chi2 err",
           "value": round(evals_per_s, 3), "unit": "chi-square evals/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
           "data": "synthetic (seeded scene builder; observations = FP64 drop-in fits + noise)",
           "config": {"workload": "large model: 163842 V / 327680 F icosphere, POS 1023^2 at 2 m/px, "
                                  "%d delay-Doppler frames 128-512 x 64-256" % a.frames,
                      "evals_per_replay": a.evals, "method": "mixed", "graph": True, "launch": tune},
           "replay_ms_p50": round(float(np.percentile(lat_m, 50)), 4),
           "replay_ms_p99": round(float(np.percentile(lat_m, 99)), 4),
           "replayed_pixels": cnt[0], "claimed_keys": fpm["claimed"]}
    if a.mixed_only:
        print(json.dumps(out), flush=True)
        return

    log("FP64 batch")
    Bf = sh.Batch([sc], evals, method=sh.METHOD_FP64, flags=sh.BATCH_GRAPH)
    obj_f, chi_f = Bf.run()
    t_f, _ = time_batch(Bf, a.fp64_steps, 1)
    Bf.close()
    out["fp64_batch_evals_per_s"] = round(a.evals * a.fp64_steps / t_f, 3)
    rel = np.abs(chi_m - chi_f) / np.abs(chi_f)
    out["chi2_rel_err_vs_fp64"] = float(np.max(rel))
    out["objective_rel_err_vs_fp64"] = float(np.max(np.abs(obj_m - obj_f) / np.abs(obj_f)))

    log("FP64 drop-in")
    t_d, _ = time_dropin(sc, a.fp64_steps)
    out["fp64_dropin_evals_per_s"] = round(1.0 / t_d, 3)
    out["speedup_vs_fp64_dropin"] = round(evals_per_s * t_d, 2)
    out["speedup_vs_fp64_batch"] = round(evals_per_s / out["fp64_batch_evals_per_s"], 2)

    log("per-kernel timing (eager replays)")
    dom, rows, fp, claimed = kernel_roofline(sc, a.evals, not a.no_tune)
    r = rows[dom]
    ach = r["GBps"]
    wl = "config4/frames=%d/evals=%d" % (a.frames, a.evals)
    out["roofline"] = {"kernel": dom, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": None if ach is None else round(ach / HBM_PEAK_GBS, 4),
                       "traffic": measured_traffic(dom, wl), "avg_ms": r["avg_ms"],
                       "algorithmic_bytes": r["bytes"]}
    # The kernel's limiter is VALU issue, not HBM: the counted VALU
    # instructions per launch over the live launch duration, against the
    # issue peak.
    vi = measured_valu(dom, wl)
    if vi is not None and r["avg_ms"]:
        va = vi / (r["avg_ms"] * 1e-3)
        out["roofline_valu"] = {"kernel": dom, "achieved": round(va / 1e9, 1), "peak": round(VALU_PEAK / 1e9, 1),
                                "unit": "G wave-instr/s", "frac": round(va / VALU_PEAK, 4),
                                "valu_instr_per_launch": vi, "source": "profiles/sq_counters.json"}
        out["limiter"] = "valu"
    out["kernels"] = rows
    out["footprint"] = fp
    pb = path_bytes(fp, a.evals, sc.nv, sc.nf)
    pach = pb / (ms_step * 1e-3) / 1e9
    out["path_roofline"] = {"bytes_per_replay": pb, "bytes_per_eval": pb // a.evals, "ms_per_replay": round(ms_step, 4),
                            "achieved": round(pach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(pach / HBM_PEAK_GBS, 4)}
    log("model images: mixed against FP64 fits, every frame")
    out["image_rel_err_vs_fp64"] = image_error(sc, evals[:1], a.frames, stride=1)
    sc.close()

    if not a.no_cpu:
        log("CPU baseline (oracle: %d x %d frames on all threads, %d x %d frames on one)"
            % (a.cpu_evals, a.cpu_frames, a.cpu1_evals, a.cpu1_frames))
        out["cpu_baseline"] = cpu_baseline(a.cpu_frames, a.cpu_evals, a.cpu1_frames, a.cpu1_evals)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
