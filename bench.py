#!/usr/bin/env python3
"""Benchmark of the MI355X SExtractor detection path.

One frame = one pass of the whole path (background, filtering, detection, deblending,
analysis, CLEAN, segmentation map) over a synthetic 4096x4096 float32 frame with ~167k
sources (the crowded single-frame workload: 60/40 Moffat stars / Sersic galaxies,
dN/dF ~ F^-2 up to 1e6 ADU, 8% in close pairs/triples, gauss_3.0_7x7 filter, BACK_SIZE 64,
BACK_FILTERSIZE 3, DETECT_THRESH 1.5, DETECT_MINAREA 5, DEBLEND_NTHRESH 32,
DEBLEND_MINCONT 0.005, CLEAN Y with CLEAN_PARAM 1.0).

A step is one batch of B such frames (distinct seeds), resident in HBM before the timed
region, run by the batched path in its two forms: W concurrent contexts (default 4), one
host thread and one HIP stream each (sxg_batch_process), and the same batch as one HIP
graph with a branch per context (sxg_batch_process_graph: count-planned runs, built once
before the timed region, replayed once per step).  value = Mpix/s of the faster of the two
over the K timed steps (config.batched_path names it).  Beside it: the same B frames run frame by frame (one context: the plain
drop-in sequence), and the single-frame loop on the default context, whose HIP-event
kernel timers give the roofline entry for the kernel with the largest share.  The CPU
baseline is the oracle (the sequential C restatement) timed on one host core on frame 0;
its segmentation map and object count are compared with the GPU's, and every batched
segmentation map is compared with the frame-by-frame one (bit-exact expected).

--workload wide8192 runs the wide-field single frame: 8192x8192 with ~400k sources,
large-scale background structure, a MAP_WEIGHT inverse-variance map (+-20% vignetting,
1% dead columns), mexhat_5.0_11x11, BACK_SIZE 128; frame by frame on the default
context, with the oracle's segmentation map compared at full size.

--workload hdr4096 runs the precision workload: a 4096x4096 high-dynamic-range frame
(200 saturated Moffat stars with beta 2.5 wings, 100k faint sources, nested groups of
5-20 components, SATUR_LEVEL 6e5); the GPU catalogue (the reference's arithmetic, bit-exact
with the oracle) is compared per column with a long-double evaluation of the same
quantities on the same pixels, and the maximum relative errors are reported.

The default run also times the 64-frame set below briefly (frame by frame, the worker
pool and the graph; config.batch64), and four dense frames (--workload dense4096's
generator: one frame, frame by frame and the graph; config.dense4096).

--workload batch64 runs the many-frame workload instead: 64 frames whose sizes are drawn
from {512^2, 1024^2, 2048x1024, 2048^2, 4096x2048} and whose source densities are
log-uniform in [1e-4, 1e-2] per pixel (the 1024^2 generator settings, default.conv,
seeds 2000+i); it reports aggregate Mpix/s and per-frame p50/p99 latency for the batched
path against frame-by-frame runs, for each worker count of --sweep.

--workload dense4096 runs the crowded generator with fluxes from 3000 ADU: 7.6 M pixels
above the threshold, most of them in large blends (the deblending-heavy regime).

--workload photom runs the crowded frame through the whole path and the catalogue
photometry of endobject_gpu (ISOCOR, APER, AUTO/Kron, CLASS_STAR, magnitudes) with the
reference's default.sex settings; the photometry columns are compared with the oracle's
at full size.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--workers W]
                       [--no-cpu-baseline] [--sweep 1,2,4,8]
                       [--workload crowded|batch64|wide8192|hdr4096|photom|dense4096]
"""
import argparse
import ctypes as C
import hashlib
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "gpu-sextractor_amd"))

import sexgpu  # noqa: E402

HBM_PEAK_GBS = 8000.0
DEFAULT_OPTIONS = {}

# Algorithmic HBM bytes per frame of each kernel family (its launches together), from the
# frame size N (4-byte pixels) and the frame's counts (sxg_frame_info):
#   meshstat/meshmode: read the frame; subback: read + write the frame; conv: read the
#   subtracted frame + write the filtered frame; det_init: read cd, write parents and
#   labels; det_fg: read the parents, write the foreground list; deb_levels: per pixel of
#   the multi-peak raw objects read its index and cd and write its kept node (12 B);
#   deb_objects: per raw pixel read its index, cd, the subtracted value and its label
#   (16 B) and write its allocation (4 B).
ALG_BYTES = {
    "meshstat": lambda N, i: 4 * N,
    "meshmode": lambda N, i: 4 * N,
    "subback": lambda N, i: 8 * N,
    "conv": lambda N, i: 8 * N,
    "det_init": lambda N, i: 12 * N,
    "det_fg": lambda N, i: 4 * N + 4 * i["npix_above"],
    "deb_levels": lambda N, i: 12 * i["deb_pixels"],
    "deb_objects": lambda N, i: 20 * i["npix_raw"],
}

# The kernel timers that bracket exactly one kernel launch, the kernel's name (as the
# rocprof kernel trace shows it) and its algorithmic HBM bytes per launch:
#   k_meshstat, k_meshmode: read the frame (4N); k_subback: read it, write the subtracted
#   frame (8N); k_convolve: read the subtracted frame, write the filtered one (8N); the
#   whole-frame runs' k_subconv_det (timed as conv, with no subback launch): read the
#   frame, write the subtracted and filtered frames and the detection parents (16N);
#   k_analyse: per raw pixel its index, value and filtered value (12 B), per object the
#   ~176 bytes of catalogue columns it writes; k_deb_tree_lds (one launch per size
#   class): per pixel of the class's objects its index and filtered value in, its kept
#   node out (12 B).
SINGLE_KERNELS = {
    "meshstat": ("k_meshstat", lambda N, i: 4 * N),
    "meshmode": ("k_meshmode", lambda N, i: 4 * N),
    "subback": ("k_subback<true>", lambda N, i: 8 * N),
    "conv": ("k_convolve<K>", lambda N, i: 8 * N),
    "analyse": ("k_analyse", lambda N, i: 12 * i["npix_raw"] + 176 * i["nobj"]),
    "deb_tree0": ("k_deb_tree_lds<256, 0>", lambda N, i: 12 * i["deb_class_pixels"][0]),
    "deb_tree1": ("k_deb_tree_lds<1024, 1>", lambda N, i: 12 * i["deb_class_pixels"][1]),
    "deb_tree2": ("k_deb_tree_lds<1024, 2>", lambda N, i: 12 * i["deb_class_pixels"][2]),
}


# HBM bytes per launch of single kernels, measured on the crowded 4096x4096 frame (seed
# 1001) with two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE), each scaled by what the
# same counter reads for a plain 256 MiB device-to-device copy (bench.py --calibrate-copy);
# see profiles/hbm_traffic_crowded4096_pmc.md.  Used for roofline.traffic on that workload.
# (k_deb_tree_lds<NT, C> carries its size class C in its name, so each class has its own
# figure).  "conv_fused" is the whole-frame runs' k_subconv_det<7>, the kernel the conv
# timer brackets when no subback runs.
MEASURED_TRAFFIC_CROWDED4096 = {
    "meshstat": 1.346e8, "meshmode": 6.859e7, "analyse": 6.599e7, "conv_fused": 2.724e8,
    "deb_tree0": 1.069e7, "deb_tree1": 1.12e6, "deb_tree2": 9.4e5,
}


def roofline_entry(ktab, N, info, traffic=None):
    """The roofline object of the dominant kernel: among the timers that bracket one
    kernel launch, the one with the largest time per frame; achieved = its algorithmic
    bytes per launch / its mean HIP-event duration (recorded on the stream it runs on)."""
    cands = [k for k in ktab if k in SINGLE_KERNELS]
    dom = max(cands, key=lambda k: ktab[k][0])
    tot, nl = ktab[dom]
    avg_s = tot / nl / 1000.0
    name, alg = SINGLE_KERNELS[dom]
    tkey = dom
    if dom == "conv" and "subback" not in ktab:
        # the fused subtraction, convolution and detection parents
        name, alg, tkey = "k_subconv_det<K>", (lambda N, i: 16 * N), "conv_fused"
    bytes_launch = int(alg(N, info))
    achieved = bytes_launch / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": (traffic or {}).get(tkey), "kernel": name, "timer": dom,
            "kernel_ms": round(avg_s * 1e3, 4), "algorithmic_bytes": bytes_launch}


def workload_params():
    mask = sexgpu.read_conv("gauss_3.0_7x7.conv")
    return sexgpu.Params(back_size=64, back_fsize=3, detect_thresh=1.5, analysis_thresh=1.5,
                         ext_minarea=5, deblend_nthresh=32, deblend_mincont=0.005, clean=1,
                         clean_param=1.0, clean_stacksize=3000, mask=mask)


def kernel_table(L):
    out = {}
    for i, name in enumerate(sexgpu.KERNELS):
        tot, n = C.c_double(), C.c_int()
        if L.sxg_kernel_times(i, C.byref(tot), C.byref(n)) == 0 and n.value > 0:
            out[name] = (tot.value, n.value)
    return out


def cpu_info():
    """The host the CPU baseline runs on: logical CPUs, the ones this process may use, and
    the CPU model."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model}


def oracle_all_core(frames, p, threads):
    """The oracle on every frame, `threads` frames at a time (one frame per thread, as
    independent frames shard on a multi-core host); returns (seconds, total pixels,
    total catalogue objects)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import sxo  # the oracle: test infrastructure, used here only as the CPU baseline
    from concurrent.futures import ThreadPoolExecutor
    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(lambda f: sxo.run(f, mask=p.mask)["nsurv"], frames))
    dt = time.perf_counter() - t
    return dt, sum(f.size for f in frames), sum(res)


def cpu_baseline(frames, p, gpu_seg, gpu_nobj):
    """The oracle (the sequential C restatement, -O3, no FP contraction) on the host: one
    core on frame 0 with its per-stage times (value), then every frame of the batch with
    one frame per core."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import sxo  # the oracle: test infrastructure, used here only as the CPU baseline
    img = frames[0]
    t = time.perf_counter()
    o = sxo.run(img, mask=p.mask)
    dt = time.perf_counter() - t
    H, W = img.shape
    parity = bool(o["nobj"] == gpu_nobj and np.array_equal(o["seg"], gpu_seg))
    info = cpu_info()
    threads = max(1, min(len(frames), info["usable_cpus"], 16))
    dta, npx, nobj = oracle_all_core(frames, p, threads)
    out = {"value": round(W * H / dt / 1e6, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
           "sample": "one full %dx%d frame of the same workload (%.2f s, single thread); "
                     "all-core: the %d frames of the batch on %d threads (%.2f s)"
                     % (W, H, dt, len(frames), threads, dta),
           "objects_per_s": round(o["nsurv"] / dt, 1),
           "stage_ms": {k: round(v, 1) for k, v in o["stage_ms"].items()},
           "all_core_mpix_s": round(npx / dta / 1e6, 2),
           "all_core_objects_per_s": round(nobj / dta, 1), "all_core_threads": threads}
    out.update(info)
    return out, parity


def timed_batches(L, sp, nworkers, steps, B):
    infos = (sexgpu.FrameInfo * B)()
    L.sxg_batch_process(C.byref(sp), nworkers, infos, None)   # warm-up (contexts, buffers)
    sexgpu.check()
    ms = (C.c_double * B)()
    t0 = time.perf_counter()
    for _ in range(steps):
        if L.sxg_batch_process(C.byref(sp), nworkers, infos, ms) < 0:
            sexgpu.check()
    dt = time.perf_counter() - t0
    return dt, list(infos), list(ms)


def timed_graph(L, sp, nworkers, steps, B):
    """The graph batch (sxg_batch_process_graph): the first call builds the graph
    (calibration, plans, warm-up, capture) and replays it once, untimed; then `steps`
    timed replays.  Returns (build s, timed s, infos, per-frame device ms, eager reruns)."""
    infos = (sexgpu.FrameInfo * B)()
    ms = (C.c_double * B)()
    t0 = time.perf_counter()
    if L.sxg_batch_process_graph(C.byref(sp), nworkers, infos, ms) < 0:
        sexgpu.check()
    tb = time.perf_counter() - t0
    reruns = 0
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        r = L.sxg_batch_process_graph(C.byref(sp), nworkers, infos, ms)
        if r < 0:
            sexgpu.check()
        reruns += r
        lat += list(ms)
    dt = time.perf_counter() - t0
    return tb, dt, list(infos), lat, reruns


def batch_segs(L, frames):
    out = []
    for i, f in enumerate(frames):
        s = np.zeros(f.shape, np.int32)
        if L.sxg_batch_segmap(i, s.ctypes.data) != 0:
            sexgpu.check()
        out.append(s)
    return out


def batch64(args):
    """The many-frame workload: aggregate throughput and per-frame latency."""
    t = time.perf_counter()
    frames = sexgpu.batch64_frames()
    npx = sum(f.size for f in frames)
    print("# 64 frames, %.1f Mpix in all, generated in %.1f s" % (npx / 1e6, time.perf_counter() - t),
          file=sys.stderr, flush=True)
    L = sexgpu.gpu()
    p = sexgpu.Params(mask=sexgpu.read_conv("default.conv"))
    sp = sexgpu.native_params(p)
    B = len(frames)
    ptrs = (C.c_void_p * B)(*[f.ctypes.data for f in frames])
    ws = (C.c_int * B)(*[f.shape[1] for f in frames])
    hs = (C.c_int * B)(*[f.shape[0] for f in frames])
    if L.sxg_batch_set_frames(B, ptrs, ws, hs) != 0:
        sexgpu.check()
    # small frames are bound by host round trips: twice the workers can pay off here
    sweep = ([int(v) for v in args.sweep.split(",")] if args.sweep
             else [1, args.workers, 2 * args.workers])
    res = {}
    ref = None
    for nw in sweep:
        dt, infos, ms = timed_batches(L, sp, nw, max(1, args.steps // 2), B)
        segs = batch_segs(L, frames)
        if ref is None:
            ref = segs
        same = all(np.array_equal(a, b) for a, b in zip(ref, segs))
        steps = max(1, args.steps // 2)
        lat = np.array(ms)
        res[nw] = {"mpix_s": round(npx * steps / dt / 1e6, 1),
                   "objects_per_s": round(sum(f.nsurv for f in infos) * steps / dt, 1),
                   "frame_ms_p50": round(float(np.percentile(lat, 50)), 3),
                   "frame_ms_p99": round(float(np.percentile(lat, 99)), 3),
                   "ms_per_batch": round(1000.0 * dt / steps, 2), "same_as_first": bool(same)}
        print("# %d workers: %s" % (nw, res[nw]), file=sys.stderr, flush=True)
    graph = graph_legs(L, sp, frames, ref, sweep, max(1, args.steps // 2))
    # and with the autotuner's branches and deblending size classes
    tune = sexgpu.autotune(p, max_branches=16, replays=2)
    tg = graph_legs(L, sp, frames, ref, [tune["branches"]], max(1, args.steps // 2))
    tg[tune["branches"]]["autotune"] = tune
    graph["tuned"] = tg[tune["branches"]]
    best = max(res, key=lambda k: res[k]["mpix_s"])
    gbest = max(graph, key=lambda k: graph[k]["mpix_s"])
    use_graph = graph[gbest]["mpix_s"] > res[best]["mpix_s"]
    top = graph[gbest] if use_graph else res[best]
    plain = res.get(1)
    line = {
        "metric": "Mpix/s + objects/s, 64 uneven frames, batched vs plain; bit-exact seg maps",
        "value": top["mpix_s"], "unit": "Mpix/s", "n_gpus": 1,
        "steps": max(1, args.steps // 2), "warmup": 1, "ms_per_step": top["ms_per_batch"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "64 frames of uneven sizes (512^2 .. 4096x2048), densities "
                               "log-U[1e-4,1e-2], default.conv; step = the 64-frame batch",
                   "path": "graph" if use_graph else "workers",
                   "workers": (graph[gbest].get("autotune", {}).get("branches", gbest)
                               if use_graph else best), "by_workers": res,
                   "graph_by_branches": graph, "plain": plain},
    }
    print(json.dumps(line), flush=True)
    L.sxg_batch_release()
    return 0


def batch64_summary(L, args):
    """The 64 uneven frames (the --workload batch64 set) in the default run, briefly: frame
    by frame on one context, the worker pool and the graph with --workers contexts; Mpix/s,
    per-frame p50/p99 and whether every map equals the frame-by-frame one."""
    t = time.perf_counter()
    frames = sexgpu.batch64_frames()
    npx = sum(f.size for f in frames)
    print("# batch64: %d frames, %.1f Mpix, generated in %.1f s" % (
        len(frames), npx / 1e6, time.perf_counter() - t), file=sys.stderr, flush=True)
    p = sexgpu.Params(mask=sexgpu.read_conv("default.conv"))
    sp = sexgpu.native_params(p)
    restore_options(DEFAULT_OPTIONS)
    sexgpu.set_batch_frames(frames)
    B = len(frames)
    out = {}
    steps = 2
    for name, nw in (("frame_by_frame", 1), ("workers", args.workers)):
        dt, infos, ms = timed_batches(L, sp, nw, steps, B)
        segs = batch_segs(L, frames)
        if name == "frame_by_frame":
            ref = segs
        lat = np.array(ms)
        out[name] = {"contexts": nw, "mpix_s": round(npx * steps / dt / 1e6, 1),
                     "frame_ms_p50": round(float(np.percentile(lat, 50)), 3),
                     "frame_ms_p99": round(float(np.percentile(lat, 99)), 3),
                     "same_as_frame_by_frame": all(np.array_equal(a, b) for a, b in zip(ref, segs))}
    tune = sexgpu.autotune(p, max_branches=8, replays=2)
    g = graph_legs(L, sp, frames, ref, [tune["branches"]], 3)[tune["branches"]]
    g["contexts"] = tune["branches"]
    g["autotune"] = tune
    out["graph"] = g
    out["graph_vs_frame_by_frame"] = round(g["mpix_s"] / out["frame_by_frame"]["mpix_s"], 2)
    L.sxg_batch_release()
    print("# batch64: %s" % out, file=sys.stderr, flush=True)
    return out


def default_options():
    """The option values as the library starts (the autotuner sets the size classes it
    picks; each leg of the default run starts from these)."""
    return {k: sexgpu.get_option(k) for k in sexgpu.OPTIONS}


def restore_options(saved):
    for k, v in saved.items():
        sexgpu.set_option(k, v)


def dense_summary(L, args, nframes=4):
    """The dense regime in the default run, briefly: the crowded generator with fluxes from
    3000 ADU (7.6 M pixels above the threshold, as the paper's 4096^2 image had 7.8 M),
    nframes frames (seeds 1001..): one frame on the default context, then the batch frame
    by frame and as a graph (autotuned branches), every map compared with the eager
    batch's."""
    t = time.perf_counter()
    frames = [sexgpu.synth("crowded4096", fmin=3000.0, seed=1001 + i) for i in range(nframes)]
    print("# dense: %d frames generated in %.1f s" % (nframes, time.perf_counter() - t),
          file=sys.stderr, flush=True)
    H, W = frames[0].shape
    N = W * H
    p = workload_params()
    sp = sexgpu.native_params(p)
    restore_options(DEFAULT_OPTIONS)
    L.sxg_set_resident_frame(frames[0].ctypes.data, W, H)
    fi = sexgpu.FrameInfo()
    L.sxg_process_resident(C.byref(sp), C.byref(fi))
    sexgpu.check()
    L.sxg_sync()
    t0 = time.perf_counter()
    for _ in range(2):
        L.sxg_process_resident(C.byref(sp), C.byref(fi))
    L.sxg_sync()
    single_ms = 1000.0 * (time.perf_counter() - t0) / 2
    sexgpu.check()
    sexgpu.set_batch_frames(frames)
    dtp, infp, _ = timed_batches(L, sp, 1, 1, nframes)
    ref = batch_segs(L, frames)
    tune = sexgpu.autotune(p, max_branches=nframes, replays=1)
    g = graph_legs(L, sp, frames, ref, [tune["branches"]], 2)[tune["branches"]]
    L.sxg_batch_release()
    nsurv = sum(f.nsurv for f in infp)
    out = {"single_frame_ms": round(single_ms, 3),
           "objects_per_frame": fi.nsurv, "deblended_per_frame": fi.nobj,
           "pixels_above_threshold": fi.npix_above,
           "frame_by_frame_mpix_s": round(nframes * N / dtp / 1e6, 1),
           "frame_by_frame_objects_per_s": round(nsurv / dtp, 1),
           "graph_mpix_s": g["mpix_s"], "graph_objects_per_s": g["objects_per_s"],
           "graph_branches": tune["branches"], "graph_frame_ms_p50": g["frame_ms_p50"],
           "graph_frame_ms_p99": g["frame_ms_p99"], "graph_equals_eager": g["same_as_eager"]}
    print("# dense: %s" % out, file=sys.stderr, flush=True)
    return out


def graph_legs(L, sp, frames, ref, sweep, steps):
    """The resident batch as one HIP graph, for each branch count of sweep: throughput,
    per-frame device latency, build time, eager reruns, and the segmentation maps
    compared with ref (the eager batch's)."""
    npx = sum(f.size for f in frames)
    B = len(frames)
    out = {}
    for nw in sweep:
        tb, dt, infos, lat, reruns = timed_graph(L, sp, nw, steps, B)
        segs = batch_segs(L, frames)
        same = all(np.array_equal(a, b) for a, b in zip(ref, segs))
        lat = np.array(lat)
        out[nw] = {"mpix_s": round(npx * steps / dt / 1e6, 1),
                   "objects_per_s": round(sum(f.nsurv for f in infos) * steps / dt, 1),
                   "frame_ms_p50": round(float(np.percentile(lat, 50)), 3),
                   "frame_ms_p99": round(float(np.percentile(lat, 99)), 3),
                   "ms_per_batch": round(1000.0 * dt / steps, 2),
                   "build_ms": round(1000.0 * tb, 1), "eager_reruns": int(reruns),
                   "same_as_eager": bool(same)}
        print("# graph, %d branches: %s" % (nw, out[nw]), file=sys.stderr, flush=True)
    return out


def wide8192(args):
    """The wide-field weighted frame (single context)."""
    t = time.perf_counter()
    img = sexgpu.synth("wide8192")
    wt = sexgpu.synth_weight("wide8192")
    H, W = img.shape
    N = W * H
    print("# frame %dx%d fnv64 %s, weight fnv64 %s, generated in %.1f s"
          % (W, H, sexgpu.frame_hash(img), sexgpu.frame_hash(wt), time.perf_counter() - t),
          file=sys.stderr, flush=True)
    p = sexgpu.Params(mask=sexgpu.read_conv("mexhat_5.0_11x11.conv"), back_size=128)
    sp = sexgpu.native_params(p)
    L = sexgpu.gpu()
    L.sxg_set_resident_frame(img.ctypes.data, W, H)
    sexgpu.set_weight(wt, p)
    sexgpu.check()
    fi = sexgpu.FrameInfo()
    for _ in range(args.warmup):
        L.sxg_process_resident(C.byref(sp), C.byref(fi))
        sexgpu.check()
    L.sxg_sync()
    L.sxg_kernel_timing(1)
    steps = max(1, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(steps):
        L.sxg_process_resident(C.byref(sp), C.byref(fi))
    L.sxg_sync()
    dt = time.perf_counter() - t0
    sexgpu.check()
    ktab = kernel_table(L)
    L.sxg_kernel_timing(0)
    info = {k: getattr(fi, k) for k, _ in sexgpu.FrameInfo._fields_}
    seg = np.zeros((H, W), np.int32)
    L.sxg_get_segmap(seg.ctypes.data)
    sexgpu.check()
    print("# %.3f ms per frame; objects: raw %d, deblended %d, catalogue %d; pixels above "
          "threshold %d" % (1000 * dt / steps, info["nraw"], info["nobj"], info["nsurv"],
                            info["npix_above"]), file=sys.stderr)
    print("# kernel time per frame (ms): " + ", ".join(
        "%s %.3f" % (k, v[0] / steps) for k, v in sorted(ktab.items(), key=lambda kv: -kv[1][0])),
        file=sys.stderr, flush=True)
    sexgpu.set_weight(None, p)
    cpu = parity = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import sxo  # the oracle: test infrastructure, used here only as the CPU baseline
        t = time.perf_counter()
        o = sxo.run(img, mask=p.mask, weight=wt, back_size=128)
        dtc = time.perf_counter() - t
        parity = bool(o["nobj"] == info["nobj"] and np.array_equal(o["seg"], seg))
        cpu = {"value": round(N / dtc / 1e6, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
               "sample": "one full %dx%d weighted frame (%.1f s, single thread)" % (W, H, dtc)}
    line = {
        "metric": "Mpix/s + objects/s at 8192^2 with MAP_WEIGHT (wide field); bit-exact seg map",
        "value": round(N * steps / dt / 1e6, 2), "unit": "Mpix/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(1000.0 * dt / steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "wide-field 8192x8192, ~400k sources, weight map, "
                               "mexhat_5.0_11x11, BACK_SIZE 128",
                   "objects_per_s": round(info["nsurv"] * steps / dt, 1),
                   "objects": info["nsurv"], "pixels_above_threshold": info["npix_above"],
                   "parity_with_oracle": parity},
        "roofline": roofline_entry(ktab, N, info),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


# HBM bytes per frame of the photometry kernels (label map, labels of the raw pixels,
# order, k_endobject, count), from profiles/photometry_crowded4096_trace_and_traffic.md
PHOT_TRAFFIC_CROWDED4096 = 3.533e8

PHOT_ORACLE = {"flux_isocor": "flux_isocor", "flux_aper": "flux_aper",
               "fluxerr_aper": "fluxerr_aper", "mag_aper": "mag_aper",
               "kron_radius": "kronfactor", "flux_auto": "flux_auto",
               "fluxerr_auto": "fluxerr_auto", "mag_auto": "mag_auto",
               "class_star": "sprob", "mag_iso": "mag_iso"}


def photom(args):
    """The crowded frame through the whole path plus the catalogue photometry of
    endobject_gpu (ISOCOR, APER, AUTO/Kron, CLASS_STAR, magnitudes), default.sex settings."""
    img = sexgpu.synth("crowded4096")
    H, W = img.shape
    N = W * H
    p = workload_params()
    sp = sexgpu.native_params(p)
    L = sexgpu.gpu()
    L.sxg_set_resident_frame(img.ctypes.data, W, H)
    sexgpu.check()
    br = sexgpu.read_nnw("default.nnw")
    sexgpu.set_photometry()
    fi = sexgpu.FrameInfo()

    def step():
        L.sxg_process_resident(C.byref(sp), C.byref(fi))
        L.endobject_gpu(None, C.byref(br), None, None, None, 0, H, W, H, fi.backsig, fi.thresh,
                        p.satur_level, 1.0, 0.0, p.gain, fi.nobj)

    for _ in range(args.warmup):
        step()
        sexgpu.check()
    L.sxg_sync()
    L.sxg_kernel_timing(1)
    steps = args.steps
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    L.sxg_sync()
    dt = time.perf_counter() - t0
    sexgpu.check()
    ktab = kernel_table(L)
    L.sxg_kernel_timing(0)
    info = {k: getattr(fi, k) for k, _ in sexgpu.FrameInfo._fields_}
    L.sxg_get_frame_info(C.byref(fi))
    apix = fi.phot_pixels
    ptot, pn = ktab["phot"]
    phot_s = ptot / pn / 1000.0
    print("# %.3f ms per frame with photometry; photometry kernels %.3f ms over %d objects, "
          "%d aperture pixels" % (1000 * dt / steps, phot_s * 1e3, info["nobj"], apix),
          file=sys.stderr, flush=True)
    nobj = info["nobj"]
    cols = {k: sexgpu.get_photometry(k, nobj) for k in sexgpu.PHOT_COLUMNS}
    seg = np.zeros((H, W), np.int32)
    L.sxg_get_segmap(seg.ctypes.data)
    sexgpu.check()
    cpu = parity = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import sxo  # the oracle: test infrastructure, used here only as the CPU baseline
        t = time.perf_counter()
        o = sxo.run(img, mask=p.mask, photom={"nnw": sxo.read_nnw(os.path.join(
            REPO, "gpu-sextractor_amd", "config", "default.nnw"))})
        dtc = time.perf_counter() - t
        parity = bool(o["nobj"] == nobj and np.array_equal(o["seg"], seg) and all(
            np.array_equal(cols[g], o["phot"][k]) for g, k in PHOT_ORACLE.items()))
        cpu = {"value": round(N / dtc / 1e6, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
               "sample": "one full 4096x4096 frame with photometry (%.1f s, single thread)"
                         % dtc}
    # photometry B_alg: write and read the label map once (8 B/pixel), and per aperture
    # pixel read the value and the label (8 B); measured traffic of the same kernels
    # (FETCH_SIZE + WRITE_SIZE, copy-calibrated): PHOT_TRAFFIC_CROWDED4096
    balg = 8 * N + 8 * apix
    achieved = balg / phot_s / 1e9
    line = {
        "metric": "Mpix/s + objects/s at 4096^2 (crowded), whole path + catalogue photometry "
                  "(ISOCOR/APER/AUTO/CLASS_STAR); bit-exact columns",
        "value": round(N * steps / dt / 1e6, 2), "unit": "Mpix/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(1000.0 * dt / steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "crowded 4096x4096, 167000 sources (seed 1001), gauss_3.0_7x7, "
                               "PHOT_APERTURES 5, PHOT_AUTOPARAMS 2.5,3.5, MASK_TYPE CORRECT, "
                               "default.nnw",
                   "objects_per_s": round(nobj * steps / dt, 1), "objects": nobj,
                   "photometry_ms": round(phot_s * 1e3, 4),
                   "photometry_objects_per_s": round(nobj / phot_s, 1),
                   "aperture_pixels": apix, "parity_with_oracle": parity},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": PHOT_TRAFFIC_CROWDED4096,
                     "kernel": "phot", "kernel_ms": round(phot_s * 1e3, 4),
                     "algorithmic_bytes": balg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def dense4096(args):
    """The crowded generator with the faint end raised to 3000 ADU: 7.6 M pixels above the
    threshold (the paper's 4096^2 image had 7.8 M), most of them in large blends."""
    img = sexgpu.synth("crowded4096", fmin=3000.0)
    H, W = img.shape
    N = W * H
    p = workload_params()
    sp = sexgpu.native_params(p)
    L = sexgpu.gpu()
    L.sxg_set_resident_frame(img.ctypes.data, W, H)
    sexgpu.check()
    fi = sexgpu.FrameInfo()
    for _ in range(args.warmup):
        L.sxg_process_resident(C.byref(sp), C.byref(fi))
        sexgpu.check()
    L.sxg_sync()
    L.sxg_kernel_timing(1)
    steps = max(1, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(steps):
        L.sxg_process_resident(C.byref(sp), C.byref(fi))
    L.sxg_sync()
    dt = time.perf_counter() - t0
    sexgpu.check()
    ktab = kernel_table(L)
    L.sxg_kernel_timing(0)
    info = {k: getattr(fi, k) for k, _ in sexgpu.FrameInfo._fields_}
    seg = np.zeros((H, W), np.int32)
    L.sxg_get_segmap(seg.ctypes.data)
    sexgpu.check()
    print("# %.3f ms per frame; raw %d, deblended %d, catalogue %d; pixels above threshold %d, "
          "in multi-peak raw objects %d" % (1000 * dt / steps, info["nraw"], info["nobj"],
                                             info["nsurv"], info["npix_above"],
                                             info["deb_pixels"]), file=sys.stderr)
    print("# kernel time per frame (ms): " + ", ".join(
        "%s %.3f" % (k, v[0] / steps) for k, v in sorted(ktab.items(), key=lambda kv: -kv[1][0])),
        file=sys.stderr, flush=True)
    cpu = parity = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import sxo  # the oracle: test infrastructure, used here only as the CPU baseline
        t = time.perf_counter()
        o = sxo.run(img, mask=p.mask)
        dtc = time.perf_counter() - t
        parity = bool(o["nobj"] == info["nobj"] and np.array_equal(o["seg"], seg))
        cpu = {"value": round(N / dtc / 1e6, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
               "sample": "one full 4096x4096 dense frame (%.1f s, single thread)" % dtc}
    line = {
        "metric": "Mpix/s + objects/s at 4096^2 dense (7.6M detected pixels); bit-exact seg map",
        "value": round(N * steps / dt / 1e6, 2), "unit": "Mpix/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(1000.0 * dt / steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "crowded 4096x4096, 167000 sources (seed 1001) with fluxes from "
                               "3000 ADU, gauss_3.0_7x7",
                   "objects_per_s": round(info["nsurv"] * steps / dt, 1),
                   "objects": info["nsurv"], "pixels_above_threshold": info["npix_above"],
                   "deblending_pixels": info["deb_pixels"], "parity_with_oracle": parity},
        "roofline": roofline_entry(ktab, N, info),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


PRECISION_COLUMNS = ["flux", "mx", "my", "mx2", "my2", "mxy", "a", "b", "theta", "cxx", "cyy",
                     "cxy"]


def precision_table(cols, truth):
    """Per-column max and median errors of the catalogue against the long-double truth:
    relative (|d|/|truth|) except XY and CXY (|d| over the scale sqrt(X2*Y2), resp.
    sqrt(CXX*CYY)) and THETA (absolute, degrees)."""
    out = {}
    for k, name in enumerate(PRECISION_COLUMNS):
        got = np.asarray(cols[name], np.float64)
        tr = truth[:, k]
        if name == "theta":
            err = np.abs(got - tr)
            err = np.minimum(err, 180.0 - err)
            out["theta_deg_abs"] = {"max": float(err.max()), "median": float(np.median(err))}
            continue
        scale = np.abs(tr)
        if name == "mxy":
            scale = np.sqrt(np.abs(truth[:, 3] * truth[:, 4]))
        if name == "cxy":
            scale = np.sqrt(np.abs(truth[:, 9] * truth[:, 10]))
        rel = np.abs(got - tr) / np.maximum(scale, 1e-300)
        out[name] = {"max": float(rel.max()), "median": float(np.median(rel))}
    return out


def hdr4096(args):
    """The precision workload: the catalogue in each precision mode against the truth."""
    t = time.perf_counter()
    img = sexgpu.synth("hdr4096")
    H, W = img.shape
    N = W * H
    print("# frame %dx%d fnv64 %s, generated in %.1f s" % (W, H, sexgpu.frame_hash(img),
                                                          time.perf_counter() - t),
          file=sys.stderr, flush=True)
    p = sexgpu.Params(mask=sexgpu.read_conv("gauss_3.0_7x7.conv"), satur_level=6e5)
    sp = sexgpu.native_params(p)
    L = sexgpu.gpu()
    L.sxg_set_resident_frame(img.ctypes.data, W, H)
    sexgpu.check()
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import sxo  # the oracle: test infrastructure, the checker and the FP64 reference here
    t = time.perf_counter()
    o = sxo.run(img, mask=p.mask, satur_level=6e5, truth=True)
    dtc = time.perf_counter() - t
    fi = sexgpu.FrameInfo()
    steps = max(1, args.steps // 2)
    modes, seg0, parity = {}, None, None
    for mode, label in ((0, "reference"), (1, "fp64_two_pass"), (2, "fp32_compensated")):
        sexgpu.set_option("precision", mode)
        for _ in range(args.warmup):
            L.sxg_process_resident(C.byref(sp), C.byref(fi))
            sexgpu.check()
        L.sxg_sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            nobj = L.sxg_process_resident(C.byref(sp), C.byref(fi))
        L.sxg_sync()
        dt = time.perf_counter() - t0
        sexgpu.check()
        cols = {name: sexgpu.get_column(name, nobj) for name in PRECISION_COLUMNS}
        seg = np.zeros((H, W), np.int32)
        L.sxg_get_segmap(seg.ctypes.data)
        sexgpu.check()
        if mode == 0:
            seg0 = seg
            parity = bool(o["nobj"] == nobj and np.array_equal(o["seg"], seg) and
                          all(np.array_equal(np.asarray(cols[c]), o[c])
                              for c in PRECISION_COLUMNS if c != "theta"))
        prec = precision_table(cols, o["truth"])
        modes[label] = {"ms_per_frame": round(1000.0 * dt / steps, 3),
                        "seg_identical_to_reference": bool(np.array_equal(seg, seg0)),
                        "errors": prec}
        print("# precision %s: %.3f ms/frame" % (label, 1000.0 * dt / steps), file=sys.stderr)
        for k, v in prec.items():
            print("#   %-14s max %.3e  median %.3e" % (k, v["max"], v["median"]), file=sys.stderr)
    sexgpu.set_option("precision", 0)
    ref_ms = modes["reference"]["ms_per_frame"]
    line = {
        "metric": "per-column error vs an all-FP64 oracle, 4096^2 HDR frame; bit-exact seg map",
        "value": round(N / (ref_ms / 1000.0) / 1e6, 2), "unit": "Mpix/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": ref_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "HDR 4096x4096: 200 saturated stars, 100k faint sources, nested "
                               "groups of 5-20; gauss_3.0_7x7, SATUR_LEVEL 6e5",
                   "objects": info_n(fi), "precision_modes": modes, "parity_with_oracle": parity},
        "roofline": None,
        "cpu_baseline": {"value": round(N / dtc / 1e6, 3), "unit": "Mpix/s", "cores": 1,
                         "kind": "port", "sample": "one full frame (%.1f s)" % dtc},
    }
    print(json.dumps(line), flush=True)
    return 0


def info_n(fi):
    return {"raw": fi.nraw, "deblended": fi.nobj, "catalogue": fi.nsurv}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=8, help="frames per step (default 8)")
    ap.add_argument("--workers", type=int, default=4,
                    help="concurrent contexts (default 4: the process's four hardware queues; "
                         "measured 11.9-12.8 Gpix/s against 10.1-10.3 with 8)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--size", type=int, default=4096, help="frame side (default 4096)")
    ap.add_argument("--workload", default="crowded",
                    choices=["crowded", "batch64", "wide8192", "hdr4096", "photom",
                             "dense4096"])
    ap.add_argument("--sweep", default="", help="worker counts to time, e.g. 1,2,4,8")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 256 MiB device to device (traffic counter calibration)")
    ap.add_argument("--no-end-to-end", action="store_true",
                    help="skip the PCIe-inclusive legs")
    ap.add_argument("--no-batch64", action="store_true",
                    help="skip the 64-frame legs of the default run")
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the dense-frame legs of the default run")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the graph legs (for kernel-trace runs: rocprofv3's tracer does "
                         "not follow graph launches here)")
    args = ap.parse_args()
    global DEFAULT_OPTIONS
    DEFAULT_OPTIONS = default_options()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (a frame or a batch "
              "of frames runs on one MI355X); run with --gpus 1", file=sys.stderr)
        return 2
    if args.calibrate_copy:
        calibrate_copy()
    if args.workload == "batch64":
        return batch64(args)
    if args.workload == "wide8192":
        return wide8192(args)
    if args.workload == "hdr4096":
        return hdr4096(args)
    if args.workload == "photom":
        return photom(args)
    if args.workload == "dense4096":
        return dense4096(args)

    t = time.perf_counter()
    nsrc = int(round(167000 * (args.size / 4096.0) ** 2))
    B = args.batch
    frames = [sexgpu.synth("crowded4096", width=args.size, height=args.size, nsrc=nsrc,
                           seed=1001 + i) for i in range(B)]
    img = frames[0]
    H, W = img.shape
    N = W * H
    print("# %d frames %dx%d, %d sources each, frame 0 fnv64 %s sha256 %s, generated in %.1f s"
          % (B, W, H, nsrc, sexgpu.frame_hash(img), hashlib.sha256(img.tobytes()).hexdigest()[:16],
             time.perf_counter() - t), file=sys.stderr, flush=True)
    p = workload_params()
    sp = sexgpu.native_params(p)
    L = sexgpu.gpu()

    # ---- single frame on the default context, with kernel timers
    L.sxg_set_resident_frame(img.ctypes.data, W, H)
    sexgpu.check()
    fi = sexgpu.FrameInfo()
    for _ in range(args.warmup):
        L.sxg_process_resident(C.byref(sp), C.byref(fi))
        sexgpu.check()
    L.sxg_sync()
    L.sxg_kernel_timing(1)
    L.sxg_sync()
    t0 = time.perf_counter()
    nobj = 0
    for _ in range(args.steps):
        nobj = L.sxg_process_resident(C.byref(sp), C.byref(fi))
    L.sxg_sync()
    dt1 = time.perf_counter() - t0
    sexgpu.check()
    ktab = kernel_table(L)
    L.sxg_kernel_timing(0)
    info = {k: getattr(fi, k) for k, _ in sexgpu.FrameInfo._fields_}
    seg0 = np.zeros((H, W), np.int32)
    L.sxg_get_segmap(seg0.ctypes.data)
    sexgpu.check()
    single_ms = 1000.0 * dt1 / args.steps
    print("# single frame: %.3f ms; objects: raw %d, deblended %d, catalogue %d; pixels above "
          "threshold %d, in raw objects %d" % (single_ms, info["nraw"], info["nobj"],
                                               info["nsurv"], info["npix_above"],
                                               info["npix_raw"]), file=sys.stderr)
    print("# kernel time per frame (ms): " + ", ".join(
        "%s %.3f" % (k, v[0] / args.steps) for k, v in sorted(ktab.items(), key=lambda kv: -kv[1][0])),
        file=sys.stderr, flush=True)

    # ---- the batch: frame by frame (one context), then W concurrent contexts
    ptrs = (C.c_void_p * B)(*[f.ctypes.data for f in frames])
    ws = (C.c_int * B)(*[W] * B)
    hs = (C.c_int * B)(*[H] * B)
    if L.sxg_batch_set_frames(B, ptrs, ws, hs) != 0:
        sexgpu.check()
    dtp, infp, msp = timed_batches(L, sp, 1, max(1, args.steps // 4), B)
    plain_steps = max(1, args.steps // 4)
    segs_plain = batch_segs(L, frames)
    dtb, infb, msb = timed_batches(L, sp, args.workers, args.steps, B)
    segs_batch = batch_segs(L, frames)
    batch_same = all(np.array_equal(a, b) for a, b in zip(segs_plain, segs_batch))
    batch_same = batch_same and np.array_equal(segs_plain[0], seg0)
    # the same batch as one HIP graph, its branches and deblending size classes chosen by
    # the autotuner (results checked equal across its candidates)
    tune, tgb, reruns, graph_same, mpix_graph, latg = None, 0.0, 0, None, 0.0, [0.0]
    if not args.no_graph:
        t = time.perf_counter()
        tune = sexgpu.autotune(p, max_branches=min(8, B), replays=2)
        tune["tune_s"] = round(time.perf_counter() - t, 2)
        print("# autotune: %s" % tune, file=sys.stderr, flush=True)
        tgb, dtg, infg, latg, reruns = timed_graph(L, sp, tune["branches"], args.steps, B)
        segs_graph = batch_segs(L, frames)
        graph_same = all(np.array_equal(a, b) for a, b in zip(segs_plain, segs_graph))
        mpix_graph = B * N * args.steps / dtg / 1e6
        print("# graph batch: %.1f Mpix/s, build %.0f ms, eager reruns %d, frame device time "
              "p50 %.2f ms p99 %.2f ms, maps equal %s" % (mpix_graph, 1e3 * tgb, reruns,
                                                          np.percentile(latg, 50),
                                                          np.percentile(latg, 99), graph_same),
              file=sys.stderr, flush=True)
    L.sxg_batch_release()
    nsurv = sum(f.nsurv for f in infb)
    mpix_plain = B * N * plain_steps / dtp / 1e6
    mpix_workers = B * N * args.steps / dtb / 1e6
    # the batched path's rate: the faster of the worker pool and the graph
    use_graph = mpix_graph > mpix_workers
    dtb = dtg if use_graph else dtb
    mpix_batch = B * N * args.steps / dtb / 1e6
    objs_batch = nsurv * args.steps / dtb
    objs_plain = nsurv * plain_steps / dtp
    lat = np.array(latg if use_graph else msb)
    print("# batch of %d: frame by frame %.1f Mpix/s, %d workers %.1f Mpix/s (%.2fx); frame "
          "latency p50 %.2f ms p99 %.2f ms" % (B, mpix_plain, args.workers, mpix_batch,
                                                 mpix_batch / mpix_plain, np.percentile(lat, 50),
                                                 np.percentile(lat, 99)), file=sys.stderr, flush=True)

    # ---- end to end from host memory (the options as the library starts)
    restore_options(DEFAULT_OPTIONS)
    e2e = None if args.no_end_to_end else end_to_end(L, frames, p, sp, args, seg0, segs_plain)

    # ---- the 64 uneven frames: frame by frame, worker pool and graph
    b64 = None if args.no_batch64 else batch64_summary(L, args)
    # ---- the dense regime
    dense = None if args.no_dense else dense_summary(L, args)

    # ---- roofline of the dominant kernel (single-frame run)
    traffic = MEASURED_TRAFFIC_CROWDED4096 if args.size == 4096 else None
    roofline = roofline_entry(ktab, N, info, traffic)
    roofline["share_of_frame"] = round(ktab[roofline["timer"]][0] / (dt1 * 1e3), 3)
    # whole path against the per-frame algorithmic bytes 12N + 12 P_all
    balg = 12 * N + 12 * info["npix_above"]
    cpu, parity = (None, None)
    if not args.no_cpu_baseline:
        cpu, parity = cpu_baseline(frames, p, seg0, nobj)
    line = {
        "metric": "Mpix/s + objects/s at 4096^2 (~167k obj), batched vs plain; bit-exact seg map",
        "value": round(mpix_batch, 2), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000.0 * dtb / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "crowded %dx%d frames, %d sources each (seeds 1001..%d), "
                               "gauss_3.0_7x7, resident in HBM; step = batch of %d frames on %d "
                               "concurrent contexts" % (W, H, nsrc, 1000 + B, B, args.workers),
                   "frame": [W, H], "batch": B, "workers": args.workers,
                   "batched_path": "graph" if use_graph else "workers",
                   "workers_mpix_s": round(mpix_workers, 2), "graph_mpix_s": round(mpix_graph, 2),
                   "graph_build_ms": round(1e3 * tgb, 1), "graph_eager_reruns": int(reruns),
                   "graph_equals_plain": graph_same, "autotune": tune,
                   "objects_per_s": round(objs_batch, 1),
                   "plain_mpix_s": round(mpix_plain, 2), "plain_objects_per_s": round(objs_plain, 1),
                   "single_frame_ms": round(single_ms, 3),
                   "frame_latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
                   "frame_latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
                   "objects_per_frame": info["nsurv"], "deblended_per_frame": info["nobj"],
                   "pixels_above_threshold": info["npix_above"],
                   "whole_path_hbm_frac": round(balg / (dtb / args.steps / B) / 1e12 / 8.0, 5),
                   "end_to_end": e2e, "batch64": b64, "dense4096": dense,
                   "parity_with_oracle": parity, "batched_equals_plain": bool(batch_same)},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def end_to_end(L, frames, p, sp, args, seg0, segs_ref):
    """End-to-end rates, PCIe included: the frame copied in from host memory, every kernel,
    the segmentation map and the catalogue rows copied out.
      plain: the reference's thirteen calls on frame 0 in scan.c order (init_device ...
             clear_device, which also copies the subtracted frame out), plus sxg_get_segmap
             and sxg_get_catalog; with pageable host buffers, as the reference front-end
             passes them, and with pinned ones;
      batched: sxg_batch_process_host over the batch with W workers, each copying its
             frame in and its results out on its own stream (pinned buffers)."""
    H, W = frames[0].shape
    N = W * H
    B = len(frames)
    cap = N // 64
    res = {}
    steps = max(2, args.steps // 2)
    for label, pinned in (("plain_pageable", False), ("plain_pinned", True)):
        keep = []

        def buf(shape, dtype):
            if not pinned:
                return np.zeros(shape, dtype)
            a = sexgpu.PinnedArray(shape, dtype)
            keep.append(a)
            return a.array

        img, seg, sub = buf((H, W), np.float32), buf((H, W), np.int32), buf((H, W), np.float32)
        img[...] = frames[0]
        master, rows = buf(cap, np.uint32), buf(cap, sexgpu.ISOROW_DTYPE)
        sexgpu.run_drop_in(img, p, seg, sub, master, rows)
        t0 = time.perf_counter()
        for _ in range(steps):
            nobj, nsurv = sexgpu.run_drop_in(img, p, seg, sub, master, rows)
        dt = (time.perf_counter() - t0) / steps
        res[label] = {"ms_per_frame": round(dt * 1e3, 3), "mpix_s": round(N / dt / 1e6, 1),
                      "objects_per_s": round(nsurv / dt, 1),
                      "seg_equals_resident": bool(np.array_equal(seg, seg0))}
        print("# end to end, %s: %.3f ms per frame" % (label, dt * 1e3), file=sys.stderr,
              flush=True)
    # the batched path from pinned host memory
    keep = [sexgpu.PinnedArray(f.shape, np.float32) for f in frames]
    for a, f in zip(keep, frames):
        a.array[...] = f
    psegs = [sexgpu.PinnedArray(f.shape, np.int32) for f in frames]
    prows = [sexgpu.PinnedArray(cap, sexgpu.ISOROW_DTYPE) for _ in frames]
    fin = [a.array for a in keep]
    segs = [a.array for a in psegs]
    rows = [a.array for a in prows]
    sexgpu.process_batch_host(fin, p, args.workers, segs, rows)
    ms_all = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        infos, ms, _, _ = sexgpu.process_batch_host(fin, p, args.workers, segs, rows)
        ms_all += ms
    dt = time.perf_counter() - t0
    same = all(np.array_equal(a, b) for a, b in zip(segs, segs_ref))
    nsurv = sum(fi.nsurv for fi in infos)
    res["batched_pinned"] = {"workers": args.workers,
                             "ms_per_frame": round(dt / args.steps / B * 1e3, 3),
                             "mpix_s": round(B * N * args.steps / dt / 1e6, 1),
                             "objects_per_s": round(nsurv * args.steps / dt, 1),
                             "frame_latency_ms_p50": round(float(np.percentile(ms_all, 50)), 3),
                             "frame_latency_ms_p99": round(float(np.percentile(ms_all, 99)), 3),
                             "seg_equals_resident": bool(same)}
    print("# end to end, batched (%d workers, pinned): %.1f Mpix/s" %
          (args.workers, res["batched_pinned"]["mpix_s"]), file=sys.stderr, flush=True)
    L.sxg_batch_release()
    return res


def calibrate_copy():
    """One device-to-device copy of 256 MiB (the HIP runtime's copy kernel): the dispatch
    whose FETCH_SIZE and WRITE_SIZE calibrate the traffic counters (known bytes each way)."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    n = 256 << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), n) != 0 or hip.hipMalloc(C.byref(b), n) != 0:
        raise RuntimeError("calibration buffers")
    hip.hipMemcpy(b, a, n, 3)
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print("# calibration copy: 268435456 bytes device to device", file=sys.stderr, flush=True)


if __name__ == "__main__":
    sys.exit(main())
