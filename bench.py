#!/usr/bin/env python3
"""Benchmark of the MI355X SExtractor detection path.

One step = one pass of the whole path (background, filtering, detection, deblending,
analysis, CLEAN, segmentation map) over one synthetic 4096x4096 float32 frame with
~167k sources (the crowded single-frame workload: 60/40 Moffat stars / Sersic galaxies,
dN/dF ~ F^-2 up to 1e6 ADU, 8% in close pairs/triples, gauss_3.0_7x7 filter,
BACK_SIZE 64, BACK_FILTERSIZE 3, DETECT_THRESH 1.5, DETECT_MINAREA 5,
DEBLEND_NTHRESH 32, DEBLEND_MINCONT 0.005, CLEAN Y with CLEAN_PARAM 1.0).

The frame is resident in HBM before the timed region (sxg_set_resident_frame); each step
runs sxg_process_resident, the C-ABI entry points in the order of the reference's
scan.c.  value = Mpix/s over the K timed steps; objects/s is reported beside it.
The roofline entry is for the kernel with the largest share of the timed region,
measured with HIP events on the library's stream.  The CPU baseline is the oracle (the
sequential C restatement) timed on one host core on the same frame; its segmentation
map and object count are also compared with the GPU's (bit-exact expected).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
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

# Algorithmic HBM bytes per launch of each kernel, in units of the frame size N (4-byte
# pixels), the number of detected pixels P_all, and raw-object pixels P0:
#   copy: read + write the frame; meshstat/meshmode: read the frame; subback: read +
#   write the frame; conv: read the subtracted frame + write the filtered frame;
#   det_init: read cd, write parents and labels; deb_levels/deb_objects/analyse: per raw
#   pixel read its index, cd, the subtracted value and label/level entries (16 B) and
#   write its label/allocation (4 B).
ALG_BYTES = {
    "copy": lambda N, P, P0: 8 * N,
    "meshstat": lambda N, P, P0: 4 * N,
    "meshmode": lambda N, P, P0: 4 * N,
    "subback": lambda N, P, P0: 8 * N,
    "conv": lambda N, P, P0: 8 * N,
    "det_init": lambda N, P, P0: 12 * N,
    "det_fg": lambda N, P, P0: 8 * N + 4 * P,
    "deb_levels": lambda N, P, P0: 20 * P0,
    "deb_objects": lambda N, P, P0: 20 * P0,
    "analyse": lambda N, P, P0: 16 * P0,
}


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


def cpu_baseline(img, p, gpu_seg, gpu_nobj):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import sxo  # the oracle: test infrastructure, used here only as the CPU baseline
    t = time.perf_counter()
    o = sxo.run(img, mask=p.mask)
    dt = time.perf_counter() - t
    H, W = img.shape
    parity = bool(o["nobj"] == gpu_nobj and np.array_equal(o["seg"], gpu_seg))
    return {"value": round(W * H / dt / 1e6, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
            "sample": "one full %dx%d frame of the same workload (%.1f s, single thread)"
                      % (W, H, dt),
            "objects_per_s": round(o["nsurv"] / dt, 1)}, parity


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--size", type=int, default=4096, help="frame side (default 4096)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (a frame or a batch "
              "of frames runs on one MI355X); run with --gpus 1", file=sys.stderr)
        return 2

    t = time.perf_counter()
    nsrc = int(round(167000 * (args.size / 4096.0) ** 2))
    img = sexgpu.synth("crowded4096", width=args.size, height=args.size, nsrc=nsrc)
    H, W = img.shape
    N = W * H
    print("# frame %dx%d, %d sources, fnv64 %s, sha256 %s, generated in %.1f s"
          % (W, H, nsrc, sexgpu.frame_hash(img),
             hashlib.sha256(img.tobytes()).hexdigest()[:16], time.perf_counter() - t),
          file=sys.stderr, flush=True)
    p = workload_params()
    L = sexgpu.gpu()
    L.sxg_set_resident_frame(img.ctypes.data, W, H)
    sexgpu.check()
    sp = sexgpu.native_params(p)
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
    t1 = time.perf_counter()
    sexgpu.check()
    dt = t1 - t0
    ktab = kernel_table(L)
    L.sxg_kernel_timing(0)
    info = {k: getattr(fi, k) for k, _ in sexgpu.FrameInfo._fields_}
    ms_step = 1000.0 * dt / args.steps
    mpix = N * args.steps / dt / 1e6
    objs = info["nsurv"] * args.steps / dt
    print("# objects: raw %d, deblended %d, catalogue %d; pixels above threshold %d, in raw "
          "objects %d" % (info["nraw"], info["nobj"], info["nsurv"], info["npix_above"],
                          info["npix_raw"]), file=sys.stderr)
    print("# kernel time per step (ms): " + ", ".join(
        "%s %.3f" % (k, v[0] / args.steps) for k, v in sorted(ktab.items(), key=lambda kv: -kv[1][0])),
        file=sys.stderr, flush=True)
    # dominant kernel with an algorithmic byte model
    dom = max((k for k in ktab if k in ALG_BYTES), key=lambda k: ktab[k][0])
    tot, nl = ktab[dom]
    avg_s = tot / nl / 1000.0
    bytes_launch = ALG_BYTES[dom](N, info["npix_above"], info["npix_raw"])
    achieved = bytes_launch / avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": dom, "kernel_ms": round(avg_s * 1e3, 4),
                "share_of_step": round(tot / (dt * 1e3), 3)}
    seg = np.zeros((H, W), np.int32)
    # the segmentation map of the last timed step
    L.sxg_get_segmap(seg.ctypes.data)
    sexgpu.check()
    cpu, parity = (None, None)
    if not args.no_cpu_baseline:
        cpu, parity = cpu_baseline(img, p, seg, nobj)
    line = {
        "metric": "Mpix/s + objects/s at 4096^2 (~167k obj), batched vs plain; bit-exact seg map",
        "value": round(mpix, 2), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "crowded %dx%d frame, %d sources, gauss_3.0_7x7, plain path "
                               "(frame resident in HBM)" % (W, H, nsrc),
                   "frame": [W, H], "objects_per_s": round(objs, 1),
                   "objects": info["nsurv"], "deblended": info["nobj"],
                   "pixels_above_threshold": info["npix_above"], "path": "plain",
                   "parity_with_oracle": parity},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
