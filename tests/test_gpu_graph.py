"""The graph batch (sxg_batch_process_graph): the resident frames replayed as one HIP
graph of count-planned runs give, frame by frame, the eager batch's frame values and
segmentation maps and the oracle's; new pixels within the plans replay the same graph,
and a frame past its plan runs again eagerly with the same results, after which the
graph is planned again; the autotuner's choice of branches and deblending size classes
leaves every result as it is."""
import ctypes as C

import numpy as np
import pytest

import sexgpu
import sxo

pytestmark = pytest.mark.gpu

SIZES = [(512, 512), (384, 640), (600, 433), (512, 1024), (256, 256), (700, 500), (1024, 1024)]
DENS = [3e-3, 1e-3, 5e-3, 2e-3, 8e-3, 3e-4, 4e-3]
INFO_FIELDS = [k for k, _ in sexgpu.FrameInfo._fields_ if k != "phot_pixels"]


def _frames(seed=0):
    out = []
    for k, ((h, w), d) in enumerate(zip(SIZES, DENS)):
        out.append(sexgpu.synth("plain1024", width=w, height=h, nsrc=max(1, int(h * w * d)),
                                seed=5000 + 100 * seed + k))
    # a crowded frame: blends in every deblending size class
    out.append(sexgpu.synth("crowded4096", width=2048, height=2048, nsrc=167000 // 4,
                            seed=7000 + seed))
    return out


def _info(fi):
    return {k: (list(getattr(fi, k)) if k == "deb_class_pixels" else getattr(fi, k))
            for k in INFO_FIELDS}


def _check_oracle(frames, infos, segs, mask):
    for i, (f, fi, seg) in enumerate(zip(frames, infos, segs)):
        o = sxo.run(f, mask=mask)
        assert (fi.nraw, fi.nobj, fi.nsurv) == (o["nraw"], o["nobj"], o["nsurv"]), i
        assert np.array_equal(seg, o["seg"]), i


def test_graph_equals_eager_and_oracle(masks):
    m = masks["default"]
    p = sexgpu.Params(mask=m)
    frames = _frames()
    shapes = [f.shape for f in frames]
    # the eager batch (this also sets the resident frames)
    infos_e, _, segs_e = sexgpu.process_batch(frames, p, nworkers=3)
    b0, l0, _ = sexgpu.graph_stats()
    for rep in range(3):
        infos_g, ms, segs_g, rerun = sexgpu.process_batch_graph(len(frames), shapes, p, nworkers=3)
        print("replay %d: frames run again eagerly %d" % (rep, rerun), flush=True)
        assert rerun == 0
        assert all(t > 0 for t in ms)
        for i in range(len(frames)):
            assert _info(infos_g[i]) == _info(infos_e[i]), i
            assert np.array_equal(segs_g[i], segs_e[i]), i
    b1, l1, _ = sexgpu.graph_stats()
    assert b1 == b0 + 1 and l1 == l0 + 3      # one build, three replays
    _check_oracle(frames, infos_g, segs_g, m)
    # the eager batch after the graph, on the same contexts, still agrees
    infos_e2, _, segs_e2 = sexgpu.process_batch(frames, p, nworkers=3)
    assert all(np.array_equal(a, b) for a, b in zip(segs_e2, segs_e))


def test_graph_new_pixels_same_plan(masks):
    """Frames of the same kind (other seeds) replay the graph built for the first ones."""
    m = masks["gauss_3p0_7x7"]
    p = sexgpu.Params(mask=m)
    first = _frames(seed=1)
    shapes = [f.shape for f in first]
    sexgpu.set_batch_frames(first)
    sexgpu.process_batch_graph(len(first), shapes, p, nworkers=2, segmaps=False)
    b0, _, _ = sexgpu.graph_stats()
    L = sexgpu.gpu()
    second = _frames(seed=2)
    for i, f in enumerate(second):
        assert L.sxg_batch_update_frame(i, f.ctypes.data) == 0
    infos, ms, segs, rerun = sexgpu.process_batch_graph(len(second), shapes, p, nworkers=2)
    print("new pixels: frames run again eagerly %d" % rerun, flush=True)
    assert sexgpu.graph_stats()[0] == b0       # no rebuild before the replay
    _check_oracle(second, infos, segs, m)


def test_graph_overflow_reruns_eagerly(masks):
    """With no margin, a frame replaced by a ten times denser one overflows its plan: it
    runs again eagerly (same results as the oracle) and the next call plans again."""
    m = masks["default"]
    p = sexgpu.Params(mask=m)
    L = sexgpu.gpu()
    frames = [sexgpu.synth("plain1024", width=512, height=512, nsrc=300, seed=8000 + k)
              for k in range(3)]
    shapes = [f.shape for f in frames]
    assert L.sxg_batch_set_plan_margin(0) == 0
    try:
        sexgpu.set_batch_frames(frames)
        _, _, _, rerun = sexgpu.process_batch_graph(3, shapes, p, nworkers=2, segmaps=False)
        assert rerun == 0
        b0, _, r0 = sexgpu.graph_stats()
        dense = sexgpu.synth("plain1024", width=512, height=512, nsrc=3000, seed=8100)
        assert L.sxg_batch_update_frame(1, dense.ctypes.data) == 0
        frames[1] = dense
        infos, _, segs, rerun = sexgpu.process_batch_graph(3, shapes, p, nworkers=2)
        assert rerun >= 1 and sexgpu.graph_stats()[2] == r0 + rerun
        _check_oracle(frames, infos, segs, m)
        infos, _, segs, rerun = sexgpu.process_batch_graph(3, shapes, p, nworkers=2)
        assert rerun == 0 and sexgpu.graph_stats()[0] == b0 + 1
        _check_oracle(frames, infos, segs, m)
    finally:
        assert L.sxg_batch_set_plan_margin(25) == 0


def test_graph_rejects_bad_arguments():
    L = sexgpu.gpu()
    sp = sexgpu.native_params(sexgpu.Params())
    assert L.sxg_batch_process_graph(C.byref(sp), 0, None, None) < 0
    assert L.sxg_batch_process_graph(None, 2, None, None) < 0
    assert L.sxg_batch_set_plan_margin(-1) < 0
    assert L.sxg_batch_update_frame(10 ** 6, None) < 0


def test_autotune_choice_never_changes_results(masks):
    """The autotuner times graph branches and deblending size classes on the resident
    batch (checking that every candidate gives the same counts and maps); replaying with
    its choice gives the eager batch's results, and the same batch again hits its table."""
    m = masks["default"]
    p = sexgpu.Params(mask=m)
    L = sexgpu.gpu()
    frames = _frames(seed=3)[:4] + [sexgpu.synth("crowded4096", width=1024, height=1024,
                                                 nsrc=167000 // 16, seed=7100)]
    shapes = [f.shape for f in frames]
    keep = {k: sexgpu.get_option(k) for k in ("deb_lds_tiny", "deb_lds_small", "deb_lds_large")}
    L.sxg_autotune_forget()
    try:
        infos_e, _, segs_e = sexgpu.process_batch(frames, p, nworkers=2)
        ch = sexgpu.autotune(p, max_branches=4, replays=1)
        print("autotune:", ch, flush=True)
        assert ch["candidates"] >= 3 and ch["branches"] in (1, 2, 4)
        infos_g, _, segs_g, rerun = sexgpu.process_batch_graph(len(frames), shapes, p,
                                                               nworkers=ch["branches"])
        assert rerun == 0
        for i in range(len(frames)):
            # the pixels per size class are the work split the classes choose, not a result
            a, b = _info(infos_g[i]), _info(infos_e[i])
            assert sum(a.pop("deb_class_pixels")) == sum(b.pop("deb_class_pixels")), i
            assert a == b, i
            assert np.array_equal(segs_g[i], segs_e[i]), i
        again = sexgpu.autotune(p, max_branches=4, replays=1)
        assert again["candidates"] == 0
        assert {k: again[k] for k in ("branches", "deb_lds_tiny", "deb_lds_small",
                                      "deb_lds_large")} == \
            {k: ch[k] for k in ("branches", "deb_lds_tiny", "deb_lds_small", "deb_lds_large")}
    finally:
        for k, v in keep.items():
            sexgpu.set_option(k, v)
        L.sxg_autotune_forget()
