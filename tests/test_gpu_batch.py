"""The batched path and contexts: frames of uneven sizes processed by concurrent
contexts (one host thread and stream each) give, frame by frame, the oracle's object
counts and segmentation maps; two contexts on one thread stay independent."""
import ctypes as C

import numpy as np
import pytest

import sexgpu
import sxo

pytestmark = pytest.mark.gpu


def _frames():
    sizes = [(512, 512), (384, 640), (600, 433), (512, 1024), (256, 256), (700, 500)]
    out = []
    for k, (h, w) in enumerate(sizes):
        nsrc = int(h * w * 3e-3)
        out.append(sexgpu.synth("plain1024", width=w, height=h, nsrc=nsrc, seed=3000 + k))
    return out


@pytest.mark.parametrize("nworkers", [1, 3])
def test_batch_matches_oracle(masks, nworkers):
    frames = _frames()
    m = masks["default"]
    infos, ms, segs = sexgpu.process_batch(frames, sexgpu.Params(mask=m), nworkers=nworkers)
    assert len(infos) == len(frames) and all(t > 0 for t in ms)
    for f, fi, seg in zip(frames, infos, segs):
        o = sxo.run(f, mask=m)
        assert fi.nraw == o["nraw"] and fi.nobj == o["nobj"] and fi.nsurv == o["nsurv"]
        assert np.array_equal(seg, o["seg"])


def test_batch_repeat_is_identical(masks):
    frames = _frames()[:3]
    p = sexgpu.Params(mask=masks["gauss_3p0_7x7"])
    _, _, a = sexgpu.process_batch(frames, p, nworkers=2)
    _, _, b = sexgpu.process_batch(frames, p, nworkers=3)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_two_contexts_on_one_thread(masks):
    L = sexgpu.gpu()
    f1, f2 = _frames()[0], _frames()[3]
    m = masks["default"]
    sp = sexgpu.native_params(sexgpu.Params(mask=m))
    c1, c2 = L.sxg_ctx_create(), L.sxg_ctx_create()
    try:
        L.sxg_ctx_bind(c1)
        assert L.sxg_set_resident_frame(f1.ctypes.data, f1.shape[1], f1.shape[0]) == 0
        L.sxg_ctx_bind(c2)
        assert L.sxg_set_resident_frame(f2.ctypes.data, f2.shape[1], f2.shape[0]) == 0
        fi = sexgpu.FrameInfo()
        n2 = L.sxg_process_resident(C.byref(sp), C.byref(fi))
        L.sxg_ctx_bind(c1)
        n1 = L.sxg_process_resident(C.byref(sp), C.byref(fi))
        s1 = np.zeros(f1.shape, np.int32)
        assert L.sxg_get_segmap(s1.ctypes.data) >= 0
        L.sxg_ctx_bind(c2)
        s2 = np.zeros(f2.shape, np.int32)
        assert L.sxg_get_segmap(s2.ctypes.data) >= 0
        o1, o2 = sxo.run(f1, mask=m), sxo.run(f2, mask=m)
        assert n1 == o1["nobj"] and n2 == o2["nobj"]
        assert np.array_equal(s1, o1["seg"]) and np.array_equal(s2, o2["seg"])
    finally:
        L.sxg_ctx_bind(None)
        L.sxg_ctx_destroy(c1)
        L.sxg_ctx_destroy(c2)


def test_batch_rejects_bad_worker_count():
    L = sexgpu.gpu()
    sp = sexgpu.native_params(sexgpu.Params())
    assert L.sxg_batch_process(C.byref(sp), 0, None, None) < 0
    assert L.sxg_batch_process(C.byref(sp), 99, None, None) < 0


def _oracle_rows(o):
    """The oracle's catalogue in the layout of sxg_isorow: the survivors of CLEAN in
    object order."""
    keep = o["master"] == 0xFFFFFFFF
    f4 = np.float32
    return {"number": np.arange(1, keep.sum() + 1), "flags": o["flag"][keep],
            "isoarea": o["npix"][keep], "fdnpix": o["fdnpix"][keep],
            "x": o["mx"][keep] + 1.0, "y": o["my"][keep] + 1.0, "x2": o["mx2"][keep],
            "y2": o["my2"][keep], "xy": o["mxy"][keep], "a": o["a"][keep], "b": o["b"][keep],
            "cxx": o["cxx"][keep], "cyy": o["cyy"][keep], "cxy": o["cxy"][keep],
            "flux_iso": o["flux"][keep], "fluxerr_iso": np.sqrt(o["fluxerr"][keep], dtype=f4),
            "peak": o["peak"][keep], "bkg": o["bkg"][keep], "theta": o["theta"][keep]}


def _rows_equal(rows, o):
    ref = _oracle_rows(o)
    assert len(rows) == len(ref["number"]) == o["nsurv"]
    for k, v in ref.items():
        if k == "theta":
            assert np.allclose(rows[k], v, rtol=0, atol=1e-5)
        else:
            assert np.array_equal(rows[k], v), k


def test_catalog_rows_match_oracle(frame1024, masks):
    m = masks["default"]
    g = sexgpu.run_frame(frame1024, sexgpu.Params(mask=m), outputs=("seg",))
    o = sxo.run(frame1024, mask=m)
    assert g["nsurv"] == o["nsurv"]
    # the rows are read after run_frame's clear_device: the analysis stays on the device
    # until the next frame
    rows = sexgpu.get_catalog(o["nsurv"] + 10)
    _rows_equal(rows, o)
    # a smaller buffer gets the first rows only
    part = sexgpu.get_catalog(7)
    assert len(part) == 7 and np.array_equal(part, rows[:7])


@pytest.mark.parametrize("pinned", [False, True])
def test_batch_from_host_memory(masks, pinned):
    """sxg_batch_process_host: frames copied in and results copied out by the workers
    themselves (pinned or pageable host buffers) give the oracle's maps and rows."""
    frames = _frames()
    m = masks["default"]
    keep = []
    if pinned:
        pf = []
        for f in frames:
            a = sexgpu.PinnedArray(f.shape, np.float32)
            a.array[...] = f
            keep.append(a)
            pf.append(a.array)
        frames_in = pf
        psegs = [sexgpu.PinnedArray(f.shape, np.int32) for f in frames]
        keep += psegs
        segs = [a.array for a in psegs]
    else:
        frames_in, segs = frames, None
    infos, ms, segs, rows = sexgpu.process_batch_host(frames_in, sexgpu.Params(mask=m),
                                                      nworkers=3, segs=segs)
    assert all(t > 0 for t in ms)
    for f, fi, seg, r in zip(frames, infos, segs, rows):
        o = sxo.run(f, mask=m)
        assert (fi.nraw, fi.nobj, fi.nsurv) == (o["nraw"], o["nobj"], o["nsurv"])
        assert np.array_equal(seg, o["seg"])
        _rows_equal(r, o)
