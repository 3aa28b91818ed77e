"""Weighted detection (MAP_WEIGHT) on the GPU against the oracle, through the C-ABI
(sxg_set_weight): a frame with smooth vignetting of the inverse variance, dead columns
and a dead block, with relative and with absolute thresholds; every column bit-exact as
in the unweighted parity tests."""
import numpy as np
import pytest

import sexgpu
import sxo
from test_gpu_pipeline import compare

pytestmark = pytest.mark.gpu


def weighted_frame(seed=21, h=768, w=1024):
    img = sexgpu.synth("plain1024", width=w, height=h, nsrc=2500, seed=seed)
    yy, xx = np.mgrid[0:h, 0:w]
    wt = ((1.0 + 0.2 * np.sin(xx / 97.0) * np.cos(yy / 71.0)) / 100.0).astype(np.float32)
    rng = np.random.default_rng(seed)
    dead = rng.choice(w, size=w // 100, replace=False)
    wt[:, dead] = 0.0
    wt[100:140, 300:420] = -1.0
    img[:, dead] = 0.0                   # dead columns read zero
    return img, wt


@pytest.mark.parametrize("kw", [dict(), dict(thresh_type=1, detect_thresh=25.0,
                                             analysis_thresh=25.0),
                                dict(back_size=32, deblend_mincont=0.001)])
def test_weighted_frame_matches_oracle(masks, kw):
    img, wt = weighted_frame()
    m = masks["default"]
    p = sexgpu.Params(mask=m, **kw)
    try:
        g = sexgpu.run_frame(img, p, weight=wt)
    finally:
        sexgpu.set_weight(None, p)
    o = sxo.run(img, mask=m, weight=wt, **kw)
    compare(g, o)
    assert o["nobj"] > 100
    if not kw.get("thresh_type"):
        # nothing is detected where the weight is not positive
        assert not (g["seg"][wt <= 0] > 0).any()


def test_weight_size_must_match(masks):
    img, wt = weighted_frame()
    p = sexgpu.Params(mask=masks["default"])
    try:
        sexgpu.set_weight(wt[:, :-1], p)
        with pytest.raises(RuntimeError):
            sexgpu.run_frame(img, p, weight=wt[:, :-1])
    finally:
        sexgpu.set_weight(None, p)


def _rms_var_maps(wt):
    """A MAP_VAR map equal to the weight map's variances, and a MAP_RMS map whose squares
    are exact (dyadic rms values), with bad pixels at BIG."""
    with np.errstate(divide="ignore"):
        var = np.where(wt > 0, (1.0 / wt.astype(np.float64)).astype(np.float32),
                       np.float32(1e30))
        rms = np.where(wt > 0, np.float32(2.0) ** np.round(np.log2(np.sqrt(var))),
                       np.float32(1e30))
    return var.astype(np.float32), rms.astype(np.float32)


@pytest.mark.parametrize("wtype", ["MAP_RMS", "MAP_VAR"])
@pytest.mark.parametrize("kw", [dict(), dict(thresh_type=1, detect_thresh=25.0,
                                             analysis_thresh=25.0)])
def test_rms_and_var_maps_match_oracle(masks, wtype, kw):
    img, wt = weighted_frame(seed=23)
    var, rms = _rms_var_maps(wt)
    m = masks["default"]
    p = sexgpu.Params(mask=m, **kw)
    wmap = rms if wtype == "MAP_RMS" else var
    try:
        g = sexgpu.run_frame(img, p, weight=wmap, weight_type=wtype)
    finally:
        sexgpu.set_weight(None, p)
    o = sxo.run(img, mask=m, weight=wmap, weight_type=sexgpu.WEIGHT_TYPES[wtype], **kw)
    compare(g, o)
    assert o["nobj"] > 100


@pytest.mark.parametrize("kw", [dict(), dict(back_size=32), dict(gain=2.0)])
def test_background_weighting_matches_oracle(masks, kw):
    """WEIGHT_TYPE BACKGROUND: the variance is the square of the interpolated background
    rms, so the thresholds follow the local noise."""
    img = sexgpu.synth("wide8192", width=1536, height=1024, nsrc=9000, seed=31)
    m = masks["default"]
    p = sexgpu.Params(mask=m, **kw)
    try:
        g = sexgpu.run_frame(img, p, weight_type="BACKGROUND")
        rms = sexgpu.back_rms(img.shape)
    finally:
        sexgpu.set_weight(None, p)
    o = sxo.run(img, mask=m, weight_type=3, **kw)
    compare(g, o)
    ov = np.asarray(o["var"])
    assert np.array_equal(rms * rms, ov)
    assert o["nobj"] > 100


@pytest.mark.parametrize("wtype", ["MAP_WEIGHT", "BACKGROUND"])
def test_weight_gain_matches_oracle(masks, wtype):
    img, wt = weighted_frame(seed=25)
    m = masks["default"]
    p = sexgpu.Params(mask=m, gain=3.0)
    w = wt if wtype == "MAP_WEIGHT" else None
    try:
        g = sexgpu.run_frame(img, p, weight=w, weight_type=wtype, weight_gain=1)
    finally:
        sexgpu.set_weight(None, p)
    o = sxo.run(img, mask=m, weight=w, weight_type=sexgpu.WEIGHT_TYPES[wtype], weight_gain=1,
                gain=3.0)
    compare(g, o)


def test_back_rms_map_matches_oracle(masks):
    img = sexgpu.synth("plain1024", width=1000, height=700, nsrc=2000, seed=41)
    p = sexgpu.Params(mask=masks["default"], back_size=48)
    g = sexgpu.run_frame(img, p, outputs=("columns",))
    sexgpu.run_background(img, p)
    rms = sexgpu.back_rms(img.shape)
    o = sxo.run(img, mask=masks["default"], back_size=48)
    L = sxo.lib()
    nbx, nby = o["nbx"], o["nby"]
    ref = np.zeros(img.shape, np.float32)
    sig, dsig = np.ascontiguousarray(o["sigma"]), np.ascontiguousarray(o["dsigma"])
    for y in range(img.shape[0]):
        L.sxo_backrmsline(sig.ctypes.data, dsig.ctypes.data, nbx, nby, 48, img.shape[1], y,
                          ref[y].ctypes.data)
    assert np.array_equal(rms, ref)
    assert g["nobj"] == o["nobj"]
