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
