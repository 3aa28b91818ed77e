"""The whole-frame runs' fused kernel (SXG_OPT_FUSED_FILTER): background subtraction,
convolution and detection parents in one pass.  The subtracted frame, the filtered frame,
the objects and the segmentation map equal the oracle's and those of the three-kernel
path (the reference's entry-point order), for every mask size the kernel takes, with a
weight map, and for a frame whose width takes the three-kernel path anyway."""
import numpy as np
import pytest

import sexgpu
import sxo
from test_gpu_weights import weighted_frame

pytestmark = pytest.mark.gpu


def _resident(img, p):
    L = sexgpu.gpu()
    _, info = sexgpu.process_resident(img, p)
    H, W = img.shape
    cd = np.zeros((H, W), np.float32)
    seg = np.zeros((H, W), np.int32)
    sub = np.zeros((H, W), np.float32)
    assert L.sxg_get_filtered(cd.ctypes.data) == 0
    assert L.sxg_get_segmap(seg.ctypes.data) >= 0
    L.clear_device(sub.ctypes.data)
    sexgpu.check()
    info = {k: (list(v) if k == "deb_class_pixels" else v) for k, v in info.items()}
    return info, sub, cd, seg


def _both(img, p):
    """(fused, three kernels) results of the resident path on img"""
    fused = _resident(img, p)
    sexgpu.set_option("fused_filter", 0)
    try:
        plain = _resident(img, p)
    finally:
        sexgpu.set_option("fused_filter", 1)
    return fused, plain


@pytest.mark.parametrize("mask", ["default", "gauss_3p0_7x7", "mexhat_5p0_11x11", "gauss_5p0_9x9"])
def test_fused_equals_oracle_and_three_kernels(masks, mask):
    img = sexgpu.synth("plain1024", width=1024, height=640, nsrc=2500, seed=4100)
    m = masks[mask]
    (fi, sub, cd, seg), (fi2, sub2, cd2, seg2) = _both(img, sexgpu.Params(mask=m))
    o = sxo.run(img, mask=m)
    assert np.array_equal(sub, o["sub"]) and np.array_equal(cd, o["cd"])
    assert (fi["nraw"], fi["nobj"], fi["nsurv"]) == (o["nraw"], o["nobj"], o["nsurv"])
    assert np.array_equal(seg, o["seg"])
    assert np.array_equal(sub, sub2) and np.array_equal(cd, cd2) and np.array_equal(seg, seg2)
    assert fi == fi2


def test_fused_weighted(masks):
    img, wt = weighted_frame()
    m = masks["gauss_3p0_7x7"]
    p = sexgpu.Params(mask=m)
    sexgpu.set_weight(wt, p)
    try:
        (fi, sub, cd, seg), (fi2, sub2, cd2, seg2) = _both(img, p)
    finally:
        sexgpu.set_weight(None, p)
    o = sxo.run(img, mask=m, weight=wt)
    assert np.array_equal(cd, o["cd"]) and np.array_equal(seg, o["seg"])
    assert fi["nobj"] == o["nobj"]
    assert np.array_equal(seg, seg2) and np.array_equal(sub, sub2)


def test_width_not_multiple_of_64_takes_three_kernels(masks):
    img = sexgpu.synth("plain1024", width=700, height=500, nsrc=1000, seed=4200)
    m = masks["default"]
    (fi, sub, cd, seg), (fi2, sub2, cd2, seg2) = _both(img, sexgpu.Params(mask=m))
    o = sxo.run(img, mask=m)
    assert np.array_equal(seg, o["seg"]) and np.array_equal(cd, o["cd"])
    assert np.array_equal(seg, seg2) and fi == fi2
