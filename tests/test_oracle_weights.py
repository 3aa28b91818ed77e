"""Weighted detection (MAP_WEIGHT) in the oracle.

The reference GPU path has no weight maps and its CPU weight code fixes sigfac = 1
(weight.c:80), so the semantics are this build's, written down here and in the oracle:
var = 1/w for w > 0, else BIG (weight.c:111-116); pixels of non-positive weight stay out
of the mesh statistics; relative thresholds are per pixel, DETECT_THRESH * sqrt(var) and
ANALYSIS_THRESH * sqrt(var); a raw object's deblending base and detection threshold is
the lowest threshold among its pixels; FLUXERR_ISO and the position errors use the
per-pixel variance.  These tests pin each rule with a check that does not reuse the
oracle's code.
"""
import numpy as np

import sexgpu
import sxo

BIG = np.float32(1e30)


def frame(seed=5, h=256, w=320, nsrc=300):
    return sexgpu.synth("plain1024", width=w, height=h, nsrc=nsrc, seed=seed)


def test_unit_weights_with_absolute_thresholds_change_only_the_errors(masks):
    img = frame()
    m = masks["default"]
    kw = dict(thresh_type=1, detect_thresh=30.0, analysis_thresh=30.0, back_size=64)
    a = sxo.run(img, mask=m, **kw)
    b = sxo.run(img, mask=m, weight=np.ones_like(img), **kw)
    assert a["nobj"] == b["nobj"] > 10
    assert np.array_equal(a["seg"], b["seg"])
    for k in ("mx", "my", "a", "b", "flux", "npix", "fdflux", "iso", "mthresh", "level"):
        assert np.array_equal(a[k], b[k]), k
    # FLUXERR_ISO: sum of the per-pixel variance (1.0 here) in raster order
    for o in range(b["nobj"]):
        n = int((b["alloc"] == o).sum())
        assert b["fluxerr"][o] == np.float32(float(n))


def test_variance_map_definition():
    img = frame()
    w = np.random.default_rng(1).uniform(0.0, 2.0, img.shape).astype(np.float32)
    w[:, 17] = 0.0
    w[5, :] = -1.0
    r = sxo.run(img, weight=w, back_size=64)
    want = np.where(w > 0, np.float32(1.0) / np.where(w > 0, w, np.float32(1.0)), BIG)
    assert np.array_equal(r["var"], want.astype(np.float32))


def test_zero_weight_column_is_never_detected():
    img = frame(seed=9)
    # a bright source straddling x = 160
    yy, xx = np.mgrid[0:img.shape[0], 0:img.shape[1]]
    img = img + (5000.0 * np.exp(-((yy - 128.0) ** 2 + (xx - 160.0) ** 2) / 18.0)).astype(np.float32)
    w = np.full(img.shape, 1.0 / 100.0, np.float32)     # sigma = 10 ADU
    w[:, 160] = 0.0
    r = sxo.run(img, weight=w, back_size=64)
    assert (r["seg"][:, 160] == 0).all()
    assert (r["alloc"][:, 160] == -1).all()
    # the source is cut in two by the dead column
    left, right = r["seg"][128, 157], r["seg"][128, 163]
    assert left > 0 and right > 0 and left != right


def test_detection_uses_the_local_noise():
    img = frame(seed=11)
    rng = np.random.default_rng(2)
    # smooth +-20% vignetting of the inverse variance
    yy, xx = np.mgrid[0:img.shape[0], 0:img.shape[1]]
    w = ((1.0 + 0.2 * np.sin(xx / 40.0) * np.cos(yy / 30.0)) / 100.0).astype(np.float32)
    r = sxo.run(img, mask=None, weight=w, back_size=64, detect_thresh=1.5, analysis_thresh=1.5)
    var = r["var"]
    thr = (1.5 * np.sqrt(var.astype(np.float64))).astype(np.float32)
    assert r["npix_above"] == int((r["cd"] >= thr).sum())
    # every detected pixel is at or above its own threshold
    det = r["alloc"] >= 0
    assert (r["cd"][det] >= thr[det]).all()
    # ISOAREA: pixels above the per-pixel analysis threshold
    for o in range(min(r["nobj"], 40)):
        pix = r["alloc"] == o
        assert r["npix"][o] == int((r["sub"][pix] > thr[pix]).sum())
        assert r["thresh_obj"][o] == thr[pix].min()
        assert r["dthresh_obj"][o] <= thr[pix].min()


def test_masked_pixels_leave_the_mesh_statistics():
    img = frame(seed=13)
    w = np.ones_like(img)
    w[40:90, 100:180] = 0.0
    img2 = img.copy()
    img2[40:90, 100:180] = 1e6          # garbage under the dead region
    r = sxo.run(img2, weight=w, back_size=64)
    masked = np.where(w > 0, img2, -BIG).astype(np.float32)
    bg = sxo.background(masked, bs=64, fsize=3)
    assert np.array_equal(r["back"], bg["back"]) and np.array_equal(r["sigma"], bg["sigma"])
    # the subtraction applies to the real frame: the garbage is subtracted, not -BIG
    assert np.isfinite(r["sub"]).all() and r["sub"][60, 120] > 1e5
