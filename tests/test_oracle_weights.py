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


def test_backrms_equals_the_reference_code():
    """backrmsline of the oracle equals the reference's own back.c:938-1051 compiled from
    the reference tree (skipped where that build is absent)."""
    import pytest
    R = sxo.ref()
    if R is None:
        pytest.skip("reference back.c not built (oracle/_ref)")
    L = sxo.lib()
    for (h, w, bs, seed) in ((300, 257, 64, 3), (200, 640, 32, 4), (90, 80, 40, 5), (64, 600, 64, 6)):
        img = frame(seed=seed, h=h, w=w, nsrc=200)
        o = sxo.run(img, mask=None, back_size=bs)
        nbx, nby = o["nbx"], o["nby"]
        sig, dsig = np.ascontiguousarray(o["sigma"]), np.ascontiguousarray(o["dsigma"])
        ref = np.zeros((h, w), np.float32)
        R.ref_backrms(sig.ctypes.data, dsig.ctypes.data, nbx, nby, bs, w, h, ref.ctypes.data)
        mine = np.zeros((h, w), np.float32)
        for y in range(h):
            L.sxo_backrmsline(sig.ctypes.data, dsig.ctypes.data, nbx, nby, bs, w, y,
                              mine[y].ctypes.data)
        assert np.array_equal(mine, ref)


def test_var_map_equals_the_weight_map_it_comes_from(masks):
    """weight_to_var: a MAP_VAR map holding 1/w (float) gives the MAP_WEIGHT result."""
    img = frame(seed=7)
    wt = np.full(img.shape, 0.01, np.float32)
    wt[:, ::37] = 0.0
    wt[40:60, 100:140] = 0.02
    with np.errstate(divide="ignore"):
        var = np.where(wt > 0, np.float32(1.0) / wt, BIG).astype(np.float32)
    a = sxo.run(img, mask=masks["default"], weight=wt)
    b = sxo.run(img, mask=masks["default"], weight=var, weight_type=2)
    assert a["nobj"] == b["nobj"] > 10
    assert np.array_equal(a["seg"], b["seg"])
    for k in ("flux", "fluxerr", "poserr_mx2", "thresh_obj", "backsig", "dthresh"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_rms_map_squares_below_big_only(masks):
    """MAP_RMS: var = rms^2 for rms < BIG, else the value itself (weight.c:104-108); a map of
    exactly representable squares gives the MAP_VAR result of those squares."""
    img = frame(seed=8)
    rng = np.random.default_rng(1)
    rms = (np.float32(2.0) ** rng.integers(1, 4, size=img.shape)).astype(np.float32)
    rms[:, ::41] = BIG
    with np.errstate(over="ignore"):
        var = np.where(rms < BIG, rms * rms, rms).astype(np.float32)
    a = sxo.run(img, mask=masks["default"], weight=rms, weight_type=1)
    b = sxo.run(img, mask=masks["default"], weight=var, weight_type=2)
    assert np.array_equal(np.asarray(a["var"]), var)
    assert np.array_equal(a["seg"], b["seg"]) and a["nobj"] == b["nobj"]
    assert not (a["seg"][:, ::41] > 0).any()       # BIG variance: never above threshold


def test_background_weighting_is_the_var_map_of_the_rms(masks):
    """WEIGHT_TYPE BACKGROUND equals MAP_VAR with the square of the interpolated
    background rms of the same frame."""
    img = frame(seed=9, h=320, w=384, nsrc=400)
    a = sxo.run(img, mask=masks["default"], weight_type=3)
    var = np.asarray(a["var"]).copy()
    b = sxo.run(img, mask=masks["default"], weight=var, weight_type=2)
    assert a["nobj"] == b["nobj"] > 10
    assert np.array_equal(a["seg"], b["seg"])
    assert np.array_equal(a["fluxerr"], b["fluxerr"])
    # the rms follows the mesh sigma map: its mean over the frame is close to backsig
    assert abs(np.sqrt(var).mean() / a["backsig"] - 1) < 0.1


def test_weight_gain_counts_positive_pixels_only(masks):
    """WEIGHT_GAIN: the per-pixel flux variance grows by pix/gain * var/backsig^2 for
    positive pixels, instead of adding FLUX/gain once (cudaanalyse.cu:2354-2355,
    :2464-2465).  With var = backsig^2 everywhere, objects whose pixels are all positive
    get the same error either way; objects with negative pixels get a larger one."""
    img = frame(seed=10)
    base = sxo.run(img, mask=masks["default"])
    v = np.full(img.shape, np.float32(base["backsig"]) ** 2, np.float32)
    a = sxo.run(img, mask=masks["default"], weight=v, weight_type=2, gain=4.0)
    b = sxo.run(img, mask=masks["default"], weight=v, weight_type=2, gain=4.0, weight_gain=1)
    assert np.array_equal(a["seg"], b["seg"])
    sub = np.asarray(a["sub"])
    allpos = np.array([(sub[a["alloc"] == o] > 0).all() for o in range(a["nobj"])])
    assert allpos.any() and (~allpos).any()
    fa, fb = a["fluxerr"].astype(np.float64), b["fluxerr"].astype(np.float64)
    assert np.allclose(fa[allpos], fb[allpos], rtol=1e-5)
    assert (fb[~allpos] > fa[~allpos]).all()
