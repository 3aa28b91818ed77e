"""CPU tests that pin the oracle before it is trusted as the GPU's checker.

- background and convolution: equal, bit for bit, to the reference's own back.c and
  filter.c (golden vectors written by tests/golden/make_golden.py);
- hmedian: the reference's known answers;
- canonical log/exp/pow: within a few ulp of libm (they only need to be deterministic
  and accurate; bit-equality with the GPU is tested on the GPU);
- labelling: equal to scipy.ndimage.label (8-connectivity) canonicalised by smallest
  raster index, with DETECT_MINAREA pruning;
- moments: closed forms on single pixels and 2x2 blocks (singularity branch, theta = 45),
  a sampled elliptical Gaussian (variances and angle);
- mthresh: the 5th order statistic; CLEAN: brute force on a two-object frame.
"""
import math
import os

import numpy as np
import pytest
import scipy.ndimage as ndi

import sxo
from conftest import GOLDEN


@pytest.fixture(scope="module")
def golden():
    d = np.load(os.path.join(GOLDEN, "back_golden.npz"), allow_pickle=False)
    return {k: d[k] for k in d.files}


def _bits(a):
    return np.ascontiguousarray(a).view(np.int32)


@pytest.mark.parametrize("key", ["a", "b", "c", "d"])
def test_background_equals_reference(golden, key):
    img = golden[key + "_img"]
    bs, fs = (int(v) for v in golden[key + "_geom"])
    r = sxo.background(img, bs=bs, fsize=fs)
    for k in ("back", "sigma", "dback", "dsigma", "sub"):
        assert np.array_equal(_bits(r[k]), _bits(golden[key + "_" + k])), k
    assert (r["backmean"], r["backsig"]) == tuple(float(v) for v in golden[key + "_medians"])


@pytest.mark.parametrize("key", ["a", "b"])
@pytest.mark.parametrize("mask", ["default", "gauss_3p0_7x7", "mexhat_5p0_11x11"])
def test_convolution_equals_reference(golden, masks, key, mask):
    sub = golden[key + "_sub"]
    out = sxo.convolve(sub, masks[mask])
    assert np.array_equal(_bits(out), _bits(golden[key + "_cd_" + mask]))


def test_convolution_brute_force_and_scipy(masks):
    rng = np.random.default_rng(1)
    img = rng.normal(0, 10, (37, 53)).astype(np.float32)
    m = masks["gauss_3p0_7x7"]
    out = sxo.convolve(img, m)
    # brute force restatement of the mask-row-major float32 accumulation
    ref = np.zeros_like(img)
    H, W = img.shape
    for y in range(H):
        for x in range(W):
            acc = np.float32(0)
            for i in range(7):
                for j in range(7):
                    yy, xx = y + i - 3, x + j - 3
                    if 0 <= yy < H and 0 <= xx < W:
                        acc = np.float32(acc + np.float32(m[i, j] * img[yy, xx]))
            ref[y, x] = acc
    assert np.array_equal(_bits(out), _bits(ref))
    sp = ndi.correlate(img.astype(np.float64), m.astype(np.float64), mode="constant", cval=0.0)
    assert np.allclose(out, sp, rtol=0, atol=1e-4)


def test_hmedian_known_answers(golden):
    vals, lens, outs = golden["hmedian_in"], golden["hmedian_len"], golden["hmedian_out"]
    pos = 0
    for n, expect in zip(lens, outs):
        v = vals[pos:pos + n].copy()
        pos += n
        got = sxo.lib().sxo_hmedian(v.ctypes.data, int(n))
        assert np.float32(got) == expect
        s = np.sort(vals[pos - n:pos])
        tb = s[n // 2] if n % 2 else np.float32((s[n // 2 - 1] + s[n // 2]) / 2.0)
        assert np.float32(got) == np.float32(tb)


def _ulps(a, b):
    a, b = float(a), float(b)
    if a == b:
        return 0
    return abs(a - b) / math.ulp(max(abs(a), abs(b)))


def test_canonical_math_against_libm():
    rng = np.random.default_rng(5)
    L = sxo.lib()
    xs = np.concatenate([10 ** rng.uniform(-300, 300, 20000), rng.uniform(0.5, 2.0, 20000),
                         rng.uniform(0.99, 1.01, 5000)])
    worst = max(_ulps(L.sxo_log(float(x)), math.log(x)) for x in xs)
    assert worst <= 2
    es = rng.uniform(-700, 700, 20000)
    assert max(_ulps(L.sxo_exp(float(x)), math.exp(x)) for x in es) <= 2
    # pow on the decision-path domain: x in [1, 1e7], y = i/N
    worst = 0
    for x in 10 ** rng.uniform(0, 7, 4000):
        for y in (1 / 32, 7 / 32, 31 / 32, 0.125, 0.875):
            worst = max(worst, _ulps(L.sxo_pow(float(x), y), x ** y))
    assert worst <= 16
    assert L.sxo_pow(3.7, 1.0) == 3.7 and L.sxo_pow(2.5, 0.0) == 1.0
    # float rounding of the thresholds agrees with libm on these inputs
    t0 = np.float32(15.3)
    for pk in (np.float32(v) for v in 10 ** rng.uniform(1.2, 6, 500)):
        for i in range(1, 32):
            a = np.float32(float(t0) * L.sxo_pow(float(pk) / float(t0), i / 32))
            b = np.float32(float(t0) * (float(pk) / float(t0)) ** (i / 32))
            assert a == b


def _flat_frame(h, w, blobs, value=0.0):
    img = np.full((h, w), value, np.float32)
    for (y, x, v) in blobs:
        img[y, x] = v
    return img


def _run_abs(img, **kw):
    """Oracle with no filter, MANUAL background of 0 and absolute thresholds."""
    base = dict(back_type=2, back_value=0.0, thresh_type=1, detect_thresh=1.0,
                analysis_thresh=1.0, back_size=min(64, *img.shape))
    base.update(kw)
    return sxo.run(img, mask=None, **base)


def test_labelling_matches_scipy():
    rng = np.random.default_rng(11)
    for trial in range(30):
        h, w = rng.integers(8, 80, 2)
        mask = rng.random((h, w)) < rng.uniform(0.1, 0.6)
        img = np.where(mask, 100.0, 0.0).astype(np.float32)
        minarea = int(rng.integers(1, 6))
        r = _run_abs(img, ext_minarea=minarea, deblend_nthresh=1, clean=0)
        lab, n = ndi.label(mask, structure=np.ones((3, 3)))
        first = []
        for k in range(1, n + 1):
            idx = np.flatnonzero(lab.ravel() == k)
            if idx.size >= minarea:
                first.append((idx.min(), idx.size))
        first.sort()
        assert r["nraw"] == len(first)
        assert np.array_equal(r["raw_first"], [f for f, _ in first])
        assert np.array_equal(r["raw_count"], [c for _, c in first])
        assert r["npix_above"] == int(mask.sum())


def test_labelling_special_cases():
    # checkerboard: 8-connectivity joins everything into one component
    cb = (np.indices((20, 30)).sum(0) % 2 == 0)
    r = _run_abs(np.where(cb, 50.0, 0.0).astype(np.float32), deblend_nthresh=1, clean=0)
    assert r["nraw"] == 1 and r["raw_count"][0] == cb.sum()
    # full frame
    r = _run_abs(np.full((16, 16), 50.0, np.float32), deblend_nthresh=1, clean=0)
    assert r["nraw"] == 1 and r["raw_first"][0] == 0
    # spiral of width 1
    sp = np.zeros((21, 21), bool)
    y0, x0, y1, x1 = 0, 0, 20, 20
    while y0 <= y1 and x0 <= x1:
        sp[y0, x0:x1 + 1] = sp[y0:y1 + 1, x1] = sp[y1, x0:x1 + 1] = True
        sp[y0 + 2:y1 + 1, x0] = True
        y0, x0, y1, x1 = y0 + 2, x0 + 2, y1 - 2, x1 - 2
    lab, n = ndi.label(sp, structure=np.ones((3, 3)))
    sizes = np.bincount(lab.ravel())[1:]
    r = _run_abs(np.where(sp, 50.0, 0.0).astype(np.float32), deblend_nthresh=1, clean=0)
    assert r["nraw"] == int((sizes >= 5).sum())


def test_moments_closed_forms():
    # a single bright pixel: variances 0 -> singular branch adds 1/12
    img = _flat_frame(9, 9, [(4, 5, 100.0)])
    r = _run_abs(img, ext_minarea=1, deblend_nthresh=1, clean=0)
    assert r["nobj"] == 1
    assert r["mx"][0] == 5.0 and r["my"][0] == 4.0
    assert r["mx2"][0] == 0.0833333 and r["my2"][0] == 0.0833333 and r["mxy"][0] == 0.0
    assert r["a"][0] == np.float32(math.sqrt(0.0833333)) and r["theta"][0] == np.float32(45.0)
    # a 2x2 block: variances 1/4, covariance 0, theta = 45 (xm2 == ym2)
    img = _flat_frame(9, 9, [(3, 3, 80.0), (3, 4, 80.0), (4, 3, 80.0), (4, 4, 80.0)])
    r = _run_abs(img, ext_minarea=1, deblend_nthresh=1, clean=0)
    assert r["mx"][0] == 3.5 and r["my"][0] == 3.5
    assert r["mx2"][0] == 0.25 and r["my2"][0] == 0.25 and r["mxy"][0] == 0.0
    assert r["a"][0] == np.float32(0.5) and r["b"][0] == np.float32(0.5)
    assert r["cxx"][0] == np.float32(4.0) and r["cxy"][0] == 0.0
    assert r["fdflux"][0] == 320.0 and r["flux"][0] == 320.0
    # a 3x1 bar: singular (ym2 = 0) -> 1/12 added to both variances
    img = _flat_frame(9, 9, [(4, 2, 10.0), (4, 3, 10.0), (4, 4, 10.0)])
    r = _run_abs(img, ext_minarea=1, deblend_nthresh=1, clean=0)
    assert r["mx2"][0] == pytest.approx(2.0 / 3.0 + 0.0833333, abs=1e-12)
    assert r["my2"][0] == 0.0833333
    assert r["theta"][0] == 0.0


def test_moments_elliptical_gaussian():
    h = w = 81
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    sx, sy, ang = 6.0, 3.0, math.radians(30.0)
    c, s = math.cos(ang), math.sin(ang)
    xr, yr = (x - 40) * c + (y - 40) * s, -(x - 40) * s + (y - 40) * c
    img = (1000.0 * np.exp(-0.5 * ((xr / sx) ** 2 + (yr / sy) ** 2))).astype(np.float32)
    r = _run_abs(img, ext_minarea=1, deblend_nthresh=1, clean=0, detect_thresh=1e-3,
                 analysis_thresh=1e-3)
    assert r["nobj"] == 1
    assert r["mx"][0] == pytest.approx(40.0, abs=1e-4)
    assert r["theta"][0] == pytest.approx(30.0, abs=0.05)
    assert r["a"][0] == pytest.approx(sx, rel=2e-3) and r["b"][0] == pytest.approx(sy, rel=2e-3)
    # FWHM of the profile's log fit: 2.3548 sigma along the mean radius, corrected for sampling
    assert 2.0 * 3.0 < r["fwhm"][0] < 2.6 * 6.0


def test_mthresh_is_fifth_largest():
    vals = [50.0, 90.0, 20.0, 70.0, 60.0, 80.0, 30.0, 40.0]
    img = _flat_frame(5, 12, [(2, 2 + i, v) for i, v in enumerate(vals)])
    r = _run_abs(img, ext_minarea=1, deblend_nthresh=1, clean=1)
    assert r["mthresh"][0] == np.float32(sorted(vals, reverse=True)[4] - 1.0)


def test_clean_two_objects_brute_force():
    # a bright object and a faint, well separated neighbour inside its CLEAN wing
    img = np.zeros((60, 80), np.float32)
    y, x = np.mgrid[0:60, 0:80]
    img += (5000.0 * np.exp(-0.5 * ((x - 30) ** 2 + (y - 30) ** 2) / 2.0 ** 2)).astype(np.float32)
    img += (30.0 * np.exp(-0.5 * ((x - 42) ** 2 + (y - 30) ** 2) / 1.0 ** 2)).astype(np.float32)
    kw = dict(ext_minarea=3, deblend_nthresh=1, detect_thresh=5.0, analysis_thresh=5.0,
              clean_param=3.0)
    r_keep = _run_abs(img, clean=0, **kw)
    assert r_keep["nobj"] == 2 and np.all(r_keep["master"] == 0xffffffff)
    o = _run_abs(img, clean=1, **kw)
    # brute-force restatement of cudaclean.cu:10-29 and :140-171 for the single pair
    i, j = 0, 1
    bright, faint = (j, i) if o["fdflux"][i] < o["fdflux"][j] else (i, j)
    f32 = np.float32
    ub = f32(math.pi * float(o["a"][bright]) * float(o["b"][bright]))
    amp = float(f32(o["fdflux"][bright] / f32(f32(2 * ub) * o["abcor"][bright])))
    pw = sxo.lib().sxo_pow(amp / float(o["dthresh"]), 1.0 / 3.0)
    alpha = (pw - 1) * float(ub) / int(o["fdnpix"][bright])
    dx = f32(o["mx"][i] - o["mx"][j])
    dy = f32(o["my"][i] - o["my"][j])
    poly = f32(f32(f32(o["cxx"][bright] * dx) * dx) + f32(f32(o["cyy"][bright] * dy) * dy))
    poly = f32(poly + f32(f32(o["cxy"][bright] * dx) * dy))
    val = 1 + alpha * float(poly)
    victim = val > 1.0 and f32(amp * (1.0 / val)) > o["mthresh"][faint]
    assert victim            # the faint neighbour sits inside the bright wing
    assert o["master"][faint] == bright and o["master"][bright] == 0xffffffff
    assert o["nsurv"] == 1
    # without the wing (CLEAN_PARAM 1) the same pair is not cleaned
    o1 = _run_abs(img, clean=1, **dict(kw, clean_param=1.0))
    assert np.all(o1["master"] == 0xffffffff)


def test_fp64_truth_closed_form():
    """The long-double evaluation of the graded columns on a hand-computed object: a
    3x4 block with known values (one raw object, no deblending)."""
    img = np.zeros((20, 24), np.float32)
    vals = np.array([[10, 20, 30, 5], [40, 80, 60, 10], [15, 25, 35, 8]], np.float32)
    img[6:9, 9:13] = vals
    r = sxo.run(img, mask=None, back_type=2, back_value=0.0, thresh_type=1, detect_thresh=1.0,
                analysis_thresh=1.0, back_size=20, clean=0, truth=True)
    assert r["nobj"] == 1
    t = r["truth"][0]
    ys, xs = np.mgrid[6:9, 9:13]
    w = vals.astype(np.float64)
    s = w.sum()
    xb, yb = (w * xs).sum() / s, (w * ys).sum() / s
    x2 = (w * (xs - xb) ** 2).sum() / s
    y2 = (w * (ys - yb) ** 2).sum() / s
    xy = (w * (xs - xb) * (ys - yb)).sum() / s
    d = x2 - y2
    tmp = math.sqrt(0.25 * d * d + xy * xy)
    det = x2 * y2 - xy * xy
    want = [s, xb, yb, x2, y2, xy, math.sqrt(0.5 * (x2 + y2) + tmp),
            math.sqrt(0.5 * (x2 + y2) - tmp), math.degrees(math.atan2(2 * xy, d) / 2),
            y2 / det, x2 / det, -2 * xy / det]
    assert np.allclose(t, want, rtol=1e-12, atol=1e-12)
    # the float catalogue is within the float/double rounding of the truth
    assert abs(r["mx"][0] - xb) < 1e-6 and abs(r["mx2"][0] - x2) / x2 < 1e-5


def test_global_sigma_rule_for_flat_meshes():
    """back.c:532-556: the global background and sigma are medians of the filtered maps
    (mean of the two middle values for an even count); a non-positive sigma median is
    replaced by the median of the positive mesh sigmas, or 1 when there is none.  Checked
    with numpy on the oracle's own maps."""
    def med(v):
        v = np.sort(v.astype(np.float32))
        n = len(v)
        return v[n // 2] if n % 2 else np.float32((np.float64(v[n // 2 - 1]) + v[n // 2]) / 2.0)
    for shape, patch in (((512, 512), (0, 192, 0, 512)), ((448, 320), (64, 256, 0, 128))):
        img = np.full(shape, 100.0, np.float32)
        y0, y1, x0, x1 = patch
        img[y0:y1, x0:x1] += np.random.default_rng(11).normal(0, 2, (y1 - y0, x1 - x0)).astype(np.float32)
        o = sxo.background(img, bs=32, fsize=3)
        sig = o["sigma"].ravel()
        assert med(sig) <= 0.0
        assert o["backsig"] == med(sig[sig > 0])
        assert o["backmean"] == med(o["back"].ravel())
    o = sxo.background(np.full((256, 256), 100.0, np.float32), bs=32, fsize=3)
    assert o["backsig"] == 1.0
