"""Pins of the oracle's analysis columns that GPU parity alone cannot vouch for: each is
checked against a closed form, a library routine or an identity that a plausible
misreading (a dropped term, a swapped axis, a wrong centre or weight) would break.

- ERRX2/ERRY2/ERRXY (cudaanalyse.cu:2363-2374, :2432-2453): sum(var_p (x_p - xbar)^2)
  / F^2 over the object's pixels, with var_p = backsig^2 (+ cd_p / GAIN for cd_p > 0);
- FWHM_IMAGE (cudaanalyse.cu:2521-2580): the log-linear fit of the profile; exactly
  2 sqrt(2 ln 2) sigma (less its 1 / (4 FWHM) sampling term) on a sampled Gaussian, and
  numpy's weighted polyfit (weights = pixel value) on a non-Gaussian profile;
- the bilinear BACKGROUND at the centroid (_back, cudaanalyse.cu:2128-2172): exact on a
  mesh map linear in x and y, scipy's linear interpolation on a real map;
- the MERGED re-centring of preanalyse_robust (cudaanalyse.cu:2059-2070): X/Y are the
  barycentre of the object's level pixels, X2/Y2/XY the second moments of its final
  pixels about that point (parallel-axis identity);
- the error and ratio columns of endobject (cudaanalyse.cu:62-98, :104-287, :293-590,
  :638-700, :2778-2786): FLUXERR_ISO, MAGERR_*, FLUXERR_APER/AUTO from the aperture
  areas, FLUXERR_ISOCOR by first-order propagation through the ISOCOR polynomial,
  ELONGATION and ELLIPTICITY.
"""
import math
import os

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

import sexgpu
import sxo

NNW = os.path.join(os.path.dirname(sexgpu.__file__), "config", "default.nnw")


def _abs_run(img, **kw):
    """The oracle with the frame itself as the subtracted frame (MANUAL background of 0),
    no filter, absolute thresholds."""
    base = dict(back_type=2, back_value=0.0, thresh_type=1, back_size=32, clean=0)
    base.update(kw)
    return sxo.run(img, mask=None, **base)


def _noise_frame(h, w, sigma, seed):
    return np.random.default_rng(seed).normal(0.0, sigma, (h, w)).astype(np.float32)


def _object_pixels(o, k, frame_w):
    s, e = o["objpix_start"][k], o["objpix_start"][k + 1]
    p = o["objpix"][s:e].astype(np.int64)
    return p % frame_w, p // frame_w, p


@pytest.mark.parametrize("gain", [0.0, 4.0])
def test_position_errors_closed_form(gain):
    h, w = 96, 96
    img = _noise_frame(h, w, 5.0, 21)
    # an asymmetric, tilted blob well above the absolute threshold of 60
    blob = np.array([[90, 140, 120, 0, 0],
                     [150, 310, 260, 110, 0],
                     [0, 220, 400, 230, 95],
                     [0, 0, 180, 210, 150]], np.float32)
    ys, xs = np.nonzero(blob)
    img[40 + ys, 50 + xs] = blob[ys, xs]
    o = _abs_run(img, detect_thresh=60.0, analysis_thresh=60.0, ext_minarea=3,
                 deblend_nthresh=1, gain=gain)
    assert o["nobj"] == 1
    x, y, p = _object_pixels(o, 0, w)
    assert len(p) == len(ys)
    v = img.ravel()[p].astype(np.float64)
    F = float(np.float32(o["fdflux"][0]))
    xb, yb = (v * x).sum() / v.sum(), (v * y).sum() / v.sum()
    var = np.full_like(v, float(o["backsig"]) ** 2)
    if gain > 0:
        var += np.where(v > 0, v / gain, 0.0)
    f2 = float(np.float32(F * F))
    want = [(var * (x - xb) ** 2).sum() / f2, (var * (y - yb) ** 2).sum() / f2,
            (var * (x - xb) * (y - yb)).sum() / f2]
    got = [o["poserr_mx2"][0], o["poserr_my2"][0], o["poserr_mxy"][0]]
    assert np.allclose(got, want, rtol=2e-6, atol=0), (got, want)
    assert abs(want[2]) > 1e-3 * math.sqrt(want[0] * want[1])   # the cross term is tested


def _gauss_frame(h, w, yc, xc, amp, s):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    return amp * np.exp(-((xx - xc) ** 2 + (yy - yc) ** 2) / (2 * s * s))


def test_fwhm_of_a_sampled_gaussian():
    """On exact samples of a circular Gaussian centred on a pixel the log of the profile is
    exactly linear in d^2, so the fit's slope is -1 / (2 s^2) whatever its weights:
    FWHM = 1.6651 sqrt(2) s - 1 / (4 FWHM) (1.6651 = 2 sqrt(ln 2), the reference's
    literal)."""
    for s in (1.5, 2.5, 4.0):
        img = _gauss_frame(81, 81, 40.0, 40.0, 5000.0, s).astype(np.float32)
        o = _abs_run(img, detect_thresh=20.0, analysis_thresh=20.0, ext_minarea=3,
                     deblend_nthresh=1)
        # the reference forms cd * x in float: the centroid is exact to float rounding
        assert o["nobj"] == 1 and o["mx"][0] == pytest.approx(40.0, abs=1e-4)
        f0 = 1.6651 * math.sqrt(2.0) * s
        assert o["fwhm"][0] == pytest.approx(f0 - 1.0 / (4.0 * f0), rel=2e-6), s


def test_fwhm_is_the_weighted_log_fit():
    """On a Moffat profile (not log-linear in d^2) the slope depends on the weights: it is
    numpy's weighted least-squares line of ln(pix) against d^2 with residual weights
    pix (squared weights pix^2), over the pixels above max(peak / 5, threshold)."""
    h = w = 81
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    d2 = (xx - 40.0) ** 2 + (yy - 40.0) ** 2
    img = (8000.0 * (1 + d2 / 9.0) ** -2.5).astype(np.float32)
    thr = 15.0
    o = _abs_run(img, detect_thresh=thr, analysis_thresh=thr, ext_minarea=3, deblend_nthresh=1)
    assert o["nobj"] == 1
    x, y, p = _object_pixels(o, 0, w)
    v = img.ravel()[p].astype(np.float64)
    t0 = max(float(o["peak"][0]) / 5.0, float(o["thresh_obj"][0]))
    sel = v > t0
    dd = (x[sel] - o["mx"][0]) ** 2 + (y[sel] - o["my"][0]) ** 2
    slope, _ = np.polyfit(dd, np.log(v[sel]), 1, w=v[sel])
    bb = max(-slope, 1.0 / (13.0 * float(o["a"][0]) * float(o["b"][0])))
    f0 = float(np.float32(1.6651 / math.sqrt(bb)))
    assert o["fwhm"][0] == pytest.approx(f0 - 1.0 / (4.0 * f0), rel=1e-5)
    # the Gaussian with the same core would give a different width: the fit is not trivial
    assert abs(o["fwhm"][0] - 2 * 3.0 * math.sqrt(2 ** (1 / 2.5) - 1)) > 0.05


def test_back_at_is_exact_on_a_linear_map():
    """Mesh (i, j) holds c0 + cx i + cy j; the bilinear value at pixel (x, y) with meshes
    of bw x bh is the same plane at mesh coordinates (x / bw - 0.5, y / bh - 0.5),
    extrapolated linearly beyond the outer mesh centres; a single mesh along an axis
    makes the map constant along it."""
    lib = sxo.lib()
    for nx, ny, bw, bh in ((7, 5, 32, 16), (1, 6, 64, 20), (5, 1, 10, 64), (2, 2, 8, 8)):
        c0, cx, cy = 1000.0, 3.25 if nx > 1 else 0.0, -1.5 if ny > 1 else 0.0
        j, i = np.mgrid[0:ny, 0:nx]
        m = (c0 + cx * i + cy * j).astype(np.float32)
        for x in (0, 3, bw // 2, bw, nx * bw // 2 + 1, nx * bw - 1):
            for y in (0, 5, bh // 2, ny * bh - 1, bh * (ny - 1)):
                got = lib.sxo_back_at(m.ctypes.data, x, y, nx, ny, bw, bh)
                want = c0 + cx * (x / bw - 0.5) + cy * (y / bh - 0.5)
                assert got == pytest.approx(want, rel=1e-6, abs=1e-4), (nx, ny, x, y)


def test_background_column_is_the_bilinear_map_value():
    """BACKGROUND of each object = scipy's linear interpolation (linear extrapolation at the
    edges) of the oracle's filtered mesh map, at the rounded centroid."""
    img = sexgpu.synth("plain1024", width=700, height=520, nsrc=600, seed=33)
    o = sxo.run(img, mask=None, back_size=64)
    nby, nbx = o["back"].shape
    f = RegularGridInterpolator((np.arange(nby), np.arange(nbx)), o["back"].astype(np.float64),
                                method="linear", bounds_error=False, fill_value=None)
    xi = np.floor(o["mx"] + 0.5)
    yi = np.floor(o["my"] + 0.5)
    want = f(np.stack([yi / 64 - 0.5, xi / 64 - 0.5], axis=1))
    assert o["nobj"] > 200
    assert np.allclose(o["bkg"], want, rtol=1e-6, atol=1e-3)


def test_merged_recentring_parallel_axis():
    """Deblended objects (level > 0, MERGED) keep the barycentre of their level pixels and
    take the second moments of their final pixels about it."""
    h, w = 90, 120
    img = (_gauss_frame(h, w, 44.0, 52.0, 3000.0, 3.0) +
           _gauss_frame(h, w, 47.0, 63.0, 2200.0, 2.6)).astype(np.float32)
    img += _noise_frame(h, w, 2.0, 8)
    o = _abs_run(img, detect_thresh=10.0, analysis_thresh=10.0, ext_minarea=5,
                 deblend_nthresh=32, deblend_mincont=0.005)
    merged = np.flatnonzero(o["level"] > 0)
    assert len(merged) >= 2 and np.all(o["flag"][merged] & 2)
    cd = img.ravel().astype(np.float64)
    for k in merged:
        ls, le = o["levpix_start"][k], o["levpix_start"][k + 1]
        lp = o["levpix"][ls:le].astype(np.int64)
        lv = cd[lp]
        xn, yn = (lv * (lp % w)).sum() / lv.sum(), (lv * (lp // w)).sum() / lv.sum()
        x, y, p = _object_pixels(o, k, w)
        v = cd[p]
        assert len(p) > len(lp)          # gatherup added pixels: the two centres differ
        # float products cd * x (cudaanalyse.cu:1830-1834): exact to float rounding
        assert o["mx"][k] == pytest.approx(xn, abs=2e-5)
        assert o["my"][k] == pytest.approx(yn, abs=2e-5)
        x2 = (v * (x - xn) ** 2).sum() / v.sum()
        y2 = (v * (y - yn) ** 2).sum() / v.sum()
        xy = (v * (x - xn) * (y - yn)).sum() / v.sum()
        assert o["mx2"][k] == pytest.approx(x2, rel=1e-5)
        assert o["my2"][k] == pytest.approx(y2, rel=1e-5)
        assert o["mxy"][k] == pytest.approx(xy, rel=1e-4, abs=1e-5 * math.sqrt(x2 * y2))
        # about the final pixels' own barycentre the second moments are smaller
        xf = (v * x).sum() / v.sum()
        assert abs(xf - xn) > 1e-3 and (v * (x - xf) ** 2).sum() / v.sum() < x2


def test_error_and_ratio_columns():
    """FLUXERR_ISO = sqrt(npix sigma^2 + F / GAIN); APER and AUTO errors from the summed
    area (pi r^2, pi k^2 a b) within the pixel sampling; MAGERR = 1.086 err / flux;
    FLUXERR_ISOCOR by first-order propagation of the relative errors of F and of
    ati = npix t / F through 1 - 0.196099 ati - 0.751208 ati^2; ELONGATION = A / B,
    ELLIPTICITY = 1 - B / A."""
    gain = 2.0
    srcs = [(40.0, 40.0, 3.0e4, 1.8), (40.0, 120.0, 8.0e4, 2.4), (110.0, 80.0, 1.5e4, 1.5)]
    img = np.zeros((150, 160))
    for (yc, xc, f, s) in srcs:
        img += f / (2 * math.pi * s * s) * np.exp(
            -((np.mgrid[0:150, 0:160][1] - xc) ** 2 + (np.mgrid[0:150, 0:160][0] - yc) ** 2)
            / (2 * s * s))
    img = (img + 1000.0 + np.random.default_rng(5).normal(0, 4.0, img.shape)).astype(np.float32)
    o = sxo.run(img, mask=None, back_size=32, clean=0, gain=gain, detect_thresh=5.0,
                analysis_thresh=5.0, photom={"nnw": sxo.read_nnw(NNW), "apert": 10.0, "autoparam": (2.5, 3.5)})
    ph = o["phot"]
    sig2 = float(o["backsig"]) ** 2
    assert o["nobj"] == 3
    for k in range(3):
        F = float(o["flux"][k])
        n = int(o["fdnpix"][k])
        assert ph["fluxerr_iso"][k] == pytest.approx(math.sqrt(n * sig2 + F / gain), rel=2e-6)
        assert ph["magerr_iso"][k] == pytest.approx(1.086 * ph["fluxerr_iso"][k] / F, rel=2e-6)
        fa = float(ph["flux_aper"][k])
        assert ph["fluxerr_aper"][k] == pytest.approx(
            math.sqrt(sig2 * math.pi * 25.0 + fa / gain), rel=0.01)
        assert ph["magerr_aper"][k] == pytest.approx(1.086 * ph["fluxerr_aper"][k] / fa, rel=2e-6)
        fu, kf = float(ph["flux_auto"][k]), float(ph["kronfactor"][k])
        area = math.pi * kf * kf * float(o["a"][k]) * float(o["b"][k])
        assert ph["fluxerr_auto"][k] == pytest.approx(math.sqrt(sig2 * area + fu / gain),
                                                      rel=0.03)
        assert ph["magerr_auto"][k] == pytest.approx(1.086 * ph["fluxerr_auto"][k] / fu,
                                                     rel=2e-6)
        # ISOCOR: relative error^2 = (dF/F)^2 + (dD)^2 with dD = D'(ati) dati,
        # D(a) = 0.196099 a + 0.751208 a^2, dati = ati sqrt((dF/F)^2 + 1 / npix)
        rel2 = float(o["fluxerr"][k]) / F ** 2
        ati = min(max(n * float(o["dthresh_obj"][k]) / F, 0.0), 1.0)
        dati = ati * math.sqrt(rel2 + 1.0 / n)
        h = 1e-6
        dD = ((0.196099 * (ati + h) + 0.751208 * (ati + h) ** 2) -
              (0.196099 * (ati - h) + 0.751208 * (ati - h) ** 2)) / (2 * h)
        assert ph["fluxerr_isocor"][k] == pytest.approx(
            math.sqrt(rel2 + (dD * dati) ** 2) * F, rel=1e-5)
        a, b = float(o["a"][k]), float(o["b"][k])
        assert ph["elong"][k] == np.float32(np.float32(a) / np.float32(b))
        assert ph["ellip"][k] == np.float32(1 - np.float32(b) / np.float32(a))
