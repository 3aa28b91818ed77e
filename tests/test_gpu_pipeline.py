"""GPU parity of the whole path (background -> filter -> detection -> deblending ->
analysis -> CLEAN -> segmentation) against the oracle, through the C-ABI.

Bar: bit-exact for every integer column, the segmentation and object-index maps, the
CLEAN masters, and the float/double columns computed with IEEE basic operations.  THETA
uses atan2 (libm on the host, the device library on the GPU) and is compared to 1e-5
degrees.
"""
import numpy as np
import pytest

import sexgpu
import sxo

pytestmark = pytest.mark.gpu

EXACT_INT = ["level", "root", "firstpix", "flag", "fdnpix", "dnpix", "npix", "xmin", "xmax",
             "ymin", "ymax", "peakx", "peaky"]
EXACT_FLOAT = ["mx", "my", "mx2", "my2", "mxy", "poserr_mx2", "poserr_my2", "poserr_mxy", "a",
               "b", "cxx", "cyy", "cxy", "abcor", "fdflux", "dflux", "flux", "fluxerr",
               "fdpeak", "dpeak", "peak", "bkg", "thresh_obj", "mthresh", "fwhm"]


def _params(mask, **kw):
    return sexgpu.Params(mask=mask, **kw)


def _oracle(img, mask, **kw):
    return sxo.run(img, mask=mask, **kw)


def compare(g, o):
    for k in ("backmean", "backsig", "dthresh", "thresh"):
        assert np.float32(g[k]) == np.float32(o[k]), k
    assert g["npix_above"] == o["npix_above"]
    assert g["nraw"] == o["nraw"]
    assert g["npix_raw"] == o["npix_raw"]
    assert g["nobj"] == o["nobj"]
    for k in EXACT_INT:
        assert np.array_equal(g[k], o[k]), k
    assert np.array_equal(g["iso"], o["iso"])
    for k in EXACT_FLOAT:
        a, b = np.asarray(g[k]), np.asarray(o[k])
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        assert same.all(), (k, np.flatnonzero(~same)[:5], a[~same][:5], b[~same][:5])
    assert np.allclose(g["theta"], o["theta"], rtol=0, atol=1e-5)
    assert np.array_equal(g["master"], o["master"])
    assert np.array_equal(g["alloc"], o["alloc"])
    assert g["nsurv"] == o["nsurv"]
    assert np.array_equal(g["seg"], o["seg"])


def test_plain1024_default_filter(frame1024, masks):
    m = masks["default"]
    g = sexgpu.run_frame(frame1024, _params(m))
    o = _oracle(frame1024, m)
    compare(g, o)
    assert o["nobj"] > 500


@pytest.mark.parametrize("mask", ["gauss_3p0_7x7", "mexhat_5p0_11x11"])
def test_plain1024_other_filters(frame1024, masks, mask):
    m = masks[mask]
    compare(sexgpu.run_frame(frame1024, _params(m)), _oracle(frame1024, m))


def test_crowded_2048_crop(masks):
    img = sexgpu.synth("crowded4096", width=2048, height=2048, nsrc=167000 // 4, seed=1101)
    m = masks["gauss_3p0_7x7"]
    g = sexgpu.run_frame(img, _params(m))
    o = _oracle(img, m)
    compare(g, o)
    assert (o["level"] > 0).sum() > 100    # deblending exercised


@pytest.mark.parametrize("kw", [dict(deblend_nthresh=1), dict(deblend_nthresh=64),
                                dict(clean=0), dict(ext_minarea=1), dict(ext_minarea=12),
                                dict(deblend_mincont=0.0), dict(deblend_mincont=1.0),
                                dict(clean_param=2.5, clean_stacksize=100),
                                dict(back_size=32, back_fsize=5), dict(gain=2.0)])
def test_parameter_variants(frame1024, masks, kw):
    m = masks["default"]
    compare(sexgpu.run_frame(frame1024, _params(m, **kw)), _oracle(frame1024, m, **kw))


def test_no_filter(frame1024):
    compare(sexgpu.run_frame(frame1024, _params(None)), _oracle(frame1024, None))


def test_empty_frame():
    img = np.full((300, 200), 100.0, np.float32)
    img += np.random.default_rng(3).normal(0, 1e-3, img.shape).astype(np.float32)
    p = _params(None, thresh_type=1, detect_thresh=1e6, analysis_thresh=1e6)
    g = sexgpu.run_frame(img, p)
    o = _oracle(img, None, thresh_type=1, detect_thresh=1e6, analysis_thresh=1e6)
    assert g["nraw"] == o["nraw"] == 0
    assert g["nobj"] == 0
    assert not g["seg"].any()


def test_catalog_rows_match_columns(frame1024, masks):
    g = sexgpu.run_frame(frame1024, _params(masks["default"]))
    rows = g["catalog"]
    n = g["nobj"]
    sposx = np.array([rows[i].sposx for i in range(n)], np.float32)
    assert np.array_equal(sposx, (g["mx"] + 1.0).astype(np.float32))
    fe = np.array([rows[i].fluxerr_iso for i in range(n)], np.float32)
    assert np.array_equal(fe, np.sqrt(g["fluxerr"].astype(np.float64)).astype(np.float32))


KNOBS = ("deb_lds_tiny", "deb_lds_small", "deb_lds_large", "deb_table", "deb_glb_levels",
         "deb_order_full")


@pytest.mark.parametrize("caps", [(0, 0, 0, 0, 1, 4096), (0, 0, 0, 0, 0, 4096),
                                  (0, 64, 256, 0, 1, 4096), (0, 6400, 6400, 0, 1, 4096),
                                  (256, 1024, 6144, 0, 1, 4096), (6400, 6400, 6400, 0, 1, 4096),
                                  (256, 1024, 6144, 1, 1, 4096), (256, 1024, 6144, 5, 1, 4096),
                                  (0, 0, 0, 3, 1, 4096), (0, 0, 0, 3, 0, 4096),
                                  (0, 64, 256, 0, 2000, 4096), (0, 64, 256, 0, 2000, 0),
                                  (0, 0, 0, 0, 1, 0), (0, 0, 0, 0, 0, 0)])
def test_deblend_kernel_variants(masks, caps):
    """Every deblending kernel variant (LDS 256- and 1024-lane workgroups, global
    scratch with the level components built level by level or, from a size on, all
    levels in parallel, launched largest first or only the level-parallel ones ordered;
    the tree's node table in LDS or moved to global memory when it outgrows its LDS
    share) gives the oracle's result: the options move raw objects between them."""
    img = sexgpu.synth("crowded4096", width=1024, height=1024, nsrc=167000 // 16, seed=1201)
    m = masks["gauss_3p0_7x7"]
    old = [sexgpu.get_option(k) for k in KNOBS]
    try:
        for k, v in zip(KNOBS, caps):
            sexgpu.set_option(k, v)
        g = sexgpu.run_frame(img, _params(m))
    finally:
        for k, v in zip(KNOBS, old):
            sexgpu.set_option(k, v)
    compare(g, _oracle(img, m))
    assert (g["level"] > 0).sum() > 20


@pytest.mark.parametrize("kw", [dict(ext_minarea=1), dict(deblend_nthresh=64),
                                dict(deblend_mincont=0.0)])
def test_level_parallel_parameter_variants(frame1024, masks, kw):
    """Every multi-peak object built level-parallel (one workgroup per object and level)
    under the parameters that stress that build: one-pixel nodes (as many node records
    as pixels per level), 64 levels, every node significant."""
    m = masks["default"]
    old = [sexgpu.get_option(k) for k in KNOBS]
    try:
        for k, v in zip(KNOBS, (0, 0, 0, 0, 1, 4096)):
            sexgpu.set_option(k, v)
        g = sexgpu.run_frame(frame1024, _params(m, **kw))
    finally:
        for k, v in zip(KNOBS, old):
            sexgpu.set_option(k, v)
    compare(g, _oracle(frame1024, m, **kw))


def test_dense_blends_default_paths(masks):
    """A 2048^2 frame of the dense regime (fluxes from 3000 ADU; blends of 10k-20k pixels)
    gives the oracle's result with the default options, where the largest trees are built
    one workgroup per level and the other global-memory trees one workgroup each, and
    with every global-memory tree built one workgroup each."""
    img = sexgpu.synth("crowded4096", width=2048, height=2048, nsrc=167000 // 4, seed=1301,
                       fmin=3000.0)
    m = masks["gauss_3p0_7x7"]
    ref = _oracle(img, m)
    lvmin = sexgpu.get_option("deb_glb_levels")
    sizes = np.asarray(ref["raw_count"])
    assert (sizes >= lvmin).sum() >= 4 and ((sizes > 6144) & (sizes < lvmin)).sum() >= 10
    try:
        for v in (lvmin, 0):
            sexgpu.set_option("deb_glb_levels", v)
            compare(sexgpu.run_frame(img, _params(m)), ref)
    finally:
        sexgpu.set_option("deb_glb_levels", lvmin)


def test_option_range():
    with pytest.raises(ValueError):
        sexgpu.set_option("deb_lds_large", 7000)
    with pytest.raises(ValueError):
        sexgpu.set_option("deb_lds_small", -1)
