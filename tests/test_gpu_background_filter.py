"""GPU parity: background estimation/subtraction and convolution through the C-ABI.

Bar: bit-exact.  Expected values come from the reference's own back.c/filter.c (golden
vectors in tests/golden/back_golden.npz) and from the oracle on synthetic frames.
"""
import numpy as np
import pytest

import sexgpu
import sxo

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.int32)


def _gpu_background(img, bs, fs, back_type=0, back_value=0.0):
    p = sexgpu.Params(back_size=bs, back_fsize=fs, back_type=back_type, back_value=back_value)
    bm, bsig = sexgpu.run_background(img, p)
    H, W = img.shape
    nbx, nby = (W - 1) // bs + 1, (H - 1) // bs + 1
    back, sigma, dback, dsigma = sexgpu.back_maps(nbx, nby)
    return dict(backmean=bm, backsig=bsig, back=back, sigma=sigma, dback=dback, dsigma=dsigma)


@pytest.fixture(scope="module")
def golden():
    import os
    from conftest import GOLDEN
    d = np.load(os.path.join(GOLDEN, "back_golden.npz"), allow_pickle=False)
    return {k: d[k] for k in d.files}


@pytest.mark.parametrize("key", ["a", "b", "c", "d"])
def test_background_matches_reference_golden(golden, key):
    img = golden[key + "_img"]
    bs, fs = (int(v) for v in golden[key + "_geom"])
    g = _gpu_background(img, bs, fs)
    assert np.array_equal(_bits(g["back"]), _bits(golden[key + "_back"]))
    assert np.array_equal(_bits(g["sigma"]), _bits(golden[key + "_sigma"]))
    assert np.array_equal(_bits(g["dback"]), _bits(golden[key + "_dback"]))
    assert np.array_equal(_bits(g["dsigma"]), _bits(golden[key + "_dsigma"]))
    assert (g["backmean"], g["backsig"]) == tuple(float(v) for v in golden[key + "_medians"])
    sub = sexgpu.subtracted(img.shape)
    assert np.array_equal(_bits(sub), _bits(golden[key + "_sub"]))


@pytest.mark.parametrize("key", ["a", "b"])
@pytest.mark.parametrize("mask", ["default", "gauss_3p0_7x7", "mexhat_5p0_11x11"])
def test_convolution_matches_reference_golden(golden, masks, key, mask):
    img = golden[key + "_img"]
    bs, fs = (int(v) for v in golden[key + "_geom"])
    _gpu_background(img, bs, fs)
    sexgpu.apply_filter(masks[mask])
    cd = sexgpu.filtered(img.shape)
    sexgpu.subtracted(img.shape)
    assert np.array_equal(_bits(cd), _bits(golden[key + "_cd_" + mask]))


@pytest.mark.parametrize("mask", ["default", "gauss_3p0_7x7", "mexhat_5p0_11x11",
                                  "gauss_5p0_9x9", "tophat_3p0_3x3"])
def test_synthetic_1024_background_and_filter_match_oracle(frame1024, masks, mask):
    img = frame1024
    o = sxo.background(img, bs=64, fsize=3)
    g = _gpu_background(img, 64, 3)
    for k in ("back", "sigma", "dback", "dsigma"):
        assert np.array_equal(_bits(g[k]), _bits(o[k])), k
    assert (g["backmean"], g["backsig"]) == (o["backmean"], o["backsig"])
    m = masks[mask]
    sexgpu.apply_filter(m)
    cd = sexgpu.filtered(img.shape)
    sub = sexgpu.subtracted(img.shape)
    assert np.array_equal(_bits(sub), _bits(o["sub"]))
    assert np.array_equal(_bits(cd), _bits(sxo.convolve(o["sub"], m)))


def test_manual_background_subtracts_global_median(frame1024):
    """BACK_TYPE MANUAL as the reference's device code runs it (cudaback.cu:1336-1345):
    the constant subtracted is the global median of the filtered mesh map, which
    back_estimate also returns; BACK_VALUE plays no part in the subtraction."""
    img = frame1024[:300, :257].copy()
    auto = sxo.background(img, bs=64, fsize=3)
    o = sxo.background(img, bs=64, fsize=3, back_type=1, back_value=997.5)
    assert o["backmean"] == auto["backmean"]
    assert np.array_equal(_bits(o["sub"]), _bits(img - np.float32(auto["backmean"])))
    g = _gpu_background(img, 64, 3, back_type=1, back_value=997.5)
    assert g["backmean"] == o["backmean"] and g["backsig"] == o["backsig"]
    sub = sexgpu.subtracted(img.shape)
    assert np.array_equal(_bits(sub), _bits(o["sub"]))


def test_manual_background_reports_back_value_for_the_frame(frame1024):
    """back.c:152-153: the whole-frame run reports BACK_VALUE as the frame's backmean."""
    img = frame1024[:256, :256].copy()
    p = sexgpu.Params(back_type=1, back_value=997.5, mask=None)
    g = sexgpu.run_frame(img, p, outputs=("seg",))
    o = sxo.run(img, mask=None, back_type=1, back_value=997.5)
    assert g["backmean"] == np.float32(997.5) and o["backmean"] == np.float32(997.5)
    assert np.array_equal(g["seg"], o["seg"])


def test_stock_manual_background_subtracts_back_value(frame1024):
    """back_type 2, stock SExtractor's MANUAL (back.c:148-149, :807-813): subtract the
    constant BACK_VALUE, which becomes backmean."""
    img = frame1024[:300, :257].copy()
    o = sxo.background(img, bs=64, fsize=3, back_type=2, back_value=997.5)
    assert np.array_equal(_bits(o["sub"]), _bits(img - np.float32(997.5)))
    g = _gpu_background(img, 64, 3, back_type=2, back_value=997.5)
    assert g["backmean"] == np.float32(997.5)
    assert g["backsig"] == o["backsig"]
    sub = sexgpu.subtracted(img.shape)
    assert np.array_equal(_bits(sub), _bits(o["sub"]))


@pytest.mark.parametrize("shape,bs", [((1, 1), 64), ((37, 5), 4), ((1000, 17), 16),
                                      ((129, 130), 64)])
def test_odd_geometries_match_oracle(shape, bs):
    rng = np.random.default_rng(7)
    img = (100.0 + rng.normal(0, 3, shape)).astype(np.float32)
    H, W = shape
    bw, bh = min(bs, W), min(bs, H)
    if bw != bh:
        pytest.skip("non-square meshes are rejected as in back.c:99-103")
    o = sxo.background(img, bs=bw, fsize=3)
    g = _gpu_background(img, bw, 3)
    assert np.array_equal(_bits(g["back"]), _bits(o["back"]))
    assert (g["backmean"], g["backsig"]) == (o["backmean"], o["backsig"])
    sub = sexgpu.subtracted(img.shape)
    assert np.array_equal(_bits(sub), _bits(o["sub"]))


def _flat_with_patch(shape, patch, seed=11):
    img = np.full(shape, 100.0, np.float32)
    if patch:
        y0, y1, x0, x1 = patch
        rng = np.random.default_rng(seed)
        img[y0:y1, x0:x1] += rng.normal(0, 2, (y1 - y0, x1 - x0)).astype(np.float32)
    return img


@pytest.mark.parametrize("shape,patch", [((512, 512), (0, 192, 0, 512)),
                                         ((448, 320), (64, 256, 0, 128)),
                                         ((256, 256), None)])
def test_zero_sigma_meshes_use_the_positive_median(shape, patch):
    """Most meshes flat: the global sigma is the median of the positive mesh sigmas, or 1
    with none (back.c:542-555); odd and even mesh counts."""
    img = _flat_with_patch(shape, patch)
    o = sxo.background(img, bs=32, fsize=3)
    g = _gpu_background(img, 32, 3)
    assert np.array_equal(_bits(g["sigma"]), _bits(o["sigma"]))
    assert (g["backmean"], g["backsig"]) == (o["backmean"], o["backsig"])
    if patch is None:
        assert g["backsig"] == 1.0
    else:
        assert g["backsig"] > 0.0
