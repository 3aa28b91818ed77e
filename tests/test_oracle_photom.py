"""Pinning the oracle's catalogue photometry (endobject_gpu, the reference's
src/cuda/cudaanalyse.cu:62-700, :2778-2921, :3320-3342) with checks that do not restate
its formulas: the published behaviour of each estimator on sources whose answer is known
(isolated and paired circular Gaussians integrated exactly over the pixels), an
independent evaluation of the classifier network, the weight table's own numbers, and the
label rules on constructed frames.
"""
import math
import os

import numpy as np
import pytest
from scipy.special import erf

import sexgpu
import sxo

NNW = os.path.join(os.path.dirname(sexgpu.__file__), "config", "default.nnw")
INT_MAX = 2 ** 31 - 1


def _gauss(h, w, sources, sky=1000.0, noise=1.0, seed=3):
    """A frame with circular Gaussians (y, x, flux, sigma) integrated over each pixel."""
    rng = np.random.default_rng(seed)
    img = sky + noise * rng.standard_normal((h, w))
    ey = np.arange(h + 1) - 0.5
    ex = np.arange(w + 1) - 0.5
    for (y0, x0, f, s) in sources:
        cy = np.diff(0.5 * erf((ey - y0) / (s * math.sqrt(2))))
        cx = np.diff(0.5 * erf((ex - x0) / (s * math.sqrt(2))))
        img += f * np.outer(cy, cx)
    return img.astype(np.float32)


def _run(img, **phot):
    base = {"nnw": sxo.read_nnw(NNW), "autoparam": (2.5, 3.5)}
    base.update(phot)
    return sxo.run(img, mask=None, back_size=32, clean=0, photom=base)


def _nearest(o, y, x):
    return int(np.argmin((o["mx"] - x) ** 2 + (o["my"] - y) ** 2))


def test_nnw_table_numbers():
    net = sxo.read_nnw(NNW)
    assert net["nn"] == [10, 10, 1]
    assert net["inbias"][0] == -1.56604 and net["inscale"][9] == 3.27734
    assert net["w"][0, 0, 0] == -0.32248 and net["b"][0, 0] == -5.30052
    assert net["w"][1, 9, 0] == -0.92225 and net["b"][1, 0] == -2.32028
    assert net["outbias"] == [0.0] and net["outscale"] == [1.0]
    # the product's parser reads the same numbers
    br = sexgpu.read_nnw(NNW)
    assert [br.nn[i] for i in range(3)] == [10, 10, 1] and br.layersnb == 3
    assert br.w[0][3][7] == net["w"][0, 3, 7] and br.b[1][0] == net["b"][1, 0]
    assert br.inscale[4] == net["inscale"][4] and br.outscale[0] == 1.0


def test_network_against_matrix_form():
    net = sxo.read_nnw(NNW)
    br = sxo.brain_struct(net)
    rng = np.random.default_rng(11)
    for _ in range(200):
        x = rng.uniform(-2, 3, 10)
        h = x * np.array(net["inscale"]) + np.array(net["inbias"])
        h1 = 1.0 / (1.0 + np.exp(-(net["w"][0].T @ h + net["b"][0])))
        y = 1.0 / (1.0 + np.exp(-(net["w"][1][:, 0] @ h1 + net["b"][1, 0])))
        got = sxo.lib().sxo_neurresp(br, x.ctypes.data)
        assert abs(got - y) < 1e-12


def test_log10_against_libm():
    L = sxo.lib()
    rng = np.random.default_rng(2)
    for x in np.concatenate([10 ** rng.uniform(-30, 30, 5000), rng.uniform(0.5, 20, 5000)]):
        a, b = L.sxo_log10(float(x)), math.log10(x)
        assert abs(a - b) <= 4 * np.spacing(abs(b)) + 1e-300


def _pixel_weights_flux(sub, mx, my, r, n=50):
    """Independent APER: each pixel weighted by the fraction of it inside the circle,
    measured on an n x n grid."""
    tot = 0.0
    g = (np.arange(n) + 0.5) / n - 0.5
    for y in range(int(my - r) - 1, int(my + r) + 2):
        for x in range(int(mx - r) - 1, int(mx + r) + 2):
            X, Y = np.meshgrid(x + g - mx, y + g - my)
            tot += ((X * X + Y * Y) < r * r).mean() * float(sub[y, x])
    return tot


def test_isolated_gaussian_estimators():
    """APER inside radius r holds F (1 - exp(-r^2 / 2 s^2)) when the edge is far out in
    the wings, and otherwise the pixel-area-weighted sum; the Kron radius of a Gaussian is
    2.5 sqrt(pi/2) s in units of the isophotal a; AUTO within it holds
    F (1 - exp(-(k a)^2 / 2 s^2)); magnitudes follow."""
    F, s = 2.0e5, 2.0
    img = _gauss(160, 160, [(80.3, 79.6, F, s)])
    zp = 25.0
    o = _run(img, apert=12.0, autoparam=(2.5, 0.0), mag_zeropoint=zp)
    i = _nearest(o, 80.3, 79.6)
    ph = o["phot"]
    assert abs(ph["flux_aper"][i] / (F * (1 - math.exp(-36.0 / (2 * s * s)))) - 1) < 0.005
    a = float(o["a"][i])
    kron = float(ph["kronfactor"][i])
    assert abs(kron * a / (2.5 * math.sqrt(math.pi / 2) * s) - 1) < 0.03
    frac = 1 - math.exp(-(kron * a) ** 2 / (2 * s * s))
    assert abs(ph["flux_auto"][i] / (F * frac) - 1) < 0.01
    assert abs(ph["mag_auto"][i] - (zp - 2.5 * math.log10(ph["flux_auto"][i]))) < 1e-5
    assert abs(ph["mag_aper"][i] - (zp - 2.5 * math.log10(ph["flux_aper"][i]))) < 1e-5
    assert ph["magerr_auto"][i] > 0 and ph["flag"][i] & 0x11 == 0
    # a small aperture: the 5x5 oversampling against a 50x50 one
    o = _run(img, apert=6.0)
    i = _nearest(o, 80.3, 79.6)
    ref = _pixel_weights_flux(o["sub"], o["mx"][i], o["my"][i], 3.0)
    assert abs(o["phot"]["flux_aper"][i] / ref - 1) < 0.015


def test_isocor_recovers_a_gaussian():
    """With the isophote at 30 ADU the isophotal flux of this Gaussian misses about 4 % of
    it; the ISOCOR correction (Gaussian profile assumption) brings it back to 1 %."""
    F, s = 2.0e4, 2.0
    img = _gauss(128, 128, [(64.2, 63.7, F, s)], noise=0.05)
    o = sxo.run(img, mask=None, back_size=32, clean=0, thresh_type=1, detect_thresh=30.0,
                analysis_thresh=30.0, photom={"nnw": None, "do_class": 0})
    i = _nearest(o, 64.2, 63.7)
    ph = o["phot"]
    assert ph["flux_iso"][i] / F < 0.975
    assert abs(ph["flux_isocor"][i] / F - 1) < 0.01
    assert ph["fluxerr_isocor"][i] > 0


def test_kron_minimum_radius_and_edge_flag():
    F, s = 5.0e4, 1.0
    img = _gauss(128, 128, [(64.0, 64.0, F, s), (3.2, 60.0, F, s)])
    o = _run(img, autoparam=(2.5, 3.5))
    ph = o["phot"]
    i, j = _nearest(o, 64.0, 64.0), _nearest(o, 3.2, 60.0)
    assert ph["kronfactor"][i] == np.float32(3.5)       # 2.5 sqrt(pi/2) s/a < 3.5
    assert ph["flag"][i] & 0x10 == 0
    assert ph["flag"][j] & 0x10                         # its ellipse crosses the edge


def test_neighbour_masking_and_mirror():
    """A sharp 3x3 neighbour 6-8 px from a Gaussian lies inside the Gaussian's Kron
    ellipse and is deblended as an object of its own.  With MASK_TYPE CORRECT its pixels
    take the mirror values through the Gaussian's centroid, so the Gaussian's flux
    matches the isolated source's; without it they count as zero and that flux is lost."""
    F, s = 1.0e5, 3.0
    base = _gauss(160, 160, [(80.0, 80.0, F, s)])
    iso = _run(base, autoparam=(2.5, 3.5))
    fi = iso["phot"]["flux_auto"][_nearest(iso, 80.0, 80.0)]
    pair = base.copy()
    pair[79:82, 86:89] += 1000.0
    lost = float((base[79:82, 86:89] - 1000.0).sum())
    for corr in (1, 0):
        o = _run(pair, mask_correct=corr, autoparam=(2.5, 3.5))
        k, j = _nearest(o, 80.0, 80.0), _nearest(o, 80.0, 87.0)
        assert k != j and (o["alloc"][79:82, 86:89] == j).all()
        rel = o["phot"]["flux_auto"][k] / fi
        if corr:
            assert abs(rel - 1) < 0.003
        else:
            assert abs(rel - (1 - lost / fi)) < 0.003 and lost / fi > 0.01


def test_class_star_separates_stars_from_galaxies():
    seeing = 3.0
    sig_star = seeing / (2 * math.sqrt(2 * math.log(2)))
    img = _gauss(200, 200, [(60.0, 60.0, 5e4, sig_star), (140.0, 140.0, 5e4, 6.0)], noise=2.0)
    o = _run(img, seeing_fwhm=seeing)
    ph = o["phot"]
    assert ph["sprob"][_nearest(o, 60.0, 60.0)] > 0.8
    assert ph["sprob"][_nearest(o, 140.0, 140.0)] < 0.2
    # and the index falls as the source broadens
    prev = 2.0
    for sg in (sig_star, 2.5, 4.0, 6.0, 8.0):
        o = _run(_gauss(120, 120, [(60.0, 60.0, 5e4, sg)], noise=2.0), seeing_fwhm=seeing)
        v = float(o["phot"]["sprob"][_nearest(o, 60.0, 60.0)])
        assert v < prev
        prev = v


def test_photometry_labels():
    """Detected pixels of no object (a blob below DETECT_MINAREA) and the pixels of objects
    removed by CLEAN (except object 0, whose label stays 0 as in the reference) become
    INT_MAX; other pixels keep their object index or -1."""
    img = _gauss(128, 128, [(40.0, 40.0, 3e4, 1.5), (90.0, 90.0, 3e4, 1.5)])
    img[10, 100] += 50.0                       # one bright pixel: detected, below minarea
    o = _run(img)
    lab = o["phot"]["labels"]
    assert lab[10, 100] == INT_MAX
    al = o["alloc"]
    assert np.array_equal(lab[al >= 0], al[al >= 0])
    assert ((lab == -1) | (lab == INT_MAX) | (al >= 0)).all()
    assert (o["cd"][lab == -1] < o["dthresh"]).all()
    # CLEAN: a faint source on a bright one's wing is removed; its pixels become INT_MAX
    img2 = _gauss(128, 128, [(64.0, 64.0, 5e6, 3.0), (64.0, 76.0, 400.0, 1.0)])
    r = sxo.run(img2, mask=None, back_size=32, clean=1, clean_param=1.0,
                photom={"nnw": None, "do_class": 0})
    lab2, al2 = r["phot"]["labels"], r["alloc"]
    for k in np.flatnonzero(r["master"] != 0xFFFFFFFF):
        sel = al2 == k
        expect = 0 if k == 0 else INT_MAX
        assert (lab2[sel] == expect).all()
