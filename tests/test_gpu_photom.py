"""GPU parity of the catalogue photometry (endobject_gpu: ISOCOR, APER, AUTO, CLASS_STAR,
magnitudes, ELONGATION/ELLIPTICITY and the APERT_PB/CROWDED/SATUR flags) against the
oracle, through the C-ABI.  Bar: bit-exact for every column and the flags: the kernels
keep the reference's raster-order sums and operand types, and both sides use the
canonical log10/exp.
"""
import os

import numpy as np
import pytest

import sexgpu
import sxo

pytestmark = pytest.mark.gpu

NNW = os.path.join(os.path.dirname(sexgpu.__file__), "config", "default.nnw")
MAP = {"flux_iso": "flux_iso", "fluxerr_iso": "fluxerr_iso", "mag_iso": "mag_iso",
       "magerr_iso": "magerr_iso", "flux_isocor": "flux_isocor",
       "fluxerr_isocor": "fluxerr_isocor", "flux_aper": "flux_aper",
       "fluxerr_aper": "fluxerr_aper", "mag_aper": "mag_aper", "magerr_aper": "magerr_aper",
       "kron_radius": "kronfactor", "flux_auto": "flux_auto", "fluxerr_auto": "fluxerr_auto",
       "mag_auto": "mag_auto", "magerr_auto": "magerr_auto", "elongation": "elong",
       "ellipticity": "ellip", "class_star": "sprob"}
PHOT_KEYS = ("apert", "autoparam", "autoaper", "seeing_fwhm", "mag_zeropoint", "mask_correct",
             "photo", "do_aper", "do_auto", "do_class")


@pytest.fixture(scope="module")
def crowded1024():
    return sexgpu.synth("crowded4096", width=1024, height=1024, nsrc=167000 // 16, seed=1201)


def run_both(img, mask, phot=None, ngamma=0.0, pixscale=1.0, **kw):
    phot = dict(phot or {})
    sexgpu.set_photometry(**phot)
    try:
        g = sexgpu.run_frame(img, sexgpu.Params(mask=mask, **kw),
                             outputs=("columns", "phot"), brain=sexgpu.read_nnw(NNW),
                             ngamma=ngamma, pixscale=pixscale)
    finally:
        sexgpu.set_photometry()
    okw = {"nnw": sxo.read_nnw(NNW), "ngamma": ngamma, "pixscale": pixscale}
    okw.update(phot)
    o = sxo.run(img, mask=mask, photom=okw, **kw)
    return g, o


def compare(g, o):
    assert g["nobj"] == o["nobj"]
    for gk, ok in MAP.items():
        a, b = np.asarray(g["phot"][gk]), np.asarray(o["phot"][ok])
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        assert same.all(), (gk, np.flatnonzero(~same)[:5], a[~same][:5], b[~same][:5])
    assert np.array_equal(g["flag_phot"], o["phot"]["flag"])


def test_plain1024_defaults(frame1024, masks):
    g, o = run_both(frame1024, masks["default"])
    compare(g, o)
    assert o["nobj"] > 500
    ph = o["phot"]
    assert (ph["flux_auto"] != 0).mean() > 0.9 and (ph["sprob"] > 0.5).any()


def test_crowded_defaults(crowded1024, masks):
    g, o = run_both(crowded1024, masks["gauss_3p0_7x7"])
    compare(g, o)
    flags = o["phot"]["flag"]
    assert (flags & 0x01).any() and (flags & 0x10).any()   # CROWDED and APERT_PB occur


@pytest.mark.parametrize("variant", [
    {"phot": {"mask_correct": 0}},
    {"phot": {"photo": 1}, "ngamma": 2.5},
    {"phot": {"autoaper": (12.0, 16.0)}},
    {"phot": {"apert": 12.0, "autoparam": (2.0, 2.5)}},
    {"phot": {"mag_zeropoint": 25.0, "seeing_fwhm": 2.0}, "pixscale": 0.5},
    {"gain": 4.0},
    {"clean": 0},
    {"phot": {"do_aper": 0, "do_class": 0}},
])
def test_crowded_variants(crowded1024, masks, variant):
    v = dict(variant)
    g, o = run_both(crowded1024, masks["gauss_3p0_7x7"], phot=v.pop("phot", None),
                    ngamma=v.pop("ngamma", 0.0), pixscale=v.pop("pixscale", 1.0), **v)
    compare(g, o)


def test_rows_carry_the_photometry(frame1024, masks):
    sexgpu.set_photometry()
    g = sexgpu.run_frame(frame1024, sexgpu.Params(mask=masks["default"]),
                         outputs=("columns", "phot"), brain=sexgpu.read_nnw(NNW))
    rows = g["catalog"]
    n = g["nobj"]
    get = lambda f: np.array([getattr(rows[i], f) for i in range(n)])
    assert np.array_equal(get("number"), np.arange(1, n + 1))
    assert np.array_equal(get("flux_auto").astype(np.float32), g["phot"]["flux_auto"])
    assert np.array_equal(np.array([rows[i].flux_aper[0] for i in range(n)], np.float32),
                          g["phot"]["flux_aper"])
    assert np.array_equal(get("sprob").astype(np.float32), g["phot"]["class_star"])
    assert np.array_equal(get("kronfactor").astype(np.float32), g["phot"]["kron_radius"])
    assert np.array_equal(get("flag"), g["flag_phot"].astype(np.int16))
    assert np.array_equal(get("mag_iso").astype(np.float32), g["phot"]["mag_iso"])


def test_empty_frame():
    img = np.full((256, 256), 100.0, np.float32)
    sexgpu.set_photometry()
    g = sexgpu.run_frame(img, sexgpu.Params(thresh_type=1, detect_thresh=1e6,
                                            analysis_thresh=1e6),
                         outputs=("columns", "phot"), brain=sexgpu.read_nnw(NNW))
    assert g["nobj"] == 0 and all(len(v) == 0 for v in g["phot"].values())


def _reference_structs(apert, autoparam, seeing, zp, mask_correct, groups, naper=1):
    """A prefstruct and an obj2struct of flags in the reference's layout, as its front-end
    fills them (prefs.c, and FLAG(obj2.x) set for the .param groups)."""
    import struct
    lay = sexgpu.reference_layout()
    prefs = bytearray(lay["prefs_size"])
    struct.pack_into("<i", prefs, lay["naper"], naper)
    struct.pack_into("<d", prefs, lay["apert"], apert)
    struct.pack_into("<2d", prefs, lay["autoparam"], *autoparam)
    struct.pack_into("<2d", prefs, lay["autoaper"], 0.0, 0.0)
    struct.pack_into("<d", prefs, lay["seeing_fwhm"], seeing)
    struct.pack_into("<d", prefs, lay["mag_zeropoint"], zp)
    struct.pack_into("<i", prefs, lay["mask_type"], lay["MASK_CORRECT"] if mask_correct else 0)
    struct.pack_into("<i", prefs, lay["detect_type"], 0)
    flags = bytearray(lay["obj2_size"])
    for g in groups:
        flags[lay[g]] = 1
    return prefs, flags


@pytest.mark.parametrize("groups", [("flux_aper", "flux_auto", "sprob"), ("flux_auto",)])
def test_reference_prefs_and_flags(frame1024, masks, groups):
    """endobject_gpu as scan.c:124-143 calls it, with the reference's prefs and obj2 flags
    (no sxg_set_photometry): the settings read from them give the oracle's columns for the
    same settings, and the groups the flags leave off are not measured."""
    prefs, flags = _reference_structs(8.0, (3.0, 2.0), 1.0, 25.0, False, groups)
    g = sexgpu.run_frame(frame1024, sexgpu.Params(mask=masks["default"]),
                         outputs=("columns", "phot"), brain=sexgpu.read_nnw(NNW), prefs=prefs,
                         flagobj2=flags)
    o = sxo.run(frame1024, mask=masks["default"], photom={
        "nnw": sxo.read_nnw(NNW), "apert": 8.0, "autoparam": (3.0, 2.0), "seeing_fwhm": 1.0,
        "mag_zeropoint": 25.0, "mask_correct": 0, "do_aper": int("flux_aper" in groups),
        "do_auto": int("flux_auto" in groups), "do_class": int("sprob" in groups)})
    compare(g, o)
    if "flux_aper" not in groups:
        assert not np.any(g["phot"]["flux_aper"]) and not np.any(g["phot"]["class_star"])


def test_reference_prefs_with_other_naper_is_an_error(frame1024, masks):
    prefs, flags = _reference_structs(5.0, (2.5, 3.5), 1.2, 0.0, True, ("flux_aper",), naper=3)
    with pytest.raises(RuntimeError, match="naper"):
        sexgpu.run_frame(frame1024, sexgpu.Params(mask=masks["default"]),
                         outputs=("columns", "phot"), prefs=prefs, flagobj2=flags)
