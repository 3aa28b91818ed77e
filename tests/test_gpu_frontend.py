"""The C++ `sex` front-end (bin/sex) on a FITS frame: it reads the .sex, .param, .conv and
.nnw files, drives the C-ABI in the reference's scan.c order and writes the ASCII_HEAD
catalogue and the SEGMENTATION check-image.  Its rows must carry the library's values
(to the printed precision) for the surviving objects, photometry columns included."""
import os
import subprocess

import numpy as np
import pytest

import sexgpu

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(sexgpu.__file__), "bin", "sex")
CONF = os.path.join(os.path.dirname(sexgpu.__file__), "config")
PARAMS = ["NUMBER", "X_IMAGE", "Y_IMAGE", "FLUX_ISO", "FLUX_ISOCOR", "FLUX_APER(1)",
          "MAG_APER(1)", "FLUX_AUTO", "MAG_AUTO", "KRON_RADIUS", "CLASS_STAR", "FLAGS"]


def test_sex_binary_catalogue_and_segmentation(tmp_path, frame1024, masks):
    img_path = str(tmp_path / "frame.fits")
    sexgpu.write_fits(img_path, frame1024)
    (tmp_path / "cat.param").write_text("\n".join(PARAMS) + "\n")
    cfg = {"CATALOG_NAME": str(tmp_path / "out.cat"), "CATALOG_TYPE": "ASCII_HEAD",
           "PARAMETERS_NAME": str(tmp_path / "cat.param"),
           "FILTER_NAME": os.path.join(CONF, "default.conv"),
           "STARNNW_NAME": os.path.join(CONF, "default.nnw"),
           "CHECKIMAGE_TYPE": "SEGMENTATION", "CHECKIMAGE_NAME": str(tmp_path / "seg.fits"),
           "VERBOSE_TYPE": "QUIET"}
    (tmp_path / "run.sex").write_text("".join("%s %s\n" % kv for kv in cfg.items()))
    r = subprocess.run([BIN, img_path, "-c", str(tmp_path / "run.sex")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = (tmp_path / "out.cat").read_text().splitlines()
    head = [ln for ln in lines if ln.startswith("#")]
    assert [h.split()[2] for h in head] == [p.split("(")[0] for p in PARAMS]
    rows = np.array([[float(v) for v in ln.split()] for ln in lines if not ln.startswith("#")])

    sexgpu.set_photometry()
    g = sexgpu.run_frame(frame1024, sexgpu.Params(mask=masks["default"]),
                         outputs=("columns", "phot", "seg"),
                         brain=sexgpu.read_nnw(os.path.join(CONF, "default.nnw")))
    keep = g["master"] == 0xFFFFFFFF
    assert rows.shape[0] == keep.sum() == g["nsurv"]
    ph = {k: v[keep] for k, v in g["phot"].items()}
    assert np.array_equal(rows[:, 0], np.arange(1, rows.shape[0] + 1))
    assert np.allclose(rows[:, 1], (g["mx"][keep] + 1.0), atol=6e-4)
    for col, key in ((3, "flux_iso"), (4, "flux_isocor"), (5, "flux_aper"), (7, "flux_auto")):
        assert np.allclose(rows[:, col], ph[key], rtol=1e-6, atol=1e-3), key
    for col, key, tol in ((6, "mag_aper", 6e-5), (8, "mag_auto", 6e-5), (9, "kron_radius", 6e-3),
                          (10, "class_star", 6e-3)):
        assert np.allclose(rows[:, col], ph[key], atol=tol), key
    assert np.array_equal(rows[:, 11].astype(int), g["flag_phot"][keep])
    seg = sexgpu.read_fits(str(tmp_path / "seg.fits"))
    assert np.array_equal(seg.astype(np.int64), g["seg"].astype(np.int64))


@pytest.mark.parametrize("wtype", ["MAP_WEIGHT", "BACKGROUND"])
def test_sex_binary_weighted(tmp_path, masks, wtype):
    """WEIGHT_TYPE MAP_WEIGHT (a FITS weight map) and BACKGROUND, with WEIGHT_GAIN and a
    gain: the segmentation check-image equals the library's map for the same weighting;
    BACKGROUND_RMS writes the interpolated background rms."""
    img = sexgpu.synth("plain1024", width=640, height=512, nsrc=1500, seed=61)
    yy, xx = np.mgrid[0:512, 0:640]
    wt = ((1.0 + 0.3 * np.sin(xx / 50.0)) / 100.0).astype(np.float32)
    wt[:, 200:204] = 0.0
    img_path, wt_path = str(tmp_path / "f.fits"), str(tmp_path / "w.fits")
    sexgpu.write_fits(img_path, img)
    sexgpu.write_fits(wt_path, wt)
    (tmp_path / "p.param").write_text("NUMBER\nFLUX_ISO\nFLUXERR_ISO\n")
    base = {"PARAMETERS_NAME": str(tmp_path / "p.param"), "VERBOSE_TYPE": "QUIET",
            "FILTER_NAME": os.path.join(CONF, "default.conv"), "WEIGHT_TYPE": wtype,
            "WEIGHT_IMAGE": wt_path, "WEIGHT_GAIN": "Y", "GAIN": "2.0"}
    outs = {}
    for check in ("SEGMENTATION", "BACKGROUND_RMS"):
        cfg = dict(base, CATALOG_NAME=str(tmp_path / ("c_%s.cat" % check)),
                   CHECKIMAGE_TYPE=check, CHECKIMAGE_NAME=str(tmp_path / (check + ".fits")))
        (tmp_path / "w.sex").write_text("".join("%s %s\n" % kv for kv in cfg.items()))
        r = subprocess.run([BIN, img_path, "-c", str(tmp_path / "w.sex")], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs[check] = sexgpu.read_fits(str(tmp_path / (check + ".fits")))
    p = sexgpu.Params(mask=masks["default"], gain=2.0)
    try:
        g = sexgpu.run_frame(img, p, outputs=("columns", "seg"),
                             weight=wt if wtype == "MAP_WEIGHT" else None, weight_type=wtype,
                             weight_gain=1)
        rms = sexgpu.back_rms(img.shape)
    finally:
        sexgpu.set_weight(None, p)
    assert np.array_equal(outs["SEGMENTATION"].astype(np.int64), g["seg"].astype(np.int64))
    assert np.array_equal(outs["BACKGROUND_RMS"], rms)
    rows = np.array([[float(v) for v in ln.split()] for ln in
                     (tmp_path / "c_SEGMENTATION.cat").read_text().splitlines()
                     if not ln.startswith("#")])
    keep = g["master"] == 0xFFFFFFFF
    assert np.allclose(rows[:, 2], np.sqrt(g["fluxerr"][keep].astype(np.float64)), rtol=1e-6)


def _read_fits_tables(path):
    """A minimal FITS reader for the test: every HDU's header (dict) and raw data bytes."""
    raw = open(path, "rb").read()
    pos, hdus = 0, []
    while pos < len(raw):
        head = {}
        while True:
            block = raw[pos:pos + 2880]
            pos += 2880
            cards = [block[i:i + 80].decode("ascii") for i in range(0, 2880, 80)]
            done = False
            for c in cards:
                k = c[:8].strip()
                if k == "END":
                    done = True
                    break
                if c[8:10] == "= ":
                    v = c[10:].split(" / ")[0].strip()
                    head[k] = v.strip("'").strip() if v.startswith("'") else v
            if done:
                break
        n = int(head.get("NAXIS1", 0)) * int(head.get("NAXIS2", 0)) if int(head["NAXIS"]) else 0
        hdus.append((head, raw[pos:pos + n]))
        pos += (n + 2879) // 2880 * 2880
    return hdus


def test_sex_binary_fits_ldac(tmp_path, frame1024, masks):
    """CATALOG_TYPE FITS_LDAC: an empty primary HDU, LDAC_IMHEAD with the frame's header
    cards, and LDAC_OBJECTS whose big-endian columns hold the library's values exactly."""
    img_path = str(tmp_path / "frame.fits")
    sexgpu.write_fits(img_path, frame1024)
    names = ["NUMBER", "X_IMAGE", "FLUX_AUTO", "X2_IMAGE", "FLAGS", "ISO0", "CLASS_STAR"]
    (tmp_path / "l.param").write_text("\n".join(names) + "\n")
    cfg = {"CATALOG_NAME": str(tmp_path / "out.ldac"), "CATALOG_TYPE": "FITS_LDAC",
           "PARAMETERS_NAME": str(tmp_path / "l.param"),
           "FILTER_NAME": os.path.join(CONF, "default.conv"),
           "STARNNW_NAME": os.path.join(CONF, "default.nnw"), "VERBOSE_TYPE": "QUIET"}
    (tmp_path / "l.sex").write_text("".join("%s %s\n" % kv for kv in cfg.items()))
    r = subprocess.run([BIN, img_path, "-c", str(tmp_path / "l.sex")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    hdus = _read_fits_tables(str(tmp_path / "out.ldac"))
    assert len(hdus) == 3 and hdus[0][0]["NAXIS"] == "0"
    imh, imd = hdus[1]
    assert imh["EXTNAME"] == "LDAC_IMHEAD" and imd[:6] == b"SIMPLE"
    assert b"NAXIS1  =                 1024" in imd
    th, td = hdus[2]
    assert th["EXTNAME"] == "LDAC_OBJECTS"
    assert [th["TTYPE%d" % (k + 1)] for k in range(len(names))] == names
    code = {"J": ">i4", "I": ">i2", "E": ">f4", "D": ">f8"}
    dt = np.dtype([(n, code[th["TFORM%d" % (k + 1)][-1]]) for k, n in enumerate(names)])
    assert dt.itemsize == int(th["NAXIS1"])
    t = np.frombuffer(td, dtype=dt)

    sexgpu.set_photometry()
    g = sexgpu.run_frame(frame1024, sexgpu.Params(mask=masks["default"]),
                         outputs=("columns", "phot"),
                         brain=sexgpu.read_nnw(os.path.join(CONF, "default.nnw")))
    keep = g["master"] == 0xFFFFFFFF
    assert len(t) == keep.sum()
    assert np.array_equal(t["NUMBER"], np.arange(1, len(t) + 1))
    assert np.array_equal(t["X_IMAGE"], (g["mx"][keep] + 1.0).astype(np.float32))
    assert np.array_equal(t["FLUX_AUTO"], g["phot"]["flux_auto"][keep])
    assert np.array_equal(t["X2_IMAGE"], g["mx2"][keep])
    assert np.array_equal(t["FLAGS"], g["flag_phot"][keep].astype(np.int16))
    assert np.array_equal(t["ISO0"], g["iso"][keep][:, 0])
    assert np.array_equal(t["CLASS_STAR"], g["phot"]["class_star"][keep])


# the settings of the reference's tests/default.sex that the detection path reads
GALAXIES_SEX = {"DETECT_MINAREA": "2", "DETECT_THRESH": "1.5", "ANALYSIS_THRESH": "2",
                "FILTER": "Y", "DEBLEND_NTHRESH": "32", "DEBLEND_MINCONT": "0.005",
                "CLEAN": "Y", "CLEAN_PARAM": "1.0", "MASK_TYPE": "CORRECT",
                "PHOT_APERTURES": "5", "PHOT_AUTOPARAMS": "2.5, 3.5", "SATUR_LEVEL": "50000.0",
                "MAG_ZEROPOINT": "0.0", "GAIN": "0.0", "PIXEL_SCALE": "1.0",
                "SEEING_FWHM": "1.2", "BACK_SIZE": "64", "BACK_FILTERSIZE": "3",
                "MEMORY_OBJSTACK": "3000", "WEIGHT_TYPE": "NONE"}


def test_reference_test_frame_against_oracle(tmp_path):
    """The reference's own test run (tests/efigi1.test: tests/galaxies.fits, BITPIX 16
    with BSCALE/BZERO, with its tests/default.sex) through bin/sex, writing the rows of
    the reference's printobj (CATALOG_ROWS REFERENCE): the segmentation check-image equals
    the oracle's on the same decoded frame, and every row field equals the oracle's value
    to its printed precision (survivors numbered from 0, as scan.c:150-156 does)."""
    from conftest import GOLDEN
    import sxo
    path = os.path.join(GOLDEN, "galaxies.fits")
    (tmp_path / "p.param").write_text("NUMBER\nX_IMAGE\nY_IMAGE\n")
    cfg = dict(GALAXIES_SEX, CATALOG_NAME=str(tmp_path / "g.cat"), CATALOG_TYPE="ASCII",
               CATALOG_ROWS="REFERENCE", PARAMETERS_NAME=str(tmp_path / "p.param"),
               FILTER_NAME=os.path.join(CONF, "default.conv"),
               STARNNW_NAME=os.path.join(CONF, "default.nnw"), CHECKIMAGE_TYPE="SEGMENTATION",
               CHECKIMAGE_NAME=str(tmp_path / "seg.fits"), VERBOSE_TYPE="QUIET")
    (tmp_path / "g.sex").write_text("".join("%s %s\n" % kv for kv in cfg.items()))
    r = subprocess.run([BIN, path, "-c", str(tmp_path / "g.sex")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = np.array([[float(v) for v in ln.split()]
                     for ln in (tmp_path / "g.cat").read_text().splitlines()])
    img = sexgpu.read_fits(path)
    o = sxo.run(img, mask=sexgpu.read_conv("default.conv"), ext_minarea=2, analysis_thresh=2.0,
                photom={"nnw": sxo.read_nnw(os.path.join(CONF, "default.nnw")), "apert": 5.0,
                        "autoparam": (2.5, 3.5)})
    seg = sexgpu.read_fits(str(tmp_path / "seg.fits"))
    assert np.array_equal(seg.astype(np.int64), o["seg"].astype(np.int64))
    keep = o["master"] == 0xFFFFFFFF
    ph = {k: v[keep] for k, v in o["phot"].items() if k != "labels"}
    n = int(keep.sum())
    assert rows.shape == (n, 31) and n == o["nsurv"] > 100
    f4 = np.float32
    want = {0: np.arange(n), 3: (o["mx"][keep] + 1.0).astype(f4), 4: (o["my"][keep] + 1.0).astype(f4),
            5: ph["flux_iso"], 6: ph["fluxerr_iso"], 7: ph["mag_iso"], 8: ph["magerr_iso"],
            9: ph["flux_isocor"], 10: ph["fluxerr_isocor"], 11: ph["flux_aper"],
            12: ph["fluxerr_aper"], 13: ph["mag_aper"], 14: ph["magerr_aper"],
            15: ph["flux_auto"], 16: ph["fluxerr_auto"], 17: ph["mag_auto"],
            18: ph["magerr_auto"], 19: o["theta"][keep], 20: o["poserr_mx2"][keep],
            21: o["poserr_my2"][keep], 22: o["poserr_mxy"][keep], 23: ph["flag"],
            24: ph["elong"], 25: ph["ellip"], 26: ph["sprob"], 27: o["peak"][keep],
            28: ph["kronfactor"], 29: o["bkg"][keep], 30: o["fwhm"][keep]}
    # the printed precision of each field (scan.c:331-406)
    fixed = {3: 6e-4, 4: 6e-4, 7: 6e-5, 8: 6e-5, 13: 6e-5, 14: 6e-5, 17: 6e-5, 18: 6e-5,
             19: 0.06, 24: 6e-4, 25: 6e-4, 26: 6e-3, 28: 6e-3, 30: 6e-3}
    for c, v in want.items():
        v = np.asarray(v, np.float64)
        if c in (0, 23):
            assert np.array_equal(rows[:, c], v), c
        elif c in fixed:
            assert np.allclose(rows[:, c], v, rtol=0, atol=fixed[c]), c
        elif c in (20, 21, 22):
            assert np.allclose(rows[:, c], v, rtol=1e-9, atol=0), c
        else:
            assert np.allclose(rows[:, c], v, rtol=1e-6, atol=1e-30), c
    assert np.all(rows[:, 1] == 0.0) and np.all(rows[:, 2] == 0.0)
