"""GPU world coordinates (ALPHA_J2000, DELTA_J2000, X_WORLD, Y_WORLD) of the catalogue
positions, through the C-ABI, against the oracle's restatement (pinned against the
reference's own compiled astrometry in tests/test_oracle_wcs.py) and, when oracle/_ref is
built, against the reference's raw_to_wcs/precess on a wcsstruct the reference itself built
and handed to endobject_gpu.  Bar: 1e-10 degree on the sky (1e-8 with precession: the
reference's precession constants are rounded differently), device and host libm apart."""
import math
import os

import numpy as np
import pytest

import sexgpu
import sxo

pytestmark = pytest.mark.gpu


def _sep(a1, d1, a2, d2):
    r = np.pi / 180.0
    h = np.sin((d2 - d1) * r / 2) ** 2 + np.cos(d1 * r) * np.cos(d2 * r) * np.sin((a2 - a1) * r / 2) ** 2
    return 2 * np.arcsin(np.sqrt(np.minimum(h, 1.0))) / r


def _frame():
    return sexgpu.synth("plain1024", width=900, height=700, nsrc=1500, seed=91)


def _cd(s, ang_deg, flip=-1.0):
    a = math.radians(ang_deg)
    return [flip * s * math.cos(a), -s * math.sin(a), flip * s * math.sin(a), s * math.cos(a)]


CASES = [("TAN", (150.1, 2.2), 2000.0, False), ("SIN", (10.0, -75.0), 2000.0, False),
         ("ARC", (300.0, 45.0), 2000.0, True), ("STG", (45.0, 89.8), 2000.0, False),
         ("ZEA", (200.0, -20.0), 1950.0, False), ("TAN", (359.9, -60.0), 1950.0, True)]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_world_coordinates_match_oracle(masks, case):
    proj, (a0, d0), eq, swap = CASES[case]
    ct, crval = ("RA---" + proj, "DEC--" + proj), [a0, d0]
    if swap:
        ct, crval = ct[::-1], crval[::-1]
    cd, crpix = _cd(2.7e-4, 17.0 + 40 * case), [451.5, 349.0]
    img = _frame()
    sexgpu.set_wcs(ct, crval, crpix, cd, equinox=eq)
    try:
        g = sexgpu.run_frame(img, sexgpu.Params(mask=masks["default"]),
                             outputs=("columns", "catalog"))
    finally:
        sexgpu.set_wcs(None)
    xw, yw, ra, dec = g["world"]
    pix = np.stack([g["mx"] + 1.0, g["my"] + 1.0], axis=1)
    w, r = sxo.wcs_world(ct, crval, crpix, cd, pix, equinox=eq)
    tol = 1e-10 if eq == 2000.0 else 1e-9
    assert g["nobj"] > 300
    assert _sep(ra, dec, r[:, 0], r[:, 1]).max() < tol
    lo, la = (0, 1) if not swap else (1, 0)
    gw = np.stack([xw, yw], axis=1)
    assert _sep(gw[:, lo], gw[:, la], w[:, lo], w[:, la]).max() < 1e-10
    # the catalogue rows carry the same values
    rows = g["catalog"]
    assert np.array_equal(np.array([rows[i].alpha2000 for i in range(g["nobj"])]), ra)
    assert np.array_equal(np.array([rows[i].delta2000 for i in range(g["nobj"])]), dec)


def test_no_wcs_gives_zeros(masks):
    g = sexgpu.run_frame(_frame(), sexgpu.Params(mask=masks["default"]),
                         outputs=("columns", "catalog"))
    assert all(not np.any(a) for a in g["world"])


@pytest.mark.skipif(not os.path.exists(sxo.REF_PATH), reason="oracle/_ref is not built")
@pytest.mark.parametrize("eq", [2000.0, 1950.0])
def test_reference_wcsstruct_through_endobject(masks, eq):
    """The reference's create_wcs builds its wcsstruct; endobject_gpu reads it as the
    reference front-end passes field->wcs (scan.c:124-143) and the positions equal the
    reference's own raw_to_wcs (and precess) of the same 1-based positions."""
    ct, crval, crpix, cd = ("RA---TAN", "DEC--TAN"), [83.6, -5.4], [450.0, 350.5], _cd(3e-4, 33.0)
    _, _, wp = sxo.ref_wcs_world(ct, crval, crpix, cd, np.zeros((1, 2)) + 400.0, equinox=eq)
    g = sexgpu.run_frame(_frame(), sexgpu.Params(mask=masks["default"]),
                         outputs=("columns", "catalog"), wcs=wp)
    xw, yw, ra, dec = g["world"]
    pix = np.stack([g["mx"] + 1.0, g["my"] + 1.0], axis=1)
    w2, r2, _ = sxo.ref_wcs_world(ct, crval, crpix, cd, pix, equinox=eq)
    assert _sep(ra, dec, r2[:, 0], r2[:, 1]).max() < (1e-10 if eq == 2000.0 else 1e-8)
    assert _sep(xw, yw, w2[:, 0], w2[:, 1]).max() < 1e-10


def _write_fits_wcs(path, img, cards):
    head = ["SIMPLE  = %20s" % "T", "BITPIX  = %20d" % -32, "NAXIS   = %20d" % 2,
            "NAXIS1  = %20d" % img.shape[1], "NAXIS2  = %20d" % img.shape[0]]
    for k, v in cards:
        head.append("%-8s= %s" % (k, ("'%-8s'" % v) if isinstance(v, str) else "%20.12E" % v))
    head.append("END")
    blob = "".join(c.ljust(80) for c in head)
    blob += " " * (-len(blob) % 2880)
    body = img.astype(">f4").tobytes()
    body += b"\0" * (-len(body) % 2880)
    with open(path, "wb") as f:
        f.write(blob.encode("ascii") + body)


def test_sex_binary_world_columns(tmp_path):
    """bin/sex reads the header WCS (CTYPE, CRVAL, CRPIX, CD, EQUINOX) and writes
    ALPHA_J2000/DELTA_J2000/X_WORLD/Y_WORLD equal to the oracle's positions through the
    oracle's WCS, to the printed precision."""
    import subprocess
    img = _frame()
    cd = _cd(2.9e-4, 25.0)
    cards = [("CTYPE1", "RA---TAN"), ("CTYPE2", "DEC--TAN"), ("CRVAL1", 210.25),
             ("CRVAL2", 54.3), ("CRPIX1", 450.0), ("CRPIX2", 351.0), ("CD1_1", cd[0]),
             ("CD1_2", cd[1]), ("CD2_1", cd[2]), ("CD2_2", cd[3]), ("EQUINOX", 1950.0)]
    path = str(tmp_path / "w.fits")
    _write_fits_wcs(path, img, cards)
    conf = os.path.join(os.path.dirname(sexgpu.__file__), "config")
    (tmp_path / "p.param").write_text("NUMBER\nALPHA_J2000\nDELTA_J2000\nX_WORLD\nY_WORLD\n")
    cfg = {"CATALOG_NAME": str(tmp_path / "w.cat"), "CATALOG_TYPE": "ASCII",
           "PARAMETERS_NAME": str(tmp_path / "p.param"),
           "FILTER_NAME": os.path.join(conf, "default.conv"), "VERBOSE_TYPE": "QUIET"}
    (tmp_path / "w.sex").write_text("".join("%s %s\n" % kv for kv in cfg.items()))
    binp = os.path.join(os.path.dirname(sexgpu.__file__), "bin", "sex")
    r = subprocess.run([binp, path, "-c", str(tmp_path / "w.sex")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = np.array([[float(v) for v in ln.split()]
                     for ln in (tmp_path / "w.cat").read_text().splitlines()])
    o = sxo.run(sexgpu.read_fits(path), mask=sexgpu.read_conv("default.conv"))
    keep = o["master"] == 0xFFFFFFFF
    pix = np.stack([o["mx"][keep] + 1.0, o["my"][keep] + 1.0], axis=1)
    w, rd = sxo.wcs_world(("RA---TAN", "DEC--TAN"), [210.25, 54.3], [450.0, 351.0], cd, pix,
                          equinox=1950.0)
    assert rows.shape == (keep.sum(), 5) and rows.shape[0] > 300
    assert np.allclose(rows[:, 1], rd[:, 0], rtol=0, atol=6e-8)
    assert np.allclose(rows[:, 2], rd[:, 1], rtol=0, atol=6e-8)
    assert np.allclose(rows[:, 3], w[:, 0], rtol=1e-6, atol=0)
    assert np.allclose(rows[:, 4], w[:, 1], rtol=1e-6, atol=0)
    # precession moved the positions: J2000 differs from the 1950 world coordinates
    assert np.abs(rows[:, 1] - rows[:, 3]).min() > 0.1
