"""The FITS reader of the front-end (host/fitsio.c) against the conversions of the
reference's FITS library (src/fits/fitsbody.c:195-345, src/fits/fitshead.c:148-160),
restated with numpy float32 arithmetic: BSCALE and BZERO rounded to float, integer pixels
scaled in float (BITPIX 8 signed unless BITSGN = 0), BLANK and NaN/Inf pixels -> -BIG,
BITPIX -64 scaled in double.  Also the reference's own test frame, tests/galaxies.fits
(BITPIX 16 with BSCALE/BZERO), read bit for bit as numpy decodes it."""
import os

import numpy as np
import pytest

import sexgpu
from conftest import GOLDEN

BIG = np.float32(-1e30)


def _write(path, data, bitpix, **cards):
    h, w = data.shape
    head = ["SIMPLE  = %20s" % "T", "BITPIX  = %20d" % bitpix, "NAXIS   = %20d" % 2,
            "NAXIS1  = %20d" % w, "NAXIS2  = %20d" % h]
    for k, v in cards.items():
        head.append("%-8s= %20s" % (k, repr(float(v)) if isinstance(v, float) else v))
    head.append("END")
    blob = "".join(c.ljust(80) for c in head)
    blob += " " * (-len(blob) % 2880)
    dt = {8: ">u1", 16: ">i2", 32: ">i4", -32: ">f4", -64: ">f8"}[bitpix]
    body = np.ascontiguousarray(data).astype(dt).tobytes()
    body += b"\0" * (-len(body) % 2880)
    with open(path, "wb") as f:
        f.write(blob.encode("ascii") + body)


def _f32(bs, bz, vals):
    """numpy's float32 restatement of `pixel * (float)BSCALE + (float)BZERO`"""
    return (np.asarray(vals).astype(np.float32) * np.float32(bs)) + np.float32(bz)


@pytest.mark.parametrize("bitpix", [16, 32])
def test_scaled_integers(tmp_path, bitpix):
    rng = np.random.default_rng(bitpix)
    lo, hi = (-32768, 32767) if bitpix == 16 else (-2 ** 31, 2 ** 31 - 1)
    raw = rng.integers(lo, hi, (37, 53), dtype=np.int64)
    if bitpix == 32:
        raw[0, :5] = [16777217, -16777219, 2 ** 31 - 1, -2 ** 31, 123456789]  # float rounding
    bs, bz = 2.92273835509, 3713.66692596
    p = str(tmp_path / "i.fits")
    _write(p, raw, bitpix, BSCALE=bs, BZERO=bz)
    got = sexgpu.read_fits(p)
    assert got.dtype == np.float32 and got.shape == raw.shape
    assert np.array_equal(got.view(np.int32), _f32(bs, bz, raw).view(np.int32))


def test_blank_pixels(tmp_path):
    raw = np.arange(-20, 20, dtype=np.int64).reshape(5, 8) * 100
    raw[2, 3] = -999
    p = str(tmp_path / "b.fits")
    _write(p, raw, 16, BSCALE=0.5, BZERO=10.0, BLANK="-999")
    got = sexgpu.read_fits(p)
    want = _f32(0.5, 10.0, raw)
    want[2, 3] = BIG
    assert np.array_equal(got.view(np.int32), want.view(np.int32))


def test_bytes_are_signed_unless_bitsgn_zero(tmp_path):
    raw = np.array([[0, 1, 127, 128, 200, 255]], np.int64)
    p = str(tmp_path / "c.fits")
    _write(p, raw, 8, BSCALE=2.0, BZERO=1.0)
    got = sexgpu.read_fits(p)
    assert np.array_equal(got, _f32(2.0, 1.0, raw.astype(np.uint8).view(np.int8)))
    _write(p, raw, 8, BSCALE=2.0, BZERO=1.0, BITSGN="0")
    assert np.array_equal(sexgpu.read_fits(p), _f32(2.0, 1.0, raw))


@pytest.mark.parametrize("bitpix", [-32, -64])
def test_floats_with_nan_and_inf(tmp_path, bitpix):
    rng = np.random.default_rng(7)
    data = rng.normal(100.0, 30.0, (9, 11))
    data[1, 2], data[3, 4], data[5, 6] = np.nan, np.inf, -np.inf
    data[0, 0] = -0.0
    p = str(tmp_path / "f.fits")
    _write(p, data, bitpix)
    got = sexgpu.read_fits(p)
    if bitpix == -32:
        want = data.astype(np.float32) * np.float32(1.0) + np.float32(0.0)
    else:
        want = (data * 1.0 + 0.0).astype(np.float32)
    bad = ~np.isfinite(data)
    want[bad] = BIG
    assert np.array_equal(got.view(np.int32), want.view(np.int32))
    # -0 * 1 + 0 is +0 in the reference's arithmetic
    assert got.view(np.int32)[0, 0] == 0


def test_reference_test_frame():
    """tests/galaxies.fits of the reference (1024 x 1024, BITPIX 16, BSCALE 2.92273835509,
    BZERO 3713.66692596) through the reader equals numpy's float32 decoding."""
    path = os.path.join(GOLDEN, "galaxies.fits")
    with open(path, "rb") as f:
        blob = f.read()
    cards = []
    for i in range(0, len(blob), 80):
        cards.append(blob[i:i + 80].decode("ascii"))
        if cards[-1].startswith("END "):
            break
    end = len(cards) - 1
    kv = {c[:8].strip(): c[10:].split("/")[0].strip() for c in cards[:end] if c[8:10] == "= "}
    assert kv["BITPIX"] == "16" and kv["NAXIS1"] == "1024" and kv["NAXIS2"] == "1024"
    off = -(-(end + 1) * 80 // 2880) * 2880
    raw = np.frombuffer(blob, ">i2", 1024 * 1024, off).reshape(1024, 1024)
    want = _f32(float(kv["BSCALE"]), float(kv["BZERO"]), raw)
    got = sexgpu.read_fits(path)
    assert np.array_equal(got.view(np.int32), want.view(np.int32))
    assert 3700.0 < float(np.median(got)) < 10000.0 and float(got.max()) < 98529.0


def test_header_wcs_keywords(tmp_path):
    """The WCS keywords as read_wcs reads them (fitswcs.c:320-470): the CD matrix from
    CDi_j; else PC00i00j times CDELTi; else CDELTi with CROTA2 (cd = R(CROTA2) diag(CDELT));
    EQUINOX, else EPOCH; CRPIX defaults to 1, LONPOLE/LATPOLE to 999."""
    import math
    data = np.zeros((4, 5))
    p = str(tmp_path / "w.fits")
    base = dict(CTYPE1="'RA---TAN'", CTYPE2="'DEC--TAN'", CRVAL1=12.5, CRVAL2=-33.0)
    _write(p, data, -32, CD1_1=-1e-4, CD1_2=2e-5, CD2_1=3e-5, CD2_2=1.5e-4, EQUINOX=1950.0,
           LONPOLE=170.0, **base)
    w = sexgpu.read_fits_wcs(p)
    assert w["ctype"] == ("RA---TAN", "DEC--TAN") and w["crval"] == [12.5, -33.0]
    assert w["cd"] == [-1e-4, 2e-5, 3e-5, 1.5e-4] and w["crpix"] == [1.0, 1.0]
    assert (w["equinox"], w["longpole"], w["latpole"]) == (1950.0, 170.0, 999.0)
    _write(p, data, -32, CDELT1=-2e-4, CDELT2=3e-4, PC001001=0.8, PC001002=-0.6,
           PC002001=0.6, PC002002=0.8, EPOCH=1985.0, CRPIX1=7.0, **base)
    w = sexgpu.read_fits_wcs(p)
    assert np.allclose(w["cd"], [-2e-4 * 0.8, -2e-4 * -0.6, 3e-4 * 0.6, 3e-4 * 0.8], rtol=1e-15)
    assert w["equinox"] == 1985.0 and w["crpix"][0] == 7.0
    _write(p, data, -32, CDELT1=-2e-4, CDELT2=3e-4, CROTA2=30.0, **base)
    w = sexgpu.read_fits_wcs(p)
    c, s = math.cos(math.radians(30.0)), math.sin(math.radians(30.0))
    assert np.allclose(w["cd"], [c * -2e-4, -s * 3e-4, s * -2e-4, c * 3e-4], rtol=1e-15)
    assert w["equinox"] == 2000.0
    _write(p, data, -32)
    assert sexgpu.read_fits_wcs(p) is None
