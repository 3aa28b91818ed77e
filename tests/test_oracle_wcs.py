"""Pinning the oracle's world coordinates (sxo_wcs.c): the reference's own astrometry
compiled here (create_wcs, raw_to_wcs with its WCSLIB and precess, oracle/_ref), the
textbook projection radii, and a published precession example."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import sxo

PROJ = ["TAN", "SIN", "ARC", "STG", "ZEA"]
REF = os.path.exists(sxo.REF_PATH)


def _sep(a1, d1, a2, d2):
    """great-circle separation in degrees (haversine)"""
    r = math.pi / 180.0
    h = (math.sin((d2 - d1) * r / 2) ** 2 +
         math.cos(d1 * r) * math.cos(d2 * r) * math.sin((a2 - a1) * r / 2) ** 2)
    return 2 * math.asin(math.sqrt(h)) / r


def _cases():
    rng = np.random.default_rng(17)
    out = []
    for k, p in enumerate(PROJ):
        for d0 in (-89.2, -30.0, 0.0, 45.0, 89.95):
            ang = rng.uniform(0, 2 * math.pi)
            s = 10 ** rng.uniform(-4.5, -3.0)
            flip = -1.0 if k % 2 == 0 else 1.0
            cd = [flip * s * math.cos(ang), -s * math.sin(ang) * 1.01, flip * s * math.sin(ang),
                  s * math.cos(ang)]
            out.append((p, float(rng.uniform(0, 360)), d0, cd, bool(k % 2)))
    return out


@pytest.mark.skipif(not REF, reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("case", range(25))
def test_world_coordinates_equal_the_reference(case):
    p, a0, d0, cd, swap = _cases()[case]
    ct = ("RA---" + p, "DEC--" + p)
    crval = [a0, d0]
    if swap:
        # declination first: the axes and their CD rows swap places
        ct, crval = ct[::-1], crval[::-1]
    rng = np.random.default_rng(case)
    pix = rng.uniform(-500, 4600, (200, 2))
    crpix = [2048.5, 2047.0]
    for eq, lp, tol in ((2000.0, 999.0, 1e-10), (1950.0, 999.0, 1e-8), (2000.0, 123.0, 1e-10)):
        w, r = sxo.wcs_world(ct, crval, crpix, cd, pix, longpole=lp, equinox=eq)
        w2, r2, _ = sxo.ref_wcs_world(ct, crval, crpix, cd, pix, longpole=lp, equinox=eq)
        ok = np.isfinite(w[:, 0])
        assert ok.mean() > 0.5
        # separations on the sky (right ascension differences shrink by cos(dec))
        la, lo = (1, 0) if not swap else (0, 1)
        dw = max(_sep(w[k, lo], w[k, la], w2[k, lo], w2[k, la]) for k in np.flatnonzero(ok))
        dr = max(_sep(r[k, 0], r[k, 1], r2[k, 0], r2[k, 1]) for k in np.flatnonzero(ok))
        assert dw < tol and dr < tol, (p, d0, eq, dw, dr)


@pytest.mark.parametrize("proj", PROJ)
def test_projection_radii(proj):
    """The reference pixel maps to CRVAL; a pixel at projection-plane distance R (degrees)
    lies at angular distance rho from it with R = (180/pi) tan rho (TAN), sin rho (SIN),
    rho (ARC), 2 tan(rho/2) (STG), 2 sin(rho/2) (ZEA)."""
    crval, crpix, s = [210.5, -37.25], [100.0, 200.0], 2.5e-3
    cd = [-s, 0.0, 0.0, s]
    pix = np.array([[100.0, 200.0], [100.0, 2200.0], [1500.0, 900.0], [-800.0, -300.0]])
    w, r = sxo.wcs_world(("RA---" + proj, "DEC--" + proj), crval, crpix, cd, pix)
    assert abs(r[0, 0] - crval[0]) < 1e-12 and abs(r[0, 1] - crval[1]) < 1e-12
    g = 180.0 / math.pi
    for k in range(1, len(pix)):
        R = math.hypot((pix[k, 0] - crpix[0]) * s, (pix[k, 1] - crpix[1]) * s)
        rho = _sep(crval[0], crval[1], r[k, 0], r[k, 1]) / g
        want = {"TAN": g * math.tan(rho), "SIN": g * math.sin(rho), "ARC": g * rho,
                "STG": 2 * g * math.tan(rho / 2), "ZEA": 2 * g * math.sin(rho / 2)}[proj]
        assert want == pytest.approx(R, rel=1e-9), (proj, k)
    # +y with this CD is north: the declination grows, the right ascension stays
    assert r[1, 1] > crval[1] and abs(r[1, 0] - crval[0]) < 1e-9


def test_precession_known_answer():
    """Meeus, Astronomical Algorithms (2nd ed.), example 21.b: theta Persei from J2000.0
    (with its proper motion applied) to JD 2462088.69 gives 41.547214, +49.348483."""
    a, d = C.c_double(), C.c_double()
    year = 2000.0 + (2462088.69 - 2451545.0) / 365.25
    sxo.lib().sxo_precess(2000.0, 41.054063, 49.227750, year, C.byref(a), C.byref(d))
    assert abs(a.value - 41.547214) < 2e-6 and abs(d.value - 49.348483) < 2e-6
    # there and back again
    a2, d2 = C.c_double(), C.c_double()
    sxo.lib().sxo_precess(year, a.value, d.value, 2000.0, C.byref(a2), C.byref(d2))
    assert abs(a2.value - 41.054063) < 1e-9 and abs(d2.value - 49.227750) < 1e-9
