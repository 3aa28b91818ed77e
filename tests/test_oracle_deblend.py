"""Pins the oracle's deblending (multi-threshold tree, MINCONT pruning, gatherup) and the
isophotal counts against checks that do not share its code:

* blends whose outcome the mathematics fixes: a single Gaussian, two Gaussians closer
  than 2 sigma (one maximum: no split), two or three well separated Gaussians (split),
  a companion below / above DEBLEND_MINCONT;
* a brute-force restatement of the tree written from the reference's definitions
  (cudadeblend.cu:199-483, :1425-1615) with scipy labelling of each level's set,
  on random small blends;
* gatherup (cudadeblend.cu:1262-1360) as a brute-force argmin over the kept objects;
* ISOAREA and the isophotal areas as direct counts.
"""
import math

import numpy as np
import pytest
import scipy.ndimage as ndi

import sxo

EIGHT = np.ones((3, 3), bool)


def gaussians(h, w, comps, sky=0.0):
    yy, xx = np.mgrid[0:h, 0:w]
    img = np.full((h, w), sky, np.float64)
    for (y, x, amp, sig) in comps:
        img += amp * np.exp(-((yy - y) ** 2 + (xx - x) ** 2) / (2.0 * sig * sig))
    return img.astype(np.float32)


def run_abs(img, **kw):
    """No filter, MANUAL background 0, absolute thresholds."""
    base = dict(back_type=2, back_value=0.0, thresh_type=1, detect_thresh=5.0,
                analysis_thresh=5.0, back_size=min(64, *img.shape), clean=0)
    base.update(kw)
    return sxo.run(img, mask=None, **base)


# ------------------------------------------------------------ known outcomes ----
def test_single_gaussian_is_one_object():
    r = run_abs(gaussians(48, 48, [(24, 24, 1000.0, 2.0)]))
    assert r["nraw"] == 1 and r["nobj"] == 1 and r["level"][0] == 0


def test_unimodal_pair_is_not_split():
    # two equal Gaussians 2 px apart with sigma 2: the sum has a single maximum
    r = run_abs(gaussians(48, 48, [(24, 23, 1000.0, 2.0), (24, 25, 1000.0, 2.0)]))
    assert r["nobj"] == 1


@pytest.mark.parametrize("sep", [10, 14, 20])
def test_separated_pair_is_split(sep):
    # the saddle (about 2000 exp(-sep^2/32)) is far below the peaks, and the wings join
    # the two objects above the detection threshold for these separations
    img = gaussians(48, 64, [(24, 32 - sep / 2, 1000.0, 3.0), (24, 32 + sep / 2, 1000.0, 3.0)])
    r = run_abs(img)
    assert r["nraw"] == 1
    assert r["nobj"] == 2
    assert (r["level"] > 0).all()
    assert (r["flag"] & 2).all()            # MERGED
    xs = sorted(r["mx"])
    assert abs(xs[0] - (32 - sep / 2)) < 1.0 and abs(xs[1] - (32 + sep / 2)) < 1.0


def test_three_in_a_row():
    img = gaussians(40, 80, [(20, 20, 800.0, 3.0), (20, 38, 900.0, 3.0), (20, 56, 1000.0, 3.0)])
    r = run_abs(img)
    assert r["nraw"] == 1 and r["nobj"] == 3


def test_mincont_decides_faint_companion():
    # companion with 0.2% of the main flux: kept as a separate object only when
    # DEBLEND_MINCONT is below its (above-threshold) flux fraction
    img = gaussians(64, 64, [(32, 24, 20000.0, 3.0), (32, 40, 60.0, 3.0)])
    assert run_abs(img, deblend_mincont=0.005)["nobj"] == 1
    assert run_abs(img, deblend_mincont=0.0001)["nobj"] == 2


# ---------------------------------------------------------- brute-force tree ----
def f32_seqsum(v):
    """float32 sum in the given order (numpy's accumulate is sequential)."""
    if len(v) == 0:
        return np.float32(0.0)
    return np.add.accumulate(np.asarray(v, np.float32), dtype=np.float32)[-1]


def brute_deblend(cd, dthresh, minarea, nthresh, mincont, with_pixels=False):
    """Kept objects [(raw index, level, first pixel)] from the reference's definitions
    (with their level pixels and the raw object's pixels when with_pixels)."""
    H, W = cd.shape
    det, n = ndi.label(cd >= dthresh, structure=EIGHT)
    raws = []
    for k in range(1, n + 1):
        idx = np.flatnonzero(det.ravel() == k)
        if idx.size >= minarea:
            raws.append(idx)
    raws.sort(key=lambda a: a[0])
    debarea = min(minarea, 3)
    out = []
    flat = cd.ravel()
    for ri, pix in enumerate(raws):
        t0 = float(np.float32(dthresh))
        pk = float(flat[pix].max())
        # nodes: level, parent, pixel set (sorted raster indices), threshold
        nodes = [dict(level=0, parent=-1, pix=pix, t=np.float32(t0))]
        prev = [0]
        for i in range(1, nthresh):
            ti = np.float32(t0 * (pk / t0) ** (i / nthresh))
            cur = []
            mask = np.zeros(H * W, bool)
            for c in prev:
                p = nodes[c]["pix"]
                mask[p[flat[p] >= ti]] = True
            lab, m = ndi.label(mask.reshape(H, W), structure=EIGHT)
            comps = []
            for j in range(1, m + 1):
                q = np.flatnonzero(lab.ravel() == j)
                if q.size >= debarea:
                    par = next(c for c in prev if np.isin(q[0], nodes[c]["pix"]))
                    comps.append((q[0], q, par))
            comps.sort(key=lambda t: t[0])
            for first, q, par in comps:
                nodes.append(dict(level=i, parent=par, pix=q, t=ti))
                cur.append(len(nodes) - 1)
            if not cur:
                break
            prev = cur
        nn = len(nodes)
        nchild = [0] * nn
        for c in range(1, nn):
            nchild[nodes[c]["parent"]] += 1
        if max(nchild) < 2:
            out.append((ri, 0, int(pix[0])) + ((pix, pix) if with_pixels else ()))
            continue
        rootflux = f32_seqsum(flat[pix])
        value0 = float(rootflux) * mincont
        sig = [False] * nn
        nson = [0] * nn
        for c in range(nn - 1, 0, -1):
            nd = nodes[c]
            if nchild[nd["parent"]] >= 2:
                f = f32_seqsum(flat[nd["pix"]])
                lhs = np.float32(f - np.float32(nd["t"] * np.float32(nd["pix"].size)))
                if float(lhs) > value0:
                    sig[c] = True
                    nson[nd["parent"]] += 1
        oka = [True] * nn
        for c in range(nn - 1, -1, -1):
            if nson[c] > 1:
                oka[c] = False
            if c > 0 and not oka[c]:
                oka[nodes[c]["parent"]] = False
        for c in range(nn):
            kept = oka[0] if c == 0 else (nson[nodes[c]["parent"]] > 1 and oka[c] and sig[c])
            if kept:
                out.append((ri, nodes[c]["level"], int(nodes[c]["pix"][0])) +
                           ((nodes[c]["pix"], pix) if with_pixels else ()))
    return out


@pytest.mark.parametrize("seed", range(12))
def test_tree_matches_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    h, w = 40, 56
    comps = []
    for _ in range(int(rng.integers(2, 6))):
        comps.append((rng.uniform(8, h - 8), rng.uniform(8, w - 8), rng.uniform(50, 3000),
                      rng.uniform(1.2, 3.0)))
    img = gaussians(h, w, comps)
    img += rng.normal(0.0, 0.5, img.shape).astype(np.float32)
    nthresh = int(rng.choice([4, 16, 32]))
    mincont = float(rng.choice([0.0, 0.005, 0.05]))
    r = run_abs(img, deblend_nthresh=nthresh, deblend_mincont=mincont)
    want = brute_deblend(img, 5.0, 5, nthresh, mincont)
    got = list(zip(r["root"].tolist(), r["level"].tolist(), r["firstpix"].tolist()))
    assert got == want


# ------------------------------------------------------------------ gatherup ----
def full_moments(pix, cd, sub, W, dthresh):
    """FULL preanalysis of a kept component over its level pixels (raster order), from
    cudaanalyse.cu:1770-1898: (mx, my, cxx, cyy, cxy, abcor) with its float/double
    types.  The log on the abcor path is the canonical one of the build, pinned against
    libm elsewhere."""
    y, x = np.divmod(pix, W)
    cval = cd.ravel()[pix].astype(np.float32)
    val = sub.ravel()[pix].astype(np.float32)
    peak = float(val.max())
    rv = float(f32_seqsum(cval))
    xmin, ymin = int(x.min()), int(y.min())
    xr, yr = (x - xmin).astype(np.float32), (y - ymin).astype(np.float32)
    dthresh2 = (dthresh + peak) / 2.0
    dnpix = int((val > dthresh).sum())
    area2 = int((val > dthresh2).sum())
    mx = my = mx2 = my2 = mxy = 0.0
    for c, xx, yy in zip(cval, xr, yr):
        mx += float(np.float32(c * xx))
        my += float(np.float32(c * yy))
        mx2 += float(np.float32(np.float32(c * xx) * xx))
        my2 += float(np.float32(np.float32(c * yy) * yy))
        mxy += float(np.float32(np.float32(c * xx) * yy))
    xm, ym = mx / rv, my / rv
    xm2, ym2, xym = mx2 / rv - xm * xm, my2 / rv - ym * ym, mxy / rv - xm * ym
    temp2 = xm2 * ym2 - xym * xym
    if temp2 < 0.00694:
        xm2 += 0.0833333
        ym2 += 0.0833333
        temp2 = xm2 * ym2 - xym * xym
    temp = xm2 - ym2
    temp = math.sqrt(0.25 * temp * temp + xym * xym)
    pmx2 = pmy2 = 0.5 * (xm2 + ym2)
    pmx2 += temp
    pmy2 -= temp
    a, b = np.float32(math.sqrt(pmx2)), np.float32(math.sqrt(pmy2))
    cxx, cyy, cxy = np.float32(ym2 / temp2), np.float32(xm2 / temp2), np.float32(-2 * xym / temp2)
    darea = float(area2) - dnpix
    t1t2 = dthresh / dthresh2
    if t1t2 > 0.0:
        lg = sxo.lib().sxo_log(t1t2 if t1t2 < 1.0 else 0.99)
        abcor = np.float32((darea if darea < 0.0 else -1.0) /
                           (2 * 3.1415926535898 * lg * float(a) * float(b)))
        if abcor > 1.0:
            abcor = np.float32(1.0)
    else:
        abcor = np.float32(1.0)
    return xm + xmin, ym + ymin, cxx, cyy, cxy, abcor


def test_gatherup_is_the_brute_force_argmin():
    rng = np.random.default_rng(7)
    img = gaussians(64, 96, [(32, 30, 1500.0, 3.0), (30, 52, 1200.0, 3.5), (36, 70, 900.0, 2.5)])
    img += rng.normal(0.0, 0.5, img.shape).astype(np.float32)
    r = run_abs(img)
    kept = brute_deblend(img, 5.0, 5, 32, 0.005, with_pixels=True)
    assert [(k[0], k[1], k[2]) for k in kept] == list(
        zip(r["root"].tolist(), r["level"].tolist(), r["firstpix"].tolist()))
    assert r["nobj"] >= 3
    W = img.shape[1]
    pk = float(img.ravel()[kept[0][4]].max())
    mom = []
    for (_, lev, _, lpix, _) in kept:
        t = float(np.float32(5.0 * (pk / 5.0) ** (lev / 32)))
        mom.append(full_moments(lpix, img, img, W, t))
    want = np.full(img.size, -1, np.int64)
    for j, (_, _, _, lpix, _) in enumerate(kept):
        want[lpix] = j
    for (ri, _, _, _, rawpix) in kept:
        cand = [i for i, k in enumerate(kept) if k[0] == ri]
        for p in rawpix:
            if want[p] >= 0:
                continue
            y, x = divmod(int(p), W)
            best, dmin = cand[0], np.float32(1e31)
            for i in cand:
                mx, my, cxx, cyy, cxy, abcor = mom[i]
                dx = np.float32(x - mx)
                dy = np.float32(y - my)
                q = np.float32(np.float32(np.float32(cxx * dx) * dx) +
                               np.float32(np.float32(cyy * dy) * dy))
                q = np.float32(q + np.float32(np.float32(cxy * dx) * dy))
                d = np.float32(0.5 * float(q) / float(abcor))
                if d < dmin:
                    dmin, best = d, i
            want[p] = best
    assert np.array_equal(r["alloc"].ravel(), want)


# ---------------------------------------------------------------- isophotal ----
def test_isoarea_is_a_direct_count():
    rng = np.random.default_rng(3)
    img = gaussians(64, 64, [(20, 20, 500.0, 2.5), (44, 40, 2000.0, 3.0)])
    img += rng.normal(0.0, 1.0, img.shape).astype(np.float32)
    r = run_abs(img, analysis_thresh=12.0)
    sub = r["sub"].ravel()
    seg_alloc = r["alloc"].ravel()
    for i in range(r["nobj"]):
        pix = np.flatnonzero(seg_alloc == i)
        assert r["npix"][i] == int((sub[pix] > np.float32(12.0)).sum())
        thr, pk = np.float32(12.0), float(sub[pix].max())
        iso = [0] * 8
        for v in sub[pix]:
            for k in range(8):
                t = np.float32(float(thr) * (pk / float(thr)) ** (k / 8))
                if not v > t:
                    break
                iso[k] += 1
        assert list(r["iso"][i]) == iso
