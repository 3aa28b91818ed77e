"""Write the golden vectors of tests/golden/ from the reference's own CPU code.

Runs in the build container only (it needs /root/reference and oracle/_ref/libref.so,
built by ``make -C oracle ref`` from the reference's src/back.c, src/filter.c,
src/misc.c, src/globals.c and src/fits/fitsmisc.c).  Outputs:
  masks.npz        the convolution masks of config/{default,gauss_3.0_7x7,
                   mexhat_5.0_11x11,gauss_5.0_9x9,tophat_3.0_3x3}.conv as normalised by
                   the reference's getconv (filter.c:101-186)
  back_golden.npz  for seeded frames: the reference's mesh statistics (backstat),
                   mode maps (backguess), filtered maps and global medians (filterback),
                   spline derivatives (makebackspline), the subtracted frame
                   (subbackline) and convolutions (convolve)
The inputs are seeded numpy frames, stored in the archive with the outputs.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import sxo  # noqa: E402

CONF = "/root/reference/config"
MASKS = ["default", "gauss_3.0_7x7", "mexhat_5.0_11x11", "gauss_5.0_9x9", "tophat_3.0_3x3"]


def ref_getconv(ref, path):
    buf = np.zeros(1024, np.float32)
    mw, mh = C.c_int(), C.c_int()
    rc = ref.ref_getconv(path.encode(), buf.ctypes.data, 1024, C.byref(mw), C.byref(mh))
    assert rc == 0, path
    return buf[: mw.value * mh.value].reshape(mh.value, mw.value).copy()


def frames():
    """Seeded inputs: sky + gradient + noise + a few Gaussian sources."""
    out = {}
    rng = np.random.default_rng(20240611)

    def sky(h, w, n_src, bkg=1000.0, sig=10.0):
        y, x = np.mgrid[0:h, 0:w]
        img = bkg * (1 + 0.05 * (2 * x / (w - 1) - 1)) + rng.normal(0, sig, (h, w))
        for _ in range(n_src):
            cx, cy = rng.uniform(0, w), rng.uniform(0, h)
            amp, s = 10 ** rng.uniform(1.5, 4), rng.uniform(0.8, 3.0)
            img += amp * np.exp(-0.5 * ((x - cx) ** 2 + (y - cy) ** 2) / s ** 2)
        return img.astype(np.float32)

    out["a"] = (sky(128, 192, 40), 64, 3)       # whole meshes
    out["b"] = (sky(150, 230, 50), 48, 3)       # partial last meshes in x and y
    c = sky(160, 192, 40)
    c[0:32, 0:32] = -1e30                       # a bad mesh (all pixels flagged)
    c[40:80, 100:104] = -1e30                   # a partly bad mesh
    out["c"] = (c, 32, 5)
    d = sky(128, 128, 10, bkg=50.0, sig=0.5)
    out["d"] = (d, 32, 1)                       # narrow histograms, no median filter
    return out


def main():
    ref = sxo.ref()
    if ref is None:
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    masks = {}
    for name in MASKS:
        m_ref = ref_getconv(ref, os.path.join(CONF, name + ".conv"))
        m_orc = sxo.getconv(os.path.join(CONF, name + ".conv"))
        assert np.array_equal(m_ref.view(np.int32), m_orc.view(np.int32)), name
        masks[name.replace(".", "p")] = m_ref
    np.savez(os.path.join(HERE, "masks.npz"), **masks)

    arrays = {}
    for key, (img, bs, fs) in frames().items():
        H, W = img.shape
        nbx, nby = (W - 1) // bs + 1, (H - 1) // bs + 1
        stats = np.zeros((nby * nbx, 8), np.float32)
        back = np.zeros((nby, nbx), np.float32)
        sig = np.zeros_like(back)
        ref.ref_makeback(img.ctypes.data, W, H, bs, stats.ctypes.data, back.ctypes.data,
                         sig.ctypes.data)
        fback, fsig = back.copy(), sig.copy()
        bm, bsg = C.c_float(), C.c_float()
        ref.ref_filterback(fback.ctypes.data, fsig.ctypes.data, nbx, nby,
                           fs if nbx > 1 else 1, fs if nby > 1 else 1, 0.0,
                           C.byref(bm), C.byref(bsg))
        dback = np.zeros_like(back)
        dsig = np.zeros_like(back)
        ref.ref_makebackspline(fback.ctypes.data, nbx, nby, dback.ctypes.data)
        ref.ref_makebackspline(fsig.ctypes.data, nbx, nby, dsig.ctypes.data)
        sub = img.copy()
        ref.ref_subback(fback.ctypes.data, dback.ctypes.data, nbx, nby, bs, W, H,
                        sub.ctypes.data)
        arrays.update({key + "_img": img, key + "_stats": stats, key + "_mode": back,
                       key + "_modesig": sig, key + "_back": fback, key + "_sigma": fsig,
                       key + "_dback": dback, key + "_dsigma": dsig, key + "_sub": sub,
                       key + "_medians": np.array([bm.value, bsg.value], np.float32),
                       key + "_geom": np.array([bs, fs], np.int32)})
        if key in ("a", "b"):
            for name in ("default", "gauss_3.0_7x7", "mexhat_5.0_11x11"):
                m = masks[name.replace(".", "p")]
                cd = np.zeros_like(sub)
                ref.ref_convolve(sub.ctypes.data, W, H, m.ctypes.data, m.shape[1], m.shape[0],
                                 cd.ctypes.data)
                arrays[key + "_cd_" + name.replace(".", "p")] = cd
    # hmedian known answers on odd/even lengths and ties
    med_in = [np.array(v, np.float32) for v in ([3.0], [2.0, 1.0], [5.0, 1.0, 3.0],
                                                 [4.0, 1.0, 3.0, 2.0],
                                                 [1.5, -2.0, 7.25, 7.25, 0.0, 1e6],
                                                 list(np.linspace(-3, 9, 17)))]
    meds = []
    for v in med_in:
        w = v.copy()
        meds.append(ref.ref_hmedian(w.ctypes.data, len(w)))
    arrays["hmedian_in"] = np.concatenate(med_in)
    arrays["hmedian_len"] = np.array([len(v) for v in med_in], np.int32)
    arrays["hmedian_out"] = np.array(meds, np.float32)
    np.savez_compressed(os.path.join(HERE, "back_golden.npz"), **arrays)
    print("wrote", sorted(arrays))


if __name__ == "__main__":
    main()
