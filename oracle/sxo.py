"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by the
CPU-baseline leg of ``bench.py``, never by the product.  The oracle restates the
reference's CPU background/filter code and its CUDA detection algorithms (see the
headers of oracle/sxo_*.c for the file:line each function follows).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = C.CDLL(LIB_PATH)
        _lib.sxo_run.restype = C.POINTER(SxoResult)
        _lib.sxo_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(SxoConfig)]
        _lib.sxo_free.argtypes = [C.POINTER(SxoResult)]
        _lib.sxo_measure_fp64.argtypes = [C.POINTER(SxoResult), C.c_void_p, C.c_int]
        _lib.sxo_log.restype = C.c_double
        _lib.sxo_log.argtypes = [C.c_double]
        _lib.sxo_exp.restype = C.c_double
        _lib.sxo_exp.argtypes = [C.c_double]
        _lib.sxo_log10.restype = C.c_double
        _lib.sxo_log10.argtypes = [C.c_double]
        _lib.sxo_pow.restype = C.c_double
        _lib.sxo_pow.argtypes = [C.c_double, C.c_double]
        _lib.sxo_hmedian.restype = C.c_float
        _lib.sxo_hmedian.argtypes = [C.c_void_p, C.c_int]
        _lib.sxo_back_at.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int]
        _lib.sxo_back_at.restype = C.c_float
        _lib.sxo_wcs_setup.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_double, C.c_double]
        _lib.sxo_wcs_world.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_void_p,
                                       C.c_void_p]
        _lib.sxo_precess.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib.sxo_backguess.restype = C.c_float
        _lib.sxo_getconv.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        _lib.sxo_convolve.argtypes = [vp, i, i, vp, i, i, vp]
        _lib.sxo_background.argtypes = [vp, i, i, i, i, i, d, i, d, vp, vp, vp, vp, vp, vp, vp]
        _lib.sxo_makeback.argtypes = [vp, i, i, i, i, i, d, vp, vp, vp, vp]
        _lib.sxo_filterback.argtypes = [vp, vp, i, i, i, i, d, vp, vp]
        _lib.sxo_makebackspline.argtypes = [vp, i, i, vp]
        _lib.sxo_subbackline.argtypes = [vp, vp, i, i, i, i, i, vp, vp]
        _lib.sxo_backrmsline.argtypes = [vp, vp, i, i, i, i, i, vp]
        _lib.sxo_backstat_mesh.argtypes = [vp, i, i, i, i, i, vp]
        _lib.sxo_backhisto_mesh.argtypes = [vp, i, i, i, i, i, vp, vp]
        _lib.sxo_backguess.argtypes = [vp, vp, vp, vp]
        _lib.sxo_endobject.argtypes = [C.POINTER(SxoResult), C.POINTER(SxoPhotomConfig),
                                       C.POINTER(SxoPhotom)]
        _lib.sxo_photom_labels.argtypes = [C.POINTER(SxoResult), vp]
        _lib.sxo_neurresp.restype = C.c_double
        _lib.sxo_neurresp.argtypes = [C.POINTER(SxoBrain), vp]
    return _lib


def ref():
    """The reference's own back.c/filter.c compiled from the reference tree, or None."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        _ref = C.CDLL(REF_PATH)
        _ref.ref_hmedian.restype = C.c_float
        _ref.ref_hmedian.argtypes = [C.c_void_p, C.c_int]
        _ref.ref_filterback.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_double, C.c_void_p, C.c_void_p]
        _ref.ref_getconv.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]
        vp, i = C.c_void_p, C.c_int
        _ref.ref_makeback.argtypes = [vp, i, i, i, vp, vp, vp]
        _ref.ref_makebackspline.argtypes = [vp, i, i, vp]
        _ref.ref_subback.argtypes = [vp, vp, i, i, i, i, i, vp]
        _ref.ref_backrms.argtypes = [vp, vp, i, i, i, i, i, vp]
        _ref.ref_convolve.argtypes = [vp, i, i, vp, i, i, vp]
        _ref.ref_wcs_new.argtypes = [C.c_char_p, C.c_char_p, vp, vp, vp, C.c_double,
                                     C.c_double, C.c_double]
        _ref.ref_wcs_new.restype = vp
        _ref.ref_wcs_world.argtypes = [vp, i, vp, vp, vp]
        _ref.ref_wcs_layout.argtypes = [vp, i]
    return _ref


class SxoConfig(C.Structure):
    _fields_ = [
        ("back_size", C.c_int), ("back_fsize", C.c_int), ("back_fthresh", C.c_double),
        ("back_type", C.c_int), ("back_value", C.c_double), ("thresh_type", C.c_int),
        ("detect_thresh", C.c_double), ("analysis_thresh", C.c_double),
        ("ext_minarea", C.c_int), ("deblend_nthresh", C.c_int),
        ("deblend_mincont", C.c_double), ("clean", C.c_int), ("clean_param", C.c_double),
        ("clean_stacksize", C.c_int), ("satur_level", C.c_double), ("gain", C.c_double),
        ("mask", C.c_void_p), ("mask_w", C.c_int), ("mask_h", C.c_int),
        ("weight", C.c_void_p), ("weight_thresh", C.c_double), ("weight_type", C.c_int),
        ("weight_gain", C.c_int),
    ]


_IP = C.POINTER(C.c_int)
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)


class SxoResult(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("nbx", C.c_int), ("nby", C.c_int),
        ("backmean", C.c_float), ("backsig", C.c_float), ("dthresh", C.c_float),
        ("thresh", C.c_float),
        ("back", _FP), ("sigma", _FP), ("dback", _FP), ("dsigma", _FP), ("sub", _FP),
        ("cd", _FP), ("var", _FP),
        ("npix_above", C.c_long), ("npix_raw", C.c_long), ("nraw", C.c_int),
        ("raw_first", _IP), ("raw_count", _IP), ("raw_fdflux", _FP), ("raw_fdpeak", _FP),
        ("nobj", C.c_int),
        ("level", _IP), ("root", _IP), ("firstpix", _IP), ("flag", _IP), ("fdnpix", _IP),
        ("dnpix", _IP), ("npix", _IP), ("xmin", _IP), ("xmax", _IP), ("ymin", _IP),
        ("ymax", _IP), ("peakx", _IP), ("peaky", _IP), ("iso", _IP),
        ("mx", _DP), ("my", _DP), ("mx2", _DP), ("my2", _DP), ("mxy", _DP),
        ("poserr_mx2", _DP), ("poserr_my2", _DP), ("poserr_mxy", _DP),
        ("a", _FP), ("b", _FP), ("theta", _FP), ("cxx", _FP), ("cyy", _FP), ("cxy", _FP),
        ("abcor", _FP), ("fdflux", _FP), ("dflux", _FP), ("flux", _FP), ("fluxerr", _FP),
        ("fdpeak", _FP), ("dpeak", _FP), ("peak", _FP), ("bkg", _FP), ("thresh_obj", _FP),
        ("mthresh", _FP), ("fwhm", _FP), ("dthresh_obj", _FP),
        ("master", C.POINTER(C.c_uint32)), ("alloc", _IP), ("seg", _IP), ("nsurv", C.c_int),
        ("objpix_start", _IP), ("objpix", _IP), ("levpix_start", _IP), ("levpix", _IP),
        ("stage_ms", C.c_double * 7),
    ]


class SxoBrain(C.Structure):
    _fields_ = [("layersnb", C.c_int), ("nn", C.c_int * 3),
                ("inbias", C.c_double * 10), ("inscale", C.c_double * 10),
                ("outbias", C.c_double * 10), ("outscale", C.c_double * 10),
                ("w", C.c_double * 10 * 10 * 2), ("b", C.c_double * 10 * 2)]


class SxoPhotomConfig(C.Structure):
    _fields_ = [("apert", C.c_double), ("autoparam", C.c_double * 2),
                ("autoaper", C.c_double * 2), ("seeing_fwhm", C.c_double),
                ("pixscale", C.c_double), ("mag_zeropoint", C.c_double),
                ("ngamma", C.c_double), ("gain", C.c_double), ("satur_level", C.c_double),
                ("backsig", C.c_float), ("thresh", C.c_float), ("mask_correct", C.c_int),
                ("photo", C.c_int), ("do_aper", C.c_int), ("do_auto", C.c_int),
                ("do_class", C.c_int), ("brain", C.POINTER(SxoBrain))]


PHOT_FLT = ["flux_iso", "fluxerr_iso", "mag_iso", "magerr_iso", "flux_isocor",
            "fluxerr_isocor", "flux_aper", "fluxerr_aper", "mag_aper", "magerr_aper",
            "kronfactor", "flux_auto", "fluxerr_auto", "mag_auto", "magerr_auto", "elong",
            "ellip", "sprob"]


class SxoPhotom(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in PHOT_FLT] + [("flag", C.c_void_p)]


def read_nnw(path):
    """The classifier weights of a NNW table, parsed as getnnw does (src/neurro.c:96-180):
    the layer sizes, the input biases and scales, per layer and neuron its input weights
    followed by its bias, then the output biases and scales."""
    lines = []
    with open(path) as f:
        head = f.readline()
        if not head.startswith("NNW"):
            raise ValueError(path + " is not a NNW table")
        for ln in f:
            t = ln.lstrip(" \t")
            if t and t[0] not in "#\n":
                lines.append([float(v) for v in t.split()])
    nn = [int(v) for v in lines[0][1:1 + int(lines[0][0])]]
    L = len(nn)
    net = {"layersnb": L, "nn": nn, "inbias": lines[1][:nn[0]], "inscale": lines[2][:nn[0]]}
    w = np.zeros((2, 10, 10))
    b = np.zeros((2, 10))
    k = 3
    for layer in range(L - 1):
        for j in range(nn[layer + 1]):
            row = lines[k]
            w[layer, :nn[layer], j] = row[:nn[layer]]
            b[layer, j] = row[nn[layer]]
            k += 1
    net["w"], net["b"] = w, b
    net["outbias"] = lines[k][:nn[-1]]
    net["outscale"] = lines[k + 1][:nn[-1]]
    return net


def brain_struct(net):
    br = SxoBrain()
    br.layersnb = net["layersnb"]
    for i, v in enumerate(net["nn"]):
        br.nn[i] = v
    for name in ("inbias", "inscale", "outbias", "outscale"):
        arr = getattr(br, name)
        for i, v in enumerate(net[name]):
            arr[i] = v
    for l in range(2):
        for i in range(10):
            for j in range(10):
                br.w[l][i][j] = float(net["w"][l, i, j])
            br.b[l][i] = float(net["b"][l, i])
    return br


PHOTOM_DEFAULTS = {"apert": 5.0, "autoparam": (2.5, 3.5), "autoaper": (0.0, 0.0),
                   "seeing_fwhm": 1.2, "pixscale": 1.0, "mag_zeropoint": 0.0, "ngamma": 0.0,
                   "mask_correct": 1, "photo": 0, "do_aper": 1, "do_auto": 1, "do_class": 1,
                   "nnw": None}


def _photom(rp, r, kw, photom):
    pk = dict(PHOTOM_DEFAULTS)
    pk.update(photom)
    pc = SxoPhotomConfig()
    pc.apert = pk["apert"]
    pc.autoparam[0], pc.autoparam[1] = pk["autoparam"]
    pc.autoaper[0], pc.autoaper[1] = pk["autoaper"]
    for k in ("seeing_fwhm", "pixscale", "mag_zeropoint", "ngamma", "mask_correct", "photo",
              "do_aper", "do_auto", "do_class"):
        setattr(pc, k, pk[k])
    pc.gain = kw.get("gain", 0.0)
    pc.satur_level = kw.get("satur_level", 50000.0)
    pc.backsig, pc.thresh = r.backsig, r.thresh
    br = None
    if pk["nnw"] is not None:
        br = brain_struct(pk["nnw"])
        pc.brain = C.pointer(br)
    m = r.nobj
    arrs = {k: np.zeros(max(m, 1), np.float32) for k in PHOT_FLT}
    arrs["flag"] = np.zeros(max(m, 1), np.int32)
    ph = SxoPhotom()
    for k, a in arrs.items():
        setattr(ph, k, a.ctypes.data)
    if lib().sxo_endobject(rp, C.byref(pc), C.byref(ph)) != m:
        raise RuntimeError("sxo_endobject failed")
    out = {k: a[:m] for k, a in arrs.items()}
    lab = np.zeros(r.width * r.height, np.int32)
    lib().sxo_photom_labels(rp, lab.ctypes.data)
    out["labels"] = lab.reshape(r.height, r.width)
    return out


OBJ_INT = ["level", "root", "firstpix", "flag", "fdnpix", "dnpix", "npix", "xmin", "xmax",
           "ymin", "ymax", "peakx", "peaky"]
OBJ_DBL = ["mx", "my", "mx2", "my2", "mxy", "poserr_mx2", "poserr_my2", "poserr_mxy"]
OBJ_FLT = ["a", "b", "theta", "cxx", "cyy", "cxy", "abcor", "fdflux", "dflux", "flux",
           "fluxerr", "fdpeak", "dpeak", "peak", "bkg", "thresh_obj", "mthresh", "fwhm",
           "dthresh_obj"]


def _arr(ptr, n, dtype):
    if n <= 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def config(**kw):
    c = SxoConfig()
    c.back_size, c.back_fsize, c.back_fthresh = 64, 3, 0.0
    c.detect_thresh = c.analysis_thresh = 1.5
    c.ext_minarea, c.deblend_nthresh, c.deblend_mincont = 5, 32, 0.005
    c.clean, c.clean_param, c.clean_stacksize = 1, 1.0, 3000
    c.satur_level, c.gain = 50000.0, 0.0
    for k, v in kw.items():
        if k != "mask":
            setattr(c, k, v)
    return c


TRUTH_COLUMNS = ["flux", "mx", "my", "mx2", "my2", "mxy", "a", "b", "theta", "cxx", "cyy", "cxy"]


# the stages sxo_run times (wall clock, milliseconds)
STAGES = ("background", "filter", "detection", "deblending", "analysis", "clean", "segmap")


def run(img, mask=None, weight=None, truth=False, photom=None, **kw):
    """Run the whole oracle path on a float32 frame (and an optional MAP_WEIGHT map);
    returns a dict of numpy arrays.  truth: also the long-double evaluation of the graded
    columns on the same pixel sets (out["truth"][:, TRUTH_COLUMNS.index(name)]).
    photom: a dict of photometry settings (keys of PHOTOM_DEFAULTS; nnw = read_nnw(...)):
    also the catalogue photometry of endobject_gpu, as out["phot"]."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    H, W = img.shape
    cfg = config(**kw)
    keep = None
    if mask is not None:
        keep = np.ascontiguousarray(mask, dtype=np.float32)
        cfg.mask = keep.ctypes.data
        cfg.mask_h, cfg.mask_w = keep.shape
    wkeep = None
    if weight is not None:
        wkeep = np.ascontiguousarray(weight, dtype=np.float32)
        if wkeep.shape != img.shape:
            raise ValueError("weight map and frame differ in shape")
        cfg.weight = wkeep.ctypes.data
    rp = lib().sxo_run(img.ctypes.data, W, H, C.byref(cfg))
    if not rp:
        raise RuntimeError("oracle rejected the configuration")
    r = rp.contents
    N = W * H
    out = {k: getattr(r, k) for k in ("backmean", "backsig", "dthresh", "thresh", "npix_above",
                                      "npix_raw", "nraw", "nobj", "nsurv", "nbx", "nby")}
    nb = r.nbx * r.nby
    for k in ("back", "sigma", "dback", "dsigma"):
        out[k] = _arr(getattr(r, k), nb, np.float32).reshape(r.nby, r.nbx)
    out["sub"] = _arr(r.sub, N, np.float32).reshape(H, W)
    out["cd"] = _arr(r.cd, N, np.float32).reshape(H, W)
    out["var"] = _arr(r.var, N, np.float32).reshape(H, W) if r.var else None
    out["alloc"] = _arr(r.alloc, N, np.int32).reshape(H, W)
    out["seg"] = _arr(r.seg, N, np.int32).reshape(H, W)
    n = r.nraw
    out["raw_first"] = _arr(r.raw_first, n, np.int32)
    out["raw_count"] = _arr(r.raw_count, n, np.int32)
    out["raw_fdflux"] = _arr(r.raw_fdflux, n, np.float32)
    out["raw_fdpeak"] = _arr(r.raw_fdpeak, n, np.float32)
    m = r.nobj
    for k in OBJ_INT:
        out[k] = _arr(getattr(r, k), m, np.int32)
    for k in OBJ_DBL:
        out[k] = _arr(getattr(r, k), m, np.float64)
    for k in OBJ_FLT:
        out[k] = _arr(getattr(r, k), m, np.float32)
    out["iso"] = _arr(r.iso, 8 * m, np.int32).reshape(m, 8)
    out["master"] = _arr(r.master, m, np.uint32)
    out["stage_ms"] = dict(zip(STAGES, r.stage_ms))
    # each object's final pixels and its level pixels (the kept component before
    # gatherup), raster order: object o owns [start[o], start[o + 1])
    out["objpix_start"] = _arr(r.objpix_start, m + 1, np.int32)
    out["objpix"] = _arr(r.objpix, int(out["objpix_start"][-1]) if m else 0, np.int32)
    out["levpix_start"] = _arr(r.levpix_start, m + 1, np.int32)
    out["levpix"] = _arr(r.levpix, int(out["levpix_start"][-1]) if m else 0, np.int32)
    if truth:
        t = np.zeros((max(m, 1), 12), np.float64)
        if lib().sxo_measure_fp64(rp, t.ctypes.data, 12) != m:
            raise RuntimeError("sxo_measure_fp64 failed")
        out["truth"] = t[:m]
    if photom is not None:
        out["phot"] = _photom(rp, r, kw, photom)
    lib().sxo_free(rp)
    return out


class SxoWcs(C.Structure):
    """Mirror of sxo_wcs (sxo.h)."""
    _fields_ = [("lng", C.c_int), ("lat", C.c_int), ("proj", C.c_int),
                ("crval", C.c_double * 2), ("crpix", C.c_double * 2), ("cd", C.c_double * 4),
                ("phip", C.c_double), ("equinox", C.c_double)]


def wcs_world(ctype, crval, crpix, cd, pix, longpole=999.0, equinox=2000.0):
    """World coordinates (n x 2, the axes' order) and J2000 right ascension and
    declination (n x 2) of 1-based pixel positions pix (n x 2) by the oracle's zenithal
    WCS; rows outside the projection's domain are NaN."""
    w = SxoWcs()
    arr = lambda v: np.ascontiguousarray(v, np.float64)
    cv, cp, m = arr(crval), arr(crpix), arr(cd).ravel()
    if lib().sxo_wcs_setup(C.byref(w), ctype[0].encode(), ctype[1].encode(), cv.ctypes.data,
                           cp.ctypes.data, m.ctypes.data, longpole, equinox) != 0:
        raise ValueError("WCS not covered by the oracle: %r" % (ctype,))
    pix = arr(pix).reshape(-1, 2)
    world = np.full_like(pix, np.nan)
    radec = np.full_like(pix, np.nan)
    wo, ro = np.zeros(2), np.zeros(2)
    for k, (x, y) in enumerate(pix):
        if lib().sxo_wcs_world(C.byref(w), x, y, wo.ctypes.data, ro.ctypes.data) == 0:
            world[k], radec[k] = wo, ro
    return world, radec


def ref_wcs_world(ctype, crval, crpix, cd, pix, longpole=999.0, latpole=999.0,
                  equinox=2000.0):
    """The same through the reference's own create_wcs/raw_to_wcs/precess (oracle/_ref);
    also returns the reference's wcsstruct pointer (kept alive for the process)."""
    arr = lambda v: np.ascontiguousarray(v, np.float64)
    cv, cp, m = arr(crval), arr(crpix), arr(cd).ravel()
    wp = ref().ref_wcs_new(ctype[0].encode(), ctype[1].encode(), cv.ctypes.data,
                           cp.ctypes.data, m.ctypes.data, longpole, latpole, equinox)
    pix = arr(pix).reshape(-1, 2)
    world, radec = np.zeros_like(pix), np.zeros_like(pix)
    if ref().ref_wcs_world(wp, len(pix), pix.ctypes.data, world.ctypes.data,
                           radec.ctypes.data) != 0:
        raise ValueError("the reference found no mapping for a position")
    return world, radec, wp


def getconv(path):
    buf = np.zeros(1024, dtype=np.float32)
    mw, mh = C.c_int(), C.c_int()
    rc = lib().sxo_getconv(path.encode(), buf.ctypes.data, 1024, C.byref(mw), C.byref(mh))
    if rc:
        raise RuntimeError("cannot read convolution mask %s (%d)" % (path, rc))
    return buf[: mw.value * mh.value].reshape(mh.value, mw.value).copy()


def convolve(img, mask):
    img = np.ascontiguousarray(img, dtype=np.float32)
    mask = np.ascontiguousarray(mask, dtype=np.float32)
    out = np.empty_like(img)
    H, W = img.shape
    lib().sxo_convolve(img.ctypes.data, W, H, mask.ctypes.data, mask.shape[1], mask.shape[0],
                       out.ctypes.data)
    return out


def background(img, bs=64, fsize=3, fthresh=0.0, back_type=0, back_value=0.0):
    """Background maps and subtracted frame (back.c restatement)."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    H, W = img.shape
    nbx, nby = (W - 1) // bs + 1, (H - 1) // bs + 1
    sub = np.empty_like(img)
    maps = [np.empty((nby, nbx), np.float32) for _ in range(4)]
    bm, bsig = C.c_float(), C.c_float()
    lib().sxo_background(img.ctypes.data, W, H, bs, fsize, fsize, C.c_double(fthresh),
                         back_type, C.c_double(back_value), sub.ctypes.data,
                         maps[0].ctypes.data, maps[1].ctypes.data, maps[2].ctypes.data,
                         maps[3].ctypes.data, C.byref(bm), C.byref(bsig))
    return dict(sub=sub, back=maps[0], sigma=maps[1], dback=maps[2], dsigma=maps[3],
                backmean=bm.value, backsig=bsig.value)
