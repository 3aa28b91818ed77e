"""Python binding of the MI355X detection path.

The product is the C-ABI library lib/libsexgpu.so (HIP kernels for gfx950 behind the
entry points the reference ``sex`` program calls: init_device, back_estimate, cudaFilter,
run_detection, init_deblend, parcel_out, clear_deblend, analyse_gpu, run_clean,
endobject_gpu, clear_detection, clear_device).  This module loads it with ctypes and
offers a numpy-level view for tests and benchmarks.  It also exposes the host helper
library lib/libsxhost.so (deterministic synthetic frames, FITS I/O).

Nothing here falls back to a CPU implementation: if the HIP library is missing or no GPU
is visible, the calls raise.
"""
import ctypes as C
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GPU_LIB = os.path.join(HERE, "lib", "libsexgpu.so")
HOST_LIB = os.path.join(HERE, "lib", "libsxhost.so")
CONFIG_DIR = os.path.join(HERE, "config")

_host = None
_gpu = None


# ------------------------------------------------------------------ host side ----
class SxsParams(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("nsrc", C.c_int),
        ("star_frac", C.c_double), ("fmin", C.c_double), ("fmax", C.c_double),
        ("bkg", C.c_double), ("grad", C.c_double), ("sigma", C.c_double),
        ("moffat_beta", C.c_double), ("fwhm", C.c_double),
        ("re_min", C.c_double), ("re_max", C.c_double), ("q_min", C.c_double),
        ("group_frac", C.c_double), ("eps", C.c_double), ("maxrad", C.c_int),
        ("seed", C.c_uint64), ("nblob", C.c_int),
        ("blob_x", C.c_double * 8), ("blob_y", C.c_double * 8),
        ("blob_sigma", C.c_double), ("blob_amp", C.c_double),
        ("nsat", C.c_int), ("sat_beta", C.c_double), ("sat_peak", C.c_double),
        ("clip_level", C.c_double), ("grp_min", C.c_int), ("grp_max", C.c_int),
    ]


def host():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError("host library not built: run `make -C gpu-sextractor_amd host`")
        _host = C.CDLL(HOST_LIB)
        _host.sxs_defaults.argtypes = [C.POINTER(SxsParams)]
        _host.sxs_render.argtypes = [C.POINTER(SxsParams), C.c_void_p, C.c_void_p, C.c_int]
        _host.sxs_weight.argtypes = [C.POINTER(SxsParams), C.c_double, C.c_double, C.c_void_p]
        _host.sxs_hash.restype = C.c_uint64
        _host.sxs_hash.argtypes = [C.c_void_p, C.c_size_t]
        _host.sxf_read_image.restype = C.POINTER(C.c_float)
        _host.sxf_read_image.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_char_p)]
        _host.sxf_write_image.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
        _host.sxf_read_wcs.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
    return _host


def read_fits_wcs(path):
    """The celestial WCS keywords of a FITS header (host/fitsio.c), or None without CTYPEs:
    dict of ctype, crval, crpix, cd, longpole, latpole, equinox."""
    ct = ((C.c_char * 9) * 2)()
    crval, crpix, cd = (C.c_double * 2)(), (C.c_double * 2)(), (C.c_double * 4)()
    lp, la, eq = C.c_double(), C.c_double(), C.c_double()
    rc = host().sxf_read_wcs(path.encode(), ct, crval, crpix, cd, C.byref(lp), C.byref(la),
                             C.byref(eq))
    if rc < 0:
        raise OSError("cannot read " + path)
    if rc == 0:
        return None
    return {"ctype": (ct[0].value.decode(), ct[1].value.decode()), "crval": list(crval),
            "crpix": list(crpix), "cd": list(cd), "longpole": lp.value, "latpole": la.value,
            "equinox": eq.value}


def synth_params(preset="plain1024", **kw):
    """Parameters of the synthetic workloads.

    plain1024   1024x1024, 3000 sources (70% Moffat stars), dN/dF~F^-2 on [1e2,1e5]
    crowded4096 4096x4096, ~167k sources (60/40), F up to 1e6, 8% in close pairs/triples
    wide8192    8192x8192, ~400k sources, large-scale background structure (3 Gaussians,
                500 px, 3 sigma); weight map from synth_weight()
    hdr4096     4096x4096 high-dynamic-range: 200 saturated Moffat stars (beta 2.5,
                peak 1e6) + 100k faint sources (1-5 sigma) + groups of 5-20 components
    """
    p = SxsParams()
    host().sxs_defaults(C.byref(p))
    if preset == "plain1024":
        pass
    elif preset == "crowded4096":
        p.width = p.height = 4096
        p.nsrc = 167000
        p.star_frac = 0.6
        p.fmax = 1e6
        p.group_frac = 0.08
        p.seed = 1001
    elif preset == "wide8192":
        p.width = p.height = 8192
        p.nsrc = 400000
        p.star_frac = 0.6
        p.fmax = 1e6
        p.group_frac = 0.08
        p.seed = 1002
        p.nblob = 3
        for i, (x, y) in enumerate([(0.25, 0.3), (0.7, 0.55), (0.4, 0.8)]):
            p.blob_x[i], p.blob_y[i] = x, y
        p.blob_sigma = 500.0
        p.blob_amp = 30.0
    elif preset == "hdr4096":
        p.width = p.height = 4096
        p.nsrc = 100000
        p.star_frac = 0.5
        # faint: peaks of about 1-5 sigma for the stars (peak ~ 0.078 F at FWHM 3, beta 3.5)
        p.fmin, p.fmax = 130.0, 650.0
        # nested groups of 5-20 components for 10% of the anchors
        p.group_frac = 0.1
        p.grp_min, p.grp_max = 5, 20
        p.seed = 1004
        p.nsat = 200
        p.sat_beta = 2.5
        p.sat_peak = 1e6
    else:
        raise ValueError("unknown preset " + preset)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def synth(preset="plain1024", **kw):
    """Render a synthetic float32 frame (deterministic for a given preset and seed)."""
    p = synth_params(preset, **kw)
    img = np.empty((p.height, p.width), dtype=np.float32)
    rc = host().sxs_render(C.byref(p), img.ctypes.data, None, 0)
    if rc:
        raise RuntimeError("sxs_render failed (%d)" % rc)
    return img


BATCH64_SIZES = [(512, 512), (1024, 1024), (1024, 2048), (2048, 2048), (2048, 4096)]


def batch64_frames(n=64, seed=2000):
    """The many-frame workload: n frames whose (height, width) are drawn from
    BATCH64_SIZES and whose source densities are log-uniform in [1e-4, 1e-2] per pixel,
    each rendered with the plain1024 generator (seed 2000 + i)."""
    rng = np.random.default_rng(seed)
    frames = []
    for i in range(n):
        h, w = BATCH64_SIZES[int(rng.integers(0, len(BATCH64_SIZES)))]
        dens = float(10 ** rng.uniform(-4, -2))
        frames.append(synth("plain1024", width=w, height=h,
                            nsrc=max(1, int(round(dens * w * h))), seed=seed + i))
    return frames


def synth_weight(preset="wide8192", vign=0.2, badcol=0.01, **kw):
    p = synth_params(preset, **kw)
    w = np.empty((p.height, p.width), dtype=np.float32)
    host().sxs_weight(C.byref(p), vign, badcol, w.ctypes.data)
    return w


def frame_hash(img):
    img = np.ascontiguousarray(img, dtype=np.float32)
    return "%016x" % host().sxs_hash(img.ctypes.data, img.size)


def read_fits(path):
    W, H = C.c_int(), C.c_int()
    gain, satur = C.c_double(0.0), C.c_double(0.0)
    err = C.c_char_p()
    ptr = host().sxf_read_image(path.encode(), C.byref(W), C.byref(H), C.byref(gain),
                                C.byref(satur), C.byref(err))
    if not ptr:
        raise RuntimeError("cannot read %s: %s" % (path, err.value))
    arr = np.ctypeslib.as_array(ptr, shape=(H.value * W.value,)).copy().reshape(H.value, W.value)
    C.CDLL(None).free(C.cast(ptr, C.c_void_p))
    return arr


def write_fits(path, arr):
    arr = np.ascontiguousarray(arr)
    bitpix = -32 if arr.dtype == np.float32 else 32
    if arr.dtype not in (np.float32, np.int32):
        raise TypeError("write_fits takes float32 or int32 arrays")
    rc = host().sxf_write_image(path.encode(), arr.ctypes.data, arr.shape[1], arr.shape[0], bitpix)
    if rc:
        raise RuntimeError("cannot write %s" % path)


# ------------------------------------------------------------------- GPU side ----
class _Lazy:
    """Wraps a CDLL so that declaring the signature of a symbol the library does not
    export is deferred until the symbol is used (AttributeError at the call)."""

    def __init__(self, lib):
        object.__setattr__(self, "_lib", lib)
        object.__setattr__(self, "_sig", {})

    def __getattr__(self, name):
        fn = getattr(self._lib, name, None)
        if fn is None:
            sig = self._sig.setdefault(name, _Sig())
            return sig
        sig = self._sig.get(name)
        if sig is not None:
            if sig.argtypes is not None:
                fn.argtypes = sig.argtypes
            if sig.restype is not None:
                fn.restype = sig.restype
        return fn

    def has(self, name):
        return hasattr(self._lib, name)


class _Sig:
    def __init__(self):
        object.__setattr__(self, "argtypes", None)
        object.__setattr__(self, "restype", None)

    def __call__(self, *a):
        raise AttributeError("libsexgpu.so does not export this entry point")


class FrameInfo(C.Structure):
    _fields_ = [("backmean", C.c_float), ("backsig", C.c_float), ("dthresh", C.c_float),
                ("thresh", C.c_float), ("npix_above", C.c_int64), ("npix_raw", C.c_int64),
                ("nraw", C.c_int32), ("nobj", C.c_int32), ("nsurv", C.c_int32),
                ("nbx", C.c_int32), ("nby", C.c_int32), ("deb_pixels", C.c_int64),
                ("deb_labels", C.c_int64), ("phot_pixels", C.c_int64),
                ("deb_class_pixels", C.c_int64 * 4)]


class Brain(C.Structure):
    """The star/galaxy network (sxg_brain: the reference's brainstruct layout,
    src/types.h:930-943)."""
    _fields_ = [("layersnb", C.c_int), ("nn", C.c_int * 3),
                ("inbias", C.c_double * 10), ("inscale", C.c_double * 10),
                ("outbias", C.c_double * 10), ("outscale", C.c_double * 10),
                ("ni", C.c_double * 10), ("no", C.c_double * 10),
                ("n", C.c_double * 10 * 3), ("w", C.c_double * 10 * 10 * 2),
                ("b", C.c_double * 10 * 2)]


class PhotomParams(C.Structure):
    _fields_ = [("apert", C.c_double), ("autoparam", C.c_double * 2),
                ("autoaper", C.c_double * 2), ("seeing_fwhm", C.c_double),
                ("mag_zeropoint", C.c_double), ("mask_correct", C.c_int), ("photo", C.c_int),
                ("do_aper", C.c_int), ("do_auto", C.c_int), ("do_class", C.c_int)]


def read_nnw(path):
    """Load a NNW weight table (STARNNW_NAME) the way the reference's getnnw reads it
    (src/neurro.c:96-180): after the "NNW" line and skipping comment and blank lines, the
    number of layers and their sizes, the input biases, the input scales, for each layer
    and neuron its input weights followed by its bias, the output biases and the output
    scales."""
    if not os.path.isabs(path) and not os.path.exists(path):
        path = os.path.join(CONFIG_DIR, path)
    with open(path) as f:
        if not f.readline().startswith("NNW"):
            raise ValueError(path + " is not a NNW table")
        rows = []
        for line in f:
            t = line.lstrip(" \t")
            if t and t[0] not in "#\n":
                rows.append([float(v) for v in t.split()])
    br = Brain()
    br.layersnb = int(rows[0][0])
    if not 2 <= br.layersnb <= 3:
        raise ValueError(path + ": unsupported number of layers")
    for l in range(br.layersnb):
        br.nn[l] = int(rows[0][1 + l])
    nin, nout = br.nn[0], br.nn[br.layersnb - 1]
    for k in range(nin):
        br.inbias[k] = rows[1][k]
        br.inscale[k] = rows[2][k]
    r = 3
    for l in range(br.layersnb - 1):
        for j in range(br.nn[l + 1]):
            for k in range(br.nn[l]):
                br.w[l][k][j] = rows[r][k]
            br.b[l][j] = rows[r][br.nn[l]]
            r += 1
    for k in range(nout):
        br.outbias[k] = rows[r][k]
        br.outscale[k] = rows[r + 1][k]
    return br


PHOT_COLUMNS = ["flux_iso", "fluxerr_iso", "mag_iso", "magerr_iso", "flux_isocor",
                "fluxerr_isocor", "flux_aper", "fluxerr_aper", "mag_aper", "magerr_aper",
                "kron_radius", "flux_auto", "fluxerr_auto", "mag_auto", "magerr_auto",
                "elongation", "ellipticity", "class_star"]


def set_photometry(apert=5.0, autoparam=(2.5, 3.5), autoaper=(0.0, 0.0), seeing_fwhm=1.2,
                   mag_zeropoint=0.0, mask_correct=1, photo=0, do_aper=1, do_auto=1,
                   do_class=1):
    """Photometry settings of endobject_gpu (defaults: the reference's default.sex)."""
    pp = PhotomParams(apert, (C.c_double * 2)(*autoparam), (C.c_double * 2)(*autoaper),
                      seeing_fwhm, mag_zeropoint, mask_correct, photo, do_aper, do_auto,
                      do_class)
    if gpu().sxg_set_photometry(C.byref(pp)) != 0:
        check()


def get_photometry(name, n):
    out = np.zeros(max(n, 1), np.float32)
    if gpu().sxg_get_photometry(PHOT_COLUMNS.index(name), out.ctypes.data, out.size) != 0:
        check()
    return out[:n]


def gpu():
    """Load lib/libsexgpu.so (raises if it is missing: there is no CPU fallback)."""
    global _gpu
    if _gpu is None:
        if not os.path.exists(GPU_LIB):
            raise RuntimeError("HIP library not built: run `make -C gpu-sextractor_amd gpu`")
        L = _Lazy(C.CDLL(GPU_LIB))
        vp, i, f, d, sz = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_size_t
        L.init_device.argtypes = [i, i, vp]
        L.back_estimate.argtypes = [i, f, C.POINTER(f), C.POINTER(f), i, i, i]
        L.init_detection.argtypes = []
        L.cudaFilter.argtypes = [vp, C.c_uint]
        L.run_detection.argtypes = [f, f, i, C.POINTER(sz), C.POINTER(sz)]
        L.init_deblend.argtypes = [f, i, sz, sz]
        L.parcel_out.argtypes = [f, i, i, sz, sz, d]
        L.parcel_out.restype = i
        L.clear_deblend.argtypes = [i]
        L.analyse_gpu.argtypes = [d, i, i, f, f, d, d, f, f, d, i, i, i, i, i, i, i, i, i, i,
                                  i, i]
        L.run_clean.argtypes = [vp, d, i, i]
        L.endobject_gpu.argtypes = [vp, vp, vp, vp, vp, i, i, i, i, f, f, d, d, d, d, i]
        L.clear_detection.argtypes = []
        L.clear_device.argtypes = [vp]
        L.sxg_last_error.restype = C.c_char_p
        L.sxg_set_back_value.argtypes = [f]
        L.sxg_get_segmap.argtypes = [vp]
        L.sxg_get_column.argtypes = [i, vp, sz]
        L.sxg_get_frame_info.argtypes = [C.POINTER(FrameInfo)]
        L.sxg_get_back_maps.argtypes = [vp, vp, vp, vp]
        L.sxg_get_filtered.argtypes = [vp]
        L.sxg_get_alloc_map.argtypes = [vp]
        L.sxg_device_frame.argtypes = [i]
        L.sxg_device_frame.restype = vp
        L.sxg_stage_times.argtypes = [vp, i]
        L.sxg_release.argtypes = []
        L.sxg_set_resident_frame.argtypes = [vp, i, i]
        L.sxg_process_resident.argtypes = [vp, vp]
        L.sxg_sync.argtypes = []
        L.sxg_kernel_timing.argtypes = [i]
        L.sxg_kernel_times.argtypes = [i, C.POINTER(d), C.POINTER(i)]
        L.sxg_set_mag_zeropoint.argtypes = [d]
        L.sxg_ctx_create.restype = vp
        L.sxg_ctx_create.argtypes = []
        L.sxg_ctx_destroy.argtypes = [vp]
        L.sxg_ctx_bind.argtypes = [vp]
        L.sxg_batch_set_frames.argtypes = [i, vp, vp, vp]
        L.sxg_batch_process.argtypes = [vp, i, vp, vp]
        L.sxg_batch_segmap.argtypes = [i, vp]
        L.sxg_batch_release.argtypes = []
        L.sxg_set_weight.argtypes = [vp, i, i, i, d, d]
        L.sxg_set_weight_map.argtypes = [vp, i, i, i, i, d, d, i]
        L.sxg_get_back_rms.argtypes = [vp]
        L.sxg_set_photometry.argtypes = [C.POINTER(PhotomParams)]
        L.sxg_get_photometry.argtypes = [i, vp, sz]
        L.sxg_set_option.argtypes = [i, C.c_long]
        L.sxg_get_option.argtypes = [i]
        L.sxg_get_option.restype = C.c_long
        L.sxg_back_estimate_noalloc.argtypes = [vp, vp, i, i, i, f, i, i, i, f, C.POINTER(f),
                                                C.POINTER(f), vp]
        L.sxg_back_estimate_noalloc.restype = i
        L.sxg_get_catalog.argtypes = [vp, i]
        L.sxg_reference_layout.argtypes = [vp, i]
        L.sxg_set_wcs.argtypes = [vp]
        L.sxg_get_world.argtypes = [vp, vp, vp, vp, sz]
        L.sxg_batch_process_host.argtypes = [vp, i, i, vp, vp, vp, vp, vp, vp, vp, vp]
        L.sxg_batch_process_graph.argtypes = [vp, i, vp, vp]
        L.sxg_batch_set_plan_margin.argtypes = [i]
        L.sxg_batch_graph_stats.argtypes = [vp, vp, vp]
        L.sxg_batch_update_frame.argtypes = [i, vp]
        L.sxg_batch_autotune.argtypes = [vp, i, i, C.POINTER(TuneChoice)]
        L.sxg_autotune_forget.argtypes = []
        _gpu = L
    return _gpu


def check():
    err = gpu().sxg_last_error()
    if err:
        raise RuntimeError("libsexgpu: " + err.decode())


EXPORTED = ["init_device", "back_estimate", "init_detection", "cudaFilter", "run_detection",
            "init_deblend", "parcel_out", "clear_deblend", "analyse_gpu", "run_clean",
            "endobject_gpu", "clear_detection", "clear_device", "sxg_last_error",
            "sxg_set_back_value", "sxg_get_segmap", "sxg_get_column", "sxg_get_frame_info",
            "sxg_get_back_maps", "sxg_get_filtered", "sxg_get_alloc_map", "sxg_device_frame",
            "sxg_stage_times", "sxg_release", "sxg_set_mag_zeropoint",
            "sxg_set_resident_frame", "sxg_process_resident", "sxg_sync", "sxg_kernel_timing",
            "sxg_kernel_times", "sxg_set_option", "sxg_get_option", "sxg_ctx_create",
            "sxg_ctx_destroy", "sxg_ctx_bind", "sxg_batch_set_frames", "sxg_batch_process",
            "sxg_batch_segmap", "sxg_batch_release", "sxg_set_weight",
            "sxg_set_photometry", "sxg_get_photometry", "sxg_set_weight_map",
            "sxg_get_back_rms", "sxg_back_estimate_noalloc", "sxg_get_catalog",
            "sxg_batch_process_host", "sxg_reference_layout", "sxg_set_wcs",
            "sxg_get_world", "sxg_batch_process_graph", "sxg_batch_set_plan_margin",
            "sxg_batch_graph_stats", "sxg_batch_update_frame", "sxg_batch_autotune",
            "sxg_autotune_forget"]
KERNELS = ["copy", "meshstat", "meshmode", "filterback", "rownodes", "subback", "conv",
           "det_init", "det_fg", "det_union", "det_flatten", "det_stat", "det_sort",
           "deb_thresh", "deb_levels", "deb_objects", "analyse", "clean", "seg", "deb_lmax", "obj_sort",
           "phot", "deb_tree0", "deb_tree1", "deb_tree2", "deb_tree_glb", "deb_lvpar"]

# options (sxg_option_id): the size caps choose kernel variants and never change a result;
# "precision" chooses the arithmetic of the graded catalogue columns
OPTIONS = {"deb_lds_small": 0, "deb_lds_large": 1, "precision": 2, "deb_lds_tiny": 3,
           "deb_table": 4, "side_streams": 5, "deb_glb_levels": 6,
           "deb_order_full": 7, "mesh_rows": 8, "fused_filter": 9}


def set_option(name, value):
    if gpu().sxg_set_option(OPTIONS[name], int(value)) != 0:
        raise ValueError("option %s: value %r out of range" % (name, value))


def get_option(name):
    return gpu().sxg_get_option(OPTIONS[name])

COLUMNS = ["level", "root", "firstpix", "flag", "fdnpix", "dnpix", "npix", "xmin", "xmax",
           "ymin", "ymax", "peakx", "peaky", "iso", "mx", "my", "mx2", "my2", "mxy",
           "poserr_mx2", "poserr_my2", "poserr_mxy", "a", "b", "theta", "cxx", "cyy", "cxy",
           "abcor", "fdflux", "dflux", "flux", "fluxerr", "fdpeak", "dpeak", "peak", "bkg",
           "thresh_obj", "mthresh", "fwhm", "master"]


def _coltype(name):
    if name in ("mx", "my", "mx2", "my2", "mxy", "poserr_mx2", "poserr_my2", "poserr_mxy"):
        return np.float64
    if name == "master":
        return np.uint32
    if name in ("level", "root", "firstpix", "flag", "fdnpix", "dnpix", "npix", "xmin", "xmax",
                "ymin", "ymax", "peakx", "peaky", "iso"):
        return np.int32
    return np.float32


def frame_info():
    fi = FrameInfo()
    gpu().sxg_get_frame_info(C.byref(fi))
    check()
    return {k: getattr(fi, k) for k, _ in FrameInfo._fields_}


def get_column(name, nobj):
    t = _coltype(name)
    n = nobj * (8 if name == "iso" else 1)
    out = np.zeros(max(n, 1), dtype=t)
    rc = gpu().sxg_get_column(COLUMNS.index(name), out.ctypes.data, out.nbytes)
    if rc:
        check()
        raise RuntimeError("sxg_get_column(%s) failed" % name)
    out = out[:n]
    return out.reshape(nobj, 8) if name == "iso" else out


class Params:
    """Configuration of the path (the .sex keys the reference front-end passes on)."""

    def __init__(self, **kw):
        self.back_size = 64
        self.back_fsize = 3
        self.back_fthresh = 0.0
        self.back_type = 0
        self.back_value = 0.0
        self.thresh_type = 0
        self.detect_thresh = 1.5
        self.analysis_thresh = 1.5
        self.ext_minarea = 5
        self.deblend_nthresh = 32
        self.deblend_mincont = 0.005
        self.clean = 1
        self.clean_param = 1.0
        self.clean_stacksize = 3000
        self.satur_level = 50000.0
        self.gain = 0.0
        self.mask = None
        for k, v in kw.items():
            if not hasattr(self, k):
                raise KeyError(k)
            setattr(self, k, v)


def run_background(img, p):
    """init_device + back_estimate; returns (backmean, backsig)."""
    L = gpu()
    img = np.ascontiguousarray(img, dtype=np.float32)
    H, W = img.shape
    L.init_device(W, H, img.ctypes.data)
    check()
    bw = min(p.back_size, W)
    bh = min(p.back_size, H)
    if bw != bh:
        raise ValueError("only square background meshes are supported (back.c:99-103)")
    nbx, nby = (W - 1) // bw + 1, (H - 1) // bh + 1
    L.sxg_set_back_value(p.back_value)
    bm, bs = C.c_float(), C.c_float()
    L.back_estimate(bw, p.back_fthresh, C.byref(bm), C.byref(bs),
                    p.back_fsize if nbx > 1 else 1, p.back_fsize if nby > 1 else 1, p.back_type)
    check()
    return bm.value, bs.value


def thresholds(p, backsig):
    """back.c:155-192: detection and analysis thresholds (double product -> float)."""
    f32 = np.float32
    if p.thresh_type == 1:
        return f32(p.detect_thresh), f32(p.analysis_thresh)
    return f32(p.detect_thresh * float(np.float64(backsig))), f32(p.analysis_thresh * float(np.float64(backsig)))


def back_maps(nbx, nby):
    maps = [np.zeros((nby, nbx), np.float32) for _ in range(4)]
    gpu().sxg_get_back_maps(*[m.ctypes.data for m in maps])
    check()
    return maps


def filtered(shape):
    out = np.zeros(shape, np.float32)
    gpu().sxg_get_filtered(out.ctypes.data)
    check()
    return out


def subtracted(shape):
    out = np.zeros(shape, np.float32)
    gpu().clear_device(out.ctypes.data)
    check()
    return out


def apply_filter(mask):
    mask = np.ascontiguousarray(mask, dtype=np.float32)
    if mask.ndim != 2 or mask.shape[0] != mask.shape[1]:
        raise ValueError("the C-ABI takes square masks (cudafilter.cu:150)")
    gpu().cudaFilter(mask.ctypes.data, mask.shape[0])
    check()


# FLAG() convention of the reference (src/define.h:283): a boolean in the first byte of
# a double/float/int argument.
FLAG_TRUE_DOUBLE = struct.unpack("<d", b"\x01" + b"\x00" * 7)[0]
FLAG_TRUE_FLOAT = struct.unpack("<f", b"\x01" + b"\x00" * 3)[0]


WEIGHT_TYPES = {"MAP_WEIGHT": 0, "MAP_RMS": 1, "MAP_VAR": 2, "BACKGROUND": 3}


def set_weight(w, p, weight_type=0, weight_gain=0):
    """Install (or with w None and a map type remove) the weighting for the next frames:
    weight_type 0 MAP_WEIGHT, 1 MAP_RMS, 2 MAP_VAR (w is the map), 3 BACKGROUND (no map);
    weight_gain: WEIGHT_GAIN."""
    L = gpu()
    if isinstance(weight_type, str):
        weight_type = WEIGHT_TYPES[weight_type]
    if w is None:
        rc = L.sxg_set_weight_map(None, 0, 0, weight_type, p.thresh_type, p.detect_thresh,
                                  p.analysis_thresh, weight_gain)
    else:
        w = np.ascontiguousarray(w, dtype=np.float32)
        rc = L.sxg_set_weight_map(w.ctypes.data, w.shape[1], w.shape[0], weight_type,
                                  p.thresh_type, p.detect_thresh, p.analysis_thresh, weight_gain)
    if rc != 0:
        check()


def back_rms(shape):
    out = np.zeros(shape, np.float32)
    if gpu().sxg_get_back_rms(out.ctypes.data) != 0:
        check()
    return out


def run_frame(img, p, outputs=("columns", "seg", "alloc", "sub", "cd", "catalog"), weight=None,
              brain=None, pixscale=1.0, ngamma=0.0, weight_type=0, weight_gain=0, prefs=None,
              flagobj2=None, wcs=None):
    """Run the whole path in the order of the reference's scan.c:78-146 and collect the
    requested outputs (numpy arrays).  weight: optional MAP_WEIGHT map; brain: the
    CLASS_STAR network (read_nnw) for the catalogue; "phot" in outputs: the photometry
    columns of endobject_gpu; prefs, flagobj2: bytearrays laid out as the reference's
    prefstruct and obj2struct (reference_layout()), passed to endobject_gpu; wcs: the
    address of a wcsstruct of the reference's layout, passed to endobject_gpu."""
    L = gpu()
    img = np.ascontiguousarray(img, dtype=np.float32)
    H, W = img.shape
    set_weight(weight, p, weight_type, weight_gain)
    bm, bsig = run_background(img, p)
    if p.back_type == 1:
        # back.c:152-153: with BACK_TYPE MANUAL the frame's backmean is BACK_VALUE
        bm = float(np.float32(p.back_value))
    dthresh, thresh = thresholds(p, bsig)
    L.init_detection()
    check()
    if p.mask is not None:
        apply_filter(p.mask)
    nraw, npix = C.c_size_t(), C.c_size_t()
    L.run_detection(dthresh, thresh, p.ext_minarea, C.byref(nraw), C.byref(npix))
    check()
    L.init_deblend(dthresh, p.deblend_nthresh, nraw, npix)
    check()
    nobj = L.parcel_out(dthresh, p.deblend_nthresh, min(p.ext_minarea, 3), nraw, npix,
                        p.deblend_mincont)
    check()
    L.clear_deblend(p.deblend_nthresh)
    bw = min(p.back_size, W)
    nbx, nby = (W - 1) // bw + 1, (H - 1) // bw + 1
    L.analyse_gpu(FLAG_TRUE_DOUBLE, 1, 1, FLAG_TRUE_FLOAT, bsig, p.gain, 0.0, thresh, dthresh,
                  p.satur_level, 0, 0, 1 if weight_gain else 0, p.ext_minarea, p.clean, 0, nbx,
                  nby, bw, bw, W, nobj)
    check()
    master = np.zeros(max(nobj, 1), np.uint32)
    L.run_clean(master.ctypes.data, p.clean_param, p.clean_stacksize, nobj)
    check()
    out = {"backmean": bm, "backsig": bsig, "dthresh": dthresh, "thresh": thresh,
           "nraw": nraw.value, "npix_raw": npix.value, "nobj": nobj, "master": master[:nobj]}
    if "columns" in outputs:
        for name in COLUMNS:
            if name != "master":
                out[name] = get_column(name, nobj)
    if "seg" in outputs:
        seg = np.zeros((H, W), np.int32)
        out["nsurv"] = L.sxg_get_segmap(seg.ctypes.data)
        check()
        out["seg"] = seg
    if "alloc" in outputs:
        al = np.zeros((H, W), np.int32)
        L.sxg_get_alloc_map(al.ctypes.data)
        check()
        out["alloc"] = al
    if "catalog" in outputs or "phot" in outputs:
        rows = (OutObj * max(nobj, 1))()
        def buf(b):
            return None if b is None else C.cast((C.c_char * len(b)).from_buffer(b), C.c_void_p)
        L.endobject_gpu(buf(prefs), C.byref(brain) if brain is not None else None,
                        buf(flagobj2), C.cast(rows, C.c_void_p), wcs, 0, H, W, H, bsig,
                        thresh, p.satur_level, pixscale, ngamma, p.gain, nobj)
        check()
        out["catalog"] = rows
        out["world"] = get_world(nobj)
        if "phot" in outputs:
            out["phot"] = {k: get_photometry(k, nobj) for k in PHOT_COLUMNS}
            out["flag_phot"] = get_column("flag", nobj)
    if "cd" in outputs:
        out["cd"] = filtered((H, W))
    info = frame_info()
    out["npix_above"] = info["npix_above"]
    L.clear_detection()
    check()
    sub = np.zeros((H, W), np.float32)
    L.clear_device(sub.ctypes.data)
    check()
    if "sub" in outputs:
        out["sub"] = sub
    return out


def run_drop_in(img, p, seg, sub, master, rows):
    """One frame through the reference's thirteen calls in scan.c order, from host memory
    (init_device ... run_clean, clear_detection, clear_device), plus the segmentation map
    and the catalogue rows: the end-to-end sequence a front-end linked against the library
    runs.  img: the float32 frame; seg (int32, frame shape), sub (float32, frame shape),
    master (uint32) and rows (ISOROW_DTYPE) are preallocated outputs (pageable or pinned).
    Returns (objects, survivors)."""
    L = gpu()
    H, W = img.shape
    L.init_device(W, H, img.ctypes.data)
    bw = min(p.back_size, W)
    nbx, nby = (W - 1) // bw + 1, (H - 1) // bw + 1
    L.sxg_set_back_value(p.back_value)
    bm, bsig = C.c_float(), C.c_float()
    L.back_estimate(bw, p.back_fthresh, C.byref(bm), C.byref(bsig),
                    p.back_fsize if nbx > 1 else 1, p.back_fsize if nby > 1 else 1, p.back_type)
    dthresh, thresh = thresholds(p, bsig.value)
    L.init_detection()
    if p.mask is not None:
        L.cudaFilter(p.mask.ctypes.data, p.mask.shape[0])
    nraw, npix = C.c_size_t(), C.c_size_t()
    L.run_detection(dthresh, thresh, p.ext_minarea, C.byref(nraw), C.byref(npix))
    L.init_deblend(dthresh, p.deblend_nthresh, nraw, npix)
    nobj = L.parcel_out(dthresh, p.deblend_nthresh, min(p.ext_minarea, 3), nraw, npix,
                        p.deblend_mincont)
    L.clear_deblend(p.deblend_nthresh)
    if nobj < 0 or nobj > master.size:
        check()
        raise RuntimeError("parcel_out: %d objects, master buffer of %d" % (nobj, master.size))
    L.analyse_gpu(FLAG_TRUE_DOUBLE, 1, 1, FLAG_TRUE_FLOAT, bsig.value, p.gain, 0.0, thresh,
                  dthresh, p.satur_level, 0, 0, 0, p.ext_minarea, p.clean, 0, nbx, nby, bw, bw, W,
                  nobj)
    L.run_clean(master.ctypes.data, p.clean_param, p.clean_stacksize, nobj)
    nsurv = L.sxg_get_segmap(seg.ctypes.data)
    L.sxg_get_catalog(rows.ctypes.data, len(rows))
    L.clear_detection()
    L.clear_device(sub.ctypes.data)
    check()
    return nobj, nsurv


REFERENCE_LAYOUT = ["prefs_size", "naper", "apert", "autoparam", "autoaper", "seeing_fwhm",
                    "mag_zeropoint", "mask_type", "detect_type", "obj2_size", "flux_aper",
                    "flux_auto", "sprob", "PHOTO", "MASK_CORRECT", "wcs_size", "wcs_naxis",
                    "wcs_ctype", "wcs_crval", "wcs_crpix", "wcs_cd", "wcs_longpole",
                    "wcs_latpole", "wcs_lat", "wcs_lng", "wcs_equinox"]


class SxgWcs(C.Structure):
    """Mirror of sxg_wcs (include/sexgpu.h)."""
    _fields_ = [("ctype", (C.c_char * 9) * 2), ("crval", C.c_double * 2),
                ("crpix", C.c_double * 2), ("cd", C.c_double * 4), ("longpole", C.c_double),
                ("latpole", C.c_double), ("equinox", C.c_double)]


def set_wcs(ctype=None, crval=(0.0, 0.0), crpix=(0.0, 0.0), cd=(1.0, 0.0, 0.0, 1.0),
            longpole=999.0, latpole=999.0, equinox=2000.0):
    """sxg_set_wcs (ctype None removes the WCS)."""
    if ctype is None:
        gpu().sxg_set_wcs(None)
        return
    w = SxgWcs()
    for k in range(2):
        w.ctype[k].value = ctype[k].encode()
        w.crval[k], w.crpix[k] = crval[k], crpix[k]
    for k in range(4):
        w.cd[k] = cd[k]
    w.longpole, w.latpole, w.equinox = longpole, latpole, equinox
    if gpu().sxg_set_wcs(C.byref(w)) != 0:
        check()


def get_world(n):
    """(X_WORLD, Y_WORLD, ALPHA_J2000, DELTA_J2000) of the last endobject_gpu call."""
    out = [np.zeros(max(n, 1)) for _ in range(4)]
    if gpu().sxg_get_world(*[a.ctypes.data for a in out], max(n, 1)) != 0:
        check()
    return [a[:n] for a in out]


def reference_layout():
    """The offsets endobject_gpu assumes in the reference's prefstruct and obj2struct."""
    out = (C.c_long * 32)()
    n = gpu().sxg_reference_layout(out, 32)
    return dict(zip(REFERENCE_LAYOUT, list(out[:n])))


def stage_times():
    ms = (C.c_float * 6)()
    gpu().sxg_stage_times(ms, 6)
    return dict(zip(["background", "filter", "detection", "deblending", "analysis", "clean"],
                    list(ms)))


class OutObj(C.Structure):
    """Catalogue row, layout of the reference's outobjstruct (src/types.h:417-455)."""
    _pack_ = 0
    _fields_ = [("number", C.c_int), ("alpha2000", C.c_double), ("delta2000", C.c_double),
                ("sposx", C.c_float), ("sposy", C.c_float), ("flux_iso", C.c_float),
                ("fluxerr_iso", C.c_float), ("mag_iso", C.c_float), ("magerr_iso", C.c_float),
                ("flux_isocor", C.c_float), ("fluxerr_isocor", C.c_float),
                ("flux_aper", C.c_float * 1), ("fluxerr_aper", C.c_float * 1),
                ("mag_aper", C.c_float * 1), ("magerr_aper", C.c_float * 1),
                ("flux_auto", C.c_float), ("fluxerr_auto", C.c_float), ("mag_auto", C.c_float),
                ("magerr_auto", C.c_float), ("theta", C.c_float), ("poserr_mx2", C.c_double),
                ("poserr_my2", C.c_double), ("poserr_mxy", C.c_double), ("flag", C.c_short),
                ("elong", C.c_float), ("ellip", C.c_float), ("sprob", C.c_float),
                ("peak", C.c_float), ("kronfactor", C.c_float), ("bkg", C.c_float),
                ("fwhm", C.c_float), ("_pad", C.c_byte * 4)]


class IsoRow(C.Structure):
    """Mirror of sxg_isorow (include/sexgpu.h): one catalogue row per survivor."""
    _fields_ = [("number", C.c_int32), ("flags", C.c_int32), ("isoarea", C.c_int32),
                ("fdnpix", C.c_int32), ("x", C.c_double), ("y", C.c_double),
                ("x2", C.c_double), ("y2", C.c_double), ("xy", C.c_double),
                ("a", C.c_float), ("b", C.c_float), ("theta", C.c_float),
                ("cxx", C.c_float), ("cyy", C.c_float), ("cxy", C.c_float),
                ("flux_iso", C.c_float), ("fluxerr_iso", C.c_float), ("peak", C.c_float),
                ("bkg", C.c_float)]


ISOROW_DTYPE = np.dtype([(name, np.dtype(t)) for name, t in
                         [("number", "<i4"), ("flags", "<i4"), ("isoarea", "<i4"),
                          ("fdnpix", "<i4"), ("x", "<f8"), ("y", "<f8"), ("x2", "<f8"),
                          ("y2", "<f8"), ("xy", "<f8"), ("a", "<f4"), ("b", "<f4"),
                          ("theta", "<f4"), ("cxx", "<f4"), ("cyy", "<f4"), ("cxy", "<f4"),
                          ("flux_iso", "<f4"), ("fluxerr_iso", "<f4"), ("peak", "<f4"),
                          ("bkg", "<f4")]])


def get_catalog(cap):
    """The catalogue rows of the last frame (after run_clean) as a numpy record array."""
    out = np.zeros(max(cap, 1), ISOROW_DTYPE)
    n = gpu().sxg_get_catalog(out.ctypes.data, cap)
    if n < 0:
        check()
    return out[:min(n, cap)]


_hip = None


def _hiprt():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
        _hip.hipHostFree.argtypes = [C.c_void_p]
    return _hip


class PinnedArray:
    """A numpy array over page-locked host memory (hipHostMalloc), freed with the object:
    copies between it and the device run asynchronously."""

    def __init__(self, shape, dtype):
        dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * dtype.itemsize
        ptr = C.c_void_p()
        if _hiprt().hipHostMalloc(C.byref(ptr), max(n, 1), 0) != 0:
            raise MemoryError("hipHostMalloc of %d bytes failed" % n)
        self.ptr = ptr.value
        buf = (C.c_byte * max(n, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    def __del__(self):
        if getattr(self, "ptr", None):
            _hiprt().hipHostFree(self.ptr)
            self.ptr = None


def process_batch_host(frames, p, nworkers=4, segs=None, rows=None, row_cap=None):
    """The batched path end to end from host memory (sxg_batch_process_host): frames are
    float32 arrays (pinned ones overlap the copies with other workers' kernels); segs and
    rows, when given, are the output arrays (int32 of the frame's shape, ISOROW_DTYPE of
    row_cap[i] rows).  Returns (infos, frame_ms, segs, rows)."""
    L = gpu()
    n = len(frames)
    if segs is None:
        segs = [np.zeros(f.shape, np.int32) for f in frames]
    if rows is None:
        row_cap = [max(1, f.size // 64) for f in frames]
        rows = [np.zeros(c, ISOROW_DTYPE) for c in row_cap]
    ptrs = (C.c_void_p * n)(*[f.ctypes.data for f in frames])
    ws = (C.c_int * n)(*[f.shape[1] for f in frames])
    hs = (C.c_int * n)(*[f.shape[0] for f in frames])
    sp = (C.c_void_p * n)(*[s.ctypes.data for s in segs])
    rp = (C.c_void_p * n)(*[r.ctypes.data for r in rows])
    caps = (C.c_int * n)(*[len(r) for r in rows])
    params = native_params(p)
    infos = (FrameInfo * n)()
    ms = (C.c_double * n)()
    if L.sxg_batch_process_host(C.byref(params), nworkers, n, ptrs, ws, hs, sp, rp, caps, infos,
                                ms) < 0:
        check()
    rows = [r[:min(fi.nsurv, len(r))] for r, fi in zip(rows, infos)]
    return list(infos), list(ms), segs, rows


class SxgParams(C.Structure):
    """Mirror of sxg_params (include/sexgpu.h)."""
    _fields_ = [("back_size", C.c_int), ("back_fsize", C.c_int), ("back_fthresh", C.c_float),
                ("back_type", C.c_int), ("back_value", C.c_float), ("thresh_type", C.c_int),
                ("detect_thresh", C.c_double), ("analysis_thresh", C.c_double),
                ("ext_minarea", C.c_int), ("deblend_nthresh", C.c_int),
                ("deblend_mincont", C.c_double), ("clean", C.c_int),
                ("clean_param", C.c_double), ("clean_stacksize", C.c_int),
                ("satur_level", C.c_double), ("gain", C.c_double),
                ("mask", C.c_void_p), ("mask_width", C.c_int)]


def native_params(p):
    """sxg_params for a Params object (keeps the mask array alive on the result)."""
    sp = SxgParams()
    for name, _ in SxgParams._fields_:
        if name in ("mask", "mask_width"):
            continue
        setattr(sp, name, getattr(p, name))
    if p.mask is not None:
        m = np.ascontiguousarray(p.mask, dtype=np.float32)
        sp._mask = m
        sp.mask = m.ctypes.data
        sp.mask_width = m.shape[0]
    return sp


def read_conv(name):
    """Parse a .conv filter file the way src/filter.c:101-186 does (float elements,
    NORM divides by the double sum of absolute values)."""
    path = name if os.path.sep in name else os.path.join(CONFIG_DIR, name)
    with open(path) as f:
        lines = f.readlines()
    if not lines or not lines[0].startswith("CONV"):
        raise ValueError("not a convolution mask: " + path)
    norm = ("NORM" in lines[0] and "NONORM" not in lines[0]) or ("NORM" not in lines[0])
    rows, width = [], 0
    for ln in lines[1:]:
        tok = ln.split()
        if not tok or tok[0].startswith("#"):
            continue
        vals = [np.float32(float(t)) for t in tok]
        width = max(width, len(vals))
        rows.extend(vals)
    m = np.array(rows, np.float32)
    if width < 1 or m.size % width:
        raise ValueError("bad mask line in " + path)
    if norm:
        s = float(np.sum(np.abs(m.astype(np.float64))))
        if s == 0.0:
            s = float(np.sqrt(np.sum(m.astype(np.float64) ** 2)))
        m = (m.astype(np.float64) / s).astype(np.float32)
    return m.reshape(-1, width)


def process_resident(img, p):
    """Upload img once and run the whole path natively from device memory."""
    L = gpu()
    img = np.ascontiguousarray(img, dtype=np.float32)
    L.sxg_set_resident_frame(img.ctypes.data, img.shape[1], img.shape[0])
    check()
    sp = native_params(p)
    fi = FrameInfo()
    n = L.sxg_process_resident(C.byref(sp), C.byref(fi))
    check()
    return n, {k: getattr(fi, k) for k, _ in FrameInfo._fields_}


def set_batch_frames(frames):
    """Copy float32 frames (uneven sizes allowed) into device memory as the resident
    batch (sxg_batch_set_frames)."""
    L = gpu()
    frames = [np.ascontiguousarray(f, dtype=np.float32) for f in frames]
    n = len(frames)
    ptrs = (C.c_void_p * n)(*[f.ctypes.data for f in frames])
    ws = (C.c_int * n)(*[f.shape[1] for f in frames])
    hs = (C.c_int * n)(*[f.shape[0] for f in frames])
    if L.sxg_batch_set_frames(n, ptrs, ws, hs) != 0:
        check()
    return frames


def batch_segmaps(shapes):
    """The segmentation maps of the resident batch's frames (their shapes given)."""
    L = gpu()
    segs = []
    for i, shp in enumerate(shapes):
        s = np.zeros(shp, np.int32)
        if L.sxg_batch_segmap(i, s.ctypes.data) != 0:
            check()
        segs.append(s)
    return segs


def process_batch(frames, p, nworkers=4, segmaps=True):
    """Run the whole path on a list of float32 frames (uneven sizes allowed) with
    nworkers concurrent contexts; returns (infos, frame_ms, seg maps or None)."""
    L = gpu()
    frames = set_batch_frames(frames)
    n = len(frames)
    sp = native_params(p)
    infos = (FrameInfo * n)()
    ms = (C.c_double * n)()
    if L.sxg_batch_process(C.byref(sp), nworkers, infos, ms) < 0:
        check()
    segs = batch_segmaps([f.shape for f in frames]) if segmaps else None
    return list(infos), list(ms), segs


def process_batch_graph(n, shapes, p, nworkers=4, segmaps=True):
    """One replay of the resident batch's graph (sxg_batch_process_graph; the frames set
    with set_batch_frames); returns (infos, frame_ms, seg maps or None, frames run again
    eagerly)."""
    L = gpu()
    sp = native_params(p)
    infos = (FrameInfo * n)()
    ms = (C.c_double * n)()
    rerun = L.sxg_batch_process_graph(C.byref(sp), nworkers, infos, ms)
    if rerun < 0:
        check()
    segs = batch_segmaps(shapes) if segmaps else None
    return list(infos), list(ms), segs, rerun


class TuneChoice(C.Structure):
    """sxg_tune_choice"""
    _fields_ = [("branches", C.c_int), ("deb_lds_tiny", C.c_long), ("deb_lds_small", C.c_long),
                ("deb_lds_large", C.c_long), ("ms_per_batch", C.c_double),
                ("candidates", C.c_int), ("size_bucket", C.c_int), ("density_bucket", C.c_int),
                ("count_bucket", C.c_int)]


def autotune(p, max_branches=8, replays=2):
    """sxg_batch_autotune on the resident batch: the chosen graph branches and deblending
    size classes (set as the options), as a dict."""
    ch = TuneChoice()
    if gpu().sxg_batch_autotune(C.byref(native_params(p)), max_branches, replays,
                                C.byref(ch)) != 0:
        check()
    return {k: getattr(ch, k) for k, _ in TuneChoice._fields_}


def graph_stats():
    """(builds, launches, eager reruns) of the graph batch."""
    b, l, r = C.c_long(), C.c_long(), C.c_long()
    gpu().sxg_batch_graph_stats(C.byref(b), C.byref(l), C.byref(r))
    return b.value, l.value, r.value
