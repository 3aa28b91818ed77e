"""GPU test of the allocation-free background entry point, through the C-ABI.

sxg_back_estimate_noalloc must give what back_estimate gives (init_device, back_estimate,
clear_device) bit for bit: the subtracted frame, backmean and backsig.  Two calls in a
row, the second on a smaller frame (the device array keeps the first call's tables), one
of them in place, plus the MANUAL types; a frame whose tables exceed the array is
rejected before any launch.  The test allocates its own device buffers with hipMalloc;
the entry point allocates nothing.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import sexgpu

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 16 * 1024 * 1024


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


class _Hip:
    """The few HIP runtime calls the test needs for its own device buffers."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def malloc(self, n):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), n) == 0
        return p.value

    def free(self, p):
        assert self.lib.hipFree(p) == 0

    def h2d(self, dst, arr):
        assert self.lib.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1) == 0

    def d2h(self, arr, src):
        assert self.lib.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2) == 0


@pytest.fixture(scope="module")
def hip():
    return _Hip()


def _lib():
    L = sexgpu.gpu()
    i, f = C.c_int, C.c_float
    L.sxg_back_estimate_noalloc.argtypes = [C.c_void_p, C.c_void_p, i, i, i, f, i, i, i, f,
                                            C.POINTER(f), C.POINTER(f), C.c_void_p]
    L.sxg_back_estimate_noalloc.restype = i
    return L


def _reference(img, bs, fs, back_type=0, back_value=0.0):
    """The existing entry point: init_device, back_estimate, clear_device."""
    p = sexgpu.Params(back_size=bs, back_fsize=fs, back_type=back_type, back_value=back_value)
    bm, bsig = sexgpu.run_background(img, p)
    return sexgpu.subtracted(img.shape), bm, bsig


def _noalloc(hip, img, bs, fs, back_type=0, back_value=0.0, in_place=False):
    L = _lib()
    H, W = img.shape
    nbx, nby = (W - 1) // bs + 1, (H - 1) // bs + 1
    src = hip.malloc(img.nbytes)
    dst = src if in_place else hip.malloc(img.nbytes)
    try:
        hip.h2d(src, img)
        bm, bsig = C.c_float(), C.c_float()
        rc = L.sxg_back_estimate_noalloc(src, dst, W, H, bs, 0.0, fs if nbx > 1 else 1,
                                         fs if nby > 1 else 1, back_type, back_value,
                                         C.byref(bm), C.byref(bsig), None)
        if rc != 0:
            raise RuntimeError((L.sxg_last_error() or b"").decode())
        out = np.empty_like(img)
        hip.d2h(out, dst)
    finally:
        hip.free(src)
        if not in_place:
            hip.free(dst)
    return out, bm.value, bsig.value


def _bits(a):
    return np.ascontiguousarray(a).view(np.int32)


def test_noalloc_equals_back_estimate_twice(hip):
    big = sexgpu.synth("plain1024", width=1536, height=1024, seed=4242)
    small = sexgpu.synth("plain1024", width=700, height=333, seed=4243)
    for name, img, bs, fs, in_place in (("1536x1024 bs64", big, 64, 3, False),
                                        ("700x333 bs32 in place", small, 32, 5, True)):
        _say("noalloc: %s" % name)
        ref, bm, bsig = _reference(img, bs, fs)
        out, bm2, bsig2 = _noalloc(hip, img, bs, fs, in_place=in_place)
        assert (bm2, bsig2) == (bm, bsig), name
        assert np.array_equal(_bits(out), _bits(ref)), name


@pytest.mark.parametrize("back_type", [1, 2])
def test_noalloc_manual_types(hip, back_type):
    img = sexgpu.synth("plain1024", width=512, height=384, seed=77)
    _say("noalloc: MANUAL back_type %d" % back_type)
    ref, bm, bsig = _reference(img, 64, 3, back_type=back_type, back_value=990.5)
    out, bm2, bsig2 = _noalloc(hip, img, 64, 3, back_type=back_type, back_value=990.5)
    assert (bm2, bsig2) == (bm, bsig)
    assert np.array_equal(_bits(out), _bits(ref))


def test_noalloc_rejects_frames_beyond_its_array(hip):
    """8192 x 8192 with BACK_SIZE 16 needs 3 * 8192 * 512 * 4 bytes of row tables alone:
    rejected by the size check, before any kernel is launched (the buffers given are
    small and never touched)."""
    L = _lib()
    p = hip.malloc(4096)
    try:
        rc = L.sxg_back_estimate_noalloc(p, p, 8192, 8192, 16, 0.0, 3, 3, 0, 0.0, None, None,
                                         None)
        assert rc == -1
        msg = L.sxg_last_error().decode()
        assert "bytes" in msg and str(SCRATCH_BYTES) in msg
    finally:
        hip.free(p)
