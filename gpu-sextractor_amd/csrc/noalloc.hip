/*
 * noalloc.hip -- the allocation-free background entry point.
 *
 * sxg_back_estimate_noalloc runs the same kernels as back_estimate (sxg_back_run:
 * mesh statistics, modes, filtering, global medians, splines, subtraction) on a frame
 * the caller already holds in device memory, and keeps every intermediate table in one
 * file-scope device array instead of the context's grow-only workspace.  It calls no
 * hipMalloc/hipFree and no allocating library routine, so that callers that cannot
 * allocate (code under stream capture, tight serving loops, harnesses that forbid it)
 * can use it.
 *
 * The array is one per GPU (each device holds its own copy at its own address, found
 * with hipGetSymbolAddress once per device, with that device current).  Every byte of
 * it that a call reads is first written by that call, so nothing is cleared between
 * calls; calls that overlap on one GPU (two streams, two host threads) share the array
 * and race.
 */
#include <mutex>

#include "sxg_internal.h"

/* the background tables of one frame: mesh statistics, filtering scratch, the four mesh
   maps, the per-row spline tables and the column tables (sxg_back_bytes) */
static __device__ __align__(256) unsigned char scratch[SXG_NOALLOC_SCRATCH_BYTES];

#define NOALLOC_MAX_DEVICES 64

/* the device address of scratch on the current device (nullptr, error recorded, when it
   cannot be found) */
static char *scratch_ptr()
{
    static std::mutex mu;
    static void *ptr[NOALLOC_MAX_DEVICES] = {};
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        sxg_set_error("sxg_back_estimate_noalloc: no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            sxg_set_error(std::string("sxg_back_estimate_noalloc: hipGetSymbolAddress: ") +
                          hipGetErrorString(e));
            return nullptr;
        }
        ptr[dev] = p;
    }
    return static_cast<char *>(ptr[dev]);
}

extern "C" int sxg_back_estimate_noalloc(const float *src, float *dst, int width, int height,
                                         int back_size, float backfthresh, int nbackfx,
                                         int nbackfy, int back_type, float back_value,
                                         float *backmean, float *backsig, void *stream)
{
    sxg_clear_error();
    if (!src || !dst || width < 1 || height < 1 || width > SXG_MAXPICSIZE ||
        height > SXG_MAXPICSIZE) {
        sxg_set_error("sxg_back_estimate_noalloc: invalid frame size or null frame pointer");
        return -1;
    }
    if (back_type < 0 || back_type > 2) {
        sxg_set_error("sxg_back_estimate_noalloc: back_type must be 0, 1 or 2");
        return -1;
    }
    BackGeom g;
    if (!sxg_back_geom(width, height, back_size, nbackfx, nbackfy, g)) return -1;
    const size_t need = sxg_back_bytes(g);
    if (need > (size_t)SXG_NOALLOC_SCRATCH_BYTES) {
        sxg_set_error("sxg_back_estimate_noalloc: the background tables of this frame need " +
                      std::to_string(need) + " bytes, more than the " +
                      std::to_string((size_t)SXG_NOALLOC_SCRATCH_BYTES) +
                      "-byte device array of this entry point (use back_estimate)");
        return -1;
    }
    char *base = scratch_ptr();
    if (!base) return -1;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const BackBufs b = sxg_back_bufs_at(base, g);
    if (!sxg_back_run(g, b, src, dst, nullptr, backfthresh, back_type, back_value, false, st))
        return -1;
    float hv[2];
    SXG_HIP_RET(hipMemcpyAsync(hv, b.outv, sizeof(hv), hipMemcpyDeviceToHost, st), -1);
    SXG_HIP_RET(hipStreamSynchronize(st), -1);
    if (backmean) *backmean = hv[0];
    if (backsig) *backsig = hv[1];
    return 0;
}
