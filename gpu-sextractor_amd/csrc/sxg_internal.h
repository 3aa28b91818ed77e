/*
 * sxg_internal.h -- shared runtime state, error recording and device helpers of the
 * MI355X detection path.  Every kernel is compiled for gfx950 with FP contraction off
 * (see the Makefile) so that FP32/FP64 expressions round exactly as written.
 */
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>

#include "sexgpu.h"
#include "sx_cmath.h"

#define SXG_BIG 1e+30
#define SXG_PI 3.1415926535898          /* src/define.h:56 */
#define SXG_OBJ_MERGED 0x0002
#define SXG_OBJ_SATUR 0x0004
#define SXG_OBJ_TRUNC 0x0008
#define SXG_NISO 8
#define SXG_MAXPICSIZE 1048576
#define SXG_NMAXLEVELS 4096              /* QUANTIF_NMAXLEVELS, src/back.h:23 */
#define SXG_MAXNTHRESH 64                /* DEBLEND_NTHRESH upper bound, preflist.h:79 */
#define SXG_MAXMASK 31                   /* largest convolution mask width handled */
#define SXG_MAXBACKF 11                  /* BACK_FILTERSIZE upper bound, preflist.h:48 */

/* ---------------------------------------------------------------- errors ---- */
void sxg_set_error(const std::string &msg);
void sxg_clear_error();
bool sxg_has_error();

#define SXG_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) {                                                        \
            sxg_set_error(std::string(#call) + ": " + hipGetErrorString(_e) + " (" +   \
                          __FILE__ + ":" + std::to_string(__LINE__) + ")");            \
            return;                                                                    \
        }                                                                              \
    } while (0)
#define SXG_HIP_RET(call, ret)                                                         \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) {                                                        \
            sxg_set_error(std::string(#call) + ": " + hipGetErrorString(_e) + " (" +   \
                          __FILE__ + ":" + std::to_string(__LINE__) + ")");            \
            return ret;                                                                \
        }                                                                              \
    } while (0)

/* ------------------------------------------------------------ workspace ---- */
/* A grow-only device buffer: reallocated only when a larger size is requested. */
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool ensure(size_t want);   /* false (error recorded) on allocation failure */
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

/* MAP_WEIGHT: var = 1/w for w > 0, else BIG (weight.c:111-116, sigfac = 1 as weight.c:80
   sets it); with relative thresholds (rel) each pixel's thresholds are kdet*sqrt(var)
   and kana*sqrt(var) */
/* w: the weight map as given (MAP_WEIGHT weights, MAP_RMS rms, MAP_VAR variances; for
   WEIGHT_TYPE BACKGROUND the frame's interpolated background rms), or nullptr;
   mask: the map the mesh statistics read, positive where a pixel takes part */
struct WeightArgs {
    const float *w = nullptr;
    const float *mask = nullptr;
    int type = SXG_WEIGHT_MAP_WEIGHT;
    int rel = 0;
    double kdet = 0.0, kana = 0.0;
};
/* the calibrated variance of pixel p (weight_to_var, weight.c:93-126, sigfac 1) */
__device__ inline float sxg_var(const WeightArgs &wa, long p)
{
    const float v = wa.w[p];
    if (wa.type == SXG_WEIGHT_MAP_WEIGHT) return v > 0.0f ? 1.0f / v : (float)SXG_BIG;
    if (wa.type == SXG_WEIGHT_MAP_VAR) return v;
    return v < SXG_BIG ? v * v : v;      /* MAP_RMS and the background rms */
}
__device__ inline float sxg_pix_thresh(double k, float var) { return (float)(k * sqrt((double)var)); }
/* the detection threshold of pixel p: per pixel with a weight map and relative
   thresholds, else the global one */
__device__ inline float sxg_det_thresh(const WeightArgs &wa, float t, long p)
{
    return (wa.w && wa.rel) ? sxg_pix_thresh(wa.kdet, sxg_var(wa, p)) : t;
}
/* the background spline at column x of a frame row: the row's nodes and their second
   derivatives (node, dnode; nbx of each) and the column tables of the cumulative dx
   recurrence (subbackline, back.c:880-908) */
__device__ inline float sxg_spline_at(const float *node, const float *dnode, const int *colseg,
                                      const float *coldx, int nbx, int x)
{
    if (nbx > 1) {
        const int s = colseg[x];
        const float dx = coldx[x];
        const float cdx = 1 - dx;
        return cdx * (node[s] + (cdx * cdx - 1) * dnode[s]) +
               dx * (node[s + 1] + (dx * dx - 1) * dnode[s + 1]);
    }
    return node[0];
}

enum StageId { ST_BACK = 0, ST_FILTER, ST_DETECT, ST_DEBLEND, ST_ANALYSE, ST_CLEAN, ST_N };

/* ------------------------------------------------------------- background ---- */
/* geometry of one background estimate: frame, square mesh size, mesh counts and the
   median-filter size along each axis */
struct BackGeom {
    int W = 0, H = 0, bs = 0, nbx = 0, nby = 0, nfx = 1, nfy = 1;
};
struct MeshStat;
/* device buffers of one background estimate (back.hip lays them out in one block) */
struct BackBufs {
    MeshStat *ms = nullptr;          /* per-mesh statistics */
    float *work = nullptr;           /* 4*n2 floats of filtering and selection scratch */
    float *outv = nullptr;           /* the two global medians: backmean, backsig */
    float *back = nullptr, *sigma = nullptr, *dback = nullptr, *dsigma = nullptr;   /* maps */
    float *nodes = nullptr, *dnodes = nullptr, *utmp = nullptr;   /* per-row spline tables */
    int *colseg = nullptr;           /* per-column left node */
    float *coldx = nullptr;          /* per-column dx of the cumulative recurrence */
};
bool sxg_back_geom(int W, int H, int back_size, int nbackfx, int nbackfy, BackGeom &g);
bool sxg_back_run(const BackGeom &g, const BackBufs &b, const float *src, float *dst,
                  const float *wmap, float backfthresh, int back_type, float back_value,
                  bool timed, hipStream_t st);
void sxg_back_estimate_from(const float *src, int back_size, float backfthresh, float *backmean,
                            float *backsig, int nbackfx, int nbackfy, int back_type);
/* whole-frame runs: subtraction of the background spline (tables of the last
   back_estimate), the K x K convolution and the detection parents of k_det_init in one
   pass over the input frame (filter.hip); false when the frame or mask does not suit it
   (nothing queued) */
bool sxg_subconv_detect(const float *h_mask, int K, float dthresh);
/* the background map of the last back_estimate on the current context (device) */
const float *sxg_back_map();
/* the device block of one background estimate of geometry g: its size, and its parts
   laid out from base */
size_t sxg_back_bytes(const BackGeom &g);
BackBufs sxg_back_bufs_at(char *base, const BackGeom &g);

/* ------------------------------------------------------ count-planned runs ---- */
/* A planned run queues a whole frame without reading a count back to the host: every
   count that sizes a launch or a buffer is replaced by a capacity fixed beforehand (the
   counts of an eager run of a frame of the same kind, with a margin), the kernels read
   the actual count on the device and stop there, and a guard after each count flags one
   above its capacity.  A flagged frame's counts are zeroed at once, so every later kernel
   of the frame finds no work and stays inside its buffers; the caller runs that frame
   again eagerly.  Such a run neither waits for the device nor allocates, so a stream
   capture can record it into a HIP graph. */
enum PlanCap {
    PL_NFG = 0,   /* detected pixels */
    PL_NRAW,      /* raw objects */
    PL_NPIX,      /* pixels of raw objects */
    PL_NTOT,      /* node capacity of the deblending trees */
    PL_C0, PL_C1, PL_C2, PL_C3,     /* raw objects per deblending size class */
    PL_BIG,       /* class-3 objects built level-parallel */
    PL_NPL,       /* pixel-levels of those */
    PL_NREC,      /* node records of those */
    PL_LDS0, PL_LDS1, PL_LDS2,      /* LDS bytes of the tree workgroups of classes 0..2 */
    PL_NOBJ,      /* deblended objects */
    PL_N
};
struct FramePlan {
    int64_t cap[PL_N] = {};
};
/* the device block of a planned run (one per context) */
struct PlanDev {
    int nfg, nraw, npix, nobj, nsurv;
    int ovf;          /* bit k: the count of PlanCap k exceeded its capacity; bit 30: the
                         detection threshold is not positive; bit 29: deblending levels
                         inconsistent */
    int pad[2];
    float dthresh, thresh, backmean, backsig;
    /* deblending work (sxg_frame_info's deb_pixels, deb_labels, deb_class_pixels) */
    long long deb_pixels, deb_labels, class_pixels[4];
};
#define SXG_PLAN_BADTHRESH 30
#define SXG_PLAN_DEBLEVELS 29
/* n, or the device count *nd when a planned run passes one */
__device__ inline int sxg_dn(const int *nd, int n) { return nd ? *nd : n; }

#define SXG_NSIDE 4
struct SxgState {
    int W = 0, H = 0;
    size_t N = 0;
    bool have_frame = false;
    /* planned run (nullptr: eager) and its device block; capturing: a stream capture is
       recording this context's launches (allocation is an error then) */
    const FramePlan *plan = nullptr;
    PlanDev *pdev = nullptr;
    bool capturing = false;
    uint64_t alloc_gen = 0;     /* bumped by every reallocation of a device buffer */
    FramePlan seen;             /* the counts of the last eager run (capacities of a plan) */
    hipStream_t stream = nullptr;
    hipEvent_t ev[2 * ST_N] = {};
    /* side streams (default priority) for independent launches inside a stage, and the
       events that join them back into stream */
    hipStream_t side[SXG_NSIDE] = {};
    hipEvent_t side_ev[SXG_NSIDE] = {};
    bool ev_used[ST_N] = {};
    float back_value = 0.0f;

    /* frame buffers */
    DevBuf img;      /* float[N]: input, background-subtracted in place */
    DevBuf cd;       /* float[N]: filtered (detection) frame */
    DevBuf aux;      /* int[N]: union-find parents, then the detection labels */

    /* background */
    int bs = 0, nbx = 0, nby = 0;
    int back_W = 0, back_H = 0;               /* frame size the mesh maps belong to */
    BackGeom bgeom;                           /* geometry of the last background */
    float backmean = 0.0f, backsig = 0.0f;
    DevBuf backbuf;  /* BackBufs of the last background (mesh maps, spline tables) */

    /* detection */
    float dthresh = 0.0f, thresh = 0.0f;
    int ext_minarea = 5;
    int64_t npix_above = 0, npix_raw = 0;
    int64_t deb_pixels = 0, deb_labels = 0;   /* level-by-level deblending work */
    int64_t deb_class_pixels[4] = {};         /* pixels per deblending size class */
    int64_t phot_pixels = 0;                  /* aperture pixels of the photometry */
    int nraw = 0;
    bool filtered = false;  /* cudaFilter ran on this frame */
    /* whole-frame runs: the background stage leaves the subtraction to the fused
       subtraction, convolution and thresholding kernel (sub_src: the input frame), which
       also writes the union-find parents (det_init_done) */
    bool defer_sub = false, det_init_done = false;
    const float *sub_src = nullptr;
    int *fpia = nullptr;      /* int[P0]: pixels grouped by raw object, raster order inside */
    int *fpia_keys = nullptr; /* int[P0]: raw object id of each fpia entry */
    DevBuf scan_tmp; /* scan scratch */

    /* deblending and objects */
    int nthresh = 0;
    int nobj = 0;
    DevBuf thr;      /* float[nraw * nthresh] sub-thresholds */
    DevBuf lvl;      /* int[nthresh * P0] component of each raw pixel per level */
    DevBuf comps;    /* component table */
    DevBuf objs;     /* object columns (SoA) */
    DevBuf rootobj;  /* per root: first object, object count */
    int *opix = nullptr;      /* int[P0]: raw pixels grouped by object, raster order inside */
    int *objoff = nullptr;    /* int[nobj + 1]: start of each object in opix */
    int nsurv = 0;
    bool cleaned = false;
    bool clean_enabled = true;
    float deb_base = 0.0f;
    double mag_zeropoint = 0.0;
    sxg_photom_params phot = {5.0, {2.5, 3.5}, {0.0, 0.0}, 1.2, 0.0, 1, 0, 1, 1, 1};
    bool phot_ready = false;               /* endobject_gpu ran on this frame */
    /* world coordinates: the WCS of sxg_set_wcs (a WcsDev, wcs.hip) and whether the
       "wcs.*" columns hold the current frame's */
    bool wcs_set = false, wcs_ready = false;
    alignas(8) unsigned char wcs_dev[128] = {};
    float stage_ms[ST_N] = {};

    /* per-context workspace */
    std::map<std::string, DevBuf> named;   /* named device buffers (sxg_take) */
    bool deb_ready = false;                /* init_deblend ran for this frame */
    DevBuf dmask;                          /* convolution mask */
    float mask_host[SXG_MAXMASK * SXG_MAXMASK] = {};   /* the mask dmask holds */
    int mask_k = 0;
    const void *mask_buf = nullptr;
    DevBuf resident;                       /* frame kept in HBM (sxg_set_resident_frame) */
    int rw = 0, rh = 0;
    bool is_default = false;               /* the process-wide default context */
    bool single_stream = false;            /* no side streams (batch workers) */
    /* MAP_WEIGHT map (sxg_set_weight) */
    DevBuf weight;
    int ww = 0, wh = 0;
    bool weighted = false;
    int wtype = SXG_WEIGHT_MAP_WEIGHT;     /* sxg_weight_type */
    bool wgain = false;                    /* WEIGHT_GAIN */
    DevBuf wmask;                          /* MAP_RMS / MAP_VAR: mesh statistics mask */
    bool wrms_frame = false;               /* BACKGROUND: the rms map of this frame is built */
    WeightArgs wargs;                      /* w filled in by sxg_weight_args() */
};
/* the weight arguments for the current frame (w = nullptr without a weight map); records
   an error and returns false when the map does not match the frame */
bool sxg_weight_args(WeightArgs &wa);

/* raw (level-0) objects, structure of arrays; t0: the object's detection threshold
   (the lowest per-pixel threshold with a weight map, else the global one) */
struct RawObj {
    int *first, *count, *off, *xmin, *xmax, *ymin, *ymax;
    float *fdpeak, *fdflux, *t0;
};


bool sxg_raw_alloc(int nraw, RawObj &ro);   /* (re)binds the named buffers */

/* kernel timers (HIP events on the library stream, enabled by sxg_kernel_timing) */

long sxg_option(int id);
bool sxg_precision_columns(int mode);   /* precision.hip */
void sxg_kt_begin(int id);
void sxg_kt_end(int id);
void sxg_kt_begin_on(int id, hipStream_t st);   /* the same, on a given stream */
void sxg_kt_end_on(int id, hipStream_t st);
bool sxg_build_segmap(int *dst = nullptr); /* device segmentation map (into dst, else the
                                              buffer "out.map") and survivor count */
/* queue the catalogue rows of the survivors and their copy to host memory (after the
   segmentation map); the number of rows copied, or -1 */
int sxg_catalog_async(sxg_isorow *out, int cap);
/* world coordinates of the current frame's objects into the "wcs.*" columns (wcs.hip) */
bool sxg_astrometry(const void *wcs);
int sxg_wcs_layout(long *out, int n);

SxgState &sxg_state();
/* the state bound to the calling thread (nullptr: the default context), and rebinding it */
SxgState *sxg_bound_state();
void sxg_bind_state(SxgState *s);
/* End of an entry point without a host output: the contexts of the batch workers
   (several sharing the device) wait for their stream there, which interleaves their
   kernels on the hardware queues; the default context returns at once and keeps its
   queue full. */
bool sxg_stage_wait();
void sxg_stage_begin(StageId s);
void sxg_stage_end(StageId s);

/* ---------------------------------------------------------- primitives ---- */
/* Exclusive prefix sum of n int32 values (out may alias in); *total (device) receives
   the sum when total != nullptr. */
/* With nd (planned runs) n is a capacity and the number of values is *nd (<= n), read
   on the device; the launches are sized for n. */
void sxg_scan_excl(const int *in, int *out, int n, int *total, hipStream_t s,
                   const int *nd = nullptr);
/* Stream compaction of the indices i in [0,n) with flag[i] != 0, in increasing order;
   *count (device) receives the number kept. */
void sxg_compact_index(const int *flag, int *out, int n, int *count, hipStream_t s,
                       const int *nd = nullptr);
/* Same, keeping the indices i with v[i] >= 0. */
/* the set bits of nw 64-bit window masks as positions 64 w + b, in increasing order */
void sxg_compact_mask(const unsigned long long *m, long nw, int *out, int *count,
                      hipStream_t s);
void sxg_compact_nonneg(const int *v, int *out, int n, int *count, hipStream_t s,
                        const int *nd = nullptr);
/* Stable LSD radix sort of (key, value) pairs on the low nbits of non-negative keys;
   returns 0 when the result is in (k0, v0), 1 when in (k1, v1). */
int sxg_radix_sort(int *k0, int *v0, int *k1, int *v1, int n, int nbits, int *hist,
                   hipStream_t s, const int *nd = nullptr);
/* Planned runs: queue the check of the device count *v against cap (bit: its PlanCap);
   above it, or after an earlier overflow of the frame, the frame's counts are zeroed
   (also the deblending class counts ccount[0..3] and [8] when ccount is given). */
void sxg_plan_guard(int *v, int64_t cap, int bit, int *ccount = nullptr);
/* the capacity of count k in a planned run, else n itself */
inline int64_t sxg_plan_cap(const SxgState &S, int k, int64_t n)
{
    return S.plan ? S.plan->cap[k] : n;
}
/* Grow-only device buffers addressed by name (detection/deblending scratch). */
DevBuf &sxg_buf(const char *name);
void sxg_release_named();
bool sxg_ensure_stream();   /* the current context's stream and events */
/* batch workers: every launch of the context on its one stream */
void sxg_ctx_single_stream(sxg_ctx *c);
SxgState &sxg_ctx_state(sxg_ctx *c);   /* the state of a context */
/* Typed view of a named buffer holding at least n elements (nullptr on failure). */
template <class T> inline T *sxg_take(const char *name, size_t n)
{
    DevBuf &b = sxg_buf(name);
    if (!b.ensure(sizeof(T) * (n + 16))) return nullptr;
    return b.as<T>();
}
/* FP32 bits of non-negative floats order like the integers. */
/* order-preserving map float -> int (monotone over all non-NaN floats) and its inverse */
__device__ inline int f2ord(float v)
{
    const int i = __float_as_int(v);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ inline float ord2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

static inline int sxg_div_up(long a, long b) { return (int)((a + b - 1) / b); }
