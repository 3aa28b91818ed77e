/*
 * sxg_state.hip -- library state, errors, workspace, stage timing, scan/compaction
 * primitives, and the frame entry points init_device / clear_device (the reference's
 * src/cuda/cudainit.cu:16 and :160).
 */
#include <cstdio>
#include <cstring>
#include <mutex>

#include "sxg_internal.h"

/* ---------------------------------------------------------------- errors ---- */
/* per calling thread: contexts may run on several host threads at once */
static thread_local std::string g_err;
static thread_local bool g_has_err = false;

void sxg_set_error(const std::string &msg)
{
    if (!g_has_err) g_err = msg;   /* keep the first failure of a call sequence */
    g_has_err = true;
}
void sxg_clear_error() { g_has_err = false; g_err.clear(); }
bool sxg_has_error() { return g_has_err; }

extern "C" const char *sxg_last_error(void) { return g_has_err ? g_err.c_str() : nullptr; }

/* ------------------------------------------------------------------ options ---- */
static long g_opt[SXG_NOPT] = {1024, 6144, 0, 512, 0, 1, 12288, 4096, 0, 1};
static const long g_opt_max[SXG_NOPT] = {6400, 6400, 2, 6400, 2048, 1, 1L << 30, 1L << 30, 8, 1};

long sxg_option(int id) { return g_opt[id]; }

extern "C" int sxg_set_option(int id, long value)
{
    if (id < 0 || id >= SXG_NOPT || value < 0 || value > g_opt_max[id]) return -1;
    g_opt[id] = value;
    return 0;
}

extern "C" long sxg_get_option(int id) { return id >= 0 && id < SXG_NOPT ? g_opt[id] : -1; }

/* ------------------------------------------------------------ workspace ---- */
bool DevBuf::ensure(size_t want)
{
    if (want <= bytes && ptr) return true;
    {
        SxgState &S = sxg_state();
        /* a graph records the buffers' addresses: a planned run that is being captured
           must find every buffer at its size already */
        if (S.capturing) {
            sxg_set_error("device buffer of " + std::to_string(want) +
                          " bytes requested while a graph is captured (plan not warmed up)");
            return false;
        }
        S.alloc_gen++;
    }
    /* grow by half again at least, so that buffers sized by per-frame counts settle
       after a few frames */
    const size_t grown = ptr ? bytes + bytes / 2 : 0;
    size_t alloc = want < 256 ? 256 : want;
    if (alloc < grown) alloc = grown;
    alloc = (alloc + 255) & ~(size_t)255;
    /* queued kernels of the context may still use the old buffer (the entry points do
       not wait for their own kernels before returning): wait for its streams, not for
       the whole device, so other contexts keep running */
    if (ptr) {
        SxgState &S = sxg_state();
        if (S.stream) {
            (void)hipStreamSynchronize(S.stream);
            for (int i = 0; i < SXG_NSIDE; i++)
                if (S.side[i]) (void)hipStreamSynchronize(S.side[i]);
        } else {
            (void)hipDeviceSynchronize();
        }
        (void)hipFree(ptr);
    }
    ptr = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&ptr, alloc);
    if (e != hipSuccess) {
        ptr = nullptr;
        sxg_set_error(std::string("device allocation of ") + std::to_string(alloc) +
                      " bytes failed: " + hipGetErrorString(e));
        return false;
    }
    bytes = alloc;
    return true;
}

void DevBuf::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

/* the context bound to the calling thread (sxg_ctx_bind), else the default one */
static thread_local SxgState *tl_ctx = nullptr;

static SxgState &default_state()
{
    static SxgState st;
    st.is_default = true;
    return st;
}

SxgState &sxg_state() { return tl_ctx ? *tl_ctx : default_state(); }
SxgState *sxg_bound_state() { return tl_ctx; }
void sxg_bind_state(SxgState *s) { tl_ctx = s; }

struct sxg_ctx {
    SxgState st;
};

extern "C" sxg_ctx *sxg_ctx_create(void)
{
    sxg_clear_error();
    return new (std::nothrow) sxg_ctx();
}

extern "C" void sxg_ctx_destroy(sxg_ctx *c)
{
    if (!c) return;
    SxgState *prev = tl_ctx;
    tl_ctx = &c->st;
    sxg_release();
    if (c->st.stream) {
        for (int i = 0; i < 2 * ST_N; i++)
            if (c->st.ev[i]) (void)hipEventDestroy(c->st.ev[i]);
        for (int i = 0; i < SXG_NSIDE; i++) {
            if (c->st.side_ev[i]) (void)hipEventDestroy(c->st.side_ev[i]);
            if (c->st.side[i]) (void)hipStreamDestroy(c->st.side[i]);
        }
        (void)hipStreamDestroy(c->st.stream);
    }
    tl_ctx = prev == &c->st ? nullptr : prev;
    delete c;
}

void sxg_ctx_single_stream(sxg_ctx *c) { c->st.single_stream = true; }
SxgState &sxg_ctx_state(sxg_ctx *c) { return c->st; }

extern "C" int sxg_ctx_bind(sxg_ctx *c)
{
    tl_ctx = c ? &c->st : nullptr;
    return 0;
}

bool sxg_ensure_stream()
{
    SxgState &S = sxg_state();
    if (S.stream) return true;
    /* default priority stream; events for per-stage timing */
    if (hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking) != hipSuccess) {
        sxg_set_error("hipStreamCreate failed");
        return false;
    }
    for (int i = 0; i < 2 * ST_N; i++)
        if (hipEventCreate(&S.ev[i]) != hipSuccess) {
            sxg_set_error("hipEventCreate failed");
            return false;
        }
    /* side streams only where a frame's independent launches may use them: a batch
       worker's context runs everything on its own stream */
    for (int i = 0; i < (S.single_stream ? 0 : SXG_NSIDE); i++)
        if (hipStreamCreateWithFlags(&S.side[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&S.side_ev[i], hipEventDisableTiming) != hipSuccess) {
            sxg_set_error("side stream creation failed");
            return false;
        }
    return true;
}

/* the stage events and waits are left out of planned runs (a capture records no timing
   and no host wait) */
void sxg_stage_begin(StageId s)
{
    SxgState &S = sxg_state();
    if (S.plan) return;
    (void)hipEventRecord(S.ev[2 * s], S.stream);
}

void sxg_stage_end(StageId s)
{
    SxgState &S = sxg_state();
    if (S.plan) return;
    (void)hipEventRecord(S.ev[2 * s + 1], S.stream);
    S.ev_used[s] = true;
}

extern "C" int sxg_stage_times(float *ms, int n)
{
    SxgState &S = sxg_state();
    int k = 0;
    for (; k < n && k < ST_N; k++) {
        ms[k] = 0.0f;
        if (S.ev_used[k]) {
            (void)hipEventSynchronize(S.ev[2 * k + 1]);
            (void)hipEventElapsedTime(&ms[k], S.ev[2 * k], S.ev[2 * k + 1]);
        }
    }
    return k;
}

bool sxg_stage_wait()
{
    SxgState &S = sxg_state();
    if (S.is_default || S.plan) return true;
    SXG_HIP_RET(hipStreamSynchronize(S.stream), false);
    return true;
}

extern "C" void sxg_release(void)
{
    SxgState &S = sxg_state();
    if (S.stream) (void)hipStreamSynchronize(S.stream);
    DevBuf *bufs[] = {&S.img, &S.cd, &S.aux, &S.backbuf, &S.scan_tmp, &S.thr, &S.lvl,
                      &S.comps, &S.objs, &S.rootobj};
    for (DevBuf *b : bufs) b->release();
    sxg_release_named();
    S.dmask.release();
    S.mask_k = 0;
    S.back_W = S.back_H = 0;
    S.resident.release();
    S.weight.release();
    S.weighted = false;
    S.have_frame = false;
}

/* ------------------------------------------------------------------ scan ---- */
#define SCAN_BLOCK 256
#define SCAN_ITEMS 16
#define SCAN_TILE (SCAN_BLOCK * SCAN_ITEMS)
/* counts up to this many are scanned by one workgroup in one launch */
#define SCAN_SMALL 32768
#define SCAN_SMALL_THREADS 1024

/* workgroup-wide exclusive scan of one value per thread (NT threads): an inclusive scan
   inside each wavefront, then the wavefront totals through LDS; returns the total */
template <int NT> __device__ inline int block_excl_scan_t(int v, int *sm, int *total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(x, d);
        if (lane >= d) x += o;
    }
    if (lane == 63) sm[wid] = x;
    __syncthreads();
    int before = 0, tot = 0;
    for (int w = 0; w < NT / 64; w++) {
        const int s = sm[w];
        before += w < wid ? s : 0;
        tot += s;
    }
    *total = tot;
    __syncthreads();
    return before + x - v;
}

__device__ inline int block_excl_scan(int v, int *sm, int *total)
{
    return block_excl_scan_t<SCAN_BLOCK>(v, sm, total);
}

/* 16 consecutive ints of a tile row (four 16-byte loads when the row is inside [0, n)) */
__device__ inline void load16(const int *in, long base, int n, int *v)
{
    if (base + SCAN_ITEMS <= n && (((uintptr_t)(in + base)) & 15) == 0) {
        const int4 *q = reinterpret_cast<const int4 *>(in + base);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int4 t = q[j];
            v[4 * j] = t.x;
            v[4 * j + 1] = t.y;
            v[4 * j + 2] = t.z;
            v[4 * j + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; i++) v[i] = base + i < n ? in[base + i] : 0;
    }
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_reduce(const int *in, int n, int *bsum,
                                                            const int *nd)
{
    __shared__ int sm[SCAN_BLOCK / 64];
    n = sxg_dn(nd, n);
    const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
    load16(in, base, n, v);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) s += v[i];
    int tot;
    block_excl_scan(s, sm, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

/* single workgroup: exclusive scan of nb block sums in place, total to *total */
__global__ void __launch_bounds__(1024) k_scan_blocks(int *bsum, int nb, int *total)
{
    __shared__ int sm[16];
    int carry = 0;
    for (int base = 0; base < nb; base += 1024) {
        const int k = base + threadIdx.x;
        const int v = k < nb ? bsum[k] : 0;
        int tot;
        const int ex = block_excl_scan_t<1024>(v, sm, &tot);
        if (k < nb) bsum[k] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_apply(const int *in, int *out, int n,
                                                           const int *bsum, const int *nd)
{
    __shared__ int sm[SCAN_BLOCK / 64];
    n = sxg_dn(nd, n);
    const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
    load16(in, base, n, v);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) s += v[i];
    int tot;
    int ex = block_excl_scan(s, sm, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const long k = base + i;
        if (k < n) out[k] = ex;
        ex += v[i];
    }
}

/* n <= SCAN_SMALL: one workgroup; the values pass through LDS (coalesced loads and
   stores), each thread scans a contiguous run of them.  The run length is odd (at most
   33), so the lanes of a wavefront, which read LDS words a run length apart, hit
   distinct banks. */
__global__ void __launch_bounds__(SCAN_SMALL_THREADS) k_scan_small(const int *in, int *out, int n,
                                                                   int *total, const int *nd)
{
    extern __shared__ int sv[];
    __shared__ int sm[SCAN_SMALL_THREADS / 64];
    n = sxg_dn(nd, n);
#pragma unroll 8
    for (int k = threadIdx.x; k < n; k += SCAN_SMALL_THREADS) sv[k] = in[k];
    __syncthreads();
    const int per = ((n + SCAN_SMALL_THREADS - 1) / SCAN_SMALL_THREADS) | 1;
    const int b = threadIdx.x * per, e = min(b + per, n);
    int s = 0;
    for (int k = b; k < e; k++) s += sv[k];
    int tot;
    int ex = block_excl_scan_t<SCAN_SMALL_THREADS>(s, sm, &tot);
    for (int k = b; k < e; k++) {
        const int v = sv[k];
        sv[k] = ex;
        ex += v;
    }
    __syncthreads();
#pragma unroll 8
    for (int k = threadIdx.x; k < n; k += SCAN_SMALL_THREADS) out[k] = sv[k];
    if (threadIdx.x == 0 && total) *total = tot;
}

__global__ void k_set_int(int *p, int v)
{
    if (threadIdx.x == 0) *p = v;
}

/* exclusive prefix sum of n ints (in == out allowed); the total to *total (device) */
void sxg_scan_excl(const int *in, int *out, int n, int *total, hipStream_t s, const int *nd)
{
    SxgState &S = sxg_state();
    if (n <= 0) {
        if (total) hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, total, 0);
        return;
    }
    if (n <= SCAN_SMALL) {
        /* the attribute is per function: set once (thread-safe static initialisation) */
        static const bool attr = [] {
            return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_scan_small),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       SCAN_SMALL * (int)sizeof(int)) == hipSuccess;
        }();
        (void)attr;
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(SCAN_SMALL_THREADS),
                           (size_t)n * sizeof(int), s, in, out, n, total, nd);
        return;
    }
    const int nb = sxg_div_up(n, SCAN_TILE);
    if (!S.scan_tmp.ensure(sizeof(int) * (nb + 64))) return;
    int *bsum = S.scan_tmp.as<int>();
    /* with a device count the blocks past it sum to zero: the same offsets and total */
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_BLOCK), 0, s, in, n, bsum, nd);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, bsum, nb, total);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_BLOCK), 0, s, in, out, n, bsum, nd);
}

/* ------------------------------------------------------------ compaction ---- */
/* Contiguous per-thread chunks keep the output in increasing index order. */
struct PredNonzero { __device__ bool operator()(int v) const { return v != 0; } };
struct PredNonneg { __device__ bool operator()(int v) const { return v >= 0; } };

template <class P>
__global__ void __launch_bounds__(SCAN_BLOCK) k_compact_count(const int *flag, int n, int *bsum,
                                                              const int *nd)
{
    __shared__ int sm[SCAN_BLOCK / 64];
    n = sxg_dn(nd, n);
    const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
    load16(flag, base, n, v);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) s += (base + i < n && P()(v[i])) ? 1 : 0;
    int tot;
    block_excl_scan(s, sm, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

template <class P>
__global__ void __launch_bounds__(SCAN_BLOCK) k_compact_write(const int *flag, int n,
                                                              const int *bsum, int *out,
                                                              const int *nd)
{
    __shared__ int sm[SCAN_BLOCK / 64];
    n = sxg_dn(nd, n);
    const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
    load16(flag, base, n, v);
    unsigned f = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++)
        if (base + i < n && P()(v[i])) {
            f |= 1u << i;
            s++;
        }
    int tot;
    int pos = block_excl_scan(s, sm, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++)
        if (f & (1u << i)) out[pos++] = (int)(base + i);
}

template <class P>
static void compact_impl(const int *flag, int *out, int n, int *count, hipStream_t s,
                         const int *nd)
{
    SxgState &S = sxg_state();
    if (n <= 0) {
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, count, 0);
        return;
    }
    const int nb = sxg_div_up(n, SCAN_TILE);
    if (!S.scan_tmp.ensure(sizeof(int) * (nb + 64))) return;
    int *bsum = S.scan_tmp.as<int>();
    hipLaunchKernelGGL(k_compact_count<P>, dim3(nb), dim3(SCAN_BLOCK), 0, s, flag, n, bsum, nd);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, bsum, nb, count);
    hipLaunchKernelGGL(k_compact_write<P>, dim3(nb), dim3(SCAN_BLOCK), 0, s, flag, n, bsum, out,
                       nd);
}

void sxg_compact_index(const int *flag, int *out, int n, int *count, hipStream_t s,
                       const int *nd)
{
    compact_impl<PredNonzero>(flag, out, n, count, s, nd);
}

void sxg_compact_nonneg(const int *v, int *out, int n, int *count, hipStream_t s, const int *nd)
{
    compact_impl<PredNonneg>(v, out, n, count, s, nd);
}

/* The positions of the set bits of nw 64-bit window masks (bit b of mask w: position
   64 w + b) in increasing order, and their number: one lane per window counts its bits;
   after the scan of the workgroup sums each lane writes its window's positions in bit
   order.  Reads the masks (8 bytes per 64 positions) instead of a flag per position. */
__global__ void __launch_bounds__(SCAN_BLOCK) k_fgm_count(const unsigned long long *m, long nw,
                                                          int *bsum)
{
    __shared__ int sm[SCAN_BLOCK / 64];
    const long w = (long)blockIdx.x * SCAN_BLOCK + threadIdx.x;
    int tot;
    block_excl_scan(w < nw ? __popcll(m[w]) : 0, sm, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_fgm_write(const unsigned long long *m, long nw,
                                                          const int *bsum, int *out)
{
    __shared__ int sm[SCAN_BLOCK / 64];
    const long w = (long)blockIdx.x * SCAN_BLOCK + threadIdx.x;
    unsigned long long b = w < nw ? m[w] : 0ull;
    int tot;
    int pos = block_excl_scan(__popcll(b), sm, &tot) + bsum[blockIdx.x];
    const int base = (int)(w * 64);
    while (b) {
        out[pos++] = base + __builtin_ctzll(b);
        b &= b - 1;
    }
}

void sxg_compact_mask(const unsigned long long *m, long nw, int *out, int *count, hipStream_t s)
{
    SxgState &S = sxg_state();
    const int nb = (int)sxg_div_up(nw, SCAN_BLOCK);
    if (nb < 1 || !S.scan_tmp.ensure(sizeof(int) * (nb + 64))) return;
    int *bsum = S.scan_tmp.as<int>();
    hipLaunchKernelGGL(k_fgm_count, dim3(nb), dim3(SCAN_BLOCK), 0, s, m, nw, bsum);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, bsum, nb, count);
    hipLaunchKernelGGL(k_fgm_write, dim3(nb), dim3(SCAN_BLOCK), 0, s, m, nw, bsum, out);
}

/* ------------------------------------------------------------ plan guards ---- */
__global__ void k_plan_guard(int *v, long long cap, int bit, PlanDev *pd, int *cc)
{
    if (threadIdx.x != 0) return;
    const int x = *v;
    if (x < 0 || (long long)x > cap) pd->ovf |= 1 << bit;
    if (pd->ovf) {
        /* the frame is run again eagerly: nothing after this point has work to do */
        *v = 0;
        pd->nfg = pd->nraw = pd->npix = pd->nobj = pd->nsurv = 0;
        if (cc) {
            for (int c = 0; c < 4; c++) cc[c] = 0;
            cc[8] = 0;
        }
    }
}

void sxg_plan_guard(int *v, int64_t cap, int bit, int *ccount)
{
    SxgState &S = sxg_state();
    if (!S.plan) return;
    hipLaunchKernelGGL(k_plan_guard, dim3(1), dim3(64), 0, S.stream, v, (long long)cap, bit,
                       S.pdev, ccount);
}

/* ------------------------------------------------------- named buffers ---- */
DevBuf &sxg_buf(const char *name) { return sxg_state().named[name]; }
void sxg_release_named()
{
    for (auto &kv : sxg_state().named) kv.second.release();
}

bool sxg_raw_alloc(int nraw, RawObj &ro)
{
    const size_t n = (size_t)(nraw > 0 ? nraw : 1);
    ro.first = sxg_take<int>("raw.first", n);
    ro.count = sxg_take<int>("raw.count", n);
    ro.off = sxg_take<int>("raw.off", n);
    ro.xmin = sxg_take<int>("raw.xmin", n);
    ro.xmax = sxg_take<int>("raw.xmax", n);
    ro.ymin = sxg_take<int>("raw.ymin", n);
    ro.ymax = sxg_take<int>("raw.ymax", n);
    ro.fdpeak = sxg_take<float>("raw.fdpeak", n);
    ro.fdflux = sxg_take<float>("raw.fdflux", n);
    ro.t0 = sxg_take<float>("raw.t0", n);
    return ro.first && ro.count && ro.off && ro.xmin && ro.xmax && ro.ymin && ro.ymax &&
           ro.fdpeak && ro.fdflux && ro.t0;
}

/* -------------------------------------------------------- frame entry points ---- */
extern "C" void init_device(int width, int height, float *imgbuf)
{
    sxg_clear_error();
    SxgState &S = sxg_state();
    if (width <= 0 || height <= 0 || (long)width * height >= 2147483647L || !imgbuf) {
        sxg_set_error("init_device: invalid frame size or null image");
        return;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        sxg_set_error("init_device: no HIP device visible");
        return;
    }
    if (!sxg_ensure_stream()) return;
    S.W = width;
    S.H = height;
    S.N = (size_t)width * height;
    S.nraw = S.nobj = S.nsurv = 0;
    S.cleaned = false;
    S.phot_ready = false;
    S.wrms_frame = false;
    S.filtered = false;
    for (int i = 0; i < ST_N; i++) S.ev_used[i] = false;
    if (!S.img.ensure(S.N * sizeof(float)) || !S.cd.ensure(S.N * sizeof(float)) ||
        !S.aux.ensure(S.N * sizeof(int)))
        return;
    SXG_HIP(hipMemcpyAsync(S.img.ptr, imgbuf, S.N * sizeof(float), hipMemcpyHostToDevice,
                           S.stream));
    SXG_HIP(hipStreamSynchronize(S.stream));
    S.have_frame = true;
}

extern "C" void clear_device(float *img)
{
    SxgState &S = sxg_state();
    if (!S.have_frame) {
        sxg_set_error("clear_device: no frame");
        return;
    }
    if (img) {
        SXG_HIP(hipMemcpyAsync(img, S.img.ptr, S.N * sizeof(float), hipMemcpyDeviceToHost,
                               S.stream));
    }
    SXG_HIP(hipStreamSynchronize(S.stream));
    S.have_frame = false;
}

extern "C" void init_detection(void)
{
    SxgState &S = sxg_state();
    if (!S.have_frame || S.W == 0 || S.H == 0) sxg_set_error("init_detection: no frame");
}

extern "C" void sxg_set_back_value(float value) { sxg_state().back_value = value; }

extern "C" void *sxg_device_frame(int which)
{
    SxgState &S = sxg_state();
    if (!S.have_frame) return nullptr;
    return which == 0 ? S.img.ptr : S.cd.ptr;
}

extern "C" int sxg_get_filtered(float *out)
{
    SxgState &S = sxg_state();
    if (!S.have_frame) return -1;
    SXG_HIP_RET(hipMemcpyAsync(out, S.cd.ptr, S.N * sizeof(float), hipMemcpyDeviceToHost,
                               S.stream), -1);
    SXG_HIP_RET(hipStreamSynchronize(S.stream), -1);
    return 0;
}

/* ------------------------------------------------------------- weight maps ---- */
/* MAP_RMS / MAP_VAR: the mesh statistics take the pixels whose variance is below BIG */
__global__ void __launch_bounds__(256) k_weight_mask(const float *map, long n, int type,
                                                     float *mask)
{
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    WeightArgs wa;
    wa.w = map;
    wa.type = type;
    mask[p] = sxg_var(wa, p) < (float)SXG_BIG ? 1.0f : 0.0f;
}

extern "C" int sxg_set_weight_map(const float *map, int width, int height, int type,
                                  int thresh_type, double detect_thresh, double analysis_thresh,
                                  int weight_gain)
{
    SxgState &S = sxg_state();
    sxg_clear_error();
    if (type < SXG_WEIGHT_MAP_WEIGHT || type > SXG_WEIGHT_BACKGROUND) {
        sxg_set_error("sxg_set_weight_map: unknown weight type");
        return -1;
    }
    S.wrms_frame = false;
    if (!map && type != SXG_WEIGHT_BACKGROUND) {
        S.weighted = false;
        S.weight.release();
        S.wmask.release();
        S.ww = S.wh = 0;
        S.wtype = SXG_WEIGHT_MAP_WEIGHT;
        S.wgain = false;
        return 0;
    }
    if (!sxg_ensure_stream()) return -1;
    if (map) {
        if (width <= 0 || height <= 0 || (long)width * height >= 2147483647L) {
            sxg_set_error("sxg_set_weight_map: bad map size");
            return -1;
        }
        const size_t n = (size_t)width * height;
        if (!S.weight.ensure(n * sizeof(float))) return -1;
        SXG_HIP_RET(hipMemcpy(S.weight.ptr, map, n * sizeof(float), hipMemcpyHostToDevice), -1);
        if (type == SXG_WEIGHT_MAP_RMS || type == SXG_WEIGHT_MAP_VAR) {
            if (!S.wmask.ensure(n * sizeof(float))) return -1;
            hipLaunchKernelGGL(k_weight_mask, dim3(sxg_div_up((long)n, 256)), dim3(256), 0,
                               S.stream, S.weight.as<float>(), (long)n, type, S.wmask.as<float>());
            SXG_HIP_RET(hipGetLastError(), -1);
        }
        S.ww = width;
        S.wh = height;
    } else {
        S.ww = S.wh = 0;    /* BACKGROUND: the map is built per frame */
    }
    S.weighted = true;
    S.wtype = type;
    S.wgain = weight_gain != 0;
    S.wargs.rel = thresh_type != 1;
    S.wargs.kdet = detect_thresh;
    S.wargs.kana = analysis_thresh;
    return 0;
}

extern "C" int sxg_set_weight(const float *w, int width, int height, int thresh_type,
                              double detect_thresh, double analysis_thresh)
{
    return sxg_set_weight_map(w, width, height, SXG_WEIGHT_MAP_WEIGHT, thresh_type,
                              detect_thresh, analysis_thresh, 0);
}

bool sxg_weight_args(WeightArgs &wa)
{
    SxgState &S = sxg_state();
    wa = WeightArgs();
    if (!S.weighted) return true;
    if (S.wtype == SXG_WEIGHT_BACKGROUND) {
        /* no mask for the mesh statistics; the rms map once back_estimate has built it */
        if (!S.wrms_frame) return true;
        wa = S.wargs;
        wa.w = S.weight.as<float>();
        wa.mask = nullptr;
        wa.type = SXG_WEIGHT_MAP_RMS;
        return true;
    }
    if (S.ww != S.W || S.wh != S.H) {
        sxg_set_error("the weight map (sxg_set_weight) does not match the frame size");
        return false;
    }
    wa = S.wargs;
    wa.w = S.weight.as<float>();
    wa.type = S.wtype;
    wa.mask = S.wtype == SXG_WEIGHT_MAP_WEIGHT ? wa.w : S.wmask.as<float>();
    return true;
}
