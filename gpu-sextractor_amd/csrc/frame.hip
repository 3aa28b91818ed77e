/*
 * frame.hip -- whole-frame driver, the batched path and kernel timers.
 *
 * process_frame runs the path in the order of the reference's scan.c:78-146
 * (back_estimate, cudaFilter, run_detection, init_deblend, parcel_out, analyse_gpu,
 * run_clean, segmentation map) on a frame already resident in device memory, on the
 * calling thread's context, so a caller that keeps frames on the GPU pays no PCIe
 * transfer and no host round trip besides the object counts the C-ABI itself returns.
 *
 * The batched path keeps many frames resident and runs them on several contexts at
 * once, one host worker thread and one stream per context: the stages of one frame are
 * mostly latency-bound (sequential per-object sums, per-level barriers, host
 * synchronisations on counts), so frames in flight together fill the GPU.
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <thread>
#include <tuple>
#include <vector>

#include "sxg_internal.h"
#include "sxg_objects.h"

/* ------------------------------------------------------------ kernel timers ---- */
struct KTimer {
    std::vector<hipEvent_t> ev;   /* begin/end pairs */
    int used = 0;
};
static KTimer g_kt[SXG_NKT];
static bool g_kt_on = false;

void sxg_kt_begin(int id)
{
    if (!g_kt_on || !sxg_state().is_default || sxg_state().plan) return;
    KTimer &t = g_kt[id];
    if ((int)t.ev.size() < 2 * (t.used + 1)) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
        t.ev.push_back(a);
        t.ev.push_back(b);
    }
    (void)hipEventRecord(t.ev[2 * t.used], sxg_state().stream);
}

void sxg_kt_end(int id)
{
    if (!g_kt_on || !sxg_state().is_default || sxg_state().plan) return;
    KTimer &t = g_kt[id];
    if ((int)t.ev.size() < 2 * (t.used + 1)) return;
    (void)hipEventRecord(t.ev[2 * t.used + 1], sxg_state().stream);
    t.used++;
}

/* the same, recorded on stream st (kernels launched on a side stream) */
void sxg_kt_begin_on(int id, hipStream_t st)
{
    if (!g_kt_on || !sxg_state().is_default || sxg_state().plan) return;
    KTimer &t = g_kt[id];
    if ((int)t.ev.size() < 2 * (t.used + 1)) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
        t.ev.push_back(a);
        t.ev.push_back(b);
    }
    (void)hipEventRecord(t.ev[2 * t.used], st);
}

void sxg_kt_end_on(int id, hipStream_t st)
{
    if (!g_kt_on || !sxg_state().is_default || sxg_state().plan) return;
    KTimer &t = g_kt[id];
    if ((int)t.ev.size() < 2 * (t.used + 1)) return;
    (void)hipEventRecord(t.ev[2 * t.used + 1], st);
    t.used++;
}

extern "C" void sxg_kernel_timing(int enable)
{
    g_kt_on = enable != 0;
    for (auto &t : g_kt) t.used = 0;
}

extern "C" int sxg_kernel_times(int id, double *total_ms, int *launches)
{
    if (id < 0 || id >= SXG_NKT) return -1;
    KTimer &t = g_kt[id];
    double tot = 0.0;
    for (int i = 0; i < t.used; i++) {
        float ms = 0.0f;
        if (hipEventSynchronize(t.ev[2 * i + 1]) != hipSuccess ||
            hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]) != hipSuccess)
            return -1;
        tot += ms;
    }
    *total_ms = tot;
    *launches = t.used;
    return 0;
}

/* ---------------------------------------------------------- resident frame ---- */
extern "C" int sxg_set_resident_frame(const float *host, int width, int height)
{
    SxgState &S = sxg_state();
    sxg_clear_error();
    if (width <= 0 || height <= 0 || (long)width * height >= 2147483647L || !host) {
        sxg_set_error("sxg_set_resident_frame: bad frame");
        return -1;
    }
    const size_t n = (size_t)width * height;
    if (!S.resident.ensure(n * sizeof(float))) return -1;
    SXG_HIP_RET(hipMemcpy(S.resident.ptr, host, n * sizeof(float), hipMemcpyHostToDevice), -1);
    S.rw = width;
    S.rh = height;
    return 0;
}

/* ------------------------------------------------------------ planned runs ---- */
/* the host copy of one planned frame: its device block and the device clock (wall_clock64)
   at its first and last kernel */
struct PlanRecord {
    PlanDev pd;
    unsigned long long t0, t1;
};

__global__ void k_plan_begin(PlanDev *pd, unsigned long long *t0)
{
    int *w = reinterpret_cast<int *>(pd);
    for (int k = threadIdx.x; k < (int)(sizeof(PlanDev) / sizeof(int)); k += blockDim.x) w[k] = 0;
    if (threadIdx.x == 0 && t0) *t0 = wall_clock64();
}

/* the thresholds from the background's medians (outv: backmean, backsig), as the host
   derives them in an eager run; manual: BACK_TYPE MANUAL reports BACK_VALUE as backmean */
__global__ void k_plan_thresh(PlanDev *pd, const float *outv, int thresh_type, double dt,
                              double at, int manual, float back_value)
{
    if (threadIdx.x != 0) return;
    const float backsig = outv[1];
    pd->backmean = manual ? back_value : outv[0];
    pd->backsig = backsig;
    float d, t;
    if (thresh_type == 1) {
        d = (float)dt;
        t = (float)at;
    } else {
        d = (float)(dt * backsig);
        t = (float)(at * backsig);
    }
    pd->dthresh = d;
    pd->thresh = t;
    /* back.c:195-197 refuses zero or negative thresholds: the eager rerun reports it */
    if (!(d > 0.0f)) pd->ovf |= 1 << SXG_PLAN_BADTHRESH;
}

__global__ void k_plan_end(const PlanDev *pd, PlanRecord *rec)
{
    if (threadIdx.x != 0) return;
    rec->pd = *pd;
    rec->t1 = wall_clock64();
}

/* resets the context's planned-run state when the frame's queueing ends, however it ends */
struct PlanScope {
    SxgState &S;
    explicit PlanScope(SxgState &s, const FramePlan *plan) : S(s) { S.plan = plan; }
    ~PlanScope() { S.plan = nullptr; }
};

/* The whole path on a device frame (not modified), on the current context.  plan: a
   planned run (counts stay on the device, sized by the plan's capacities; nothing waits
   or allocates once the context has run a frame of that plan), whose device block is
   copied to *rec (device) at the end; info is then not written. */
static int process_frame(const float *dframe, int W, int H, const sxg_params *p,
                         sxg_frame_info *info, const FramePlan *plan = nullptr,
                         PlanRecord *rec = nullptr, int *segdst = nullptr)
{
    SxgState &S = sxg_state();
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        sxg_set_error("no HIP device visible");
        return -1;
    }
    if (!sxg_ensure_stream()) return -1;
    PlanScope scope(S, plan);
    if (plan) {
        S.pdev = sxg_take<PlanDev>("plan.dev", 1);
        if (!S.pdev) return -1;
        hipLaunchKernelGGL(k_plan_begin, dim3(1), dim3(64), 0, S.stream, S.pdev,
                           rec ? &rec->t0 : nullptr);
    }
    /* init_device from device memory: the background stage reads the frame where it is
       and writes the subtracted frame into the context's buffer, so the input is never
       copied */
    {
        const size_t N = (size_t)W * H;
        S.W = W;
        S.H = H;
        S.N = N;
        S.nraw = S.nobj = S.nsurv = 0;
        S.cleaned = false;
        S.phot_ready = false;
        S.wrms_frame = false;
        S.filtered = false;
        S.deb_ready = false;
        for (int i = 0; i < ST_N; i++) S.ev_used[i] = false;
        if (!S.img.ensure(N * sizeof(float)) || !S.cd.ensure(N * sizeof(float)) ||
            !S.aux.ensure(N * sizeof(int)))
            return -1;
        S.have_frame = true;
    }
    const int bw = p->back_size < W ? p->back_size : W;
    const int bh = p->back_size < H ? p->back_size : H;
    if (bw != bh) {
        sxg_set_error("only square background meshes are supported (back.c:99-103)");
        return -1;
    }
    const int nbx = (W - 1) / bw + 1, nby = (H - 1) / bw + 1;
    float backmean = 0.0f, backsig = 0.0f;
    S.back_value = p->back_value;
    /* the spline's subtraction, the convolution and the detection parents in one pass
       (frames of a width that is a multiple of 64, masks of 3 to 11) */
    const int K = p->mask ? p->mask_width : 0;
    const bool fuse = p->back_type == 0 && (K == 3 || K == 5 || K == 7 || K == 9 || K == 11) &&
                      (W & 63) == 0 && sxg_option(SXG_OPT_FUSED_FILTER) != 0;
    S.defer_sub = fuse;
    S.det_init_done = false;
    sxg_back_estimate_from(dframe, bw, p->back_fthresh, &backmean, &backsig,
                           nbx > 1 ? p->back_fsize : 1, nby > 1 ? p->back_fsize : 1,
                           p->back_type);
    if (sxg_has_error()) return -1;
    /* back.c:152-153: with BACK_TYPE MANUAL the frame's backmean is BACK_VALUE */
    if (p->back_type == 1) S.backmean = p->back_value;
    float dthresh = 0.0f, thresh = 0.0f;
    if (plan) {
        /* the medians stay on the device: so do the thresholds */
        BackBufs b = sxg_back_bufs_at(S.backbuf.as<char>(), S.bgeom);
        hipLaunchKernelGGL(k_plan_thresh, dim3(1), dim3(64), 0, S.stream, S.pdev,
                           (const float *)b.outv, p->thresh_type, p->detect_thresh,
                           p->analysis_thresh, p->back_type == 1 ? 1 : 0, p->back_value);
    } else if (p->thresh_type == 1) {
        dthresh = (float)p->detect_thresh;
        thresh = (float)p->analysis_thresh;
    } else {
        dthresh = (float)(p->detect_thresh * backsig);
        thresh = (float)(p->analysis_thresh * backsig);
    }
    init_detection();
    if (fuse) {
        if (!sxg_subconv_detect(p->mask, K, dthresh)) {
            if (!sxg_has_error()) sxg_set_error("process_frame: the fused filter did not run");
            return -1;
        }
    } else if (p->mask && p->mask_width > 0) {
        cudaFilter(const_cast<float *>(p->mask), p->mask_width);
    }
    if (sxg_has_error()) return -1;
    size_t nraw = 0, npix = 0;
    run_detection(dthresh, thresh, p->ext_minarea, &nraw, &npix);
    if (sxg_has_error()) return -1;
    init_deblend(dthresh, p->deblend_nthresh, nraw, npix);
    const int debarea = p->ext_minarea < 3 ? p->ext_minarea : 3;
    const int nobj = parcel_out(dthresh, p->deblend_nthresh, debarea, nraw, npix, p->deblend_mincont);
    clear_deblend(p->deblend_nthresh);
    if (sxg_has_error() || nobj < 0) return -1;
    const double ftrue = [] { double d; uint64_t u = 1; memcpy(&d, &u, 8); return d; }();
    const float ftruef = [] { float f; uint32_t u = 1; memcpy(&f, &u, 4); return f; }();
    analyse_gpu(ftrue, 1, 1, ftruef, backsig, p->gain, 0.0, thresh, dthresh, p->satur_level, 0, 0,
                S.wgain ? 1 : 0, p->ext_minarea, p->clean, 0, nbx, nby, bw, bw, W, nobj);
    if (sxg_has_error()) return -1;
    run_clean(nullptr, p->clean_param, p->clean_stacksize, nobj);
    if (sxg_has_error()) return -1;
    if (!sxg_build_segmap(segdst)) return -1;
    if (plan) {
        if (rec) hipLaunchKernelGGL(k_plan_end, dim3(1), dim3(64), 0, S.stream, S.pdev, rec);
        SXG_HIP_RET(hipGetLastError(), -1);
        return 0;
    }
    if (info) sxg_get_frame_info(info);
    return nobj;
}

/* the frame's sxg_frame_info from the record of a planned run */
static void info_from_record(const PlanRecord &r, int nbx, int nby, sxg_frame_info *fi)
{
    memset(fi, 0, sizeof(*fi));
    fi->backmean = r.pd.backmean;
    fi->backsig = r.pd.backsig;
    fi->dthresh = r.pd.dthresh;
    fi->thresh = r.pd.thresh;
    fi->npix_above = r.pd.nfg;
    fi->npix_raw = r.pd.npix;
    fi->nraw = r.pd.nraw;
    fi->nobj = r.pd.nobj;
    fi->nsurv = r.pd.nsurv;
    fi->nbx = nbx;
    fi->nby = nby;
    fi->deb_pixels = r.pd.deb_pixels;
    fi->deb_labels = r.pd.deb_labels;
    for (int c = 0; c < 4; c++) fi->deb_class_pixels[c] = r.pd.class_pixels[c];
}

extern "C" int sxg_process_resident(const sxg_params *p, sxg_frame_info *info)
{
    SxgState &S = sxg_state();
    sxg_clear_error();
    if (!S.resident.ptr || !p) {
        sxg_set_error("sxg_process_resident: no resident frame");
        return -1;
    }
    return process_frame(S.resident.as<float>(), S.rw, S.rh, p, info);
}

extern "C" void sxg_sync(void)
{
    SxgState &S = sxg_state();
    if (S.stream) (void)hipStreamSynchronize(S.stream);
}

/* ---------------------------------------------------------------- batches ---- */
struct BatchFrame {
    DevBuf frame;       /* float[W*H] input, resident */
    DevBuf seg;         /* int32[W*H] segmentation map of the last run */
    int W = 0, H = 0;
};
static std::vector<BatchFrame> g_batch;
static void graph_drop();   /* forgets the batch's graph (its frames change) */
static std::vector<sxg_ctx *> g_workers;

extern "C" int sxg_batch_set_frames(int n, const float *const *host, const int *width,
                                    const int *height)
{
    sxg_clear_error();
    if (n < 0 || (n > 0 && (!host || !width || !height))) {
        sxg_set_error("sxg_batch_set_frames: bad arguments");
        return -1;
    }
    graph_drop();
    for (auto &f : g_batch) {
        f.frame.release();
        f.seg.release();
    }
    g_batch.assign(n, BatchFrame());
    for (int i = 0; i < n; i++) {
        const int W = width[i], H = height[i];
        if (W <= 0 || H <= 0 || (long)W * H >= 2147483647L || !host[i]) {
            sxg_set_error("sxg_batch_set_frames: bad frame " + std::to_string(i));
            return -1;
        }
        const size_t N = (size_t)W * H;
        if (!g_batch[i].frame.ensure(N * sizeof(float)) || !g_batch[i].seg.ensure(N * sizeof(int)))
            return -1;
        SXG_HIP_RET(hipMemcpy(g_batch[i].frame.ptr, host[i], N * sizeof(float),
                              hipMemcpyHostToDevice), -1);
        g_batch[i].W = W;
        g_batch[i].H = H;
    }
    return 0;
}

/* The worker pool of the batched paths: frames 0..n-1 are handed out in order to
   nworkers contexts, one host thread each (the caller's own binding is untouched);
   run(i) processes frame i on the calling thread's context and returns false (error
   recorded) on failure.  frame_ms[i]: wall time of run(i). */
static int run_pool(int n, int nworkers, const std::function<bool(int)> &run, double *frame_ms,
                    const char *who)
{
    if (nworkers < 1 || nworkers > SXG_MAX_WORKERS) {
        sxg_set_error(std::string(who) + ": nworkers must be 1.." + std::to_string(SXG_MAX_WORKERS));
        return -1;
    }
    while ((int)g_workers.size() < nworkers) {
        sxg_ctx *c = sxg_ctx_create();
        if (!c) {
            sxg_set_error(std::string(who) + ": context allocation failed");
            return -1;
        }
        /* one stream per worker: the frames in flight already fill the hardware queues */
        sxg_ctx_single_stream(c);
        g_workers.push_back(c);
    }
    std::atomic<int> next(0), failed(0);
    std::vector<std::string> errs(nworkers);
    auto work = [&](int w) {
        sxg_ctx_bind(g_workers[w]);
        sxg_clear_error();
        for (int i = next++; i < n && !failed; i = next++) {
            const auto t0 = std::chrono::steady_clock::now();
            if (!run(i)) {
                failed = 1;
                break;
            }
            if (frame_ms)
                frame_ms[i] = std::chrono::duration<double, std::milli>(
                                  std::chrono::steady_clock::now() - t0).count();
        }
        if (sxg_has_error()) errs[w] = sxg_last_error();
        sxg_ctx_bind(nullptr);
    };
    std::vector<std::thread> th;
    for (int w = 0; w < nworkers; w++) th.emplace_back(work, w);
    for (auto &t : th) t.join();
    if (failed) {
        for (auto &e : errs)
            if (!e.empty()) {
                sxg_set_error(e);
                break;
            }
        if (!sxg_has_error()) sxg_set_error(std::string(who) + ": a frame failed");
        return -1;
    }
    return n;
}

extern "C" int sxg_batch_process(const sxg_params *p, int nworkers, sxg_frame_info *info,
                                 double *frame_ms)
{
    sxg_clear_error();
    const int n = (int)g_batch.size();
    if (!p) {
        sxg_set_error("sxg_batch_process: no parameters");
        return -1;
    }
    auto run = [&](int i) {
        BatchFrame &f = g_batch[i];
        sxg_frame_info fi;
        if (process_frame(f.frame.as<float>(), f.W, f.H, p, &fi, nullptr, nullptr,
                          f.seg.as<int>()) < 0)
            return false;
        if (hipStreamSynchronize(sxg_state().stream) != hipSuccess) {
            sxg_set_error("sxg_batch_process: the frame's stream failed");
            return false;
        }
        if (info) info[i] = fi;
        return true;
    };
    return run_pool(n, nworkers, run, frame_ms, "sxg_batch_process");
}

extern "C" int sxg_batch_process_host(const sxg_params *p, int nworkers, int n,
                                      const float *const *frames, const int *width,
                                      const int *height, int32_t *const *segs,
                                      sxg_isorow *const *rows, const int *row_cap,
                                      sxg_frame_info *info, double *frame_ms)
{
    sxg_clear_error();
    if (!p || n < 0 || (n > 0 && (!frames || !width || !height || !segs))) {
        sxg_set_error("sxg_batch_process_host: bad arguments");
        return -1;
    }
    for (int i = 0; i < n; i++)
        if (width[i] <= 0 || height[i] <= 0 || (long)width[i] * height[i] >= 2147483647L ||
            !frames[i] || !segs[i] || (rows && rows[i] && (!row_cap || row_cap[i] < 0))) {
            sxg_set_error("sxg_batch_process_host: bad frame " + std::to_string(i));
            return -1;
        }
    auto run = [&](int i) {
        const int W = width[i], H = height[i];
        const size_t N = (size_t)W * H;
        if (!sxg_ensure_stream()) return false;
        SxgState &S = sxg_state();
        /* the worker's own input buffer: the frame is copied in on the worker's stream,
           ahead of the kernels that read it */
        float *in = sxg_take<float>("batch.in", N);
        if (!in) return false;
        if (hipMemcpyAsync(in, frames[i], N * sizeof(float), hipMemcpyHostToDevice, S.stream) !=
            hipSuccess) {
            sxg_set_error("sxg_batch_process_host: frame copy failed");
            return false;
        }
        sxg_frame_info fi;
        if (process_frame(in, W, H, p, &fi) < 0) return false;
        int *map = sxg_take<int>("out.map", S.N);
        if (!map || hipMemcpyAsync(segs[i], map, N * sizeof(int), hipMemcpyDeviceToHost,
                                   S.stream) != hipSuccess) {
            sxg_set_error("sxg_batch_process_host: segmentation copy failed");
            return false;
        }
        if (rows && rows[i] && sxg_catalog_async(rows[i], row_cap[i]) < 0) return false;
        if (hipStreamSynchronize(S.stream) != hipSuccess) {
            sxg_set_error("sxg_batch_process_host: stream wait failed");
            return false;
        }
        if (info) info[i] = fi;
        return true;
    };
    return run_pool(n, nworkers, run, frame_ms, "sxg_batch_process_host");
}

/* ---------------------------------------------------------- graph batches ---- */
/* The resident batch as one HIP graph: a branch per worker context, each running its
   frames one after the other as planned runs (no host wait, no allocation), the
   segmentation maps copied into the frames' slots and every frame's record into pinned
   host memory at the end.  Built in four steps:
     1. calibration: every frame eagerly (the worker pool), recording its counts and time;
     2. plans: each count with a margin (margin_pct percent plus a few units) as its
        capacity; branches by longest-first greedy assignment of the calibrated times;
     3. warm-up: every frame once as a planned run on its branch's context, uncaptured,
        which sizes the contexts' buffers for the plans (the graph records addresses);
     4. capture: the branches recorded from an origin stream (fork and join events) and
        the graph instantiated.
   A launch replays all of it; a frame whose record shows an overflow (counts above the
   plan, e.g. after sxg_batch_update_frame loaded a denser frame) runs again eagerly on
   its context, which may grow its buffers and so marks the graph for a rebuild. */
struct GraphBatch {
    bool built = false;
    sxg_params p{};
    std::vector<float> mask;
    int nworkers = 0, margin = 0;
    std::vector<FramePlan> plans;
    std::vector<int> branch;        /* worker context of each frame */
    std::vector<double> calib_ms;   /* calibrated time of each frame */
    std::vector<uint64_t> gen;      /* each worker's allocation generation at capture */
    hipStream_t origin = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf rec;                     /* PlanRecord of each frame (device) */
    PlanRecord *hrec = nullptr;     /* their pinned host copy */
    int nrec = 0;
    long builds = 0, reruns = 0, launches = 0;
    double clock_khz = 0.0;         /* wall_clock64 rate */
    long opts[SXG_NOPT] = {};       /* the options the graph was built with */
};
static GraphBatch g_gb;
static int g_plan_margin = 25;

static void graph_drop()
{
    GraphBatch &G = g_gb;
    if (G.exec) (void)hipGraphExecDestroy(G.exec);
    if (G.graph) (void)hipGraphDestroy(G.graph);
    G.exec = nullptr;
    G.graph = nullptr;
    G.built = false;
}

static void graph_release()
{
    GraphBatch &G = g_gb;
    graph_drop();
    if (G.origin) (void)hipStreamSynchronize(G.origin);
    for (hipEvent_t e : G.join) (void)hipEventDestroy(e);
    G.join.clear();
    if (G.fork) (void)hipEventDestroy(G.fork);
    G.fork = nullptr;
    if (G.origin) (void)hipStreamDestroy(G.origin);
    G.origin = nullptr;
    if (G.hrec) (void)hipHostFree(G.hrec);
    G.hrec = nullptr;
    G.nrec = 0;
    G.rec.release();
}

/* capacity of a count seen in calibration: the margin, and a few units of slack */
static FramePlan plan_from(const FramePlan &seen, int margin)
{
    static const int64_t slack[PL_N] = {4096, 64, 1024, 256, 16, 8, 4, 4, 2, 65536, 1024,
                                        0, 0, 0, 128};
    FramePlan pl;
    for (int k = 0; k < PL_N; k++) {
        if (k >= PL_LDS0 && k <= PL_LDS2) continue;
        pl.cap[k] = seen.cap[k] + seen.cap[k] * margin / 100 + slack[k];
    }
    /* the level-parallel objects are one grid row each (at most 65535) */
    pl.cap[PL_BIG] = std::min<int64_t>(pl.cap[PL_BIG], 65535);
    /* LDS of a class's workgroups: the class's budget (the kernel is launched with the
       smaller of the two), except that class 2 asks for 76 KiB when its largest tree
       seen, with the margin, fits there: two of its 1024-lane workgroups then share a
       CU */
    for (int c = 0; c < 3; c++) pl.cap[PL_LDS0 + c] = 0x7fffffff;
    const int64_t need2 = seen.cap[PL_LDS2] + seen.cap[PL_LDS2] * margin / 100;
    if (need2 <= 76 * 1024) pl.cap[PL_LDS2] = 76 * 1024;
    for (int k = 0; k < PL_N; k++) pl.cap[k] = std::min<int64_t>(pl.cap[k], 0x7ffffff0);
    return pl;
}

static bool same_params(const sxg_params &a, const sxg_params *b, const std::vector<float> &mask)
{
    /* field by field (the struct has padding) */
    if (a.back_size != b->back_size || a.back_fsize != b->back_fsize ||
        a.back_fthresh != b->back_fthresh || a.back_type != b->back_type ||
        a.back_value != b->back_value || a.thresh_type != b->thresh_type ||
        a.detect_thresh != b->detect_thresh || a.analysis_thresh != b->analysis_thresh ||
        a.ext_minarea != b->ext_minarea || a.deblend_nthresh != b->deblend_nthresh ||
        a.deblend_mincont != b->deblend_mincont || a.clean != b->clean ||
        a.clean_param != b->clean_param || a.clean_stacksize != b->clean_stacksize ||
        a.satur_level != b->satur_level || a.gain != b->gain || a.mask_width != b->mask_width)
        return false;
    const size_t nm = b->mask && b->mask_width > 0 ? (size_t)b->mask_width * b->mask_width : 0;
    if (nm != mask.size()) return false;
    return nm == 0 || memcmp(mask.data(), b->mask, nm * sizeof(float)) == 0;
}

static int graph_build(const sxg_params *p, int nworkers, int margin)
{
    GraphBatch &G = g_gb;
    const int n = (int)g_batch.size();
    graph_drop();
    G.builds++;
    /* 1. calibration */
    G.plans.assign(n, FramePlan());
    G.calib_ms.assign(n, 0.0);
    std::vector<FramePlan> seen(n);
    auto calib = [&](int i) {
        BatchFrame &f = g_batch[i];
        if (process_frame(f.frame.as<float>(), f.W, f.H, p, nullptr) < 0) return false;
        seen[i] = sxg_state().seen;
        return hipStreamSynchronize(sxg_state().stream) == hipSuccess;
    };
    if (run_pool(n, nworkers, calib, G.calib_ms.data(), "graph batch calibration") < 0) return -1;
    /* 2. plans and branches (longest calibrated frame first, to the least loaded branch) */
    for (int i = 0; i < n; i++) G.plans[i] = plan_from(seen[i], margin);
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return G.calib_ms[a] > G.calib_ms[b]; });
    std::vector<double> load(nworkers, 0.0);
    G.branch.assign(n, 0);
    for (int i : order) {
        const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        G.branch[i] = w;
        load[w] += G.calib_ms[i];
    }
    /* the records and the streams and events of the fork and join */
    if (G.nrec < n) {
        if (G.hrec) (void)hipHostFree(G.hrec);
        G.hrec = nullptr;
        G.nrec = 0;
        SXG_HIP_RET(hipHostMalloc((void **)&G.hrec, sizeof(PlanRecord) * (n > 0 ? n : 1)), -1);
        G.nrec = n;
    }
    if (!G.rec.ensure(sizeof(PlanRecord) * (n > 0 ? n : 1))) return -1;
    if (!G.origin) {
        SXG_HIP_RET(hipStreamCreateWithFlags(&G.origin, hipStreamNonBlocking), -1);
        SXG_HIP_RET(hipEventCreateWithFlags(&G.fork, hipEventDisableTiming), -1);
        int rate = 0;
        SXG_HIP_RET(hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, 0), -1);
        G.clock_khz = rate;
    }
    while ((int)G.join.size() < nworkers) {
        hipEvent_t e;
        SXG_HIP_RET(hipEventCreateWithFlags(&e, hipEventDisableTiming), -1);
        G.join.push_back(e);
    }
    PlanRecord *drec = G.rec.as<PlanRecord>();
    SxgState *prev = sxg_bound_state();
    auto queue_branch = [&](int w) -> bool {
        sxg_ctx_bind(g_workers[w]);
        for (int i = 0; i < n; i++) {
            if (G.branch[i] != w) continue;
            BatchFrame &f = g_batch[i];
            if (process_frame(f.frame.as<float>(), f.W, f.H, p, nullptr, &G.plans[i], drec + i,
                              f.seg.as<int>()) < 0)
                return false;
        }
        return true;
    };
    /* 3. warm-up, uncaptured: sizes every buffer; the planned frames must not overflow
       here (they are the calibrated frames) */
    for (int w = 0; w < nworkers; w++) {
        if (!queue_branch(w) || hipStreamSynchronize(sxg_state().stream) != hipSuccess) {
            if (!sxg_has_error()) sxg_set_error("graph batch: warm-up failed");
            sxg_bind_state(prev);
            return -1;
        }
    }
    sxg_bind_state(prev);
    SXG_HIP_RET(hipMemcpy(G.hrec, drec, sizeof(PlanRecord) * n, hipMemcpyDeviceToHost), -1);
    for (int i = 0; i < n; i++)
        if (G.hrec[i].pd.ovf != 0) {
            const PlanDev &d = G.hrec[i].pd;
            std::string caps;
            for (int k = 0; k < PL_N; k++)
                caps += " " + std::to_string(seen[i].cap[k]) + "/" +
                        std::to_string(G.plans[i].cap[k]);
            sxg_set_error("graph batch: calibrated frame " + std::to_string(i) +
                          " overflowed its plan (internal error): flags " +
                          std::to_string(d.ovf) + ", counts nfg " + std::to_string(d.nfg) +
                          " nraw " + std::to_string(d.nraw) + " npix " + std::to_string(d.npix) +
                          " nobj " + std::to_string(d.nobj) + "; seen/capacity" + caps);
            return -1;
        }
    /* 4. capture */
    SXG_HIP_RET(hipStreamBeginCapture(G.origin, hipStreamCaptureModeRelaxed), -1);
    bool ok = hipEventRecord(G.fork, G.origin) == hipSuccess;
    for (int w = 0; ok && w < nworkers; w++) {
        sxg_ctx_bind(g_workers[w]);
        SxgState &S = sxg_state();
        ok = hipStreamWaitEvent(S.stream, G.fork, 0) == hipSuccess;
        S.capturing = true;
        ok = ok && queue_branch(w);
        S.capturing = false;
        ok = ok && hipEventRecord(G.join[w], S.stream) == hipSuccess &&
             hipStreamWaitEvent(G.origin, G.join[w], 0) == hipSuccess;
    }
    sxg_bind_state(prev);
    ok = ok && hipMemcpyAsync(G.hrec, drec, sizeof(PlanRecord) * n, hipMemcpyDeviceToHost,
                              G.origin) == hipSuccess;
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(G.origin, &graph);
    if (!ok || ec != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        if (!sxg_has_error())
            sxg_set_error(std::string("graph batch: capture failed: ") + hipGetErrorString(ec));
        return -1;
    }
    G.graph = graph;
    SXG_HIP_RET(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0), -1);
    G.gen.assign(nworkers, 0);
    for (int w = 0; w < nworkers; w++) G.gen[w] = sxg_ctx_state(g_workers[w]).alloc_gen;
    G.p = *p;
    G.mask.assign(p->mask && p->mask_width > 0 ? p->mask : nullptr,
                  p->mask && p->mask_width > 0 ? p->mask + (size_t)p->mask_width * p->mask_width
                                               : nullptr);
    G.nworkers = nworkers;
    G.margin = margin;
    for (int k = 0; k < SXG_NOPT; k++) G.opts[k] = sxg_option(k);
    G.built = true;
    return 0;
}

extern "C" int sxg_batch_process_graph(const sxg_params *p, int nworkers, sxg_frame_info *info,
                                       double *frame_ms)
{
    sxg_clear_error();
    GraphBatch &G = g_gb;
    const int n = (int)g_batch.size();
    if (!p) {
        sxg_set_error("sxg_batch_process_graph: no parameters");
        return -1;
    }
    if (nworkers < 1 || nworkers > SXG_MAX_WORKERS) {
        sxg_set_error("sxg_batch_process_graph: nworkers must be 1.." +
                      std::to_string(SXG_MAX_WORKERS));
        return -1;
    }
    if (n == 0) return 0;
    bool fresh = G.built && G.nworkers == nworkers && G.margin == g_plan_margin &&
                 same_params(G.p, p, G.mask) && (int)g_workers.size() >= nworkers;
    for (int w = 0; fresh && w < nworkers; w++)
        fresh = sxg_ctx_state(g_workers[w]).alloc_gen == G.gen[w];
    /* the options choose launch shapes the graph has recorded */
    for (int k = 0; fresh && k < SXG_NOPT; k++) fresh = G.opts[k] == sxg_option(k);
    if (!fresh && graph_build(p, nworkers, g_plan_margin) < 0) return -1;
    SXG_HIP_RET(hipGraphLaunch(G.exec, G.origin), -1);
    SXG_HIP_RET(hipStreamSynchronize(G.origin), -1);
    G.launches++;
    /* frames past their plan run again eagerly on their own context */
    SxgState *prev = sxg_bound_state();
    int rerun = 0;
    for (int i = 0; i < n; i++) {
        const PlanRecord &r = G.hrec[i];
        BatchFrame &f = g_batch[i];
        const int nbx = (f.W - 1) / std::min(p->back_size, f.W) + 1;
        const int nby = (f.H - 1) / std::min(p->back_size, f.W) + 1;
        if (r.pd.ovf == 0) {
            if (info) info_from_record(r, nbx, nby, &info[i]);
            if (frame_ms) frame_ms[i] = (double)(r.t1 - r.t0) / G.clock_khz;
            continue;
        }
        sxg_ctx_bind(g_workers[G.branch[i]]);
        const auto t0 = std::chrono::steady_clock::now();
        sxg_frame_info fi;
        const bool ok = process_frame(f.frame.as<float>(), f.W, f.H, p, &fi, nullptr, nullptr,
                                      f.seg.as<int>()) >= 0 &&
                        hipStreamSynchronize(sxg_state().stream) == hipSuccess;
        sxg_bind_state(prev);
        if (!ok) {
            if (!sxg_has_error()) sxg_set_error("sxg_batch_process_graph: eager rerun failed");
            return -1;
        }
        if (info) info[i] = fi;
        if (frame_ms)
            frame_ms[i] = std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - t0).count();
        rerun++;
    }
    G.reruns += rerun;
    /* frames that outgrew their plans are calibrated again at the next call */
    if (rerun > 0) graph_drop();
    return rerun;
}

extern "C" int sxg_batch_graph_stats(long *builds, long *launches, long *reruns)
{
    if (builds) *builds = g_gb.builds;
    if (launches) *launches = g_gb.launches;
    if (reruns) *reruns = g_gb.reruns;
    return g_gb.built ? 1 : 0;
}

extern "C" int sxg_batch_set_plan_margin(int percent)
{
    if (percent < 0 || percent > 1000) {
        sxg_set_error("sxg_batch_set_plan_margin: 0..1000 percent");
        return -1;
    }
    g_plan_margin = percent;
    return 0;
}

extern "C" int sxg_batch_update_frame(int frame, const float *host)
{
    sxg_clear_error();
    if (frame < 0 || frame >= (int)g_batch.size() || !host) {
        sxg_set_error("sxg_batch_update_frame: bad frame index");
        return -1;
    }
    BatchFrame &f = g_batch[frame];
    /* the slot keeps its address: a built graph reads the new pixels at its next launch */
    SXG_HIP_RET(hipMemcpy(f.frame.ptr, host, (size_t)f.W * f.H * sizeof(float),
                          hipMemcpyHostToDevice), -1);
    return 0;
}

/* -------------------------------------------------------------- autotuner ---- */
/* For the resident batch, the number of graph branches and the deblending size classes
   (the options SXG_OPT_DEB_LDS_TINY/SMALL/LARGE, which route the multi-peak raw objects to
   256-lane LDS, 1024-lane LDS or global-memory trees) that replay fastest, keyed by the
   batch's mean frame size, detected-pixel density and frame count.  Coordinate descent:
   the branch counts with the current classes, then the class candidates with the best
   branch count; each candidate is a graph build and `replays` timed replays.  Every
   candidate's results are compared with the first one's (each frame's counts and a hash
   of its segmentation map): a difference is an error, since the options only choose
   kernels. */
__global__ void k_seg_hash(const int *seg, long n, unsigned long long *out)
{
    unsigned long long h = 0;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (long)gridDim.x * blockDim.x) {
        /* an order-free sum of per-pixel mixes of (position, label) */
        unsigned long long x = ((unsigned long long)p << 32) ^ (unsigned)seg[p];
        x ^= x >> 33;
        x *= 0xff51afd7ed558ccdull;
        x ^= x >> 33;
        x *= 0xc4ceb9fe1a85ec53ull;
        x ^= x >> 33;
        h += x;
    }
    for (int d = 32; d > 0; d >>= 1) h += __shfl_xor(h, d);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, h);
}

struct TuneKey {
    int size_bucket, density_bucket, count_bucket;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(size_bucket, density_bucket, count_bucket) <
               std::tie(o.size_bucket, o.density_bucket, o.count_bucket);
    }
};
static std::map<TuneKey, sxg_tune_choice> g_tune_table;

static int ilog2(double v) { return v >= 1.0 ? (int)std::floor(std::log2(v)) : 0; }

/* the results a candidate must reproduce: per frame its counts and its map's hash */
static bool batch_signature(const sxg_frame_info *info, int n, std::vector<uint64_t> &sig)
{
    DevBuf h;
    if (!h.ensure(sizeof(unsigned long long) * (n > 0 ? n : 1))) return false;
    SXG_HIP_RET(hipMemset(h.ptr, 0, sizeof(unsigned long long) * (n > 0 ? n : 1)), false);
    for (int i = 0; i < n; i++) {
        const BatchFrame &f = g_batch[i];
        hipLaunchKernelGGL(k_seg_hash, dim3(256), dim3(256), 0, 0, f.seg.as<int>(),
                           (long)f.W * f.H, h.as<unsigned long long>() + i);
    }
    std::vector<unsigned long long> hh(n);
    const hipError_t e = hipMemcpy(hh.data(), h.ptr, sizeof(unsigned long long) * n,
                                   hipMemcpyDeviceToHost);
    h.release();
    SXG_HIP_RET(e, false);
    sig.clear();
    for (int i = 0; i < n; i++) {
        sig.push_back(hh[i]);
        sig.push_back((uint64_t)info[i].nraw << 32 | (uint32_t)info[i].nobj);
        sig.push_back((uint64_t)info[i].nsurv << 32 | (uint32_t)info[i].npix_raw);
        sig.push_back((uint64_t)info[i].deb_labels);
    }
    return true;
}

extern "C" int sxg_batch_autotune(const sxg_params *p, int max_branches, int replays,
                                  sxg_tune_choice *out)
{
    sxg_clear_error();
    const int n = (int)g_batch.size();
    if (!p || !out || n == 0 || max_branches < 1 || max_branches > SXG_MAX_WORKERS ||
        replays < 1) {
        sxg_set_error("sxg_batch_autotune: parameters, an output, a resident batch, "
                      "1..16 branches and at least one replay are needed");
        return -1;
    }
    /* the key: mean frame size, detected pixels per pixel (one eager pass), frame count */
    std::vector<int64_t> nfg(n, 0);
    auto probe = [&](int i) {
        BatchFrame &f = g_batch[i];
        if (process_frame(f.frame.as<float>(), f.W, f.H, p, nullptr) < 0) return false;
        nfg[i] = sxg_state().seen.cap[PL_NFG];
        return hipStreamSynchronize(sxg_state().stream) == hipSuccess;
    };
    if (run_pool(n, std::min(max_branches, n), probe, nullptr, "sxg_batch_autotune") < 0)
        return -1;
    double pix = 0.0, det = 0.0;
    for (int i = 0; i < n; i++) {
        pix += (double)g_batch[i].W * g_batch[i].H;
        det += (double)nfg[i];
    }
    const TuneKey key{ilog2(pix / n), ilog2(1e4 * det / pix), ilog2((double)n)};
    auto hit = g_tune_table.find(key);
    if (hit != g_tune_table.end()) {
        *out = hit->second;
        out->candidates = 0;
        sxg_set_option(SXG_OPT_DEB_LDS_TINY, out->deb_lds_tiny);
        sxg_set_option(SXG_OPT_DEB_LDS_SMALL, out->deb_lds_small);
        sxg_set_option(SXG_OPT_DEB_LDS_LARGE, out->deb_lds_large);
        return 0;
    }
    std::vector<sxg_frame_info> info(n);
    std::vector<uint64_t> ref, sig;
    int tried = 0;
    /* the replay time of one candidate (ms per batch), its results checked */
    auto time_candidate = [&](int branches, const long cls[3], double &ms) -> bool {
        sxg_set_option(SXG_OPT_DEB_LDS_TINY, cls[0]);
        sxg_set_option(SXG_OPT_DEB_LDS_SMALL, cls[1]);
        sxg_set_option(SXG_OPT_DEB_LDS_LARGE, cls[2]);
        /* the first call builds the graph for this candidate */
        if (sxg_batch_process_graph(p, branches, info.data(), nullptr) < 0) return false;
        const auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < replays; r++)
            if (sxg_batch_process_graph(p, branches, info.data(), nullptr) < 0) return false;
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                 .count() / replays;
        if (!batch_signature(info.data(), n, sig)) return false;
        if (ref.empty()) {
            ref = sig;
        } else if (sig != ref) {
            sxg_set_error("sxg_batch_autotune: a candidate setting changed a result "
                          "(internal error)");
            return false;
        }
        tried++;
        return true;
    };
    const long cur[3] = {sxg_option(SXG_OPT_DEB_LDS_TINY), sxg_option(SXG_OPT_DEB_LDS_SMALL),
                         sxg_option(SXG_OPT_DEB_LDS_LARGE)};
    sxg_tune_choice best{};
    best.ms_per_batch = 1e300;
    const int bc[] = {1, 2, 4, 8, 16};
    for (int b : bc) {
        if (b > max_branches || b > n) continue;
        double ms;
        if (!time_candidate(b, cur, ms)) return -1;
        if (ms < best.ms_per_batch) {
            best.ms_per_batch = ms;
            best.branches = b;
        }
    }
    /* size classes: tiny <= small <= large <= 6400 pixels */
    static const long cls_cand[][3] = {{512, 1024, 6144}, {256, 1024, 6144}, {1024, 2048, 6400},
                                       {512, 2048, 4096}, {256, 512, 6400}};
    long bestc[3] = {cur[0], cur[1], cur[2]};
    for (const auto &c : cls_cand) {
        if (c[0] == cur[0] && c[1] == cur[1] && c[2] == cur[2]) continue;
        double ms;
        if (!time_candidate(best.branches, c, ms)) return -1;
        if (ms < best.ms_per_batch) {
            best.ms_per_batch = ms;
            for (int k = 0; k < 3; k++) bestc[k] = c[k];
        }
    }
    best.deb_lds_tiny = bestc[0];
    best.deb_lds_small = bestc[1];
    best.deb_lds_large = bestc[2];
    best.candidates = tried;
    best.size_bucket = key.size_bucket;
    best.density_bucket = key.density_bucket;
    best.count_bucket = key.count_bucket;
    sxg_set_option(SXG_OPT_DEB_LDS_TINY, bestc[0]);
    sxg_set_option(SXG_OPT_DEB_LDS_SMALL, bestc[1]);
    sxg_set_option(SXG_OPT_DEB_LDS_LARGE, bestc[2]);
    g_tune_table[key] = best;
    *out = best;
    return 0;
}

extern "C" void sxg_autotune_forget(void) { g_tune_table.clear(); }

extern "C" int sxg_batch_segmap(int frame, int32_t *out)
{
    sxg_clear_error();
    if (frame < 0 || frame >= (int)g_batch.size() || !out) {
        sxg_set_error("sxg_batch_segmap: bad frame index");
        return -1;
    }
    const BatchFrame &f = g_batch[frame];
    SXG_HIP_RET(hipMemcpy(out, f.seg.ptr, (size_t)f.W * f.H * sizeof(int), hipMemcpyDeviceToHost),
                -1);
    return 0;
}

extern "C" void sxg_batch_release(void)
{
    graph_release();
    for (auto &f : g_batch) {
        f.frame.release();
        f.seg.release();
    }
    g_batch.clear();
    for (sxg_ctx *c : g_workers) sxg_ctx_destroy(c);
    g_workers.clear();
}
