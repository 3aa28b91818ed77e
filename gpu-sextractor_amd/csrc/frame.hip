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
#include <cstring>
#include <functional>
#include <thread>
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
    if (!g_kt_on || !sxg_state().is_default) return;
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
    if (!g_kt_on || !sxg_state().is_default) return;
    KTimer &t = g_kt[id];
    if ((int)t.ev.size() < 2 * (t.used + 1)) return;
    (void)hipEventRecord(t.ev[2 * t.used + 1], sxg_state().stream);
    t.used++;
}

/* the same, recorded on stream st (kernels launched on a side stream) */
void sxg_kt_begin_on(int id, hipStream_t st)
{
    if (!g_kt_on || !sxg_state().is_default) return;
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
    if (!g_kt_on || !sxg_state().is_default) return;
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

/* the whole path on a device frame (not modified), on the current context */
static int process_frame(const float *dframe, int W, int H, const sxg_params *p,
                         sxg_frame_info *info)
{
    SxgState &S = sxg_state();
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        sxg_set_error("no HIP device visible");
        return -1;
    }
    if (!sxg_ensure_stream()) return -1;
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
            !S.lab.ensure(N * sizeof(int)) || !S.aux.ensure(N * sizeof(int)))
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
    sxg_back_estimate_from(dframe, bw, p->back_fthresh, &backmean, &backsig,
                           nbx > 1 ? p->back_fsize : 1, nby > 1 ? p->back_fsize : 1,
                           p->back_type);
    if (sxg_has_error()) return -1;
    /* back.c:152-153: with BACK_TYPE MANUAL the frame's backmean is BACK_VALUE */
    if (p->back_type == 1) S.backmean = p->back_value;
    float dthresh, thresh;
    if (p->thresh_type == 1) {
        dthresh = (float)p->detect_thresh;
        thresh = (float)p->analysis_thresh;
    } else {
        dthresh = (float)(p->detect_thresh * backsig);
        thresh = (float)(p->analysis_thresh * backsig);
    }
    init_detection();
    if (p->mask && p->mask_width > 0) cudaFilter(const_cast<float *>(p->mask), p->mask_width);
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
    if (!sxg_build_segmap()) return -1;
    if (info) sxg_get_frame_info(info);
    return nobj;
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
static std::vector<sxg_ctx *> g_workers;

extern "C" int sxg_batch_set_frames(int n, const float *const *host, const int *width,
                                    const int *height)
{
    sxg_clear_error();
    if (n < 0 || (n > 0 && (!host || !width || !height))) {
        sxg_set_error("sxg_batch_set_frames: bad arguments");
        return -1;
    }
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
        if (process_frame(f.frame.as<float>(), f.W, f.H, p, &fi) < 0) return false;
        SxgState &S = sxg_state();
        int *map = sxg_take<int>("out.map", S.N);
        if (!map || hipMemcpyAsync(f.seg.ptr, map, S.N * sizeof(int), hipMemcpyDeviceToDevice,
                                   S.stream) != hipSuccess ||
            hipStreamSynchronize(S.stream) != hipSuccess) {
            sxg_set_error("sxg_batch_process: segmentation copy failed");
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
    for (auto &f : g_batch) {
        f.frame.release();
        f.seg.release();
    }
    g_batch.clear();
    for (sxg_ctx *c : g_workers) sxg_ctx_destroy(c);
    g_workers.clear();
}
