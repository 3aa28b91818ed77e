/*
 * detect.hip -- detection (run_detection, the reference's src/cuda/cudadetection.cu:1107).
 *
 * Semantics (cudadetection.cu:36-64, :156-270, :677-1090): a pixel is detected when
 * cd >= dthresh; detected pixels form 8-connected components labelled by their smallest
 * raster index; components with at least DETECT_MINAREA pixels are the raw objects,
 * numbered in label order; the pixels of each raw object are listed in raster order.
 *
 * Kernels:
 *   k_det_init     one lane per pixel: union-find parent = the first pixel of its run
 *                  inside its 64-pixel window (a ballot), or -1; and the window's 64-bit
 *                  detection mask
 *   (compaction)   detected pixel indices in raster order (the foreground list), from
 *                  the window masks
 *   k_det_union    one lane per detected pixel: lock-free union across window edges and
 *                  with the row above, once per overlapping stretch.  A root is only ever linked under a
 *                  smaller index (atomicMin, retried with the value found), so every
 *                  component ends with its smallest index as root, in one launch, with
 *                  no host convergence loop (the reference loops with a host sync per
 *                  iteration, cudadetection.cu:610-638)
 *   k_det_flatten  label = root; roots remember their foreground position
 *   k_det_rootstat per root: size (atomics), peak of cd, bounding box; every entry's
 *                  root entry (rk), which the root flags and the pixel keys read
 *   (compaction)   roots with size >= DETECT_MINAREA, in raster order = raw objects
 *   (radix sort)   pixels of raw objects grouped by object id, stable, so each object's
 *                  pixels stay in raster order
 */
#include "sxg_internal.h"
#include "sxg_wave.h"

__device__ inline float det_thresh(const WeightArgs &wa, float t, long p)
{
    return sxg_det_thresh(wa, t, p);
}

/* union-find parent of each pixel: -1 for the background; a detected pixel points at the
   first pixel of its run (the detected pixels before it on its row, inside its aligned
   64-pixel window: a ballot), so rows never become chains.  After k_det_flatten the same
   plane is the label map: the root of each detected pixel, -1 elsewhere (one plane
   written per pixel, not two) */
__global__ void __launch_bounds__(256) k_det_init(const float *__restrict__ cd, float t, long n,
                                                  int W, int *par, unsigned long long *fgm,
                                                  WeightArgs wa, const float *td)
{
    if (td) t = *td;
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool live = p < n;
    const bool fg = live && cd[p] >= det_thresh(wa, t, p);
    const int x = live ? (int)(p % W) : 0;
    const unsigned long long F = __ballot(fg);
    const bool start = fg && (x == 0 || lane == 0 || !((F >> (lane - 1)) & 1ull));
    const unsigned long long S = __ballot(start);
    if (!live) return;
    /* the window's detection mask (workgroups start on multiples of 64 positions) */
    if (lane == 0) fgm[p >> 6] = F;
    if (!fg) {
        par[p] = -1;
        return;
    }
    const unsigned long long le = S & (~0ull >> (63 - lane));
    par[p] = (int)(p - lane + (63 - __builtin_clzll(le)));
}

/* root of x; path halving with atomicMin keeps the chains short (every pointer goes
   to a smaller-or-equal index, so the grandparent is a valid replacement) */
__device__ inline int uf_root(int *par, int x)
{
    int p = par[x];
    while (p != x) {
        const int gp = par[p];
        if (gp != p) atomicMin(&par[x], gp);
        x = p;
        p = gp;
    }
    return x;
}

/* link the two trees; every link goes from a larger root to a smaller index */
__device__ inline void uf_unite(int *par, int a, int b)
{
    while (true) {
        a = uf_root(par, a);
        b = uf_root(par, b);
        if (a == b) return;
        if (a < b) {
            int t = a;
            a = b;
            b = t;
        }
        int old = atomicMin(&par[a], b);
        if (old == a) return;
        a = old;
    }
}

/* 8-connected union with the backward neighbours.  W is joined only across a window
   edge (inside a window the run is linked already); NW and NE are already joined through
   N when N is set (they are its W and E neighbours), NW also through W; and N through W
   when W and NW are both set (W's N is this pixel's NW), so the row above is joined once
   per overlapping stretch */
__global__ void __launch_bounds__(256) k_det_union(const int *__restrict__ fg, int nfg, int W,
                                                   int *par, const int *nd)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= sxg_dn(nd, nfg)) return;
    const int p = fg[k];
    const int y = p / W, x = p - y * W;
    /* the four neighbours' flags read together, unconditionally from positions clamped
       into the frame (a detected pixel's parent stays >= 0 while other lanes link
       trees, so reading them first is safe) */
    const bool hw = x > 0, hn = y > 0, he = x < W - 1;
    const int fw = par[hw ? p - 1 : p];
    const int fn = par[hn ? p - W : p];
    const int fnw = par[hn && hw ? p - W - 1 : p];
    const int fne = par[hn && he ? p - W + 1 : p];
    const bool w = hw && fw >= 0;
    const bool nn = hn && fn >= 0;
    const bool nw = hn && hw && fnw >= 0;
    const bool ne = hn && he && fne >= 0;
    if (w && (p & 63) == 0) uf_unite(par, p, p - 1);
    if (nn) {
        if (!(w && nw)) uf_unite(par, p, p - W);
    } else {
        if (nw && !w) uf_unite(par, p, p - W - 1);
        if (ne) uf_unite(par, p, p - W + 1);
    }
}

/* Runs inside a wavefront: the foreground list is in raster order, so a lane whose
   previous lane holds its W neighbour (p - 1 on the same row) is in that lane's
   component.  run_head() gives each active lane the lane of the first pixel of its run
   within the wavefront; a value found by that lane serves the whole run. */
__device__ inline int run_head(bool act, int p, int W)
{
    const int lane = threadIdx.x & 63;
    const int pl = __shfl_up(p, 1);
    const bool cont = act && lane > 0 && pl == p - 1 && p % W != 0;
    const unsigned long long S = __ballot(act && !cont);
    const unsigned long long le = S & (~0ull >> (63 - lane));
    return le ? 63 - __builtin_clzll(le) : lane;
}

struct RootStat {
    int *cnt, *peak, *xmin, *xmax, *ymax, *t0;
};

/* full path compression: every detected pixel's parent becomes its root (a valid parent
   while other lanes still walk: the root is an ancestor with the smallest index, so the
   path halving's atomicMin never moves a pointer past it), and roots remember their
   foreground position.  Only the first pixel of each run in a wavefront walks; the
   others take its root.  The root statistics and the raw-object index of each root's
   entry start here (no fill launches). */
__global__ void __launch_bounds__(256) k_det_flatten(const int *__restrict__ fg, int nfg, int W,
                                                     int *par, int *rootpos, RootStat rs,
                                                     int *rawid_of_pos, const int *nd)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    const bool act = k < sxg_dn(nd, nfg);
    const int p = act ? fg[k] : -1;
    const int h = run_head(act, p, W);
    int r = -1;
    if (act && h == (int)(threadIdx.x & 63)) r = uf_root(par, p);
    r = __shfl(r, h);
    if (!act) return;
    par[p] = r;
    if (r != p) return;
    /* a root: its position, and the statistics and raw-object index of its entry (only
       root entries are ever read: the atomics and k_det_rawinfo address them, the pixel
       keys look up the root entry, the root flags use a count only at a root) */
    rootpos[p] = k;
    rs.cnt[k] = 0;
    rs.peak[k] = (int)0x80000000;   /* below f2ord of every float */
    rs.xmin[k] = 0x7fffffff;
    rs.xmax[k] = -1;
    rs.ymax[k] = -1;
    rs.t0[k] = 0x7fffffff;          /* above f2ord of every float */
    rawid_of_pos[k] = -1;
}


/* per-root size, peak and bounding box, in one pass: the pixel list is in raster order,
   so a run (consecutive positions of one row, all of one root) is a contiguous range of
   lanes; a segmented scan over the run gives its last lane the run's peak (and lowest
   pixel threshold), and that lane issues the run's atomics (integer add, min and max:
   their order does not matter).  The root is the component's smallest position, so the
   smallest row is the root's and needs no atomic (k_det_rawinfo). */
__global__ void __launch_bounds__(256) k_det_rootstat(const int *__restrict__ fg, int nfg, int W,
                                                      const int *__restrict__ lab,
                                                      const int *__restrict__ rootpos,
                                                      const float *__restrict__ cd, RootStat rs,
                                                      int *rkout, WeightArgs wa, const int *nd)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool perpix = wa.w && wa.rel;
    const int n = sxg_dn(nd, nfg);
    const bool act = k < n;
    const int pp = act ? fg[k] : -1;
    /* the root's position, looked up by the first pixel of each run */
    const int h = run_head(act, pp, W);
    int rk = -1;
    if (act && h == lane) rk = rootpos[lab[pp]];
    rk = __shfl(rk, h);
    /* every entry's root entry, for the root flags and the pixel keys */
    if (act) rkout[k] = rk;
    int pv = act ? f2ord(cd[pp]) : (int)0x80000000;
    int tv = (act && perpix) ? f2ord(det_thresh(wa, 0.0f, pp)) : 0x7fffffff;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(pv, d);
        if (lane - d >= h) pv = max(pv, o);
    }
    if (perpix) {
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(tv, d);
            if (lane - d >= h) tv = min(tv, o);
        }
    }
    /* the run's last lane: the next lane starts another run or is past the list */
    const int hn = __shfl_down(h, 1);
    if (!act || (lane < 63 && k + 1 < n && hn == h)) return;
    const int y = pp / W, x1 = pp - y * W, x0 = x1 - (lane - h);
    atomicAdd(&rs.cnt[rk], lane - h + 1);
    atomicMax(&rs.peak[rk], pv);
    atomicMin(&rs.xmin[rk], x0);
    atomicMax(&rs.xmax[rk], x1);
    atomicMax(&rs.ymax[rk], y);
    /* the lowest pixel threshold of the object */
    if (perpix) atomicMin(&rs.t0[rk], tv);
}

/* keep[k] = k for kept roots (size >= minarea), -1 otherwise; entry k is a root when its
   root entry rk[k] is itself */
__global__ void __launch_bounds__(256) k_det_rootflag(const int *__restrict__ rk, int nfg,
                                                      const int *__restrict__ cnt, int minarea,
                                                      int *keep, const int *nd)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= sxg_dn(nd, nfg)) return;
    const int r = rk[k], c = cnt[k];
    keep[k] = (r == k && c >= minarea) ? k : -1;
}

__global__ void __launch_bounds__(256) k_det_rawinfo(const int *__restrict__ rawpos, int nraw,
                                                     const int *__restrict__ fg, int W,
                                                     RootStat rs,
                                                     int *rawid_of_pos, RawObj ro, float t,
                                                     int perpix, const int *nd, const float *td)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= sxg_dn(nd, nraw)) return;
    if (td) t = *td;
    const int k = rawpos[i];
    rawid_of_pos[k] = i;
    ro.first[i] = fg[k];
    ro.count[i] = rs.cnt[k];
    ro.fdpeak[i] = ord2f(rs.peak[k]);
    ro.xmin[i] = rs.xmin[k];
    ro.xmax[i] = rs.xmax[k];
    ro.ymin[i] = fg[k] / W;   /* the root: the component's first pixel in raster order */
    ro.ymax[i] = rs.ymax[k];
    ro.t0[i] = perpix ? ord2f(rs.t0[k]) : t;
}


/* object id of each detected pixel (or -1 when its component is too small) */
__global__ void __launch_bounds__(256) k_det_pixkey(const int *__restrict__ rk, int nfg,
                                                    const int *__restrict__ rawid_of_pos,
                                                    int *key, const int *nd)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= sxg_dn(nd, nfg)) return;
    key[k] = rawid_of_pos[rk[k]];
}

__global__ void __launch_bounds__(256) k_det_gather(const int *__restrict__ idx, int n,
                                                    const int *__restrict__ key,
                                                    const int *__restrict__ fg, int *k0, int *v0,
                                                    const int *nd)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= sxg_dn(nd, n)) return;
    const int k = idx[j];
    k0[j] = key[k];
    v0[j] = fg[k];
}

__global__ void __launch_bounds__(256) k_det_copy(const float *a, float *b, long n)
{
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    if (k < n) b[k] = a[k];
}

static int read_count(const int *d, hipStream_t s)
{
    int h = 0;
    if (hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        sxg_set_error("run_detection: count readback failed");
        return -1;
    }
    return h;
}

extern "C" void run_detection(float dthreshold, float thresh1, int ext_minarea,
                              size_t *numValidObj, size_t *numValidPix)
{
    SxgState &S = sxg_state();
    if (!S.have_frame) {
        sxg_set_error("run_detection: no frame");
        return;
    }
    /* a planned run reads the thresholds on the device (PlanDev) and its counts stay
       there: every size below is then the plan's capacity */
    PlanDev *pd = S.plan ? S.pdev : nullptr;
    if (!pd && !(dthreshold > 0.0f)) {
        /* back.c:195-197 refuses zero or negative thresholds */
        sxg_set_error("run_detection: the detection threshold must be > 0");
        return;
    }
    const int W = S.W;
    const long N = (long)S.N;
    hipStream_t st = S.stream;
    S.dthresh = dthreshold;
    S.thresh = thresh1;
    S.ext_minarea = ext_minarea;
    if (!S.filtered) {
        /* FILTER N: detection on the subtracted frame itself */
        hipLaunchKernelGGL(k_det_copy, dim3(sxg_div_up(N, 256)), dim3(256), 0, st,
                           S.img.as<float>(), S.cd.as<float>(), N);
    }
    sxg_stage_begin(ST_DETECT);
    /* par: the union-find parents, then (after k_det_flatten) the labels */
    int *par = S.aux.as<int>(), *lab = par;
    int *rootpos = sxg_take<int>("det.rootpos", S.N);
    int *cnt_d = sxg_take<int>("det.counts", 8);
    if (!rootpos || !cnt_d) return;
    const float *td = pd ? &pd->dthresh : nullptr;
    int *nfg_d = pd ? &pd->nfg : cnt_d, *nraw_d = pd ? &pd->nraw : cnt_d;
    int *npix_d = pd ? &pd->npix : cnt_d;
    /* the device counts the kernels read (planned runs) */
    const int *nfg_k = pd ? nfg_d : nullptr, *nraw_k = pd ? nraw_d : nullptr;
    const int *npix_k = pd ? npix_d : nullptr;
    WeightArgs wa;
    if (!sxg_weight_args(wa)) return;
    /* the detection masks of the 64-position windows, written with the parents */
    unsigned long long *fgm = sxg_take<unsigned long long>("det.fgmask", sxg_div_up(N, 64));
    if (!fgm) return;
    /* the parents are written already when the frame's convolution was fused with them */
    if (!S.det_init_done) {
        sxg_kt_begin(SXG_KT_DET_INIT);
        hipLaunchKernelGGL(k_det_init, dim3(sxg_div_up(N, 256)), dim3(256), 0, st,
                           S.cd.as<float>(), dthreshold, N, W, par, fgm, wa, td);
        sxg_kt_end(SXG_KT_DET_INIT);
    }
    S.det_init_done = false;
    /* the foreground list has at most N entries; size it after counting */
    int *fg = sxg_take<int>("det.fg", S.N);
    if (!fg) return;
    sxg_kt_begin(SXG_KT_DET_FG);
    sxg_compact_mask(fgm, sxg_div_up(N, 64), fg, nfg_d, st);
    sxg_kt_end(SXG_KT_DET_FG);
    int nfg;
    if (pd) {
        nfg = (int)S.plan->cap[PL_NFG];
        sxg_plan_guard(nfg_d, nfg, PL_NFG);
    } else {
        nfg = read_count(cnt_d, st);
        if (nfg < 0) return;
    }
    S.npix_above = nfg;
    const int gfg = sxg_div_up(nfg > 0 ? nfg : 1, 256);
    RootStat rs;
    rs.cnt = sxg_take<int>("det.cnt", nfg);
    rs.peak = sxg_take<int>("det.peak", nfg);
    rs.xmin = sxg_take<int>("det.xmin", nfg);
    rs.xmax = sxg_take<int>("det.xmax", nfg);
    rs.ymax = sxg_take<int>("det.ymax", nfg);
    rs.t0 = sxg_take<int>("det.t0", nfg);
    int *keep = sxg_take<int>("det.keep", nfg);
    int *rawpos = sxg_take<int>("det.rawpos", nfg);
    int *rawid_of_pos = sxg_take<int>("det.rawid", nfg);
    int *rk = sxg_take<int>("det.rk", nfg);
    if (!rs.cnt || !rs.peak || !rs.xmin || !rs.xmax || !rs.ymax || !rs.t0 || !keep ||
        !rawpos || !rawid_of_pos || !rk)
        return;
    if (nfg > 0) {
        sxg_kt_begin(SXG_KT_DET_UNION);
        hipLaunchKernelGGL(k_det_union, dim3(gfg), dim3(256), 0, st, fg, nfg, W, par, nfg_k);
        sxg_kt_end(SXG_KT_DET_UNION);
        sxg_kt_begin(SXG_KT_DET_FLATTEN);
        hipLaunchKernelGGL(k_det_flatten, dim3(gfg), dim3(256), 0, st, fg, nfg, W, par,
                           rootpos, rs, rawid_of_pos, nfg_k);
        sxg_kt_end(SXG_KT_DET_FLATTEN);
        sxg_kt_begin(SXG_KT_DET_STAT);
        hipLaunchKernelGGL(k_det_rootstat, dim3(gfg), dim3(256), 0, st, fg, nfg, W, lab, rootpos,
                           S.cd.as<float>(), rs, rk, wa, nfg_k);
        sxg_kt_end(SXG_KT_DET_STAT);
        hipLaunchKernelGGL(k_det_rootflag, dim3(gfg), dim3(256), 0, st, rk, nfg, rs.cnt,
                           ext_minarea, keep, nfg_k);
    }
    sxg_compact_nonneg(keep, rawpos, nfg, nraw_d, st, nfg_k);
    int nraw;
    if (pd) {
        nraw = (int)S.plan->cap[PL_NRAW];
        sxg_plan_guard(nraw_d, nraw, PL_NRAW);
    } else {
        nraw = read_count(cnt_d, st);
        if (nraw < 0) return;
    }
    S.nraw = nraw;
    RawObj ro;
    if (!sxg_raw_alloc(nraw, ro)) return;
    if (nfg > 0) {
        if (nraw > 0)
            hipLaunchKernelGGL(k_det_rawinfo, dim3(sxg_div_up(nraw, 256)), dim3(256), 0, st, rawpos,
                               nraw, fg, W, rs, rawid_of_pos, ro, dthreshold,
                               (int)(wa.w && wa.rel),
                               nraw_k, td);
        /* pixel keys, kept pixels, stable sort by object */
        hipLaunchKernelGGL(k_det_pixkey, dim3(gfg), dim3(256), 0, st, rk, nfg, rawid_of_pos,
                           keep, nfg_k);
    }
    int *idx = rawpos;   /* reuse: indices of kept pixels in the foreground list */
    int *k0 = sxg_take<int>("det.k0", nfg), *v0 = sxg_take<int>("det.v0", nfg);
    int *k1 = sxg_take<int>("det.k1", nfg), *v1 = sxg_take<int>("det.v1", nfg);
    int *hist = sxg_take<int>("det.hist", 256 * (size_t)(sxg_div_up(nfg + 1, 4096) + 1) + 64);
    if (!k0 || !v0 || !k1 || !v1 || !hist) return;
    sxg_compact_nonneg(keep, idx, nfg, npix_d, st, nfg_k);
    int npix;
    if (pd) {
        npix = (int)S.plan->cap[PL_NPIX];
        sxg_plan_guard(npix_d, npix, PL_NPIX);
    } else {
        npix = read_count(cnt_d, st);
        if (npix < 0) return;
    }
    S.npix_raw = npix;
    if (!pd) {
        S.seen.cap[PL_NFG] = nfg;
        S.seen.cap[PL_NRAW] = nraw;
        S.seen.cap[PL_NPIX] = npix;
    }
    if (npix > 0) {
        hipLaunchKernelGGL(k_det_gather, dim3(sxg_div_up(npix, 256)), dim3(256), 0, st, idx, npix,
                           keep, fg, k0, v0, npix_k);
        /* enough key bits for every object id (for the capacity in a planned run: more
           passes over zero digits leave a stable sort's order as it is) */
        int nbits = 0;
        while ((1L << nbits) < (long)nraw) nbits++;
        if (nbits == 0) nbits = 1;
        sxg_kt_begin(SXG_KT_DET_SORT);
        int where = sxg_radix_sort(k0, v0, k1, v1, npix, nbits, hist, st, npix_k);
        sxg_kt_end(SXG_KT_DET_SORT);
        int *ks = where ? k1 : k0, *vs = where ? v1 : v0;
        S.fpia_keys = ks;
        S.fpia = vs;
        /* offsets of each object's pixel run */
        sxg_scan_excl(ro.count, ro.off, nraw, nullptr, st, nraw_k);
    } else {
        S.fpia_keys = k0;
        S.fpia = v0;
    }
    sxg_stage_end(ST_DETECT);
    /* the counts were read above; the sort and scan queued since run on the stream the
       next entry points use */
    SXG_HIP(hipGetLastError());
    if (!sxg_stage_wait()) return;
    if (numValidObj) *numValidObj = (size_t)nraw;
    if (numValidPix) *numValidPix = (size_t)npix;
}

extern "C" void clear_detection(void)
{
    SxgState &S = sxg_state();
    if (S.stream) (void)hipStreamSynchronize(S.stream);
}
