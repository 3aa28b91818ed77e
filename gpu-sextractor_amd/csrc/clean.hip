/*
 * clean.hip -- CLEAN (run_clean, the reference's src/cuda/cudaclean.cu:260), catalogue
 * rows are built by photom.hip (endobject_gpu); here are the outputs the
 * reference never wrote: the SEGMENTATION map and per-object columns.
 *
 * CLEAN semantics (cudaclean.cu:10-29, :120-174): amp = fdflux/(2 pi a b abcor),
 * alpha = (pow(amp/dthresh, 1/CLEAN_PARAM) - 1) pi a b / fdnpix; for every pair (i, j)
 * with 0 < j - i <= 16*ceil(MEMORY_OBJSTACK/16) whose centres are closer than
 * CLEAN_ZONE (a_i + a_j), the fainter object is a victim of the brighter when the
 * brighter's profile amp/val, val = 1 + alpha (cxx dx^2 + cyy dy^2 + cxy dx dy) > 1,
 * exceeds the fainter's mthresh; its master is the smallest such brighter index.
 * Victims are dropped (the reference disabled merging, :311-368).
 *
 * Layout: one lane per object i scans its window of following objects, staged through
 * LDS per workgroup of 256 objects; masters are combined with atomicMin, so the result
 * does not depend on scheduling.  Every 64 consecutive objects have a bounding box of
 * their centres and their largest a (k_clean_prep); a wavefront skips the 64 staged
 * objects whose box is farther from the box of its own 64 than CLEAN_ZONE times the sum
 * of the two largest a, with a margin above the rounding of the pair test, since no
 * pair of them can pass it (the objects follow the raster order of their roots, so most
 * of a window is far away).
 */
#include <cmath>
#include <algorithm>
#include <cstring>
#include <vector>

#include "sxg_internal.h"
#include "sxg_objects.h"

#define CLEAN_ZONE 10.0

/* the box of each group of 64 objects: min and max of mx and my, the largest a */
struct CleanBox {
    double x0, x1, y0, y1, amax;
};

__device__ inline double wave_dmax(double v)
{
    for (int d = 32; d > 0; d >>= 1) v = fmax(v, __shfl_xor(v, d));
    return v;
}

__global__ void __launch_bounds__(256) k_clean_prep(ObjCols o, int nobj, double clean_param,
                                                    double *amp, double *alpha, unsigned *master,
                                                    CleanBox *box, const int *nd)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool live = i < sxg_dn(nd, nobj);
    double x = -INFINITY, y = -INFINITY, nx = -INFINITY, ny = -INFINITY, am = -INFINITY;
    if (live) {
        const float unitarea = (float)(SXG_PI * o.a[i] * o.b[i]);
        const double ampin = amp[i] = o.fdflux[i] / (2 * unitarea * o.abcor[i]);
        alpha[i] = (sx_pow(ampin / o.dthresh[i], 1.0 / clean_param) - 1) * unitarea / o.fdnpix[i];
        master[i] = 0xffffffffu;
        x = o.mx[i];
        y = o.my[i];
        nx = -x;
        ny = -y;
        am = o.a[i];
    }
    /* the group is the wavefront: 64 consecutive objects */
    x = wave_dmax(x);
    y = wave_dmax(y);
    nx = wave_dmax(nx);
    ny = wave_dmax(ny);
    am = wave_dmax(am);
    if ((threadIdx.x & 63) == 0 && i < sxg_dn(nd, nobj)) {
        CleanBox b;
        b.x0 = -nx; b.x1 = x; b.y0 = -ny; b.y1 = y; b.amax = am;
        box[i >> 6] = b;
    }
}

/* no pair of the two groups of objects can pass the CLEAN_ZONE test: the squared
   distance of their boxes exceeds (a_i + a_j)^2 CLEAN_ZONE^2 for the largest a of each by
   more than the pair test's float rounding (a few 1e-7 relative) */
__device__ inline bool clean_far(const CleanBox &p, const CleanBox &q)
{
    const double dx = fmax(0.0, fmax(q.x0 - p.x1, p.x0 - q.x1));
    const double dy = fmax(0.0, fmax(q.y0 - p.y1, p.y0 - q.y1));
    const double r = (p.amax + q.amax) * CLEAN_ZONE;
    return dx * dx + dy * dy > r * r * 1.0001 + 1e-9;
}

/* one workgroup per 256 consecutive objects i; the window of following objects is
   staged through LDS 256 at a time and read by broadcast */
#define CP_T 256
__global__ void __launch_bounds__(CP_T) k_clean_pairs(ObjCols o, int nobj, int win,
                                                      const double *__restrict__ amp,
                                                      const double *__restrict__ alpha,
                                                      unsigned *master,
                                                      const CleanBox *__restrict__ box,
                                                      const int *nd)
{
    __shared__ double smx[CP_T], smy[CP_T], samp[CP_T], salpha[CP_T];
    __shared__ float sa[CP_T], sf[CP_T], scxx[CP_T], scyy[CP_T], scxy[CP_T], smt[CP_T];
    const int t = threadIdx.x;
    const int i0 = blockIdx.x * CP_T;
    nobj = sxg_dn(nd, nobj);
    if (i0 >= nobj) return;
    const int i = i0 + t;
    const bool live = i < nobj;
    const int ii = live ? i : nobj - 1;
    const double mxi = o.mx[ii], myi = o.my[ii];
    const float ai = o.a[ii], fi = o.fdflux[ii], mti = o.mthresh[ii];
    const float cxxi = o.cxx[ii], cyyi = o.cyy[ii], cxyi = o.cxy[ii];
    const double ampi = amp[ii], alphai = alpha[ii];
    const int jmax = live ? min(i + win, nobj - 1) : -1;
    unsigned mine = 0xffffffffu;
    const int jlast = min(i0 + CP_T - 1 + win, nobj - 1);
    /* the box of this wavefront's 64 objects (a wavefront past the last object reads
       the last group's box and tests nothing) */
    const CleanBox bi = box[min((i0 >> 6) + (t >> 6), (nobj - 1) >> 6)];
    /* the window's tiles are shared out over gridDim.y workgroups */
    for (int j0 = i0 + 1 + blockIdx.y * CP_T; j0 <= jlast; j0 += CP_T * gridDim.y) {
        __syncthreads();
        const int jl = j0 + t;
        if (jl <= jlast) {
            smx[t] = o.mx[jl];
            smy[t] = o.my[jl];
            samp[t] = amp[jl];
            salpha[t] = alpha[jl];
            sa[t] = o.a[jl];
            sf[t] = o.fdflux[jl];
            scxx[t] = o.cxx[jl];
            scyy[t] = o.cyy[jl];
            scxy[t] = o.cxy[jl];
            smt[t] = o.mthresh[jl];
        }
        __syncthreads();
        for (int s0 = j0; s0 < j0 + CP_T && s0 <= jlast; s0 += 64) {
            /* the 64 staged objects from s0: the union of the two boxes covering them */
            const int s1 = min(s0 + 63, jlast);
            CleanBox bj = box[s0 >> 6];
            if ((s1 >> 6) != (s0 >> 6)) {
                const CleanBox b2 = box[s1 >> 6];
                bj.x0 = fmin(bj.x0, b2.x0); bj.x1 = fmax(bj.x1, b2.x1);
                bj.y0 = fmin(bj.y0, b2.y0); bj.y1 = fmax(bj.y1, b2.y1);
                bj.amax = fmax(bj.amax, b2.amax);
            }
            if (clean_far(bi, bj)) continue;
            const int jb = max(s0, i + 1), je = min(s1, jmax);
            for (int j = jb; j <= je; j++) {
                const int q = j - j0;
                const float dx = (float)(mxi - smx[q]);
                const float dy = (float)(myi - smy[q]);
                float rlim = ai + sa[q];
                rlim *= rlim;
                if (dx * dx + dy * dy < rlim * CLEAN_ZONE * CLEAN_ZONE) {
                    double val;
                    if (fi < sf[q]) {
                        val = 1 + salpha[q] * (scxx[q] * dx * dx + scyy[q] * dy * dy + scxy[q] * dx * dy);
                        if (val > 1.0 && ((float)(val < 1e10 ? samp[q] * (1.0 / val) : 0.0) > mti))
                            mine = min(mine, (unsigned)j);
                    } else {
                        val = 1 + alphai * (cxxi * dx * dx + cyyi * dy * dy + cxyi * dx * dy);
                        if (val > 1.0 && ((float)(val < 1e10 ? ampi * (1.0 / val) : 0.0) > smt[q]))
                            atomicMin(&master[j], (unsigned)i);
                    }
                }
            }
        }
    }
    if (live && mine != 0xffffffffu) atomicMin(&master[i], mine);
}

extern "C" void run_clean(int *masterIndex, double clean_param, int clean_stacksize, int numobj)
{
    SxgState &S = sxg_state();
    if (!S.have_frame || numobj != S.nobj) {
        sxg_set_error("run_clean: numobj differs from parcel_out");
        return;
    }
    if (clean_stacksize < 1 || !(clean_param > 0.0)) {
        sxg_set_error("run_clean: MEMORY_OBJSTACK and CLEAN_PARAM must be positive");
        return;
    }
    ObjCols o;
    if (!sxg_obj_view(o)) return;
    const int n = numobj;   /* a capacity in a planned run (the count is on the device) */
    const int *nd = S.plan ? &S.pdev->nobj : nullptr;
    double *amp = sxg_take<double>("clean.amp", n), *alpha = sxg_take<double>("clean.alpha", n);
    CleanBox *box = sxg_take<CleanBox>("clean.box", (size_t)sxg_div_up(n > 0 ? n : 1, 64));
    if (!amp || !alpha || !box) return;
    sxg_stage_begin(ST_CLEAN);
    if (n > 0) {
        hipLaunchKernelGGL(k_clean_prep, dim3(sxg_div_up(n, 256)), dim3(256), 0, S.stream, o, n,
                           clean_param, amp, alpha, o.master, box, nd);
        /* CLEAN N: keep everything (the reference cleans regardless of the flag) */
        if (S.clean_enabled) {
            const int win = ((clean_stacksize - 1) / 16 + 1) * 16;
            sxg_kt_begin(SXG_KT_CLEAN);
            /* enough workgroups for every CU: split each window into up to 8 slices */
            const int ntile = (win + CP_T - 1) / CP_T + 1;
            const int slices = std::max(1, std::min(8, ntile));
            hipLaunchKernelGGL(k_clean_pairs, dim3(sxg_div_up(n, 256), slices), dim3(256), 0, S.stream, o,
                               n, win, amp, alpha, o.master, box, nd);
            sxg_kt_end(SXG_KT_CLEAN);
        }
    }
    sxg_stage_end(ST_CLEAN);
    S.cleaned = true;
    /* the catalogue's graded columns in the chosen precision (after CLEAN: the
       segmentation map and the CLEAN decisions keep the reference's arithmetic) */
    if (!sxg_precision_columns((int)sxg_option(SXG_OPT_PRECISION))) return;
    SXG_HIP(hipGetLastError());
    /* the only host output is masterIndex: wait only when it is asked for */
    if (masterIndex && n > 0) {
        SXG_HIP(hipMemcpyAsync(masterIndex, o.master, sizeof(unsigned) * n, hipMemcpyDeviceToHost,
                               S.stream));
        SXG_HIP(hipStreamSynchronize(S.stream));
    } else {
        sxg_stage_wait();
    }
}

/* ------------------------------------------------------------- segmentation ---- */
__global__ void __launch_bounds__(256) k_keepflag(const unsigned *master, int n, int *keep,
                                                  const int *nd)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < sxg_dn(nd, n)) keep[i] = master[i] == 0xffffffffu ? 1 : 0;
}

/* a[0..n) = v, four per lane with 16-byte stores (a is 16-byte aligned) */
__global__ void __launch_bounds__(256) k_fill_i32(int *a, long n, int v)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x, k = 4 * q;
    if (k + 4 <= n)
        *reinterpret_cast<int4 *>(a + k) = make_int4(v, v, v, v);
    else
        for (long j = k; j < n; j++) a[j] = v;
}

/* seg[p] = 1-based rank of the surviving object of p (0 for cleaned objects) */
__global__ void __launch_bounds__(256) k_seg_scatter(const int *fpia, long P0, const int *alloc,
                                                     const int *keep, const int *rank, int *seg,
                                                     int mode, const int *nd)
{
    const long g = (long)blockIdx.x * 256 + threadIdx.x;
    if (nd) P0 = *nd;
    if (g >= P0) return;
    const int p = fpia[g];
    const int o = alloc[p];
    seg[p] = mode ? o : (keep[o] ? rank[o] + 1 : 0);
}

/* the survivors of CLEAN: keep[i], rank[i] (catalogue NUMBER - 1 of a survivor) and their
   number in tot[0] (in a planned run the device's nsurv), all on the device (queued on
   the context's stream) */
static bool survivor_ranks(const ObjCols &o, int *&keep, int *&rank, int *&tot)
{
    SxgState &S = sxg_state();
    keep = sxg_take<int>("out.keep", S.nobj);
    rank = sxg_take<int>("out.rank", S.nobj);
    tot = S.plan ? &S.pdev->nsurv : sxg_take<int>("out.total", 4);
    if (!keep || !rank || !tot) return false;
    const int *nd = S.plan ? &S.pdev->nobj : nullptr;
    hipLaunchKernelGGL(k_keepflag, dim3(sxg_div_up(S.nobj, 256)), dim3(256), 0, S.stream,
                       o.master, S.nobj, keep, nd);
    sxg_scan_excl(keep, rank, S.nobj, tot, S.stream, nd);
    return true;
}

/* 0: segmentation map, 1: object index map; into dst (N ints) or, without one, the
   device buffer "out.map" */
static int *build_map(int mode, int *dst = nullptr)
{
    SxgState &S = sxg_state();
    int *map = dst ? dst : sxg_take<int>("out.map", S.N);
    if (!map) return nullptr;
    const long N = (long)S.N;
    hipLaunchKernelGGL(k_fill_i32, dim3(sxg_div_up(sxg_div_up(N, 4), 256)), dim3(256), 0,
                       S.stream, map, N, mode ? -1 : 0);
    if (S.nobj == 0 || S.npix_raw == 0) {
        if (mode == 0) S.nsurv = 0;
        if (mode == 0 && S.plan)
            hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, S.stream, &S.pdev->nsurv, 1L,
                               0);
        return map;
    }
    ObjCols o;
    if (!sxg_obj_view(o)) return nullptr;
    int *keep = nullptr, *rank = nullptr, *tot = nullptr;
    if (mode == 0) {
        if (!survivor_ranks(o, keep, rank, tot)) return nullptr;
    } else {
        keep = sxg_take<int>("out.keep", S.nobj);
        rank = sxg_take<int>("out.rank", S.nobj);
        if (!keep || !rank) return nullptr;
    }
    sxg_kt_begin(SXG_KT_SEG);
    hipLaunchKernelGGL(k_seg_scatter, dim3(sxg_div_up(S.npix_raw, 256)), dim3(256), 0, S.stream,
                       S.fpia, (long)S.npix_raw, sxg_take<int>("deb.alloc", S.N), keep, rank, map,
                       mode, (const int *)(S.plan ? &S.pdev->npix : nullptr));
    sxg_kt_end(SXG_KT_SEG);
    /* a planned run leaves the survivor count on the device */
    if (mode == 0 && !S.plan) {
        int ns = 0;
        if (hipMemcpyAsync(&ns, tot, sizeof(int), hipMemcpyDeviceToHost, S.stream) == hipSuccess &&
            hipStreamSynchronize(S.stream) == hipSuccess)
            S.nsurv = ns;
    }
    return map;
}

bool sxg_build_segmap(int *dst) { return build_map(0, dst) != nullptr; }

extern "C" int sxg_get_segmap(int32_t *out)
{
    SxgState &S = sxg_state();
    if (!S.have_frame || !S.cleaned) {
        sxg_set_error("sxg_get_segmap: call after run_clean");
        return -1;
    }
    int *map = build_map(0);
    if (!map) return -1;
    SXG_HIP_RET(hipMemcpyAsync(out, map, sizeof(int) * S.N, hipMemcpyDeviceToHost, S.stream), -1);
    SXG_HIP_RET(hipStreamSynchronize(S.stream), -1);
    return S.nsurv;
}

extern "C" int sxg_get_alloc_map(int32_t *out)
{
    SxgState &S = sxg_state();
    if (!S.have_frame) return -1;
    int *map = build_map(1);
    if (!map) return -1;
    SXG_HIP_RET(hipMemcpyAsync(out, map, sizeof(int) * S.N, hipMemcpyDeviceToHost, S.stream), -1);
    SXG_HIP_RET(hipStreamSynchronize(S.stream), -1);
    return 0;
}

/* ---------------------------------------------------------- catalogue rows ---- */
/* one row per survivor of CLEAN, at its catalogue NUMBER - 1 */
__global__ void __launch_bounds__(256) k_isorows(ObjCols o, int nobj, const int *keep,
                                                 const int *rank, sxg_isorow *rows)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nobj || !keep[i]) return;
    sxg_isorow r;
    r.number = rank[i] + 1;
    r.flags = o.flag[i];
    r.isoarea = o.npix[i];
    r.fdnpix = o.fdnpix[i];
    /* X_IMAGE, Y_IMAGE are 1-based (cudaanalyse.cu:2715-2716) */
    r.x = o.mx[i] + 1.0;
    r.y = o.my[i] + 1.0;
    r.x2 = o.mx2[i];
    r.y2 = o.my2[i];
    r.xy = o.mxy[i];
    r.a = o.a[i];
    r.b = o.b[i];
    r.theta = o.theta[i];
    r.cxx = o.cxx[i];
    r.cyy = o.cyy[i];
    r.cxy = o.cxy[i];
    r.flux_iso = o.flux[i];
    r.fluxerr_iso = sqrtf(o.fluxerr[i]);
    r.peak = o.peak[i];
    r.bkg = o.bkg[i];
    rows[rank[i]] = r;
}

/* Queue the rows of the survivors and their copy to out[0..min(nsurv, cap)) on the
   context's stream (S.nsurv must be current: after the segmentation map).  Returns the
   number of rows copied or -1. */
int sxg_catalog_async(sxg_isorow *out, int cap)
{
    SxgState &S = sxg_state();
    const int n = S.nsurv < cap ? S.nsurv : cap;
    if (n <= 0 || S.nobj <= 0) return 0;
    ObjCols o;
    if (!sxg_obj_view(o)) return -1;
    int *keep = nullptr, *rank = nullptr, *tot = nullptr;
    if (!survivor_ranks(o, keep, rank, tot)) return -1;
    sxg_isorow *rows = sxg_take<sxg_isorow>("out.rows", (size_t)S.nsurv);
    if (!rows) return -1;
    hipLaunchKernelGGL(k_isorows, dim3(sxg_div_up(S.nobj, 256)), dim3(256), 0, S.stream, o,
                       S.nobj, keep, rank, rows);
    SXG_HIP_RET(hipGetLastError(), -1);
    SXG_HIP_RET(hipMemcpyAsync(out, rows, sizeof(sxg_isorow) * n, hipMemcpyDeviceToHost,
                               S.stream), -1);
    return n;
}

extern "C" int sxg_get_catalog(sxg_isorow *out, int cap)
{
    SxgState &S = sxg_state();
    sxg_clear_error();
    if (!S.cleaned || (cap > 0 && !out) || cap < 0) {
        sxg_set_error("sxg_get_catalog: call after run_clean, with an output buffer");
        return -1;
    }
    if (S.nobj > 0) {
        /* the number of survivors, read back once */
        ObjCols o;
        int *keep = nullptr, *rank = nullptr, *tot = nullptr;
        if (!sxg_obj_view(o) || !survivor_ranks(o, keep, rank, tot)) return -1;
        int ns = 0;
        SXG_HIP_RET(hipMemcpyAsync(&ns, tot, sizeof(int), hipMemcpyDeviceToHost, S.stream), -1);
        SXG_HIP_RET(hipStreamSynchronize(S.stream), -1);
        S.nsurv = ns;
    } else {
        S.nsurv = 0;
    }
    if (sxg_catalog_async(out, cap) < 0) return -1;
    SXG_HIP_RET(hipStreamSynchronize(S.stream), -1);
    return S.nsurv;
}

/* ------------------------------------------------------------------ columns ---- */
bool sxg_obj_alloc(int nobj, ObjCols &o)
{
    const size_t n = (size_t)(nobj > 0 ? nobj : 1);
#define OI(f) o.f = sxg_take<int>("obj." #f, n)
#define OD(f) o.f = sxg_take<double>("obj." #f, n)
#define OF(f) o.f = sxg_take<float>("obj." #f, n)
    OI(level); OI(root); OI(firstpix); OI(flag); OI(fdnpix); OI(dnpix); OI(npix); OI(xmin);
    OI(xmax); OI(ymin); OI(ymax); OI(peakx); OI(peaky);
    o.iso = sxg_take<int>("obj.iso", n * SXG_NISO);
    OD(mx); OD(my); OD(mx2); OD(my2); OD(mxy); OD(poserr_mx2); OD(poserr_my2); OD(poserr_mxy);
    OF(a); OF(b); OF(theta); OF(cxx); OF(cyy); OF(cxy); OF(abcor); OF(fdflux); OF(dflux);
    OF(flux); OF(fluxerr); OF(fdpeak); OF(dpeak); OF(peak); OF(bkg); OF(thresh); OF(mthresh);
    OF(fwhm); OF(dthresh);
    o.master = sxg_take<unsigned>("obj.master", n);
#undef OI
#undef OD
#undef OF
    void *all[] = {o.level, o.root, o.firstpix, o.flag, o.fdnpix, o.dnpix, o.npix, o.xmin, o.xmax,
                   o.ymin, o.ymax, o.peakx, o.peaky, o.iso, o.mx, o.my, o.mx2, o.my2, o.mxy,
                   o.poserr_mx2, o.poserr_my2, o.poserr_mxy, o.a, o.b, o.theta, o.cxx, o.cyy,
                   o.cxy, o.abcor, o.fdflux, o.dflux, o.flux, o.fluxerr, o.fdpeak, o.dpeak,
                   o.peak, o.bkg, o.thresh, o.mthresh, o.fwhm, o.dthresh, o.master};
    for (void *p : all)
        if (!p) return false;
    return true;
}

bool sxg_obj_view(ObjCols &o) { return sxg_obj_alloc(sxg_state().nobj, o); }

extern "C" int sxg_get_column(int column, void *out, size_t bytes)
{
    SxgState &S = sxg_state();
    ObjCols o;
    if (column < 0 || column >= SXG_NCOL || !sxg_obj_view(o)) {
        sxg_set_error("sxg_get_column: bad column");
        return -1;
    }
    const void *src[SXG_NCOL] = {o.level, o.root, o.firstpix, o.flag, o.fdnpix, o.dnpix, o.npix,
                                 o.xmin, o.xmax, o.ymin, o.ymax, o.peakx, o.peaky, o.iso, o.mx,
                                 o.my, o.mx2, o.my2, o.mxy, o.poserr_mx2, o.poserr_my2,
                                 o.poserr_mxy, o.a, o.b, o.theta, o.cxx, o.cyy, o.cxy, o.abcor,
                                 o.fdflux, o.dflux, o.flux, o.fluxerr, o.fdpeak, o.dpeak, o.peak,
                                 o.bkg, o.thresh, o.mthresh, o.fwhm, o.master};
    size_t esize = (column >= SXG_COL_MX && column <= SXG_COL_POSERR_MXY) ? 8 : 4;
    size_t need = esize * (size_t)S.nobj * (column == SXG_COL_ISO ? SXG_NISO : 1);
    if (bytes < need) {
        sxg_set_error("sxg_get_column: output buffer too small");
        return -1;
    }
    if (need == 0) return 0;
    SXG_HIP_RET(hipMemcpyAsync(out, src[column], need, hipMemcpyDeviceToHost, S.stream), -1);
    SXG_HIP_RET(hipStreamSynchronize(S.stream), -1);
    return 0;
}

extern "C" int sxg_get_frame_info(sxg_frame_info *info)
{
    SxgState &S = sxg_state();
    info->backmean = S.backmean;
    info->backsig = S.backsig;
    info->dthresh = S.dthresh;
    info->thresh = S.thresh;
    info->npix_above = S.npix_above;
    info->npix_raw = S.npix_raw;
    info->nraw = S.nraw;
    info->nobj = S.nobj;
    info->nsurv = S.nsurv;
    info->nbx = S.nbx;
    info->nby = S.nby;
    info->deb_pixels = S.deb_pixels;
    info->deb_labels = S.deb_labels;
    for (int c = 0; c < 4; c++) info->deb_class_pixels[c] = S.deb_class_pixels[c];
    info->phot_pixels = S.phot_pixels;
    return 0;
}

extern "C" void sxg_set_mag_zeropoint(double zp) { sxg_state().mag_zeropoint = zp; }
