/*
 * deblend.hip -- multi-threshold deblending (init_deblend, parcel_out, clear_deblend;
 * the reference's src/cuda/cudadeblend.cu:508, :579, :1937).
 *
 * Semantics (cudadeblend.cu:199-483, :579-1615; cudaanalyse.cu:1686-1904):
 *   t_i = (float)(t0 * pow(fdpeak/t0, i/N)) per raw object, i = 1..N-1;
 *   level i holds the pixels of the kept level-(i-1) components with cd >= t_i, split
 *   into 8-connected components labelled by their first raster pixel; components with
 *   fewer than deb_minarea pixels are dropped; each kept component records its parent
 *   (the level-(i-1) component containing it), pixel count, and the float sum of cd over
 *   its pixels in raster order;
 *   a component is significant when fdflux - t_i*npix > fdflux_root*DEBLEND_MINCONT;
 *   A(c) = [fewer than two significant children] and A of every child;
 *   a level-0 object is kept when A(root); a deeper one when its parent has at least
 *   two significant children, it is significant, and A(c);
 *   kept objects are ordered by root, then level, then first pixel; each is preanalysed
 *   over its level pixels (FULL, cudaanalyse.cu:1686-1904), and every other pixel of the
 *   root goes to the kept object of the same root that minimises
 *   0.5*(cxx dx^2 + cyy dy^2 + cxy dx dy)/abcor (gatherup, cudadeblend.cu:1262-1360).
 *
 * Layout and kernels (one 64-lane wavefront per raw object):
 *   k_deb_thresh   one lane per (object, level): the sub-thresholds (canonical pow)
 *   k_deb_levels   per level: membership flags, lock-free union-find on the frame-wide
 *                  parent array (smallest index becomes the root), flattening, component
 *                  ids by ballot compaction in raster order, sizes by LDS-free atomics,
 *                  kept components appended to the object's region of the component
 *                  table; level labels stored as lvl[level][pixel of the raw list];
 *                  then the tree rules, with fluxes computed only for the children of
 *                  nodes that have two or more children (the only ones whose
 *                  significance can matter) and for the root
 *   k_deb_objects  kept components -> objects: preanalysis over level pixels (one lane
 *                  per object, sequential raster-order sums), then gatherup (one lane
 *                  per pixel)
 */
#include "sxg_internal.h"
#include "sxg_objects.h"

#define DW 64

__device__ inline unsigned long long lanes_below(int lane)
{
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ inline int uf_root_d(const int *par, int x)
{
    int y = par[x];
    while (y != x) {
        x = y;
        y = par[x];
    }
    return x;
}

__device__ inline void uf_unite_d(int *par, int a, int b)
{
    while (true) {
        a = uf_root_d(par, a);
        b = uf_root_d(par, b);
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

/* ---------------------------------------------------------------- thresholds ---- */
__global__ void __launch_bounds__(256) k_deb_thresh(const float *__restrict__ fdpeak, int nraw,
                                                    int nthresh, float base, float *thr)
{
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    if (k >= (long)nraw * nthresh) return;
    const int r = (int)(k / nthresh), i = (int)(k - (long)r * nthresh);
    const double t0 = (double)base, pk = (double)fdpeak[r];
    thr[k] = i == 0 ? base : (float)(t0 * sx_pow(pk / t0, (double)i / nthresh));
}

/* ------------------------------------------------------------------- levels ---- */
struct DebArgs {
    const int *fpia;        /* raw pixel list */
    const int *off, *cnt, *first;   /* raw objects */
    const float *cd;
    const int *lab;         /* detection labels (root pixel of each detected pixel) */
    int *par;               /* frame-wide union-find parents */
    int *cmap;              /* frame-wide: candidate id at a candidate's root pixel */
    int *lvl;               /* lvl[level * P0 + g] */
    long P0;
    const float *thr;       /* thr[r * nthresh + i] */
    int nthresh, debarea;
    double mincont;
    int W, H;
    Comp *comps;            /* component table */
    const long *compoff;    /* region of each raw object in the table */
    int *cand_count, *cand_first, *cand_parent, *cand_kid;   /* per raw pixel slots */
    int *needlist;          /* per component-table slot: components whose flux is needed */
    int *rootnc;            /* components per raw object */
    int *rootkept;          /* kept objects per raw object */
    float *rawflux;         /* raw object fdflux (when needed) */
};

__global__ void __launch_bounds__(DW) k_deb_levels(DebArgs a)
{
    const int r = blockIdx.x;
    const int lane = threadIdx.x;
    const int n = a.cnt[r], off = a.off[r], rootpix = a.first[r];
    const long P0 = a.P0;
    const unsigned long long below = lanes_below(lane);
    Comp *C = a.comps + a.compoff[r];
    const float t0 = a.thr[(size_t)r * a.nthresh];
    if (lane == 0) {
        Comp c0;
        c0.level = 0; c0.parent = -1; c0.first = rootpix; c0.count = n;
        c0.fdflux = 0.0f; c0.dthresh = t0; c0.nchild = 0; c0.nson = 0; c0.flags = 0;
        C[0] = c0;
    }
    int nc = 1;
    for (int i = 1; i < a.nthresh; i++) {
        const float ti = a.thr[(size_t)r * a.nthresh + i];
        int any = 0;
        for (int k = lane; k < n; k += DW) {
            const long g = off + k;
            const int p = a.fpia[g];
            const int prev = (i == 1) ? 0 : a.lvl[(size_t)(i - 1) * P0 + g];
            const bool in = prev >= 0 && a.cd[p] >= ti;
            a.par[p] = in ? p : -1;
            any |= in;
        }
        if (!__syncthreads_or(any)) break;
        for (int k = lane; k < n; k += DW) {
            const int p = a.fpia[off + k];
            if (a.par[p] < 0) continue;
            const int y = p / a.W, x = p - y * a.W;
            int q;
            if (x > 0 && a.lab[q = p - 1] == rootpix && a.par[q] >= 0) uf_unite_d(a.par, p, q);
            if (y > 0) {
                if (x > 0 && a.lab[q = p - a.W - 1] == rootpix && a.par[q] >= 0) uf_unite_d(a.par, p, q);
                if (a.lab[q = p - a.W] == rootpix && a.par[q] >= 0) uf_unite_d(a.par, p, q);
                if (x < a.W - 1 && a.lab[q = p - a.W + 1] == rootpix && a.par[q] >= 0)
                    uf_unite_d(a.par, p, q);
            }
        }
        __syncthreads();
        for (int k = lane; k < n; k += DW) {
            const int p = a.fpia[off + k];
            if (a.par[p] >= 0) a.par[p] = uf_root_d(a.par, p);
        }
        __syncthreads();
        /* candidate components: roots in raster order */
        int ncand = 0;
        for (int k0 = 0; k0 < n; k0 += DW) {
            const int k = k0 + lane;
            int p = -1;
            bool isroot = false;
            if (k < n) {
                p = a.fpia[off + k];
                isroot = a.par[p] == p;
            }
            const unsigned long long b = __ballot(isroot);
            if (isroot) {
                const int cid = ncand + __popcll(b & below);
                a.cmap[p] = cid;
                a.cand_count[off + cid] = 0;
                a.cand_first[off + cid] = p;
                a.cand_parent[off + cid] = (i == 1) ? 0 : a.lvl[(size_t)(i - 1) * P0 + off + k];
            }
            ncand += __popcll(b);
        }
        __syncthreads();
        for (int k = lane; k < n; k += DW) {
            const int p = a.fpia[off + k];
            const int rt = a.par[p];
            if (rt >= 0) atomicAdd(&a.cand_count[off + a.cmap[rt]], 1);
        }
        __syncthreads();
        int nkept = 0;
        for (int c0 = 0; c0 < ncand; c0 += DW) {
            const int c = c0 + lane;
            const bool kp = c < ncand && a.cand_count[off + c] >= a.debarea;
            const unsigned long long b = __ballot(kp);
            if (c < ncand) {
                const int kid = kp ? nc + nkept + __popcll(b & below) : -1;
                a.cand_kid[off + c] = kid;
                if (kp) {
                    Comp cc;
                    cc.level = i; cc.parent = a.cand_parent[off + c]; cc.first = a.cand_first[off + c];
                    cc.count = a.cand_count[off + c]; cc.fdflux = 0.0f; cc.dthresh = ti;
                    cc.nchild = 0; cc.nson = 0; cc.flags = 0;
                    C[kid] = cc;
                }
            }
            nkept += __popcll(b);
        }
        __syncthreads();
        for (int k = lane; k < n; k += DW) {
            const int p = a.fpia[off + k];
            const int rt = a.par[p];
            a.lvl[(size_t)i * P0 + off + k] = rt >= 0 ? a.cand_kid[off + a.cmap[rt]] : -1;
        }
        nc += nkept;
        __syncthreads();
        if (nkept == 0) break;
    }
    /* ---- the tree ---- */
    if (lane == 0)
        for (int c = 1; c < nc; c++) C[C[c].parent].nchild++;
    __syncthreads();
    int multi = 0;
    for (int c = lane; c < nc; c += DW) multi |= C[c].nchild >= 2;
    if (!__syncthreads_or(multi)) {
        /* no node has two children: the raw object is kept as it is */
        if (lane == 0) {
            C[0].flags = COMP_KEPT;
            a.rootnc[r] = nc;
            a.rootkept[r] = 1;
        }
        return;
    }
    /* fluxes of the root and of every child of a multi-child node, sequential over the
       raw pixel list (raster order) */
    int nneed = 0;
    for (int c0 = 0; c0 < nc; c0 += DW) {
        const int c = c0 + lane;
        const bool need = c < nc && (c == 0 || C[C[c].parent].nchild >= 2);
        const unsigned long long b = __ballot(need);
        if (need) a.needlist[a.compoff[r] + nneed + __popcll(b & below)] = c;
        nneed += __popcll(b);
    }
    __syncthreads();
    for (int j = lane; j < nneed; j += DW) {
        const int c = a.needlist[a.compoff[r] + j];
        const int L = C[c].level;
        float f = 0.0f;
        if (L == 0) {
            for (int k = 0; k < n; k++) f += a.cd[a.fpia[off + k]];
        } else {
            const int *lv = a.lvl + (size_t)L * P0 + off;
            for (int k = 0; k < n; k++)
                if (lv[k] == c) f += a.cd[a.fpia[off + k]];
        }
        C[c].fdflux = f;
    }
    __syncthreads();
    if (lane == 0) {
        a.rawflux[r] = C[0].fdflux;
        const double value0 = (double)C[0].fdflux * a.mincont;
        for (int c = nc - 1; c >= 1; c--) {
            Comp &cc = C[c];
            if (C[cc.parent].nchild >= 2) {
                float lhs = cc.fdflux - cc.dthresh * (float)cc.count;
                if ((double)lhs > value0) {
                    cc.flags |= COMP_SIG;
                    C[cc.parent].nson++;
                }
            }
        }
        /* A(c) = nson(c) <= 1 and A(children); stored as COMP_OKA */
        for (int c = 0; c < nc; c++) C[c].flags |= COMP_OKA;
        for (int c = nc - 1; c >= 0; c--) {
            if (C[c].nson > 1) C[c].flags &= ~COMP_OKA;
            if (c > 0 && !(C[c].flags & COMP_OKA)) C[C[c].parent].flags &= ~COMP_OKA;
        }
        int nk = 0;
        for (int c = 0; c < nc; c++) {
            const bool okA = C[c].flags & COMP_OKA;
            const bool kept = c == 0 ? okA
                                     : (C[C[c].parent].nson > 1 && okA && (C[c].flags & COMP_SIG));
            if (kept) {
                C[c].flags |= COMP_KEPT;
                nk++;
            }
        }
        a.rootnc[r] = nc;
        a.rootkept[r] = nk;
    }
}

/* ------------------------------------------------------------------ objects ---- */
struct ObjArgs {
    DebArgs d;
    const int *objbase;     /* first object index of each raw object */
    int *alloc;             /* frame-wide object index of each raw pixel */
    const float *sub;       /* background-subtracted frame */
    ObjCols o;
};

/* preanalyse over the pixels k of the raw list with member(k) (FULL pass) */
__device__ void preanalyse_full(const ObjArgs &A, int off, int n, int L, int c, int obj,
                                int root, int rawid)
{
    const int W = A.d.W, H = A.d.H;
    const int *lv = L > 0 ? A.d.lvl + (size_t)L * A.d.P0 + off : nullptr;
    float peak = (float)-SXG_BIG, cpeak = (float)-SXG_BIG, val, cval, rv = 0.0f;
    int xmin = 2 * SXG_MAXPICSIZE, ymin = 2 * SXG_MAXPICSIZE, xmax = 0, ymax = 0, cnt = 0;
    for (int k = 0; k < n; k++) {
        if (lv && lv[k] != c) continue;
        const int p = A.d.fpia[off + k];
        const int y = p / W, x = p - y * W;
        A.alloc[p] = obj;
        if (peak < (val = A.sub[p])) peak = val;
        if (cpeak < (cval = A.d.cd[p])) cpeak = cval;
        rv += cval;
        if (xmin > x) xmin = x;
        if (xmax < x) xmax = x;
        if (ymin > y) ymin = y;
        if (ymax < y) ymax = y;
        cnt++;
    }
    int flag = L > 0 ? SXG_OBJ_MERGED : 0;
    if (xmin == 0 || xmax == W - 1 || ymin == 0 || ymax == H - 1) flag |= SXG_OBJ_TRUNC;
    const double dthresh = (double)A.d.thr[(size_t)rawid * A.d.nthresh + L];
    double mx = 0.0, my = 0.0, tv = 0.0, mx2 = 0.0, my2 = 0.0, mxy = 0.0;
    const double dthresh2 = (dthresh + peak) / 2.0;
    int area2 = 0, dnpix = 0;
    for (int k = 0; k < n; k++) {
        if (lv && lv[k] != c) continue;
        const int p = A.d.fpia[off + k];
        const int y = p / W - ymin, x = p - (p / W) * W - xmin;
        cval = A.d.cd[p];
        tv += (val = A.sub[p]);
        if (val > dthresh) dnpix++;
        if (val > dthresh2) area2++;
        mx += cval * x;
        my += cval * y;
        mx2 += cval * x * x;
        my2 += cval * y * y;
        mxy += cval * x * y;
    }
    Moments m;
    sxg_moments(mx, my, mx2, my2, mxy, rv, false, 0.0, 0.0, xmin, ymin, area2, dnpix, dthresh,
                dthresh2, m);
    ObjCols &o = const_cast<ObjCols &>(A.o);
    o.level[obj] = L;
    o.root[obj] = root;
    o.firstpix[obj] = A.d.comps[A.d.compoff[rawid] + c].first;
    o.flag[obj] = flag;
    o.fdnpix[obj] = cnt;
    o.dnpix[obj] = dnpix;
    o.xmin[obj] = xmin; o.xmax[obj] = xmax; o.ymin[obj] = ymin; o.ymax[obj] = ymax;
    o.fdflux[obj] = rv;
    o.fdpeak[obj] = cpeak;
    o.dpeak[obj] = peak;
    o.dflux[obj] = (float)tv;
    o.dthresh[obj] = (float)dthresh;
    sxg_store_moments(o, obj, m);
}

__global__ void __launch_bounds__(DW) k_deb_objects(ObjArgs A)
{
    const int r = blockIdx.x;
    const int lane = threadIdx.x;
    const int n = A.d.cnt[r], off = A.d.off[r];
    const int nc = A.d.rootnc[r], base = A.objbase[r], nk = A.d.rootkept[r];
    const Comp *C = A.d.comps + A.d.compoff[r];
    const unsigned long long below = lanes_below(lane);
    for (int k = lane; k < n; k += DW) A.alloc[A.d.fpia[off + k]] = -1;
    __syncthreads();
    int done = 0;
    for (int c0 = 0; c0 < nc; c0 += DW) {
        const int c = c0 + lane;
        const bool kept = c < nc && (C[c].flags & COMP_KEPT);
        const unsigned long long b = __ballot(kept);
        if (kept) preanalyse_full(A, off, n, C[c].level, c, base + done + __popcll(b & below), r, r);
        done += __popcll(b);
    }
    __syncthreads();
    /* gatherup: remaining pixels go to the nearest kept object in the ellipse metric */
    const ObjCols &o = A.o;
    for (int k = lane; k < n; k += DW) {
        const int p = A.d.fpia[off + k];
        if (A.alloc[p] >= 0) continue;
        const int y = p / A.d.W, x = p - y * A.d.W;
        float distmin = (float)1e+31, dist, dx, dy;
        int iclst = base;
        for (int j = base; j < base + nk; j++) {
            dx = (float)(x - o.mx[j]);
            dy = (float)(y - o.my[j]);
            dist = (float)(0.5 * (o.cxx[j] * dx * dx + o.cyy[j] * dy * dy + o.cxy[j] * dx * dy) / o.abcor[j]);
            if (dist < distmin) {
                distmin = dist;
                iclst = j;
            }
        }
        A.alloc[p] = iclst;
    }
}

/* --------------------------------------------------------------- host side ---- */
static bool g_deb_ready = false;

extern "C" void init_deblend(float basethresh, int nthresh, size_t numValidObj, size_t numValidPix)
{
    SxgState &S = sxg_state();
    if (!S.have_frame) {
        sxg_set_error("init_deblend: no frame");
        return;
    }
    if (nthresh < 1 || nthresh > SXG_MAXNTHRESH) {
        sxg_set_error("init_deblend: DEBLEND_NTHRESH must be 1..64");
        return;
    }
    if ((long)numValidObj != S.nraw || (long)numValidPix != S.npix_raw) {
        sxg_set_error("init_deblend: object/pixel counts differ from run_detection");
        return;
    }
    S.nthresh = nthresh;
    S.deb_base = basethresh;
    float *thr = sxg_take<float>("deb.thr", (size_t)(S.nraw > 0 ? S.nraw : 1) * nthresh);
    if (!thr) return;
    RawObj ro;
    if (!sxg_raw_alloc(S.nraw, ro)) return;
    const long tot = (long)S.nraw * nthresh;
    if (tot > 0) {
        sxg_kt_begin(SXG_KT_DEB_THRESH);
        hipLaunchKernelGGL(k_deb_thresh, dim3(sxg_div_up(tot, 256)), dim3(256), 0, S.stream,
                           ro.fdpeak, S.nraw, nthresh, basethresh, thr);
        sxg_kt_end(SXG_KT_DEB_THRESH);
    }
    SXG_HIP(hipGetLastError());
    g_deb_ready = true;
}

__global__ void __launch_bounds__(256) k_deb_capacity(const int *cnt, int nraw, int nthresh,
                                                      int debarea, long *cap)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r < nraw) cap[r] = 1 + (long)(nthresh - 1) * (cnt[r] / debarea);
}

/* exclusive scan of nraw longs (single workgroup; the counts are small) */
__global__ void __launch_bounds__(1024) k_scan_long(long *v, int n, long *total)
{
    __shared__ long sm[1024];
    long carry = 0;
    for (int base = 0; base < n; base += 1024) {
        const int k = base + threadIdx.x;
        long x = k < n ? v[k] : 0;
        sm[threadIdx.x] = x;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            long add = threadIdx.x >= o ? sm[threadIdx.x - o] : 0;
            __syncthreads();
            sm[threadIdx.x] += add;
            __syncthreads();
        }
        if (k < n) v[k] = carry + sm[threadIdx.x] - x;
        carry += sm[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

extern "C" int parcel_out(float basethresh, int n_thresh, int deb_minarea, size_t numValidObj,
                          size_t numValidPix, double deblend_mincont)
{
    SxgState &S = sxg_state();
    if (!S.have_frame || !g_deb_ready || n_thresh != S.nthresh) {
        sxg_set_error("parcel_out: call init_deblend first with the same DEBLEND_NTHRESH");
        return -1;
    }
    if (deb_minarea < 1) {
        sxg_set_error("parcel_out: deb_minarea must be >= 1");
        return -1;
    }
    (void)basethresh; (void)numValidObj; (void)numValidPix;
    hipStream_t st = S.stream;
    const int nraw = S.nraw;
    const long P0 = S.npix_raw;
    sxg_stage_begin(ST_DEBLEND);
    RawObj ro;
    if (!sxg_raw_alloc(nraw, ro)) return -1;
    const size_t nr = (size_t)(nraw > 0 ? nraw : 1);
    long *compoff = sxg_take<long>("deb.compoff", nr + 1);
    long *dtot = sxg_take<long>("deb.total", 4);
    int *rootnc = sxg_take<int>("deb.rootnc", nr);
    int *rootkept = sxg_take<int>("deb.rootkept", nr);
    int *objbase = sxg_take<int>("deb.objbase", nr);
    int *cmap = sxg_take<int>("deb.cmap", S.N);
    int *alloc = sxg_take<int>("deb.alloc", S.N);
    if (!compoff || !dtot || !rootnc || !rootkept || !objbase || !cmap || !alloc) return -1;
    long ncomp_cap = 0;
    if (nraw > 0) {
        hipLaunchKernelGGL(k_deb_capacity, dim3(sxg_div_up(nraw, 256)), dim3(256), 0, st, ro.count,
                           nraw, n_thresh, deb_minarea, compoff);
        hipLaunchKernelGGL(k_scan_long, dim3(1), dim3(1024), 0, st, compoff, nraw, dtot);
        SXG_HIP_RET(hipMemcpyAsync(&ncomp_cap, dtot, sizeof(long), hipMemcpyDeviceToHost, st), -1);
        SXG_HIP_RET(hipStreamSynchronize(st), -1);
    }
    Comp *comps = sxg_take<Comp>("deb.comps", (size_t)ncomp_cap + 1);
    int *lvl = sxg_take<int>("deb.lvl", (size_t)(P0 > 0 ? P0 : 1) * n_thresh);
    int *cc = sxg_take<int>("deb.cand_count", (size_t)P0 + 1);
    int *cf = sxg_take<int>("deb.cand_first", (size_t)P0 + 1);
    int *cp = sxg_take<int>("deb.cand_parent", (size_t)P0 + 1);
    int *ck = sxg_take<int>("deb.cand_kid", (size_t)P0 + 1);
    int *nl = sxg_take<int>("deb.needlist", (size_t)ncomp_cap + 1);
    if (!comps || !lvl || !cc || !cf || !cp || !ck || !nl) return -1;
    DebArgs a;
    a.fpia = S.fpia; a.off = ro.off; a.cnt = ro.count; a.first = ro.first;
    a.cd = S.cd.as<float>(); a.lab = S.lab.as<int>(); a.par = S.aux.as<int>(); a.cmap = cmap;
    a.lvl = lvl; a.P0 = P0; a.thr = sxg_take<float>("deb.thr", nr * n_thresh);
    a.nthresh = n_thresh; a.debarea = deb_minarea; a.mincont = deblend_mincont;
    a.W = S.W; a.H = S.H; a.comps = comps; a.compoff = compoff;
    a.cand_count = cc; a.cand_first = cf; a.cand_parent = cp; a.cand_kid = ck; a.needlist = nl;
    a.rootnc = rootnc; a.rootkept = rootkept; a.rawflux = ro.fdflux;
    int nobj = 0;
    if (nraw > 0) {
        sxg_kt_begin(SXG_KT_DEB_LEVELS);
        hipLaunchKernelGGL(k_deb_levels, dim3(nraw), dim3(DW), 0, st, a);
        sxg_kt_end(SXG_KT_DEB_LEVELS);
        sxg_scan_excl(rootkept, objbase, nraw, reinterpret_cast<int *>(dtot), st);
        SXG_HIP_RET(hipMemcpyAsync(&nobj, dtot, sizeof(int), hipMemcpyDeviceToHost, st), -1);
        SXG_HIP_RET(hipStreamSynchronize(st), -1);
    }
    S.nobj = nobj;
    ObjCols oc;
    if (!sxg_obj_alloc(nobj, oc)) return -1;
    if (nraw > 0) {
        ObjArgs A;
        A.d = a; A.objbase = objbase; A.alloc = alloc; A.sub = S.img.as<float>(); A.o = oc;
        sxg_kt_begin(SXG_KT_DEB_OBJECTS);
        hipLaunchKernelGGL(k_deb_objects, dim3(nraw), dim3(DW), 0, st, A);
        sxg_kt_end(SXG_KT_DEB_OBJECTS);
    }
    sxg_stage_end(ST_DEBLEND);
    SXG_HIP_RET(hipGetLastError(), -1);
    SXG_HIP_RET(hipStreamSynchronize(st), -1);
    return nobj;
}

extern "C" void clear_deblend(int nthresh)
{
    (void)nthresh;
    g_deb_ready = false;
}
