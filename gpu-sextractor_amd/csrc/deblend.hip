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
 * Two facts about these level sets reduce the work without changing any result.
 * Let tau_i = max(t_1..t_i) and V_i = {p in the raw object : cd[p] >= tau_i}.
 *   (1) Every component of level i is a component of V_i, and every component of V_i
 *       with at least deb_minarea pixels is a kept component of level i (its ancestors
 *       are larger, hence kept).  So the kept components of level i are the components
 *       of V_i with >= deb_minarea pixels: no level needs the previous level's labels
 *       to decide membership, only a per-pixel level index m(p) = max{i : tau_i <= cd}.
 *   (2) The highest pixel of a component of V_i is a local maximum of cd over the raw
 *       object (no 8-neighbour in the object is higher), with value >= tau_i.  So a
 *       level with two or more components needs two local maxima >= tau_i: with v2 the
 *       second-highest local maximum, every level with tau_i > v2 holds at most one
 *       component.  Nodes at or below such levels have at most one child, which makes
 *       them irrelevant to A() and never kept; levels are processed only up to
 *       L* = (first i with tau_i > v2) - 1, and a raw object with a single local
 *       maximum (L* = 0) is kept as it is without any level work.
 *
 * Kernels:
 *   k_deb_thresh     one lane per (object, level): the sub-thresholds (canonical pow)
 *   k_deb_lmax<0,1>  one lane per raw pixel: local maxima (flagged per pixel); per object
 *                    their number, the highest (packed value and index, 64-bit
 *                    atomicMax), then, over the flagged pixels only, the second-highest
 *   k_deb_lsfirst    one lane per (object, level): L* from the second-highest maximum
 *   k_deb_nodecap    one lane per object: a bound on the tree's nodes (1 + L* times the
 *                    smaller of the local maxima and n / deb_minarea), scanned into the
 *                    node table's regions
 *   k_deb_classify   one lane per object: single-peak objects are finished; the rest are
 *                    binned by size (and node bound) into LDS classes
 *   k_deb_tree_lds   one workgroup per object (256 or 1024 lanes by size; the classes
 *                    run side by side on their own streams): cd, m(p), neighbours,
 *                    union-find parents, sizes and the node table in LDS; the tree is
 *                    built from the top level down with one union-find that only grows
 *                    (see deb_tree)
 *   k_deb_order      the objects too large for LDS, largest first
 *   k_deb_tree_glb   the same code as k_deb_tree_lds on global memory, for those of them
 *                    below SXG_OPT_DEB_GLB_LEVELS pixels
 *   k_deb_glb_prep,  the larger ones level-parallel: per-pixel setup; one workgroup per
 *   k_deb_glb_levels (object, level) finding the components of V_i by (1), numbering,
 *   k_deb_glb_tree   summing and testing the kept ones; the tree linking their records
 *   k_deb_objlist    one wavefront per raw object: its kept nodes -> objects
 *   k_deb_pre        one wavefront per object: FULL preanalysis over the node's level
 *                    pixels (kcomp names them; lanes compute the terms, the raster-order
 *                    sums are transposed folds: sxg_wave.h)
 *   k_deb_gather     one lane per raw pixel: gatherup of the pixels no kept node holds
 *   then the raw pixels are regrouped by object with the stable radix sort (opix,
 *   objoff): every later per-object pass reads one contiguous raster-ordered range
 */
#include <algorithm>
#include <cmath>

#include "sxg_internal.h"
#include "sxg_objects.h"
#include "sxg_wave.h"

#define DW 64
/* pixels per lane whose global-memory accesses the tree passes keep in flight together */
#define DEB_ILP 4
#define DEB_RUN 8
/* list chunks of NT entries whose roots an LDS tree's thread remembers per level (bits of
   a word); larger lists take the contiguous-run numbering */
#define DEB_NCH 32
/* words of the per-frame counter block (DebArgs::ccount) */
#define DEB_NCOUNT 14

/* unsigned order-preserving key of a float (0 is below every non-NaN value) */
__device__ inline unsigned uord(float v) { return (unsigned)f2ord(v) ^ 0x80000000u; }
__device__ inline float uord_inv(unsigned u) { return ord2f((int)(u ^ 0x80000000u)); }

/* loads of words that other wavefronts of the workgroup update with atomics; for global
   memory they go to the device-coherent level */
template <bool G> __device__ inline int ldw(const int *p)
{
    if (G) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return *p;
}
template <bool G> __device__ inline double ldw(const double *p)
{
    if (G) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return *p;
}

/* root of x with path halving: every parent pointer points to a smaller-or-equal index
   (links always go from a larger root to a smaller index), so replacing par[x] by its
   grandparent with atomicMin keeps every pointer on the path to the same or a merged
   root and keeps the chains short */
template <bool G> __device__ inline int uf_find(int *par, int x)
{
    int p = ldw<G>(&par[x]);
    while (p != x) {
        const int gp = ldw<G>(&par[p]);
        if (gp != p) atomicMin(&par[x], gp);
        x = p;
        p = gp;
    }
    return x;
}

/* global memory: the lanes that add to the same counter add once, through the lowest
   of them (same-address atomics serialise in L2); key < 0: nothing to add */
__device__ inline void wave_add_by_key(int *base, int key, int v)
{
    const int lane = threadIdx.x & 63;
    bool pending = key >= 0;
    unsigned long long act;
    while ((act = __ballot(pending)) != 0ull) {
        const int lead = __builtin_ctzll(act);
        const int kl = __shfl(key, lead);
        const bool mine = pending && key == kl;
        int s = mine ? v : 0;
        for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d);
        if (lane == lead) atomicAdd(&base[kl], s);
        pending = pending && !mine;
    }
}

/* the same for the two double sums of a node (their order is free: the significance
   bound allows for any summation order) */
template <class Node>
__device__ inline void wave_add2_by_key(Node *T, int key, double v, double w)
{
    const int lane = threadIdx.x & 63;
    bool pending = key >= 0;
    unsigned long long act;
    while ((act = __ballot(pending)) != 0ull) {
        const int lead = __builtin_ctzll(act);
        const int kl = __shfl(key, lead);
        const bool mine = pending && key == kl;
        double s = mine ? v : 0.0, sa = mine ? w : 0.0;
        for (int d = 32; d > 0; d >>= 1) {
            s += __shfl_xor(s, d);
            sa += __shfl_xor(sa, d);
        }
        if (lane == lead) {
            atomicAdd(&T[kl].ssum, s);
            atomicAdd(&T[kl].sabs, sa);
        }
        pending = pending && !mine;
    }
}

/* lock-free union: the larger root is linked to the smaller index; the two paths are
   walked side by side (path halving on both), so their loads overlap */
template <bool G> __device__ inline void uf_unite(int *par, int a, int b)
{
    while (true) {
        int pa = ldw<G>(&par[a]), pb = ldw<G>(&par[b]);
        while (pa != a || pb != b) {
            const int ga = ldw<G>(&par[pa]), gb = ldw<G>(&par[pb]);
            if (pa != a) {
                if (ga != pa) atomicMin(&par[a], ga);
                a = pa;
                pa = ga;
            }
            if (pb != b) {
                if (gb != pb) atomicMin(&par[b], gb);
                b = pb;
                pb = gb;
            }
        }
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicMin(&par[a], b);
        if (old == a) return;
        a = old;
    }
}

/* exclusive prefix of an int over the workgroup in thread order, and the total; all
   threads call it */
template <int NT> __device__ inline int wg_scan_excl(int v, int *wsum, int &total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(x, d);
        if (lane >= d) x += o;
    }
    if (NT == 64) {
        total = __shfl(x, 63);
        return x - v;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    int before = 0, tot = 0;
    for (int w = 0; w < NT / 64; w++) {
        const int s = wsum[w];
        before += w < wid ? s : 0;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return before + x - v;
}

/* exclusive prefix of a flag over the workgroup in thread order; all threads call it */
template <int NT> __device__ inline int wg_prefix(bool f, int *wsum, int &total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const unsigned long long b = __ballot(f);
    if (NT == 64) {
        total = __popcll(b);
        return __popcll(b & sxw_lanes_below(lane));
    }
    if (lane == 0) wsum[wid] = __popcll(b);
    __syncthreads();
    int before = 0, tot = 0;
    for (int w = 0; w < NT / 64; w++) {
        const int s = wsum[w];
        before += w < wid ? s : 0;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return before + __popcll(b & sxw_lanes_below(lane));
}

/* float sum of v[k] over start <= k < n with sel, in increasing k, computed by one
   wavefront (every lane returns the same value); sel(k) is evaluated for k < n only.
   The scan ends once `need` values are in the sum (need < 0: no limit); four 64-entry
   chunks are tested and read together, so their loads are in flight at once. */
template <class Sel>
__device__ inline float wave_ordered_sum(const float *v, int start, int n, Sel sel,
                                         int need = -1)
{
    const int lane = threadIdx.x & 63;
    float f = 0.0f;
    int got = 0;
    for (int base = start & ~63; base < n; base += 4 * 64) {
        bool in[4];
        int bits[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int k = base + u * 64 + lane;
            in[u] = k >= start && k < n && sel(k);
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
            bits[u] = in[u] ? __float_as_int(v[base + u * 64 + lane]) : 0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            unsigned long long b = __ballot(in[u]);
            got += __popcll(b);
            while (b) {
                const int j = __builtin_ctzll(b);
                b &= b - 1;
                f += __int_as_float(__builtin_amdgcn_readlane(bits[u], j));
            }
        }
        if (need >= 0 && got >= need) break;
    }
    return f;
}

/* ---------------------------------------------------------------- thresholds ---- */
/* the base of object r: basethresh, or with per-pixel thresholds (weight map) the
   lowest threshold among the object's pixels, rawt0[r] */
__global__ void __launch_bounds__(256) k_deb_thresh(const float *__restrict__ fdpeak, int nraw,
                                                    int nthresh, float basethresh,
                                                    const float *__restrict__ rawt0, int perpix,
                                                    float *thr, const PlanDev *pd)
{
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    if (pd) {
        nraw = pd->nraw;
        basethresh = pd->dthresh;
    }
    if (k >= (long)nraw * nthresh) return;
    const int r = (int)(k / nthresh), i = (int)(k - (long)r * nthresh);
    const float b = perpix ? rawt0[r] : basethresh;
    const double t0 = (double)b, pk = (double)fdpeak[r];
    thr[k] = i == 0 ? b : (float)(t0 * sx_pow(pk / t0, (double)i / nthresh));
}


/* ------------------------------------------------------------------- levels ---- */
/* node of the tree while it is built (LDS, or global memory for the largest objects);
   index = creation order (levels from the top down, raster order inside a level) */
struct DebNode {
    int level, parent, count, nchild, nson, flags;
    int kroot;              /* list index of the first pixel */
    int kanc;               /* the kept node among the node and its ancestors, -1 if none */
    double ssum, sabs;      /* sums of cd and |cd| over the pixels (any order) */
};

/* LDS bytes per pixel of k_deb_tree_lds: cd, parent, count (4 bytes each), the
   neighbour word (8 bytes), the highest node (16 bits), the level index */
#define DEB_LDS_BYTES_PER_PIX 23
/* LDS a workgroup may use (dynamic part; the static part of deb_tree is under 2 KiB) */
#define DEB_LDS_LIMIT ((size_t)160 * 1024 - 4096)

__host__ __device__ inline size_t deb_lds_pixbytes(int cap)
{
    return ((size_t)cap * DEB_LDS_BYTES_PER_PIX + 15) & ~(size_t)15;
}

/* LDS of k_deb_tree_lds for a tree of ct nodes over n pixels */
__host__ __device__ inline int deb_lds_bytes(int ct, int n)
{
    return (int)((((size_t)ct * sizeof(DebNode) + 15) & ~(size_t)15) +
                 deb_lds_pixbytes((n + 63) & ~63));
}

struct DebArgs {
    const int *fpia;        /* raw pixel list (raster order within each raw object) */
    const int *rawkey;      /* raw object id of each raw-list entry */
    const int *off, *cnt, *first;   /* raw objects */
    const float *cd;
    const int *lab;         /* detection labels (root pixel of each detected pixel) */
    int *kcomp;             /* kcomp[g]: the kept node (of a raw object with a tree) whose
                               level pixels hold the raw pixel g, or -1 */
    long P0;
    const float *thr;       /* thr[r * nthresh + i] */
    int nthresh, debarea;
    int lvmin;              /* global-memory objects of at least this many pixels are built
                               level-parallel (0: none) */
    double mincont;
    int W, H, nraw;
    int cstride;            /* entries per class in clist (nraw, or its capacity) */
    const PlanDev *pd;      /* planned run: the device counts (nraw, npix), else nullptr */
    Comp *comps;            /* output node table, region compoff[r] of each raw object */
    DebNode *nodes;         /* build table of the objects built in global memory */
    int *compoff;           /* node capacity of each raw object, then its region */
    int *rootnc;            /* nodes per raw object */
    int *rootkept;          /* kept objects per raw object */
    unsigned long long *top1;   /* highest local maximum: uord(cd) << 32 | raw-list index */
    unsigned *top2;         /* uord(cd) of the second-highest local maximum, 0 if none */
    int *nmax;              /* local maxima per raw object */
    unsigned char *lmflag;  /* per raw-list entry: 1 for a local maximum (first pass) */
    int *lstar;             /* first level above the second-highest maximum, then L* */
    int *clist;             /* objects of each size class: clist[cls * nraw + j] */
    int *ccount;            /* [0..3] objects per class, [4..6] the largest LDS need of
                               classes 0..2 (bytes), [7] consistency error flag, [8] class 3
                               objects built level-parallel, [9] order counter, [10..13]
                               pixels per class */
    unsigned long long *dwork;  /* [0] pixels, [1] pixel-levels of the multi-peak objects,
                                   [2] node capacity of the frame, [3] pixel-levels and
                                   [4] node records of the objects built in global memory */
    unsigned long long *lvbase; /* objects built in global memory: [2r] their first
                                   pixel-level entry, [2r + 1] their first node record */
    /* global scratch for the objects built in global memory, by raw-list position */
    float *g_cdv;
    unsigned char *g_m;
    int *g_nb;              /* neighbour words: 4 ints per raw-list entry */
    int *g_par, *g_deep, *g_cnt;
};

/* the raw objects and their pixels: the device counts of a planned run */
__device__ inline int deb_nraw(const DebArgs &a) { return a.pd ? a.pd->nraw : a.nraw; }
__device__ inline long deb_P0(const DebArgs &a) { return a.pd ? (long)a.pd->npix : a.P0; }
/* class-3 objects built level-parallel (none when there are more than a grid holds) */
__device__ inline int deb_nbig(const DebArgs &a)
{
    const int c8 = a.ccount[8];
    return c8 <= 65535 ? c8 : 0;
}

/* PASS 0: local maxima (no 8-neighbour of the raw object is higher): the highest of each
   raw object and their number; PASS 1: the highest of the others */
template <int PASS> __global__ void __launch_bounds__(256) k_deb_lmax(DebArgs a)
{
    const long g = (long)blockIdx.x * 256 + threadIdx.x;
    if (g >= deb_P0(a)) return;
    /* the second pass visits the local maxima the first one flagged */
    if (PASS == 1 && !a.lmflag[g]) return;
    const int p = a.fpia[g], r = a.rawkey[g];
    const float v = a.cd[p];
    if (PASS == 0) {
        const int root = a.lab[p];
        const int W = a.W, H = a.H;
        const int y = p / W, x = p - y * W;
        /* the eight neighbours' labels and values, loaded unconditionally from positions
           clamped into the frame (all sixteen loads in flight together), then compared;
           a position outside the frame takes no part */
        int lq[8], inb[8];
        float cq[8];
#pragma unroll
        for (int n = 0; n < 8; n++) {
            const int dd = n < 4 ? n : n + 1;   /* the 3x3 neighbourhood without its centre */
            const int yy = y + dd / 3 - 1, xx = x + dd % 3 - 1;
            inb[n] = (yy >= 0) & (yy < H) & (xx >= 0) & (xx < W);
            const int q = inb[n] ? yy * W + xx : p;
            lq[n] = a.lab[q];
            cq[n] = a.cd[q];
        }
        int higher = 0;
#pragma unroll
        for (int n = 0; n < 8; n++) higher |= inb[n] & (lq[n] == root) & (cq[n] > v);
        const bool top = higher == 0;
        a.lmflag[g] = top ? 1 : 0;
        if (!top) return;
    }
    const unsigned u = uord(v);
    const unsigned long long key = ((unsigned long long)u << 32) | (unsigned)g;
    if (PASS == 0) {
        atomicMax(&a.top1[r], key);
        atomicAdd(&a.nmax[r], 1);
    } else if (a.top1[r] != key) {
        atomicMax(&a.top2[r], u);
    }
}

__global__ void __launch_bounds__(256) k_deb_prep(DebArgs a)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r < DEB_NCOUNT) a.ccount[r] = 0;
    if (r < 5) a.dwork[r] = 0ull;
    if (r >= deb_nraw(a)) return;
    a.top1[r] = 0ull;
    a.top2[r] = 0u;
    a.nmax[r] = 0;
    a.lstar[r] = a.nthresh;
}

/* one lane per (raw object, level >= 1): the first level whose running-maximum
   sub-threshold tau exceeds the second-highest local maximum (tau_i > v2 iff some
   t_j > v2, j <= i; a NaN sub-threshold empties its level and ends the descent) */
__global__ void __launch_bounds__(256) k_deb_lsfirst(DebArgs a)
{
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    const int N = a.nthresh;
    if (k >= (long)deb_nraw(a) * N) return;
    const int r = (int)(k / N), i = (int)(k - (long)r * N);
    const unsigned u2 = a.top2[r];
    if (i == 0 || u2 == 0u) return;
    const float t = a.thr[k];
    if (t != t || t > uord_inv(u2)) atomicMin(&a.lstar[r], i);
}

/* node capacity of a tree with L* = ls: every node of level i holds a local maximum
   with m >= i (its highest pixel) and has at least deb_minarea pixels, and the nodes of
   a level are disjoint, so at most 1 + ls * min(local maxima, n / deb_minarea) */
__host__ __device__ inline int deb_nodecap(int nmax, int n, int ls, int debarea)
{
    if (ls == 0) return 1;
    const long per = nmax < n / debarea ? nmax : n / debarea;
    const long c = 1 + (long)ls * per;
    return c < 0x7fffffff ? (int)c : 0x7fffffff;
}

/* L* and the node capacity of each raw object (the total in 64 bits: one atomic per
   wavefront) */
__global__ void __launch_bounds__(256) k_deb_nodecap(DebArgs a)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    unsigned long long c = 0;
    if (r < deb_nraw(a)) {
        const int ls = a.top2[r] == 0u ? 0 : a.lstar[r] - 1;
        a.lstar[r] = ls;
        const int cap = deb_nodecap(a.nmax[r], a.cnt[r], ls, a.debarea);
        a.compoff[r] = cap;
        c = (unsigned long long)cap;
    }
    for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&a.dwork[2], c);
}

/* single-peak objects are kept as they are (one node); the others go to a size class (0:
   up to cap0 pixels, 256-lane LDS workgroups; 1: up to cap1, 1024 lanes; 2: up to cap2,
   1024 lanes; 3: global memory), and up a class while their pixels and node bound need
   more LDS than the class's budget (bud0..bud2 bytes).  The counters take one atomic per
   wavefront and class; [4 + class] records the largest need. */
__global__ void __launch_bounds__(256) k_deb_classify(DebArgs a, int cap0, int cap1, int cap2,
                                                      int bud0, int bud1, int bud2)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int cls = -1, n = 0, ls = 0, need = 0;
    if (r < deb_nraw(a)) {
        n = a.cnt[r];
        ls = a.lstar[r];
        if (ls == 0) {
            a.rootnc[r] = 1;
            a.rootkept[r] = 1;
        } else {
            /* LDS of the object's workgroup: its node table, then its pixels */
            need = deb_lds_bytes(deb_nodecap(a.nmax[r], n, ls, a.debarea), n);
            cls = n <= cap0 ? 0 : n <= cap1 ? 1 : n <= cap2 ? 2 : 3;
            /* a tree too large for its class's budget moves up a class */
            if (cls == 0 && need > bud0) cls = 1;
            if (cls == 1 && need > bud1) cls = 2;
            if (cls == 2 && need > bud2) cls = 3;
            /* the per-level arrays of the global-memory trees: n entries and
               n / deb_minarea + 1 node records for each level 1..L* */
            if (cls == 3 && a.lvmin > 0 && n >= a.lvmin) {
                atomicAdd(&a.ccount[8], 1);
                a.lvbase[2 * r] = atomicAdd(&a.dwork[3], (unsigned long long)n * ls);
                a.lvbase[2 * r + 1] = atomicAdd(
                    &a.dwork[4], (unsigned long long)ls * (n / max(a.debarea, 1) + 1));
            }
        }
    }
    unsigned long long w0 = cls >= 0 ? (unsigned long long)n : 0ull;
    unsigned long long w1 = cls >= 0 ? (unsigned long long)n * ls : 0ull;
    for (int d = 32; d > 0; d >>= 1) {
        w0 += __shfl_xor(w0, d);
        w1 += __shfl_xor(w1, d);
    }
    if (lane == 0 && w0) {
        atomicAdd(&a.dwork[0], w0);
        atomicAdd(&a.dwork[1], w1);
    }
    for (int c = 0; c < 4; c++) {
        const unsigned long long b = __ballot(cls == c);
        if (!b) continue;
        const int lead = __builtin_ctzll(b);
        int base = 0;
        if (lane == lead) base = atomicAdd(&a.ccount[c], __popcll(b));
        base = __shfl(base, lead);
        int mx = cls == c ? need : 0, px = cls == c ? n : 0;
        for (int d = 32; d > 0; d >>= 1) {
            mx = max(mx, __shfl_xor(mx, d));
            px += __shfl_xor(px, d);
        }
        if (lane == lead && c < 3) atomicMax(&a.ccount[4 + c], mx);
        if (lane == lead) atomicAdd(&a.ccount[10 + c], px);
        if (cls == c) a.clist[(size_t)c * a.cstride + base + __popcll(b & sxw_lanes_below(lane))] = r;
    }
}

/*
 * Significance test fdflux - t*npix > fdflux_root * DEBLEND_MINCONT, where fdflux is the
 * float sum of cd over the component's pixels in raster order, decided from bounds.
 * s and sabs are the exact-to-double sums of cd and |cd| (any order); the sequential
 * float sum differs from s by at most gamma_{n-1} * sabs, gamma_m = m u / (1 - m u),
 * u = 2^-24 (recursive summation, Higham, Accuracy and Stability of Numerical
 * Algorithms, 2nd ed., eq. 4.4), plus the double rounding of s itself.  The float
 * subtraction and the final comparison are bracketed the same way.  Returns 1 or 0 when
 * the bounds decide, -1 when the exact sequential sum is needed.
 */
__device__ inline double deb_sum_err(double sabs, int n)
{
    const double m = (double)(n > 1 ? n - 1 : 1) * 0x1p-24;
    if (m >= 0.5) return INFINITY;
    return sabs * (m / (1.0 - m) + (double)n * 0x1p-52) * 1.0001;
}

__device__ inline int deb_sig_bound(double s, double sabs, int n, float b, double v0lo,
                                    double v0hi)
{
    const double e = deb_sum_err(sabs, n);
    double lo = (s - e) - (double)b, hi = (s + e) - (double)b;
    lo -= fabs(lo) * 0x1p-23 + 1e-37;
    hi += fabs(hi) * 0x1p-23 + 1e-37;
    if (lo > v0hi) return 1;
    if (hi <= v0lo) return 0;
    return -1;
}

/* backward neighbours of a raw-list entry: the present ones among NW, N, NE are
   consecutive list entries from j0; W is the previous entry */
#define NB_W 1
#define NB_NW 2
#define NB_N 4
#define NB_NE 8
template <typename IT> struct __align__(8) NbP {
    IT j0;
    /* mask: which neighbours exist; the level indices of the pixel and of its W, NW, N
       and NE neighbours (0 when absent), so one read decides the pixel's joins */
    unsigned char mask, mself, mw, mnw, mn, mne;
};

/* a node of level i found by k_deb_glb_levels: first pixel, pixel count, its double sums
   while they are added, then the significance bound's answer (1, 0, or -1: decided by
   the node's float sum f against the object's float flux) */
struct LvNode {
    int kroot, count, sig;
    float f;
    double ssum, sabs;
};

/* what the tree of an object built level-parallel reads: node index of each pixel per
   level (nid[(i - 1) * n + k], -1 outside kept nodes), node records per level (rec[(i -
   1) * nstride + j]) and their counts (nk[i]), the object's float flux; the per-pixel
   setup writes the root flux bounds to v0 */
struct LvArgs {
    const int *nid;
    const LvNode *rec;
    const int *nk;
    int nstride;
    float fr;
    double *v0;
};

/*
 * The component tree of raw object r, built from the top level down.  Per raw-list
 * index k (LDS, or global scratch when G): cdv (cd), m (level index), nb (backward
 * neighbours), par (union-find parent), cnt (pixel count of a root, then its node's
 * index within the level, -1 if dropped), deep (the highest node holding k; 0 = the
 * root).  pixw: where to copy the pixel list first (LDS only; it shares cnt).  T: the
 * build table (at least 1 + local maxima * L* nodes).
 *
 * V_i = {k : m[k] >= i} only grows as i falls, so one union-find serves every level: at
 * level i only the pixels that enter (m == i) are joined, to each other and to what is
 * there; the paths stay short because every pass leaves each pixel pointing at its root.
 * Entering pixels first point at the head of their run of entering pixels (a ballot
 * over the 64-entry window); of the vertical links only those that the W neighbour does
 * not already imply are joined (8-connectivity: the row above is joined at the first
 * pixel of each overlapping stretch), and of those only the ones with an entering end.
 * A node's parent is the node of the level below that holds its first pixel; its
 * significance is decided as soon as it exists, when its level has two or more nodes
 * (otherwise no parent can have two children there).
 */
/* MODE 0: the whole tree; 1: only the per-pixel setup (cd values, level indices,
   neighbour words, root flux bounds), for the objects whose levels are then built by
   k_deb_glb_levels; 2: the tree of such an object, linking the node records of its
   levels (lv) instead of growing a union-find */
template <int NT, bool G, typename IT, int MODE = 0>
__device__ __forceinline__ void deb_tree(const DebArgs &a, const int r, float *cdv,
                                         unsigned char *m, NbP<IT> *nb, int *par, int *cnt,
                                         IT *deep, int *pixw, const int *pix, DebNode *T,
                                         const LvArgs *lv = nullptr)
{
    __shared__ int wsum[NT / 64];
    __shared__ int levstart[SXG_MAXNTHRESH + 1], levcnt[SXG_MAXNTHRESH + 1];
    __shared__ int vcount[SXG_MAXNTHRESH + 1];
    __shared__ float tthr[SXG_MAXNTHRESH], ttau[SXG_MAXNTHRESH];
    __shared__ int s_multi, s_nk, s_top;
    __shared__ double s_rsum[NT / 64], s_rabs[NT / 64], s_v0lo, s_v0hi;
    __shared__ int s_ctab[G ? 1 : DEB_NCH * (NT / 64)], s_nkl;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int n = a.cnt[r], off = a.off[r], ls = a.lstar[r];
    const int nch = (n + NT - 1) / NT;
    const int W = a.W;
    Comp *const C = a.comps + a.compoff[r];
    if (wid == 0) {
        /* sub-thresholds 1..L* and their running maximum tau (L* <= 63: one lane each) */
        const float v = lane < ls ? a.thr[(size_t)r * a.nthresh + lane + 1] : -INFINITY;
        float tau = v;
        for (int d = 1; d < 64; d <<= 1) {
            const float o = __shfl_up(tau, d);
            if (lane >= d) tau = fmaxf(tau, o);
        }
        if (lane < ls) {
            tthr[lane + 1] = v;
            ttau[lane + 1] = tau;
        }
        vcount[lane] = 0;
        levcnt[lane] = 0;
        if (lane == 0) {
            vcount[64] = 0;
            levcnt[64] = 0;
        }
    }
    if (tid == 0) {
        s_multi = 0;
        s_nk = 0;
        DebNode e;
        e.level = 0; e.parent = -1; e.count = n; e.nchild = 0; e.nson = 0; e.flags = 0;
        e.kroot = 0; e.kanc = -1; e.ssum = 0.0; e.sabs = 0.0;
        T[0] = e;
    }
    for (int k = tid; k < n; k += NT) {
        const int p = a.fpia[off + k];
        if (!G) pixw[k] = p;
        cdv[k] = a.cd[p];
    }
    __syncthreads();
    /* level index m = max{i <= L* : tau_i <= cd} (tau is non-decreasing), its histogram,
       bounds of the root's flux */
    double rsum = 0.0, rabs = 0.0;
    for (int k = tid; k < n; k += NT) {
        const float v = cdv[k];
        rsum += v;
        rabs += fabs(v);
        int lo = 0, hi = ls;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ttau[mid] <= v) lo = mid;
            else hi = mid - 1;
        }
        m[k] = (unsigned char)lo;
        atomicAdd(&vcount[lo], 1);
    }
    rsum = sxw_sum(rsum);
    rabs = sxw_sum(rabs);
    if (lane == 0) {
        s_rsum[wid] = rsum;
        s_rabs[wid] = rabs;
    }
    __syncthreads();
    if (tid == 0) {
        double rs = 0.0, ra = 0.0;
        for (int w = 0; w < NT / 64; w++) {
            rs += s_rsum[w];
            ra += s_rabs[w];
        }
        /* value0 = (double)fdflux_root * mincont with fdflux_root in [rs - e, rs + e] */
        const double e = deb_sum_err(ra, n);
        const double p = (rs - e) * a.mincont, q = (rs + e) * a.mincont;
        s_v0lo = fmin(p, q) - fabs(fmin(p, q)) * 0x1p-50 - 1e-300;
        s_v0hi = fmax(p, q) + fabs(fmax(p, q)) * 0x1p-50 + 1e-300;
        /* |V_i| = pixels with m >= i; levels where it is below deb_minarea hold no node */
        int top = 0, acc = 0;
        for (int i = ls; i >= 1; i--) {
            acc += vcount[i];
            if (top == 0 && acc >= a.debarea) top = i;
        }
        s_top = top;
    }
    /* backward neighbours by position in the list (binary search of the row above) */
    if (MODE == 2)
        for (int k = tid; k < n; k += NT) deep[k] = 0;
    else
    for (int k = tid; k < n; k += NT) {
        const int p = pix[k];
        const int y = p / W, x = p - y * W;
        int mask = 0, j0 = 0;
        if (x > 0 && k > 0 && pix[k - 1] == p - 1) mask |= NB_W;
        if (y > 0) {
            const int q0 = p - W - (x > 0 ? 1 : 0);
            int l = 0, h = k;
            while (l < h) {
                const int mid = (l + h) >> 1;
                if (pix[mid] < q0) l = mid + 1;
                else h = mid;
            }
            j0 = l;
            int j = l;
            if (x > 0 && j < k && pix[j] == p - W - 1) {
                mask |= NB_NW;
                j++;
            }
            if (j < k && pix[j] == p - W) {
                mask |= NB_N;
                j++;
            }
            if (x < W - 1 && j < k && pix[j] == p - W + 1) mask |= NB_NE;
        }
        NbP<IT> b;
        b.j0 = (IT)j0;
        b.mask = (unsigned char)mask;
        b.mself = m[k];
        b.mw = (mask & NB_W) ? m[k - 1] : 0;
        const int jn = j0 + ((mask & NB_NW) ? 1 : 0), jne = jn + ((mask & NB_N) ? 1 : 0);
        b.mnw = (mask & NB_NW) ? m[j0] : 0;
        b.mn = (mask & NB_N) ? m[jn] : 0;
        b.mne = (mask & NB_NE) ? m[jne] : 0;
        nb[k] = b;
        deep[k] = 0;
    }
    __syncthreads();
    if constexpr (MODE == 1) {
        if (tid == 0) {
            lv->v0[0] = s_v0lo;
            lv->v0[1] = s_v0hi;
        }
        return;
    }
    const int top = s_top;
    /* run heads of the pixels entering at the top level (every pixel of V_top) */
    auto prelink = [&](int lev, int enter_hi) {
        for (int base = 0; base < n; base += NT) {
            const int k = base + tid;
            const int mk = k < n ? (int)m[k] : -1;
            const bool ent = mk >= lev && mk < enter_hi;
            bool wen = false;
            if (ent && (nb[k].mask & NB_W)) {
                const int mw = m[k - 1];
                wen = mw >= lev && mw < enter_hi;
            }
            const unsigned long long S = __ballot(ent && !wen);
            if (ent) {
                const unsigned long long le = S & (~0ull >> (63 - lane));
                par[k] = k - lane + (le ? 63 - __builtin_clzll(le) : 0);
            }
        }
    };
    if (MODE != 2 && top > 0) prelink(top, ls + 1);
    int nc = 1, last = 0, p0 = 0, p1 = 0, prev = ls + 1;
    /* the object's float flux when a wavefront needed it (MODE 2: computed beforehand) */
    float fr_w = MODE == 2 ? lv->fr : 0.0f;
    bool fr_ok = MODE == 2;
    for (int i = top; i >= 1; i--) {
        __syncthreads();
        if constexpr (MODE == 2) {
            /* the kept nodes of level i from their records, their significance settled */
            const int nkept = lv->nk[i];
            const LvNode *rec = lv->rec + (size_t)(i - 1) * lv->nstride;
            const float ti = tthr[i];
            for (int c = tid; c < nkept; c += NT) {
                const LvNode q = rec[c];
                bool sig = false;
                if (nkept >= 2) {
                    sig = q.sig > 0;
                    if (q.sig < 0) {
                        const float lhs = q.f - ti * (float)q.count;
                        sig = (double)lhs > (double)fr_w * a.mincont;
                    }
                }
                DebNode e;
                e.level = i; e.parent = -1; e.count = q.count; e.nchild = 0; e.nson = 0;
                e.flags = sig ? COMP_SIG : 0; e.kroot = q.kroot; e.kanc = -1;
                e.ssum = 0.0; e.sabs = 0.0;
                T[nc + c] = e;
            }
            __syncthreads();
            /* parents of the level i + 1 nodes: the node that holds their first pixel */
            const int *nid = lv->nid + (size_t)(i - 1) * n;
            for (int c = p0 + tid; c < p1; c += NT) {
                const int j = nid[T[c].kroot];
                int parent = nc + j;
                if (j < 0) {
                    atomicOr(&a.ccount[7], 1);
                    parent = 0;
                }
                T[c].parent = parent;
                if (atomicAdd(&T[parent].nchild, 1) == 1) s_multi = 1;
                if (ldw<G>(&T[c].flags) & COMP_SIG) atomicAdd(&T[parent].nson, 1);
            }
            if (tid == 0) {
                levstart[i] = nc;
                levcnt[i] = nkept;
            }
            p0 = nc;
            p1 = nc + nkept;
            nc += nkept;
            if (nkept > 0 && last == 0) last = i;
            prev = i;
            continue;
        }
        /* join the entering pixels */
        auto join = [&](int k, const NbP<IT> b) {
            const int mk = b.mself;
            if (mk < i) return;
            const bool ek = mk < prev;
            const int msk = (int)b.mask, j0 = (int)b.j0;
            const int jnw = j0, jn = j0 + ((msk & NB_NW) ? 1 : 0);
            const int jne = jn + ((msk & NB_N) ? 1 : 0);
            const int mw = b.mw, mnw = b.mnw, mn = b.mn, mne = b.mne;
            const bool wl = mw >= i, nw = mnw >= i, nn = mn >= i, ne = mne >= i;
            /* runs of entering pixels are linked inside their window already */
            if (wl && (ek || mw < prev) && !(ek && mw < prev && (k & 63) != 0))
                uf_unite<G>(par, k, k - 1);
            if (nn) {
                /* the W neighbour's N is this pixel's NW */
                if (!(wl && nw) && (ek || mn < prev)) uf_unite<G>(par, k, jn);
            } else {
                if (nw && !wl && (ek || mnw < prev)) uf_unite<G>(par, k, jnw);
                if (ne && (ek || mne < prev)) uf_unite<G>(par, k, jne);
            }
        };
        if constexpr (G) {
            /* global memory: the neighbour words of DEB_ILP pixels per lane read together */
            for (int k0 = tid; k0 < n; k0 += DEB_ILP * NT) {
                NbP<IT> b[DEB_ILP];
#pragma unroll
                for (int u = 0; u < DEB_ILP; u++)
                    if (k0 + u * NT < n) b[u] = nb[k0 + u * NT];
#pragma unroll
                for (int u = 0; u < DEB_ILP; u++)
                    if (k0 + u * NT < n) {
                        cnt[k0 + u * NT] = 0;
                        join(k0 + u * NT, b[u]);
                    }
            }
        } else {
            for (int k = tid; k < n; k += NT) {
                cnt[k] = 0;
                join(k, nb[k]);
            }
        }
        __syncthreads();
        unsigned rbits = 0;
        /* roots (the smallest list index = first raster pixel) and sizes; a concurrent
           path-halving step can only lower par[k] towards the same root */
        if constexpr (G) {
            /* global memory: DEB_ILP pixels per lane walk their paths side by side, so
               their loads are in flight together */
            /* a bound the whole wavefront shares: the counts are added wavefront-wide */
            for (int b0 = 0; b0 < n; b0 += DEB_ILP * NT) {
                const int k0 = b0 + tid;
                int x[DEB_ILP], pp[DEB_ILP];
                bool act[DEB_ILP];
#pragma unroll
                for (int u = 0; u < DEB_ILP; u++) {
                    const int k = k0 + u * NT;
                    act[u] = k < n && m[k] >= i;
                    x[u] = k;
                    pp[u] = act[u] ? ldw<G>(&par[k]) : k;
                }
                bool busy = true;
                while (busy) {
                    busy = false;
                    int gp[DEB_ILP];
#pragma unroll
                    for (int u = 0; u < DEB_ILP; u++)
                        gp[u] = pp[u] != x[u] ? ldw<G>(&par[pp[u]]) : pp[u];
#pragma unroll
                    for (int u = 0; u < DEB_ILP; u++)
                        if (pp[u] != x[u]) {
                            if (gp[u] != pp[u]) atomicMin(&par[x[u]], gp[u]);
                            x[u] = pp[u];
                            pp[u] = gp[u];
                            busy = busy || pp[u] != x[u];
                        }
                }
#pragma unroll
                for (int u = 0; u < DEB_ILP; u++) {
                    if (act[u]) par[k0 + u * NT] = x[u];
                    wave_add_by_key(cnt, act[u] ? x[u] : -1, 1);
                }
            }
        } else {
            /* bit j: the thread's pixel j * NT + tid is a root of V_i */
            for (int j = 0, k = tid; k < n; j++, k += NT)
                if (m[k] >= i) {
                    const int rt = uf_find<G>(par, k);
                    par[k] = rt;
                    atomicAdd(&cnt[rt], 1);
                    if (rt == k && j < DEB_NCH) rbits |= 1u << j;
                }
        }
        __syncthreads();
        int nkept = 0;
        if (!G && nch <= DEB_NCH) {
            /* the nodes of level i in raster order of their first pixel, from the roots
               the thread found above (no pass over the other pixels): per chunk of NT
               list entries and wavefront, the kept roots are counted by a ballot, one
               wavefront turns the table into an exclusive prefix in (chunk, wavefront)
               order, and each kept root's number is its entry plus the kept roots of
               the lower lanes */
            unsigned fbits = 0;
            for (int j = 0; j < nch; j++)
                if ((rbits >> j) & 1u)
                    if (cnt[j * NT + tid] >= a.debarea) fbits |= 1u << j;
            for (int j = 0; j < nch; j++) {
                const unsigned long long b = __ballot((fbits >> j) & 1u);
                if (lane == 0) s_ctab[j * (NT / 64) + wid] = __popcll(b);
            }
            __syncthreads();
            if (wid == 0) {
                const int ne = nch * (NT / 64);
                int v[8], sl = 0;
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int e = lane * 8 + u;
                    v[u] = e < ne ? s_ctab[e] : 0;
                    sl += v[u];
                }
                int x = sl;
                for (int d = 1; d < 64; d <<= 1) {
                    const int o = __shfl_up(x, d);
                    if (lane >= d) x += o;
                }
                int run = x - sl;
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int e = lane * 8 + u;
                    if (e < ne) s_ctab[e] = run;
                    run += v[u];
                }
                if (lane == 63) s_nkl = x;
            }
            __syncthreads();
            nkept = s_nkl;
            const unsigned long long below = sxw_lanes_below(lane);
            for (int j = 0; j < nch; j++) {
                const bool f = (fbits >> j) & 1u;
                const unsigned long long b = __ballot(f);
                if ((rbits >> j) & 1u) {
                    const int k = j * NT + tid;
                    int jk = -1;
                    if (f) {
                        jk = s_ctab[j * (NT / 64) + wid] + __popcll(b & below);
                        DebNode e;
                        e.level = i; e.parent = -1; e.count = cnt[k]; e.nchild = 0; e.nson = 0;
                        e.flags = 0; e.kroot = k; e.kanc = -1; e.ssum = 0.0; e.sabs = 0.0;
                        T[nc + jk] = e;
                    }
                    cnt[k] = jk;
                }
            }
        } else {
        /* the nodes of level i in raster order of their first pixel: each thread numbers
           a contiguous run of the list after one prefix over the threads' counts */
            const int per = (n + NT - 1) / NT;
            const int kb = tid * per, ke = min(kb + per, n);
            int mine = 0;
            if constexpr (G) {
                /* global memory: the run read DEB_RUN pixels at a time, loads in flight
                   together */
                for (int k0 = kb; k0 < ke; k0 += DEB_RUN) {
                    int pk[DEB_RUN], ck[DEB_RUN];
                    bool in[DEB_RUN];
#pragma unroll
                    for (int u = 0; u < DEB_RUN; u++) in[u] = k0 + u < ke && m[k0 + u] >= i;
#pragma unroll
                    for (int u = 0; u < DEB_RUN; u++) pk[u] = in[u] ? ldw<G>(&par[k0 + u]) : -1;
#pragma unroll
                    for (int u = 0; u < DEB_RUN; u++)
                        ck[u] = pk[u] == k0 + u ? ldw<G>(&cnt[k0 + u]) : 0;
#pragma unroll
                    for (int u = 0; u < DEB_RUN; u++) mine += pk[u] == k0 + u && ck[u] >= a.debarea;
                }
            } else {
                for (int k = kb; k < ke; k++)
                    if (m[k] >= i && ldw<G>(&par[k]) == k && ldw<G>(&cnt[k]) >= a.debarea) mine++;
            }
            int j = wg_scan_excl<NT>(mine, wsum, nkept);
            auto number = [&](int k, int ck) {
                int jk = -1;
                if (ck >= a.debarea) {
                    jk = j++;
                    DebNode e;
                    e.level = i; e.parent = -1; e.count = ck; e.nchild = 0; e.nson = 0;
                    e.flags = 0; e.kroot = k; e.kanc = -1; e.ssum = 0.0; e.sabs = 0.0;
                    T[nc + jk] = e;
                }
                cnt[k] = jk;
            };
            if constexpr (G) {
                for (int k0 = kb; k0 < ke; k0 += DEB_RUN) {
                    int pk[DEB_RUN], ck[DEB_RUN];
                    bool in[DEB_RUN];
#pragma unroll
                    for (int u = 0; u < DEB_RUN; u++) in[u] = k0 + u < ke && m[k0 + u] >= i;
#pragma unroll
                    for (int u = 0; u < DEB_RUN; u++) pk[u] = in[u] ? ldw<G>(&par[k0 + u]) : -1;
#pragma unroll
                    for (int u = 0; u < DEB_RUN; u++)
                        ck[u] = pk[u] == k0 + u ? ldw<G>(&cnt[k0 + u]) : 0;
#pragma unroll
                    for (int u = 0; u < DEB_RUN; u++)
                        if (pk[u] == k0 + u) number(k0 + u, ck[u]);
                }
            } else {
                for (int k = kb; k < ke; k++)
                    if (m[k] >= i && ldw<G>(&par[k]) == k) number(k, ldw<G>(&cnt[k]));
            }
        }
        __syncthreads();
        const bool sums = nkept >= 2;
        /* parents of the level i + 1 nodes: the node that holds their first pixel */
        for (int c = p0 + tid; c < p1; c += NT) {
            const int j = ldw<G>(&cnt[ldw<G>(&par[T[c].kroot])]);
            int parent = nc + j;
            if (j < 0) {
                /* cannot happen when the level sets nest; reported, never used */
                atomicOr(&a.ccount[7], 1);
                parent = 0;
            }
            T[c].parent = parent;
            if (atomicAdd(&T[parent].nchild, 1) == 1) s_multi = 1;
            if (ldw<G>(&T[c].flags) & COMP_SIG) atomicAdd(&T[parent].nson, 1);
        }
        /* the highest node of each pixel; node sums; run heads of the level i - 1
           entries */
        if constexpr (G) {
            /* global memory: the node lookups of DEB_ILP pixels per lane in flight
               together, then the same steps as below for each of them */
            for (int b0 = 0; b0 < n; b0 += DEB_ILP * NT) {
                int mk[DEB_ILP], pk[DEB_ILP], jj[DEB_ILP];
                IT dk[DEB_ILP];
                bool need[DEB_ILP];
#pragma unroll
                for (int u = 0; u < DEB_ILP; u++) {
                    const int k = b0 + u * NT + tid;
                    mk[u] = k < n ? (int)m[k] : -1;
                    dk[u] = mk[u] >= i ? deep[k] : (IT)0;
                    need[u] = mk[u] >= i && (dk[u] == (IT)0 || sums);
                }
#pragma unroll
                for (int u = 0; u < DEB_ILP; u++)
                    pk[u] = need[u] ? ldw<G>(&par[b0 + u * NT + tid]) : 0;
#pragma unroll
                for (int u = 0; u < DEB_ILP; u++) jj[u] = need[u] ? ldw<G>(&cnt[pk[u]]) : -1;
#pragma unroll
                for (int u = 0; u < DEB_ILP; u++) {
                    const int k = b0 + u * NT + tid;
                    const int j = jj[u];
                    float v = 0.0f;
                    if (j >= 0) {
                        if (dk[u] == (IT)0) deep[k] = (IT)(nc + j);
                        v = cdv[k];
                    }
                    if (sums) wave_add2_by_key(T + nc, j, (double)v, fabs((double)v));
                    if (i > 1) {
                        const bool ent = mk[u] == i - 1;
                        bool wen = false;
                        if (ent && (nb[k].mask & NB_W)) wen = m[k - 1] == i - 1;
                        const unsigned long long S = __ballot(ent && !wen);
                        if (ent) {
                            const unsigned long long le = S & (~0ull >> (63 - lane));
                            par[k] = k - lane + (le ? 63 - __builtin_clzll(le) : 0);
                        }
                    }
                }
            }
        } else
        for (int base = 0; base < n; base += NT) {
            const int k = base + tid;
            const int mk = k < n ? (int)m[k] : -1;
            int j = -1;
            float v = 0.0f;
            if (mk >= i) {
                /* a pixel with its highest node already needs its node only for sums */
                const IT dk = deep[k];
                if (dk == (IT)0 || sums) {
                    j = ldw<G>(&cnt[ldw<G>(&par[k])]);
                    if (j >= 0) {
                        if (dk == (IT)0) deep[k] = (IT)(nc + j);
                        v = cdv[k];
                    }
                }
            }
            if (sums && j >= 0) {
                /* LDS atomics without a return value: the wavefront does not wait */
                atomicAdd(&T[nc + j].ssum, (double)v);
                atomicAdd(&T[nc + j].sabs, fabs((double)v));
            }
            if (i > 1) {
                const bool ent = mk == i - 1;
                bool wen = false;
                if (ent && (nb[k].mask & NB_W)) wen = m[k - 1] == i - 1;
                const unsigned long long S = __ballot(ent && !wen);
                if (ent) {
                    const unsigned long long le = S & (~0ull >> (63 - lane));
                    par[k] = k - lane + (le ? 63 - __builtin_clzll(le) : 0);
                }
            }
        }
        if (tid == 0) {
            levstart[i] = nc;
            levcnt[i] = nkept;
        }
        p0 = nc;
        p1 = nc + nkept;
        nc += nkept;
        if (nkept > 0 && last == 0) last = i;
        prev = i;
        if (sums) {
            __syncthreads();
            /* significance of the new nodes, one wavefront each; the exact sequential
               float sums when the bounds cannot decide */
            const float ti = tthr[i];
            for (int c = p0 + wid; c < p1; c += NT / 64) {
                const int cc = T[c].count, kr = T[c].kroot;
                const float b = ti * (float)cc;
                int sig = deb_sig_bound(ldw<G>(&T[c].ssum), ldw<G>(&T[c].sabs), cc, b, s_v0lo,
                                        s_v0hi);
                if (sig < 0) {
                    /* the node's cc pixels start at its first pixel kr */
                    const float f = wave_ordered_sum(
                        cdv, kr, n, [&](int k) { return m[k] >= i && ldw<G>(&par[k]) == kr; },
                        cc);
                    /* the object's flux, once per wavefront */
                    if (!fr_ok) {
                        fr_w = wave_ordered_sum(cdv, 0, n, [](int) { return true; });
                        fr_ok = true;
                    }
                    const float lhs = f - b;
                    sig = (double)lhs > (double)fr_w * a.mincont;
                }
                if (sig && lane == 0) atomicOr(&T[c].flags, COMP_SIG);
            }
        }
    }
    __syncthreads();
    if constexpr (MODE == 2) {
        /* the highest node of each pixel: its node on the highest level that keeps one */
        for (int k = tid; k < n; k += NT) {
            int d = 0;
            for (int i = min((int)m[k], top); i >= 1; i--) {
                const int j = lv->nid[(size_t)(i - 1) * n + k];
                if (j >= 0) {
                    d = levstart[i] + j;
                    break;
                }
            }
            deep[k] = (IT)d;
        }
        __syncthreads();
    }
    /* the level 1 nodes hang from the root */
    for (int c = p0 + tid; c < p1; c += NT) {
        T[c].parent = 0;
        if (atomicAdd(&T[0].nchild, 1) == 1) s_multi = 1;
        if (ldw<G>(&T[c].flags) & COMP_SIG) atomicAdd(&T[0].nson, 1);
    }
    __syncthreads();
    const bool multi = s_multi != 0;
    if (multi) {
        for (int c = tid; c < nc; c += NT) atomicOr(&T[c].flags, COMP_OKA);
        __syncthreads();
        /* A(c) = nson(c) <= 1 and A(children), bottom-up one level at a time */
        for (int L = last; L >= 0; L--) {
            const int c0 = L > 0 ? levstart[L] : 0, c1 = L > 0 ? c0 + levcnt[L] : 1;
            for (int c = c0 + tid; c < c1; c += NT) {
                int fl = ldw<G>(&T[c].flags);
                if (ldw<G>(&T[c].nson) > 1) {
                    atomicAnd(&T[c].flags, ~COMP_OKA);
                    fl &= ~COMP_OKA;
                }
                if (c > 0 && !(fl & COMP_OKA)) atomicAnd(&T[T[c].parent].flags, ~COMP_OKA);
            }
            __syncthreads();
        }
        for (int c = tid; c < nc; c += NT) {
            const int fl = ldw<G>(&T[c].flags);
            const bool okA = fl & COMP_OKA;
            const bool kept =
                c == 0 ? okA : (ldw<G>(&T[T[c].parent].nson) > 1 && okA && (fl & COMP_SIG));
            if (kept) {
                atomicOr(&T[c].flags, COMP_KEPT);
                atomicAdd(&s_nk, 1);
            }
        }
        __syncthreads();
        /* the kept node on each node's path to the root (kept nodes never nest) */
        for (int L = 0; L <= last; L++) {
            const int c0 = L > 0 ? levstart[L] : 0, c1 = L > 0 ? c0 + levcnt[L] : 1;
            for (int c = c0 + tid; c < c1; c += NT) {
                const int fl = ldw<G>(&T[c].flags);
                T[c].kanc = (fl & COMP_KEPT) ? c : c == 0 ? -1 : T[T[c].parent].kanc;
            }
            __syncthreads();
        }
    }
    /* output order: level, then raster order; levbase[L] = first output index of level L */
    if (tid == 0) {
        int acc = 1;
        for (int L = 1; L <= last; L++) {
            const int c = levcnt[L];
            levcnt[L] = acc - levstart[L];   /* output index = creation index + this */
            acc += c;
        }
    }
    __syncthreads();
    auto outid = [&](int c) { return c == 0 ? 0 : c + levcnt[T[c].level]; };
    if (multi)
        for (int k = tid; k < n; k += NT) {
            const int ka = T[(int)deep[k]].kanc;
            a.kcomp[off + k] = ka < 0 ? -1 : outid(ka);
        }
    /* what the object kernels read; without a node with two children the raw object is
       kept as it is */
    for (int c = tid; c < nc; c += NT) {
        Comp o;
        o.level = T[c].level;
        o.first = a.fpia[off + T[c].kroot];
        o.flags = multi ? (ldw<G>(&T[c].flags) & COMP_KEPT) : (c == 0 ? COMP_KEPT : 0);
        C[outid(c)] = o;
    }
    if (tid == 0) {
        a.rootnc[r] = nc;
        a.rootkept[r] = multi ? s_nk : 1;
    }
}

/* CLS names the size class in the kernel's symbol only (classes 1 and 2 share the code
   and the workgroup size; a kernel trace then times each class on its own) */
template <int NT, int CLS>
__global__ void __launch_bounds__(NT, NT == 256 ? 7 : 1) k_deb_tree_lds(DebArgs a, int cls)
{
    extern __shared__ __align__(16) unsigned char smem[];
    /* a planned run's grid is the class's capacity */
    if (a.pd && (int)blockIdx.x >= a.ccount[cls]) return;
    const int r = a.clist[(size_t)cls * a.cstride + blockIdx.x];
    DebNode *tl = reinterpret_cast<DebNode *>(smem);
    const int ct = deb_nodecap(a.nmax[r], a.cnt[r], a.lstar[r], a.debarea);
    const int cap = (a.cnt[r] + 63) & ~63;
    float *cdv = reinterpret_cast<float *>(smem + (((size_t)ct * sizeof(DebNode) + 15) & ~15));
    int *par = reinterpret_cast<int *>(cdv + cap);
    int *cnt = par + cap;
    NbP<unsigned short> *nb = reinterpret_cast<NbP<unsigned short> *>(cnt + cap);
    unsigned short *deep = reinterpret_cast<unsigned short *>(nb + cap);
    unsigned char *m = reinterpret_cast<unsigned char *>(deep + cap);
    /* the pixel list is only needed while the neighbours are found: it shares cnt */
    deb_tree<NT, false, unsigned short>(a, r, cdv, m, nb, par, cnt, deep, cnt, cnt, tl);
}

/* raw objects too large for LDS or with trees that may outgrow it: all state in global
   memory */
__global__ void __launch_bounds__(1024) k_deb_tree_glb(DebArgs a, int j0)
{
    if (a.pd) {
        /* planned run: the level-parallel objects come first, the grid is the class's
           capacity */
        j0 = deb_nbig(a);
        if ((int)blockIdx.x >= a.ccount[3] - j0) return;
    }
    const int r = a.clist[3 * (size_t)a.cstride + j0 + blockIdx.x];
    const int off = a.off[r];
    deb_tree<1024, true, int>(a, r, a.g_cdv + off, a.g_m + off,
                              reinterpret_cast<NbP<int> *>(a.g_nb) + off, a.g_par + off,
                              a.g_cnt + off, a.g_deep + off, nullptr, a.fpia + off,
                              a.nodes + a.compoff[r]);
}

/* the objects of class 3 into out, largest first (ties in list order), so the kernels
   start their longest trees first: with full, all of them (O(nc^2) comparisons);
   otherwise those of at least lvmin pixels (the level-parallel ones, nbig of them) in
   that order, then the others in any order from the counter *ctr (zero on entry),
   O(nbig * nc).  Either way the level-parallel objects come first.  The results do
   not depend on the order.  cc (planned runs): nc, nbig and full from the class counts
   on the device (full when nc <= fullmax). */
__global__ void __launch_bounds__(256) k_deb_order(const int *__restrict__ clist,
                                                   const int *__restrict__ cnt, int nraw,
                                                   int nc, int lvmin, int nbig, int full,
                                                   int *ctr, int *__restrict__ out,
                                                   const int *cc, int fullmax)
{
    __shared__ int sn[256];
    if (cc) {
        nc = cc[3];
        nbig = cc[8] <= 65535 ? cc[8] : 0;
        if (nbig == 0) lvmin = 0;
        full = nc <= fullmax ? 1 : 0;
    }
    const int *L = clist + 3 * (size_t)nraw;
    int *O = out + 3 * (size_t)nraw;
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int nj = j < nc ? cnt[L[j]] : 0;
    const bool big = j < nc && (full || (lvmin > 0 && nj >= lvmin));
    if (j < nc && !big) O[nbig + atomicAdd(ctr, 1)] = L[j];
    if (!__syncthreads_or(big)) return;
    int rank = 0;
    for (int t0 = 0; t0 < nc; t0 += 256) {
        __syncthreads();
        sn[threadIdx.x] = t0 + (int)threadIdx.x < nc ? cnt[L[t0 + threadIdx.x]] : -1;
        __syncthreads();
        if (big) {
            const int tn = min(256, nc - t0);
            for (int u = 0; u < tn; u++) {
                const int x = sn[u];
                rank += x > nj || (x == nj && t0 + u < j);
            }
        }
    }
    if (big) O[rank] = L[j];
}

/* float sum of v[0..n) in increasing index order by thread 0, the values staged through
   LDS in 4096-entry chunks by the other wavefronts while it adds the previous chunk
   (thread 0's return value is the sum) */
__device__ inline float block_serial_sum(const float *v, int n, float (*buf)[4096])
{
    const int tid = threadIdx.x, NS = blockDim.x - 64;
    const int nch = (n + 4095) / 4096;
    auto stage = [&](int c) {
        if (tid >= 64)
            for (int t = tid - 64; t < 4096; t += NS) {
                const int k = c * 4096 + t;
                buf[c & 1][t] = k < n ? v[k] : 0.0f;
            }
    };
    float f = 0.0f;
    stage(0);
    __syncthreads();
    for (int c = 0; c < nch; c++) {
        if (c + 1 < nch) stage(c + 1);
        if (tid == 0) {
            const int e = min(4096, n - c * 4096);
            const float4 *q = reinterpret_cast<const float4 *>(buf[c & 1]);
            int t = 0;
            for (; t + 4 <= e; t += 4) {
                const float4 x = q[t >> 2];
                f += x.x;
                f += x.y;
                f += x.z;
                f += x.w;
            }
            for (; t < e; t++) f += buf[c & 1][t];
        }
        __syncthreads();
    }
    return f;
}

/* The same objects built level-parallel: the per-pixel setup (k_deb_glb_prep), then one
   workgroup per (object, level i) finds the 8-connected components of V_i = {m >= i},
   numbers the kept ones (at least deb_minarea pixels) in raster order of their first
   pixels, adds their sums and settles their significance where the bounds decide
   (k_deb_glb_levels), then the tree links the node records level by level
   (k_deb_glb_tree).  The components of V_i do not depend on how they are found, so the
   tree is the one the growing union-find gives. */
__global__ void __launch_bounds__(1024) k_deb_glb_prep(DebArgs a, double *v0)
{
    if (a.pd && (int)blockIdx.x >= deb_nbig(a)) return;
    const int r = a.clist[3 * (size_t)a.cstride + blockIdx.x];
    const int off = a.off[r];
    LvArgs lv{};
    lv.v0 = v0 + 2 * blockIdx.x;
    deb_tree<1024, true, int, 1>(a, r, a.g_cdv + off, a.g_m + off,
                                 reinterpret_cast<NbP<int> *>(a.g_nb) + off, a.g_par + off,
                                 a.g_cnt + off, a.g_deep + off, nullptr, a.fpia + off,
                                 a.nodes + a.compoff[r], &lv);
}

__global__ void __launch_bounds__(1024) k_deb_glb_levels(DebArgs a, int *lvpar, int *lvnid,
                                                        LvNode *lvrec, int *lvk, float *frbuf,
                                                        const double *v0)
{
    __shared__ int wsum[16];
    __shared__ __align__(16) float sbuf[2][4096];
    /* blockIdx.y: the object (largest first); blockIdx.x: 0 computes the object's float
       flux (for the significance tests the bounds leave open), x >= 1 builds level x */
    const int oj = blockIdx.y;
    if (a.pd && oj >= deb_nbig(a)) return;
    const int r = a.clist[3 * (size_t)a.cstride + oj];
    const int i = blockIdx.x;
    const int n = a.cnt[r], off = a.off[r], tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const float *cdv = a.g_cdv + off;
    if (i == 0) {
        const float f = block_serial_sum(cdv, n, sbuf);
        if (tid == 0) frbuf[oj] = f;
        return;
    }
    if (i > a.lstar[r]) return;
    const unsigned char *m = a.g_m + off;
    const NbP<int> *nb = reinterpret_cast<const NbP<int> *>(a.g_nb) + off;
    int *par = lvpar + a.lvbase[2 * r] + (size_t)(i - 1) * n;
    int *cnt = lvnid + a.lvbase[2 * r] + (size_t)(i - 1) * n;
    LvNode *rec = lvrec + a.lvbase[2 * r + 1] + (size_t)(i - 1) * (n / max(a.debarea, 1) + 1);
    /* every pixel of V_i starts at the head of its run inside its 64-entry window (runs
       of W neighbours in V_i), so the unions below only join runs */
    for (int b0 = 0; b0 < n; b0 += 1024) {
        const int k = b0 + tid;
        bool in = false, wen = false;
        if (k < n) {
            const NbP<int> b = nb[k];
            in = b.mself >= i;
            wen = in && b.mw >= i;
            cnt[k] = 0;
        }
        const unsigned long long S = __ballot(in && !wen);
        if (in) {
            const unsigned long long le = S & (~0ull >> (63 - lane));
            par[k] = k - lane + (le ? 63 - __builtin_clzll(le) : 0);
        }
    }
    __syncthreads();
    /* the backward 8-neighbours in V_i, the redundant ones left out (W: only across a
       window boundary; N and NW through W: NW is W's N; NE through N: it is N's E) */
    auto join = [&](int k, const NbP<int> b) {
        if (b.mself < i) return;
        const int msk = (int)b.mask, j0 = b.j0;
        const int jn = j0 + ((msk & NB_NW) ? 1 : 0), jne = jn + ((msk & NB_N) ? 1 : 0);
        const bool wl = b.mw >= i, nw = b.mnw >= i, nn = b.mn >= i, ne = b.mne >= i;
        if (wl && (k & 63) == 0) uf_unite<true>(par, k, k - 1);
        if (nn) {
            if (!(wl && nw)) uf_unite<true>(par, k, jn);
        } else {
            if (nw && !wl) uf_unite<true>(par, k, j0);
            if (ne) uf_unite<true>(par, k, jne);
        }
    };
    for (int k0 = tid; k0 < n; k0 += DEB_ILP * 1024) {
        NbP<int> b[DEB_ILP];
#pragma unroll
        for (int u = 0; u < DEB_ILP; u++)
            if (k0 + u * 1024 < n) b[u] = nb[k0 + u * 1024];
#pragma unroll
        for (int u = 0; u < DEB_ILP; u++)
            if (k0 + u * 1024 < n) join(k0 + u * 1024, b[u]);
    }
    __syncthreads();
    /* roots and sizes: DEB_ILP pixels per lane walk their paths side by side; a loop
       bound the wavefront shares, the sizes added per wavefront */
    for (int b0 = 0; b0 < n; b0 += DEB_ILP * 1024) {
        int x[DEB_ILP], pp[DEB_ILP];
        bool act[DEB_ILP];
#pragma unroll
        for (int u = 0; u < DEB_ILP; u++) {
            const int k = b0 + u * 1024 + tid;
            act[u] = k < n && m[k] >= i;
            x[u] = k;
            pp[u] = act[u] ? ldw<true>(&par[k]) : k;
        }
        bool busy = true;
        while (busy) {
            busy = false;
            int gp[DEB_ILP];
#pragma unroll
            for (int u = 0; u < DEB_ILP; u++)
                gp[u] = pp[u] != x[u] ? ldw<true>(&par[pp[u]]) : pp[u];
#pragma unroll
            for (int u = 0; u < DEB_ILP; u++)
                if (pp[u] != x[u]) {
                    if (gp[u] != pp[u]) atomicMin(&par[x[u]], gp[u]);
                    x[u] = pp[u];
                    pp[u] = gp[u];
                    busy = busy || pp[u] != x[u];
                }
        }
#pragma unroll
        for (int u = 0; u < DEB_ILP; u++) {
            if (act[u]) par[b0 + u * 1024 + tid] = x[u];
            wave_add_by_key(cnt, act[u] ? x[u] : -1, 1);
        }
    }
    __syncthreads();
    /* the kept components in raster order of their first pixels: each thread numbers a
       contiguous run of the list after one prefix over the threads' counts; a kept root's
       parent word becomes -2 - its index, another root's -1 */
    int nkept = 0;
    {
        const int per = (n + 1023) / 1024;
        const int kb = tid * per, ke = min(kb + per, n);
        auto load = [&](int k0, int *pk, int *ck) {
#pragma unroll
            for (int u = 0; u < DEB_RUN; u++)
                pk[u] = k0 + u < ke && m[k0 + u] >= i ? ldw<true>(&par[k0 + u]) : -1;
#pragma unroll
            for (int u = 0; u < DEB_RUN; u++)
                ck[u] = pk[u] == k0 + u ? ldw<true>(&cnt[k0 + u]) : 0;
        };
        int mine = 0;
        for (int k0 = kb; k0 < ke; k0 += DEB_RUN) {
            int pk[DEB_RUN], ck[DEB_RUN];
            load(k0, pk, ck);
#pragma unroll
            for (int u = 0; u < DEB_RUN; u++) mine += pk[u] == k0 + u && ck[u] >= a.debarea;
        }
        int j = wg_scan_excl<1024>(mine, wsum, nkept);
        for (int k0 = kb; k0 < ke; k0 += DEB_RUN) {
            int pk[DEB_RUN], ck[DEB_RUN];
            load(k0, pk, ck);
#pragma unroll
            for (int u = 0; u < DEB_RUN; u++)
                if (pk[u] == k0 + u) {
                    if (ck[u] >= a.debarea) {
                        LvNode q;
                        q.kroot = k0 + u; q.count = ck[u]; q.sig = 0; q.f = 0.0f;
                        q.ssum = 0.0; q.sabs = 0.0;
                        rec[j] = q;
                        par[k0 + u] = -2 - j;
                        j++;
                    } else {
                        par[k0 + u] = -1;
                    }
                }
        }
    }
    if (tid == 0) lvk[(size_t)oj * a.nthresh + i] = nkept;
    __syncthreads();
    /* the node index of every pixel of V_i (in the count row), the node sums when the
       level keeps two nodes or more */
    const bool sums = nkept >= 2;
    for (int b0 = 0; b0 < n; b0 += DEB_ILP * 1024) {
        int v[DEB_ILP], w[DEB_ILP];
        bool in[DEB_ILP];
#pragma unroll
        for (int u = 0; u < DEB_ILP; u++) {
            const int k = b0 + u * 1024 + tid;
            in[u] = k < n && m[k] >= i;
            v[u] = in[u] ? ldw<true>(&par[k]) : -1;
        }
#pragma unroll
        for (int u = 0; u < DEB_ILP; u++) w[u] = v[u] >= 0 ? ldw<true>(&par[v[u]]) : v[u];
#pragma unroll
        for (int u = 0; u < DEB_ILP; u++) {
            const int k = b0 + u * 1024 + tid;
            const int nid = in[u] && w[u] <= -2 ? -2 - w[u] : -1;
            if (in[u]) cnt[k] = nid;
            if (sums) {
                const float x = nid >= 0 ? cdv[k] : 0.0f;
                wave_add2_by_key(rec, nid, (double)x, fabs((double)x));
            }
        }
    }
    if (!sums) return;
    __syncthreads();
    /* significance bounds, one wavefront per node; a node's exact float sum where they
       cannot decide (its count pixels start at its first pixel) */
    const float ti = a.thr[(size_t)r * a.nthresh + i];
    const double v0lo = v0[2 * oj], v0hi = v0[2 * oj + 1];
    for (int c = wid; c < nkept; c += 16) {
        const int cc = ldw<true>(&rec[c].count), kr = ldw<true>(&rec[c].kroot);
        const float b = ti * (float)cc;
        const int sig = deb_sig_bound(ldw<true>(&rec[c].ssum), ldw<true>(&rec[c].sabs), cc, b,
                                      v0lo, v0hi);
        float f = 0.0f;
        if (sig < 0)
            f = wave_ordered_sum(
                cdv, kr, n, [&](int k) { return m[k] >= i && ldw<true>(&cnt[k]) == c; }, cc);
        if (lane == 0) {
            rec[c].sig = sig;
            rec[c].f = f;
        }
    }
}

__global__ void __launch_bounds__(1024) k_deb_glb_tree(DebArgs a, const int *lvnid,
                                                      const LvNode *lvrec, const int *lvk,
                                                      const float *frbuf)
{
    if (a.pd && (int)blockIdx.x >= deb_nbig(a)) return;
    const int r = a.clist[3 * (size_t)a.cstride + blockIdx.x];
    const int off = a.off[r], n = a.cnt[r];
    LvArgs lv{};
    lv.nid = lvnid + a.lvbase[2 * r];
    lv.rec = lvrec + a.lvbase[2 * r + 1];
    lv.nk = lvk + (size_t)blockIdx.x * a.nthresh;
    lv.nstride = n / max(a.debarea, 1) + 1;
    lv.fr = frbuf[blockIdx.x];
    deb_tree<1024, true, int, 2>(a, r, a.g_cdv + off, a.g_m + off,
                                 reinterpret_cast<NbP<int> *>(a.g_nb) + off, a.g_par + off,
                                 a.g_cnt + off, a.g_deep + off, nullptr, a.fpia + off,
                                 a.nodes + a.compoff[r], &lv);
}

/* ------------------------------------------------------------------ objects ---- */
struct ObjArgs {
    DebArgs d;
    const int *objbase;     /* first object index of each raw object */
    int *objraw, *objcomp;  /* raw object and component of each object */
    int *alloc;             /* frame-wide object index of each raw pixel */
    const float *sub;       /* background-subtracted frame */
    int *okey, *oval;       /* object index / pixel of each raw-list entry (sort input) */
    ObjCols o;
    int nobj;
    const int *order;       /* objects in launch order: those of large raw objects first */
};

/* objects whose raw object has at least PRE_BIG pixels are preanalysed first (each scans
   its raw object's list) */
#define PRE_BIG 256

/* the objects: the device count of a planned run */
__device__ inline int obj_n(const ObjArgs &A) { return A.d.pd ? A.d.pd->nobj : A.nobj; }

__global__ void __launch_bounds__(256) k_pre_order(ObjArgs A, int *cnt, int *order)
{
    const int obj = blockIdx.x * 256 + threadIdx.x;
    const int nobj = obj_n(A);
    if (obj >= nobj) return;
    if (A.d.cnt[A.objraw[obj]] >= PRE_BIG)
        order[atomicAdd(&cnt[0], 1)] = obj;
    else
        order[nobj - 1 - atomicAdd(&cnt[1], 1)] = obj;
}

/* objects of each raw object, in component-table order (one wavefront per raw object) */
__global__ void __launch_bounds__(256) k_deb_objlist(ObjArgs A, int *zero2)
{
    /* the two counters of k_pre_order start at zero (no fill launch) */
    if (blockIdx.x == 0 && threadIdx.x < 2) zero2[threadIdx.x] = 0;
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= deb_nraw(A.d)) return;
    const int nc = A.d.rootnc[r], base = A.objbase[r];
    Comp *C = A.d.comps + A.d.compoff[r];
    if (A.d.lstar[r] == 0) {
        /* a single local maximum: the raw object is the object */
        if (lane == 0) {
            Comp c0;
            c0.level = 0;
            c0.first = A.d.first[r];
            c0.flags = COMP_KEPT;
            C[0] = c0;
            A.objraw[base] = r;
            A.objcomp[base] = 0;
        }
        return;
    }
    const unsigned long long below = sxw_lanes_below(lane);
    int done = 0;
    for (int c0 = 0; c0 < nc; c0 += 64) {
        const int c = c0 + lane;
        const bool kept = c < nc && (C[c].flags & COMP_KEPT);
        const unsigned long long b = __ballot(kept);
        if (kept) {
            const int j = base + done + __popcll(b & below);
            A.objraw[j] = r;
            A.objcomp[j] = c;
        }
        done += __popcll(b);
    }
}


/*
 * FULL preanalysis of each kept component over its level pixels (one wavefront per
 * object); the component's pixels get the object index.
 */
__global__ void __launch_bounds__(256) k_deb_pre(ObjArgs A)
{
    __shared__ double s_rows[4][6 * SXW_ROW];
    __shared__ float s_frows[4][SXW_ROW];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int slot = blockIdx.x * 4 + wid;
    if (slot >= obj_n(A)) return;
    const int obj = A.order[slot];
    double *rows = s_rows[wid];
    float *frows = s_frows[wid];
    const DebArgs &d = A.d;
    const int r = A.objraw[obj], c = A.objcomp[obj];
    const Comp *C = d.comps + d.compoff[r];
    const int L = C[c].level;
    const int n = d.cnt[r], off = d.off[r], W = d.W, H = d.H;
    const int *lv = L > 0 ? d.kcomp + off : nullptr;
    /* extent, peaks, count (reductions); the node's pixels get the object index */
    float peak = (float)-SXG_BIG, cpeak = (float)-SXG_BIG;
    int xmin = 2 * SXG_MAXPICSIZE, ymin = 2 * SXG_MAXPICSIZE, xmax = 0, ymax = 0, cnt = 0;
#pragma unroll 4
    for (int k = lane; k < n; k += 64) {
        if (lv && lv[k] != c) continue;
        const int p = d.fpia[off + k];
        const int y = p / W, x = p - y * W;
        A.alloc[p] = obj;
        peak = fmaxf(peak, A.sub[p]);
        cpeak = fmaxf(cpeak, d.cd[p]);
        xmin = min(xmin, x);
        xmax = max(xmax, x);
        ymin = min(ymin, y);
        ymax = max(ymax, y);
        cnt++;
    }
    peak = sxw_max(peak);
    cpeak = sxw_max(cpeak);
    xmin = sxw_min(xmin);
    xmax = sxw_max(xmax);
    ymin = sxw_min(ymin);
    ymax = sxw_max(ymax);
    cnt = sxw_sum(cnt);
    int flag = L > 0 ? SXG_OBJ_MERGED : 0;
    if (xmin == 0 || xmax == W - 1 || ymin == 0 || ymax == H - 1) flag |= SXG_OBJ_TRUNC;
    const double dthresh = (double)d.thr[(size_t)r * d.nthresh + L];
    const double dthresh2 = (dthresh + peak) / 2.0;
    /* one fold pass: rv, the raster-order float sum of cd, and the moments, double sums
       of float terms in raster order (independent chains, one lane each) */
    double acc = 0.0;
    float racc = 0.0f;
    int area2 = 0, dnpix = 0;
    /* the next chunk's membership and pixel are read while this chunk is folded */
    bool nin = lane < n && (!lv || lv[lane] == c);
    int npix_ = lane < n ? d.fpia[off + lane] : 0;
    for (int base = 0; base < n; base += 64) {
        const int k = base + lane;
        const bool in = nin;
        const int pc = npix_;
        if (k + 64 < n) {
            nin = !lv || lv[k + 64] == c;
            npix_ = d.fpia[off + k + 64];
        } else {
            nin = false;
        }
        float val = 0.0f, cval = 0.0f, tx = 0.0f, ty = 0.0f, txx = 0.0f, tyy = 0.0f, txy = 0.0f;
        if (in) {
            const int p = pc;
            const int y = p / W - ymin, x = p - (p / W) * W - xmin;
            cval = d.cd[p];
            val = A.sub[p];
            if (val > dthresh) dnpix++;
            if (val > dthresh2) area2++;
            tx = cval * x;
            ty = cval * y;
            txx = cval * x * x;
            tyy = cval * y * y;
            txy = cval * x * y;
        }
        const double t6[6] = {val, tx, ty, txx, tyy, txy};
        const float tf[1] = {cval};
        sxw_fold2<6, 1>(rows, frows, t6, tf, in, acc, racc);
    }
    const float rv = sxw_rl(racc, 6);
    const double tv = sxw_rl(acc, 0), mx = sxw_rl(acc, 1), my = sxw_rl(acc, 2);
    const double mx2 = sxw_rl(acc, 3), my2 = sxw_rl(acc, 4), mxy = sxw_rl(acc, 5);
    area2 = sxw_sum(area2);
    dnpix = sxw_sum(dnpix);
    if (lane != 0) return;
    Moments m;
    sxg_moments(mx, my, mx2, my2, mxy, rv, false, 0.0, 0.0, xmin, ymin, area2, dnpix, dthresh,
                dthresh2, m);
    ObjCols &o = const_cast<ObjCols &>(A.o);
    o.level[obj] = L;
    o.root[obj] = r;
    o.firstpix[obj] = C[c].first;
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

/* gatherup: every unassigned pixel goes to the nearest kept object of its raw object
   in the ellipse metric (first minimum in object order); then the sort keys */
__global__ void __launch_bounds__(256) k_deb_gather(ObjArgs A, int *objcnt, long nz)
{
    const long g = (long)blockIdx.x * 256 + threadIdx.x;
    /* the per-object pixel counts of k_deb_objcount start at zero (no fill launch) */
    if (g < nz) objcnt[g] = 0;
    if (g >= deb_P0(A.d)) return;
    const int p = A.d.fpia[g];
    const int r = A.d.rawkey[g];
    /* k_deb_pre gave every pixel of a raw object kept whole, and every pixel of a kept
       node, its object; the rest are the pixels of split raw objects outside all kept
       nodes */
    int a;
    if (A.d.rootkept[r] >= 2 && A.d.kcomp[g] < 0) {
        const int base = A.objbase[r], nk = A.d.rootkept[r];
        const int y = p / A.d.W, x = p - y * A.d.W;
        const ObjCols &o = A.o;
        float distmin = (float)1e+31, dist, dx, dy;
        a = base;
        for (int j = base; j < base + nk; j++) {
            dx = (float)(x - o.mx[j]);
            dy = (float)(y - o.my[j]);
            dist = (float)(0.5 * (o.cxx[j] * dx * dx + o.cyy[j] * dy * dy + o.cxy[j] * dx * dy) /
                           o.abcor[j]);
            if (dist < distmin) {
                distmin = dist;
                a = j;
            }
        }
        A.alloc[p] = a;
    } else {
        a = A.alloc[p];
    }
    A.okey[g] = a;
    A.oval[g] = p;
}

/* pixels per object; lanes with the same object (long runs: the list is grouped by raw
   object) add their count once */
__global__ void __launch_bounds__(256) k_deb_objcount(ObjArgs A, int *objcnt)
{
    const long g = (long)blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    bool act = g < deb_P0(A.d);
    const int key = act ? A.okey[g] : -1;
    unsigned long long pending = __ballot(act);
    while (pending) {
        const int leader = __builtin_ctzll(pending);
        const int k = __shfl(key, leader);
        const unsigned long long grp = __ballot(act && key == k);
        if (lane == leader) atomicAdd(&objcnt[k], __popcll(grp));
        if (act && key == k) act = false;
        pending &= ~grp;
    }
}

/*
 * The raw pixels regrouped by object, stably (each object's pixels in the raster order
 * of its raw object's list), from the object offsets: a raw object kept as one object
 * moves as one block (k_obj_place1, a lane per pixel); a raw object split into several,
 * one wavefront each (k_obj_place_multi), walks its list 64 entries at a time with the
 * next free position of each of its objects in LDS: the lanes of one object take
 * consecutive positions in lane order, and the chunks follow the list's order.
 */
__global__ void __launch_bounds__(256) k_obj_place1(ObjArgs A, const int *__restrict__ objoff,
                                                    int *opix)
{
    const long g = (long)blockIdx.x * 256 + threadIdx.x;
    if (g >= deb_P0(A.d)) return;
    const int r = A.d.rawkey[g];
    if (A.d.rootkept[r] != 1) return;
    opix[objoff[A.okey[g]] + (int)(g - A.d.off[r])] = A.d.fpia[g];
}

/* positions kept in LDS per wavefront: objects of one raw object per window */
#define OP_WIN 1024
/* list chunks of 64 entries read ahead of the one being placed */
#define OP_AHEAD 4
__global__ void __launch_bounds__(256) k_obj_place_multi(ObjArgs A,
                                                         const int *__restrict__ objoff,
                                                         int *opix)
{
    __shared__ int s_next[4][OP_WIN];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int r = blockIdx.x * 4 + wid;
    if (r >= deb_nraw(A.d)) return;
    const int nk = A.d.rootkept[r];
    if (nk < 2) return;
    const int base = A.objbase[r], off = A.d.off[r], n = A.d.cnt[r];
    const unsigned long long below = sxw_lanes_below(lane);
    int *nxt = s_next[wid];
    for (int w0 = 0; w0 < nk; w0 += OP_WIN) {
        const int wn = min(OP_WIN, nk - w0);
        for (int j = lane; j < wn; j += 64) nxt[j] = objoff[base + w0 + j];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        /* chunk c's object ids and pixels, OP_AHEAD chunks in flight */
        int qk[OP_AHEAD], qp[OP_AHEAD];
#pragma unroll
        for (int u = 0; u < OP_AHEAD; u++) {
            const int k = u * 64 + lane;
            qk[u] = k < n ? A.okey[off + k] : -1;
            qp[u] = k < n ? A.d.fpia[off + k] : 0;
        }
        for (int c0 = 0; c0 < n; c0 += 64 * OP_AHEAD) {
#pragma unroll
            for (int u = 0; u < OP_AHEAD; u++) {
                const int c = c0 + u * 64;
                if (c >= n) break;
                const int key = qk[u], pix = qp[u];
                /* the same slot's chunk of the next round */
                const int kn = c + 64 * OP_AHEAD + lane;
                qk[u] = kn < n ? A.okey[off + kn] : -1;
                qp[u] = kn < n ? A.d.fpia[off + kn] : 0;
                const int j = key - base - w0;
                bool act = c + lane < n && j >= 0 && j < wn;
                unsigned long long pending = __ballot(act);
                while (pending) {
                    const int lead = __builtin_ctzll(pending);
                    const int jl = __shfl(j, lead);
                    const unsigned long long grp = __ballot(act && j == jl);
                    const int start = nxt[jl];
                    if (act && j == jl) {
                        opix[start + __popcll(grp & below)] = pix;
                        act = false;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (lane == lead) nxt[jl] = start + __popcll(grp);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    pending &= ~grp;
                }
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_deb_zero(int *v, long n)
{
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    if (k < n) v[k] = 0;
}

/* planned runs: the node total and the level-parallel arrays against their capacities
   (the 64-bit node sum too: a total past 2^31 is an overflow); on any overflow of the
   frame the class counts are zeroed so that no tree kernel has work */
/* the capacities of the class words of the counter block: LDS bytes of classes 0..2,
   level-parallel objects, objects per class 0..3 */
struct DebGuardCaps {
    long long lds[3], big, cls[4];
};

__global__ void k_deb_plan_guard(const unsigned long long *dw, PlanDev *pd, int *cc,
                                 const int *ntot, long long cap_tot, long long cap_npl,
                                 long long cap_nrec, DebGuardCaps gc)
{
    if (threadIdx.x != 0) return;
    if (dw[2] > (unsigned long long)cap_tot || (long long)*ntot > cap_tot || *ntot < 0)
        pd->ovf |= 1 << PL_NTOT;
    if (dw[3] > (unsigned long long)cap_npl) pd->ovf |= 1 << PL_NPL;
    if (dw[4] > (unsigned long long)cap_nrec) pd->ovf |= 1 << PL_NREC;
    /* the class words against their capacities (one launch for all of them) */
    for (int c = 0; c < 3; c++)
        if (cc[4 + c] < 0 || (long long)cc[4 + c] > gc.lds[c]) pd->ovf |= 1 << (PL_LDS0 + c);
    if (cc[8] < 0 || (long long)cc[8] > gc.big) pd->ovf |= 1 << PL_BIG;
    for (int c = 0; c < 4; c++)
        if (cc[c] < 0 || (long long)cc[c] > gc.cls[c]) pd->ovf |= 1 << (PL_C0 + c);
    if (pd->ovf) {
        pd->nfg = pd->nraw = pd->npix = pd->nobj = pd->nsurv = 0;
        for (int c = 0; c < 7; c++) cc[c] = 0;
        cc[8] = 0;
    }
}

/* planned runs: the frame's deblending work into its device block, and the level check
   (the eager run's "inconsistent deblending levels") as an overflow bit */
__global__ void k_deb_plan_info(const unsigned long long *dw, const int *cc, PlanDev *pd)
{
    if (threadIdx.x != 0) return;
    pd->deb_pixels = (long long)dw[0];
    pd->deb_labels = (long long)dw[1];
    for (int c = 0; c < 4; c++) pd->class_pixels[c] = cc[10 + c];
    if (cc[7] != 0) pd->ovf |= 1 << SXG_PLAN_DEBLEVELS;
}

/* --------------------------------------------------------------- host side ---- */
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
                           ro.fdpeak, S.nraw, nthresh, basethresh, ro.t0,
                           (int)(S.weighted && S.wargs.rel), thr,
                           (const PlanDev *)(S.plan ? S.pdev : nullptr));
        sxg_kt_end(SXG_KT_DEB_THRESH);
    }
    SXG_HIP(hipGetLastError());
    S.deb_ready = true;
}

extern "C" int parcel_out(float basethresh, int n_thresh, int deb_minarea, size_t numValidObj,
                          size_t numValidPix, double deblend_mincont)
{
    SxgState &S = sxg_state();
    if (!S.have_frame || !S.deb_ready || n_thresh != S.nthresh) {
        sxg_set_error("parcel_out: call init_deblend first with the same DEBLEND_NTHRESH");
        return -1;
    }
    if (deb_minarea < 1) {
        sxg_set_error("parcel_out: deb_minarea must be >= 1");
        return -1;
    }
    (void)basethresh; (void)numValidObj; (void)numValidPix;
    hipStream_t st = S.stream;
    /* a planned run: S.nraw and S.npix_raw are capacities, the counts are on the device */
    const FramePlan *pl = S.plan;
    PlanDev *pd = pl ? S.pdev : nullptr;
    const int nraw = S.nraw;
    const long P0 = S.npix_raw;
    sxg_stage_begin(ST_DEBLEND);
    RawObj ro;
    if (!sxg_raw_alloc(nraw, ro)) return -1;
    const size_t nr = (size_t)(nraw > 0 ? nraw : 1);
    int *compoff = sxg_take<int>("deb.compoff", nr + 1);
    int *dtot = sxg_take<int>("deb.total", 4);
    int *rootnc = sxg_take<int>("deb.rootnc", nr);
    int *rootkept = sxg_take<int>("deb.rootkept", nr);
    int *objbase = sxg_take<int>("deb.objbase", nr);
    int *alloc = sxg_take<int>("deb.alloc", S.N);
    auto *top1 = sxg_take<unsigned long long>("deb.top1", nr);
    auto *top2 = sxg_take<unsigned>("deb.top2", nr);
    int *nmax = sxg_take<int>("deb.nmax", nr);
    int *lstar = sxg_take<int>("deb.lstar", nr);
    int *clist = sxg_take<int>("deb.clist", 4 * nr);
    int *ccount = sxg_take<int>("deb.ccount", DEB_NCOUNT);
    auto *dwork = sxg_take<unsigned long long>("deb.dwork", 6);
    auto *lvbase = sxg_take<unsigned long long>("deb.lvbase", 2 * (size_t)nr);
    int *kcomp = sxg_take<int>("deb.kcomp", (size_t)(P0 > 0 ? P0 : 1));
    auto *lmflag = sxg_take<unsigned char>("deb.lmflag", (size_t)(P0 > 0 ? P0 : 1));
    if (!compoff || !dtot || !rootnc || !rootkept || !objbase || !alloc || !top1 || !top2 ||
        !nmax || !lstar || !clist || !ccount || !dwork || !kcomp || !lvbase || !lmflag)
        return -1;
    DebArgs a{};
    a.fpia = S.fpia; a.rawkey = S.fpia_keys; a.off = ro.off; a.cnt = ro.count; a.first = ro.first;
    a.cd = S.cd.as<float>(); a.lab = S.aux.as<int>();   /* the detection labels */
    a.kcomp = kcomp; a.P0 = P0; a.thr = sxg_take<float>("deb.thr", nr * n_thresh);
    a.nthresh = n_thresh; a.debarea = deb_minarea; a.mincont = deblend_mincont;
    a.lvmin = (int)std::min<long>(sxg_option(SXG_OPT_DEB_GLB_LEVELS), 0x7fffffffL);
    a.W = S.W; a.H = S.H; a.nraw = nraw; a.compoff = compoff;
    a.cstride = nraw; a.pd = pd;
    a.rootnc = rootnc; a.rootkept = rootkept;
    a.top1 = top1; a.top2 = top2; a.nmax = nmax; a.lstar = lstar; a.clist = clist;
    a.lmflag = lmflag;
    a.ccount = ccount; a.dwork = dwork; a.lvbase = lvbase;
    S.deb_pixels = S.deb_labels = 0;
    for (int c = 0; c < 4; c++) S.deb_class_pixels[c] = 0;
    int nobj = 0;
    if (nraw > 0) {
        const int cap0 = (int)sxg_option(SXG_OPT_DEB_LDS_TINY);
        const int cap1 = (int)sxg_option(SXG_OPT_DEB_LDS_SMALL);
        const int cap2 = (int)sxg_option(SXG_OPT_DEB_LDS_LARGE);
        /* LDS budget of each class's workgroups: its pixel capacity and a node table of
           SXG_OPT_DEB_TABLE nodes (else 192 / 512 / 1024); an object whose own pixels and
           node bound need more goes to the global kernel */
        const long topt = sxg_option(SXG_OPT_DEB_TABLE);
        auto budget = [&](int cap, long dflt) -> int {
            const long nodes = topt > 0 ? topt : dflt;
            return (int)std::min<size_t>(DEB_LDS_LIMIT, (size_t)deb_lds_bytes((int)nodes, cap));
        };
        const int ct0 = budget(cap0, 192), ct1 = budget(cap1, 512), ct2 = budget(cap2, 1024);
        const int gr = sxg_div_up(nraw, 256);
        sxg_kt_begin(SXG_KT_DEB_LMAX);
        hipLaunchKernelGGL(k_deb_prep, dim3(sxg_div_up(nraw > 8 ? nraw : 8, 256)), dim3(256), 0,
                           st, a);
        hipLaunchKernelGGL(k_deb_lmax<0>, dim3(sxg_div_up(P0, 256)), dim3(256), 0, st, a);
        hipLaunchKernelGGL(k_deb_lmax<1>, dim3(sxg_div_up(P0, 256)), dim3(256), 0, st, a);
        hipLaunchKernelGGL(k_deb_lsfirst, dim3(sxg_div_up((long)nraw * n_thresh, 256)), dim3(256),
                           0, st, a);
        hipLaunchKernelGGL(k_deb_nodecap, dim3(gr), dim3(256), 0, st, a);
        sxg_scan_excl(compoff, compoff, nraw, dtot, st, pd ? &pd->nraw : nullptr);
        hipLaunchKernelGGL(k_deb_classify, dim3(gr), dim3(256), 0, st, a, cap0, cap1, cap2, ct0,
                           ct1, ct2);
        sxg_kt_end(SXG_KT_DEB_LMAX);
        int cc[DEB_NCOUNT] = {};
        int ntot = 0;
        unsigned long long dw[5] = {0, 0, 0, 0, 0};
        if (pl) {
            /* the capacities stand for the counts; the guards check the device counts
               against them (the node total first, then every class) */
            ntot = (int)pl->cap[PL_NTOT];
            for (int c = 0; c < 4; c++) cc[c] = (int)pl->cap[PL_C0 + c];
            /* the class's LDS: its budget, or less when the plan asks less */
            const int ctc[3] = {ct0, ct1, ct2};
            for (int c = 0; c < 3; c++)
                cc[4 + c] = (int)std::min<int64_t>(pl->cap[PL_LDS0 + c], ctc[c]);
            cc[8] = (int)pl->cap[PL_BIG];
            dw[3] = (unsigned long long)pl->cap[PL_NPL];
            dw[4] = (unsigned long long)pl->cap[PL_NREC];
            DebGuardCaps gc;
            for (int c = 0; c < 3; c++) gc.lds[c] = cc[4 + c];
            gc.big = cc[8];
            for (int c = 0; c < 4; c++) gc.cls[c] = cc[c];
            hipLaunchKernelGGL(k_deb_plan_guard, dim3(1), dim3(64), 0, st, dwork, pd, ccount,
                               dtot, (long long)pl->cap[PL_NTOT], (long long)pl->cap[PL_NPL],
                               (long long)pl->cap[PL_NREC], gc);
        } else {
            SXG_HIP_RET(hipMemcpyAsync(cc, ccount, sizeof(cc), hipMemcpyDeviceToHost, st), -1);
            SXG_HIP_RET(hipMemcpyAsync(&ntot, dtot, sizeof(int), hipMemcpyDeviceToHost, st), -1);
            SXG_HIP_RET(hipMemcpyAsync(dw, dwork, sizeof(dw), hipMemcpyDeviceToHost, st), -1);
            SXG_HIP_RET(hipStreamSynchronize(st), -1);
            if (dw[2] > (unsigned long long)0x7fffffff) {
                sxg_set_error("parcel_out: deblending trees exceed 2^31 nodes in one frame");
                return -1;
            }
            S.deb_pixels = (int64_t)dw[0];
            for (int c = 0; c < 4; c++) S.deb_class_pixels[c] = cc[10 + c];
            S.deb_labels = (int64_t)dw[1];
            S.seen.cap[PL_NTOT] = ntot;
            for (int c = 0; c < 4; c++) S.seen.cap[PL_C0 + c] = cc[c];
            for (int c = 0; c < 3; c++) S.seen.cap[PL_LDS0 + c] = cc[4 + c];
            S.seen.cap[PL_BIG] = cc[8];
            S.seen.cap[PL_NPL] = (int64_t)dw[3];
            S.seen.cap[PL_NREC] = (int64_t)dw[4];
        }
        a.comps = sxg_take<Comp>("deb.comps", (size_t)ntot + 1);
        if (!a.comps) return -1;
        sxg_kt_begin(SXG_KT_DEB_LEVELS);
        /* one launch per class; LDS: the class's node table, then its largest object's
           pixels */
        /* the classes run side by side: each kernel lasts as long as its slowest tree.
           Everything they read was finished by the synchronisation above. */
        int nside = 0;
        const bool use_side = !S.single_stream && !pl && sxg_option(SXG_OPT_SIDE_STREAMS) != 0;
        /* the dynamic-LDS limit is an attribute of the function, shared by every context:
           it is raised once, to the most any class asks for (DEB_LDS_LIMIT), never per
           launch (batch workers launch these kernels from several host threads) */
        static const bool lds_attr = [] {
            return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_deb_tree_lds<1024, 2>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)DEB_LDS_LIMIT) == hipSuccess &&
                   hipFuncSetAttribute(reinterpret_cast<const void *>(&k_deb_tree_lds<1024, 1>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)DEB_LDS_LIMIT) == hipSuccess &&
                   hipFuncSetAttribute(reinterpret_cast<const void *>(&k_deb_tree_lds<256, 0>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)DEB_LDS_LIMIT) == hipSuccess;
        }();
        if (!lds_attr) {
            sxg_set_error("parcel_out: cannot raise the LDS limit of the deblending kernels");
            return -1;
        }
        auto launch_lds = [&](const void *fn, int cls, int nt, int ct) -> bool {
            (void)ct;
            const size_t lds = (size_t)cc[4 + cls];
            if (lds > DEB_LDS_LIMIT) {
                sxg_set_error("parcel_out: LDS request above the limit (internal error)");
                return false;
            }
            void *args[] = {&a, &cls};
            const bool side = cls != 0 && use_side;
            hipStream_t ks = side ? S.side[nside] : st;
            /* each class kernel is timed on the stream it runs on */
            sxg_kt_begin_on(SXG_KT_DEB_TREE0 + cls, ks);
            SXG_HIP_RET(hipLaunchKernel(fn, dim3(cc[cls]), dim3(nt), args, lds, ks), false);
            sxg_kt_end_on(SXG_KT_DEB_TREE0 + cls, ks);
            if (side) SXG_HIP_RET(hipEventRecord(S.side_ev[nside++], ks), false);
            return true;
        };
        if (cc[3] > 0) {
            a.nodes = sxg_take<DebNode>("deb.nodes", (size_t)ntot + 1);
            a.g_cdv = sxg_take<float>("deb.g_cdv", (size_t)P0);
            a.g_m = sxg_take<unsigned char>("deb.g_m", (size_t)P0);
            a.g_nb = sxg_take<int>("deb.g_nb", 4 * (size_t)P0);
            a.g_par = sxg_take<int>("deb.g_par", (size_t)P0);
            a.g_deep = sxg_take<int>("deb.g_deep", (size_t)P0);
            a.g_cnt = sxg_take<int>("deb.g_cnt", (size_t)P0);
            if (!a.nodes || !a.g_cdv || !a.g_m || !a.g_nb || !a.g_par || !a.g_deep || !a.g_cnt)
                return -1;
            /* class 3 largest first: the objects of at least lvmin pixels (cc[8]) are
               built level-parallel, the others one workgroup each, on a stream of their
               own; both read the ordered list */
            int *clist3 = sxg_take<int>("deb.clist3", 4 * nr);
            if (!clist3) return -1;
            DebArgs a3 = a;
            a3.clist = clist3;
            hipStream_t ks = use_side ? S.side[nside] : st;
            /* planned: the capacities bound both parts (the split is on the device) */
            const int big = cc[8] <= 65535 ? cc[8] : 0, rest = pl ? cc[3] : cc[3] - big;
            const long fullmax = sxg_option(SXG_OPT_DEB_ORDER_FULL);
            hipLaunchKernelGGL(k_deb_order, dim3(sxg_div_up(cc[3], 256)), dim3(256), 0, ks,
                               (const int *)a.clist, (const int *)a.cnt, nraw, cc[3],
                               big > 0 ? a.lvmin : 0, big, cc[3] <= fullmax ? 1 : 0, ccount + 9,
                               clist3, (const int *)(pl ? ccount : nullptr),
                               (int)std::min<long>(fullmax, 0x7fffffffL));
            if (rest > 0) {
                hipStream_t ks2 = ks;
                if (use_side) {
                    ks2 = S.side[nside + 1];
                    SXG_HIP_RET(hipEventRecord(S.side_ev[nside], ks), -1);
                    SXG_HIP_RET(hipStreamWaitEvent(ks2, S.side_ev[nside], 0), -1);
                }
                sxg_kt_begin_on(SXG_KT_DEB_TREE_GLB, ks2);
                hipLaunchKernelGGL(k_deb_tree_glb, dim3(rest), dim3(1024), 0, ks2, a3, big);
                sxg_kt_end_on(SXG_KT_DEB_TREE_GLB, ks2);
                if (use_side) SXG_HIP_RET(hipEventRecord(S.side_ev[nside + 1], ks2), -1);
            }
            if (big > 0) {
                const size_t npl = dw[3] > 0 ? (size_t)dw[3] : 1, nrec = dw[4] > 0 ? dw[4] : 1;
                int *lvpar = sxg_take<int>("deb.lvpar", npl);
                int *lvnid = sxg_take<int>("deb.lvnid", npl);
                LvNode *lvrec = sxg_take<LvNode>("deb.lvrec", nrec);
                int *lvk = sxg_take<int>("deb.lvk", (size_t)big * n_thresh);
                float *frbuf = sxg_take<float>("deb.glbflux", (size_t)big);
                double *v0 = sxg_take<double>("deb.glbv0", 2 * (size_t)big);
                if (!lvpar || !lvnid || !lvrec || !lvk || !frbuf || !v0) return -1;
                sxg_kt_begin_on(SXG_KT_DEB_LVPAR, ks);
                hipLaunchKernelGGL(k_deb_glb_prep, dim3(big), dim3(1024), 0, ks, a3, v0);
                hipLaunchKernelGGL(k_deb_glb_levels, dim3(n_thresh, big), dim3(1024), 0, ks, a3,
                                   lvpar, lvnid, lvrec, lvk, frbuf, (const double *)v0);
                hipLaunchKernelGGL(k_deb_glb_tree, dim3(big), dim3(1024), 0, ks, a3,
                                   (const int *)lvnid, (const LvNode *)lvrec, (const int *)lvk,
                                   (const float *)frbuf);
                sxg_kt_end_on(SXG_KT_DEB_LVPAR, ks);
            }
            if (use_side) {
                SXG_HIP_RET(hipEventRecord(S.side_ev[nside], ks), -1);
                nside += rest > 0 ? 2 : 1;
            }
        }
        if (cc[2] > 0 &&
            !launch_lds(reinterpret_cast<const void *>(&k_deb_tree_lds<1024, 2>), 2, 1024, ct2))
            return -1;
        if (cc[1] > 0 &&
            !launch_lds(reinterpret_cast<const void *>(&k_deb_tree_lds<1024, 1>), 1, 1024, ct1))
            return -1;
        if (cc[0] > 0 &&
            !launch_lds(reinterpret_cast<const void *>(&k_deb_tree_lds<256, 0>), 0, 256, ct0))
            return -1;
        for (int j = 0; j < nside; j++) SXG_HIP_RET(hipStreamWaitEvent(st, S.side_ev[j], 0), -1);
        sxg_kt_end(SXG_KT_DEB_LEVELS);
        SXG_HIP_RET(hipGetLastError(), -1);
        if (pl) {
            sxg_scan_excl(rootkept, objbase, nraw, &pd->nobj, st, &pd->nraw);
            nobj = (int)pl->cap[PL_NOBJ];
            sxg_plan_guard(&pd->nobj, nobj, PL_NOBJ);
            /* the frame's deblending work and level check, for its record */
            hipLaunchKernelGGL(k_deb_plan_info, dim3(1), dim3(64), 0, st,
                               (const unsigned long long *)dwork, (const int *)ccount, pd);
        } else {
            sxg_scan_excl(rootkept, objbase, nraw, dtot, st);
            SXG_HIP_RET(hipMemcpyAsync(&nobj, dtot, sizeof(int), hipMemcpyDeviceToHost, st), -1);
            SXG_HIP_RET(hipMemcpyAsync(cc, ccount + 7, sizeof(int), hipMemcpyDeviceToHost, st),
                        -1);
            SXG_HIP_RET(hipStreamSynchronize(st), -1);
            if (cc[0] != 0) {
                sxg_set_error("parcel_out: inconsistent deblending levels (internal error)");
                return -1;
            }
            S.seen.cap[PL_NOBJ] = nobj;
        }
    } else if (pl) {
        hipLaunchKernelGGL(k_deb_zero, dim3(1), dim3(256), 0, st, &pd->nobj, 1L);
    } else {
        for (int k = PL_NTOT; k <= PL_NOBJ; k++) S.seen.cap[k] = 0;
    }
    S.nobj = nobj;
    S.opix = nullptr;
    S.objoff = nullptr;
    ObjCols oc;
    if (!sxg_obj_alloc(nobj, oc)) return -1;
    if (nraw > 0 && nobj > 0) {
        ObjArgs A{};
        A.d = a; A.objbase = objbase; A.alloc = alloc; A.sub = S.img.as<float>(); A.o = oc;
        A.nobj = nobj;
        A.objraw = sxg_take<int>("obj.raw", nobj);
        A.objcomp = sxg_take<int>("obj.comp", nobj);
        int *k0 = sxg_take<int>("opix.k0", P0), *v0 = sxg_take<int>("opix.v0", P0);
        int *k1 = sxg_take<int>("opix.k1", P0), *v1 = sxg_take<int>("opix.v1", P0);
        int *hist = sxg_take<int>("opix.hist", 256 * (size_t)(sxg_div_up(P0 + 1, 4096) + 1) + 64);
        int *objcnt = sxg_take<int>("obj.pixcount", (size_t)nobj + 1);
        int *objoff = sxg_take<int>("obj.pixoff", (size_t)nobj + 1);
        if (!A.objraw || !A.objcomp || !k0 || !v0 || !k1 || !v1 || !hist || !objcnt || !objoff)
            return -1;
        A.okey = k0;
        A.oval = v0;
        const int gp = sxg_div_up(P0, 256);
        sxg_kt_begin(SXG_KT_DEB_OBJECTS);
        int *order = sxg_take<int>("pre.order", (size_t)nobj + 2);
        if (!order) return -1;
        hipLaunchKernelGGL(k_deb_objlist, dim3(sxg_div_up(nraw, 4)), dim3(256), 0, st, A,
                           order + nobj);
        hipLaunchKernelGGL(k_pre_order, dim3(sxg_div_up(nobj, 256)), dim3(256), 0, st, A,
                           order + nobj, order);
        A.order = order;
        hipLaunchKernelGGL(k_deb_pre, dim3(sxg_div_up(nobj, 4)), dim3(256), 0, st, A);
        /* planned: the counts past the device's nobj stay zero, so the scan over the
           capacity puts the total at objoff[nobj] as well; the gather kernel clears them */
        const long nz = pl ? (long)nobj + 1 : (long)nobj;
        hipLaunchKernelGGL(k_deb_gather, dim3(sxg_div_up(std::max<long>(P0, nz), 256)), dim3(256),
                           0, st, A, objcnt, nz);
        sxg_kt_end(SXG_KT_DEB_OBJECTS);
        /* raw pixels regrouped by object (stable: raster order inside each object) */
        sxg_kt_begin(SXG_KT_OBJ_SORT);
        hipLaunchKernelGGL(k_deb_objcount, dim3(gp), dim3(256), 0, st, A, objcnt);
        if (pl)
            sxg_scan_excl(objcnt, objoff, nobj + 1, nullptr, st);
        else
            sxg_scan_excl(objcnt, objoff, nobj, objoff + nobj, st);
        hipLaunchKernelGGL(k_obj_place1, dim3(gp), dim3(256), 0, st, A, (const int *)objoff, v1);
        hipLaunchKernelGGL(k_obj_place_multi, dim3(sxg_div_up(nraw, 4)), dim3(256), 0, st, A,
                           (const int *)objoff, v1);
        sxg_kt_end(SXG_KT_OBJ_SORT);
        S.opix = v1;
        S.objoff = objoff;
    }
    sxg_stage_end(ST_DEBLEND);
    /* the object count was read above; the kernels queued since need no wait here (the
       next entry points run on the same stream, the getters wait for it) */
    SXG_HIP_RET(hipGetLastError(), -1);
    if (!sxg_stage_wait()) return -1;
    return nobj;
}

extern "C" void clear_deblend(int nthresh)
{
    (void)nthresh;
    sxg_state().deb_ready = false;
}
