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
 *   k_deb_lmax<0,1>  one lane per raw pixel: local maxima; per object the highest
 *                    (packed value and index, 64-bit atomicMax) then the second-highest
 *   k_deb_classify   one lane per object: L*, the level-0 node, single-peak objects
 *                    finished here; the rest binned by size into LDS classes
 *   k_deb_tree_lds   one workgroup per object whose pixel list fits in LDS: cd, m(p),
 *                    backward-neighbour indices (16-bit), union-find parents, labels
 *                    and sizes all in LDS; per level: union-find with LDS atomicMin
 *                    (smallest list index = first raster pixel becomes the root),
 *                    sizes by LDS atomics, kept components numbered in raster order by
 *                    a workgroup ballot prefix, labels written to lvl[level][pixel];
 *                    fluxes of the children of multi-child nodes by one wavefront each
 *                    (ballot + readlane keeps the raster-order float sum); then the
 *                    tree rules level by level, bottom-up, with atomics on the node table
 *   k_deb_tree_glb   the same code on global scratch for objects too large for LDS
 *   k_deb_objlist    one wavefront per raw object: its kept components -> objects
 *   k_deb_pre        one wavefront per object: FULL preanalysis over the component's
 *                    level pixels (lanes compute the terms, the raster-order sums are
 *                    transposed folds: sxg_wave.h)
 *   k_deb_gather     one lane per raw pixel: gatherup of the pixels no kept component
 *                    holds
 *   then the raw pixels are regrouped by object with the stable radix sort (opix,
 *   objoff): every later per-object pass reads one contiguous raster-ordered range
 */
#include <cmath>

#include "sxg_internal.h"
#include "sxg_objects.h"
#include "sxg_wave.h"

#define DW 64

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

/* lock-free union: the larger root is linked to the smaller index */
template <bool G> __device__ inline void uf_unite(int *par, int a, int b)
{
    while (true) {
        a = uf_find<G>(par, a);
        b = uf_find<G>(par, b);
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

/* exclusive prefix of a flag over the workgroup in thread order; all threads call it */
template <int NT> __device__ inline int wg_prefix(bool f, int *wsum, int &total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const unsigned long long b = __ballot(f);
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

/* float sum of v[k] over k < n with sel, in increasing k, computed by one wavefront
   (every lane returns the same value); sel(k) is evaluated for k < n only */
template <class Sel>
__device__ inline float wave_ordered_sum(const float *v, int n, Sel sel)
{
    const int lane = threadIdx.x & 63;
    float f = 0.0f;
    for (int base = 0; base < n; base += 64) {
        const int k = base + lane;
        const bool in = k < n && sel(k);
        const int bits = in ? __float_as_int(v[k]) : 0;
        unsigned long long b = __ballot(in);
        while (b) {
            const int j = __builtin_ctzll(b);
            b &= b - 1;
            f += __int_as_float(__builtin_amdgcn_readlane(bits, j));
        }
    }
    return f;
}

/* ---------------------------------------------------------------- thresholds ---- */
/* the base of object r: basethresh, or with per-pixel thresholds (weight map) the
   lowest threshold among the object's pixels, rawt0[r] */
__global__ void __launch_bounds__(256) k_deb_thresh(const float *__restrict__ fdpeak, int nraw,
                                                    int nthresh, float basethresh,
                                                    const float *__restrict__ rawt0, int perpix,
                                                    float *thr)
{
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    if (k >= (long)nraw * nthresh) return;
    const int r = (int)(k / nthresh), i = (int)(k - (long)r * nthresh);
    const float b = perpix ? rawt0[r] : basethresh;
    const double t0 = (double)b, pk = (double)fdpeak[r];
    thr[k] = i == 0 ? b : (float)(t0 * sx_pow(pk / t0, (double)i / nthresh));
}


/* ------------------------------------------------------------------- levels ---- */
struct DebArgs {
    const int *fpia;        /* raw pixel list (raster order within each raw object) */
    const int *rawkey;      /* raw object id of each raw-list entry */
    const int *off, *cnt, *first;   /* raw objects */
    const float *cd;
    const int *lab;         /* detection labels (root pixel of each detected pixel) */
    int *lvl;               /* lvl[level * P0 + g]: component of the raw pixel g, or -1 */
    long P0;
    const float *thr;       /* thr[r * nthresh + i] */
    int nthresh, debarea;
    double mincont;
    int W, H, nraw;
    Comp *comps;            /* component table */
    const long *compoff;    /* region of each raw object in the table */
    int *rootnc;            /* components per raw object */
    int *rootkept;          /* kept objects per raw object */
    unsigned long long *top1;   /* highest local maximum: uord(cd) << 32 | raw-list index */
    unsigned *top2;         /* uord(cd) of the second-highest local maximum, 0 if none */
    int *lstar;             /* deepest level that can hold two components */
    int *clist;             /* objects of each size class: clist[cls * nraw + j] */
    int *ccount;            /* [0..2] objects per class, [3] largest object of class 1,
                               [4] consistency error flag */
    unsigned long long *dwork;  /* [0] pixels, [1] level labels of the multi-peak objects */
    /* global scratch for objects too large for LDS, indexed by raw-list position */
    float *g_cdv;
    unsigned char *g_m;
    int *g_nb;              /* 4 planes of P0 */
    int *g_par, *g_lab, *g_sz;
};

/* PASS 0: highest local maximum of each raw object; PASS 1: the highest of the others */
template <int PASS> __global__ void __launch_bounds__(256) k_deb_lmax(DebArgs a)
{
    const long g = (long)blockIdx.x * 256 + threadIdx.x;
    if (g >= a.P0) return;
    const int p = a.fpia[g], r = a.rawkey[g];
    const int root = a.lab[p];
    const float v = a.cd[p];
    const int W = a.W, H = a.H;
    const int y = p / W, x = p - y * W;
    for (int dy = -1; dy <= 1; dy++) {
        const int yy = y + dy;
        if (yy < 0 || yy >= H) continue;
        for (int dx = -1; dx <= 1; dx++) {
            const int xx = x + dx;
            if ((dx | dy) == 0 || xx < 0 || xx >= W) continue;
            const int q = yy * W + xx;
            if (a.lab[q] == root && a.cd[q] > v) return;
        }
    }
    const unsigned u = uord(v);
    const unsigned long long key = ((unsigned long long)u << 32) | (unsigned)g;
    if (PASS == 0)
        atomicMax(&a.top1[r], key);
    else if (a.top1[r] != key)
        atomicMax(&a.top2[r], u);
}

__global__ void __launch_bounds__(256) k_deb_prep(DebArgs a)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r < 5) a.ccount[r] = 0;
    if (r < 2) a.dwork[r] = 0ull;
    if (r >= a.nraw) return;
    a.top1[r] = 0ull;
    a.top2[r] = 0u;
}

/* L*, the level-0 node, single-peak objects, size classes */
__global__ void __launch_bounds__(256) k_deb_classify(DebArgs a, int cap_small, int cap_large)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= a.nraw) return;
    const int n = a.cnt[r], N = a.nthresh;
    const float *t = a.thr + (size_t)r * N;
    int ls = 0;
    const unsigned u2 = a.top2[r];
    if (u2 != 0u) {
        const float v2 = uord_inv(u2);
        float tau = -INFINITY;
        ls = N - 1;
        for (int i = 1; i < N; i++) {
            /* a NaN sub-threshold empties its level and ends the descent */
            if (t[i] != t[i] || (tau = fmaxf(tau, t[i])) > v2) {
                ls = i - 1;
                break;
            }
        }
    }
    Comp c0;
    c0.level = 0; c0.parent = -1; c0.first = a.first[r]; c0.count = n;
    c0.fdflux = 0.0f; c0.dthresh = t[0]; c0.nchild = 0; c0.nson = 0;
    c0.flags = ls == 0 ? COMP_KEPT : 0;
    a.comps[a.compoff[r]] = c0;
    a.lstar[r] = ls;
    if (ls == 0) {
        a.rootnc[r] = 1;
        a.rootkept[r] = 1;
        return;
    }
    const int cls = n <= cap_small ? 0 : n <= cap_large ? 1 : 2;
    atomicAdd(&a.dwork[0], (unsigned long long)n);
    atomicAdd(&a.dwork[1], (unsigned long long)n * ls);
    const int j = atomicAdd(&a.ccount[cls], 1);
    a.clist[(size_t)cls * a.nraw + j] = r;
    if (cls == 1) atomicMax(&a.ccount[3], n);
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

/*
 * The component tree of raw object r, levels 1..L*.  cdv, m, par, lab, sz and the four
 * neighbour planes nb[d * nbs + k] are per raw-list index k (LDS or global scratch);
 * pix is the object's raw pixel list (pixw: where to copy it first, LDS only).
 */
template <int NT, bool G, typename NB>
__device__ __forceinline__ void deb_tree(const DebArgs &a, const int r, float *cdv,
                                         unsigned char *m, NB *nb, const long nbs, int *par,
                                         int *lab, int *sz, int *pixw, const int *pix)
{
    __shared__ int wsum[NT / 64];
    __shared__ int levstart[SXG_MAXNTHRESH + 1];
    __shared__ float tthr[SXG_MAXNTHRESH], ttau[SXG_MAXNTHRESH];
    __shared__ int s_multi, s_nk;
    __shared__ double s_rsum[NT / 64], s_rabs[NT / 64], s_v0lo, s_v0hi;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int n = a.cnt[r], off = a.off[r], ls = a.lstar[r];
    const int W = a.W;
    const long P0 = a.P0;
    Comp *C = a.comps + a.compoff[r];
    const float *t = a.thr + (size_t)r * a.nthresh;
    const NB none = (NB)-1;
    if (tid == 0) {
        float tau = -INFINITY;
        for (int i = 1; i <= ls; i++) {
            tau = fmaxf(tau, t[i]);
            tthr[i] = t[i];
            ttau[i] = tau;
        }
        s_multi = 0;
        s_nk = 0;
        levstart[0] = 0;
        levstart[1] = 1;
    }
    for (int k = tid; k < n; k += NT) {
        const int p = a.fpia[off + k];
        if (!G) pixw[k] = p;
        cdv[k] = a.cd[p];
    }
    __syncthreads();
    /* level index and backward neighbours (W, NW, N, NE) by position in the list; bounds
       of the root's flux */
    double rsum = 0.0, rabs = 0.0;
    for (int k = tid; k < n; k += NT) {
        const float v = cdv[k];
        rsum += v;
        rabs += fabs(v);
        int mk = 0;
        while (mk < ls && ttau[mk + 1] <= v) mk++;
        m[k] = (unsigned char)mk;
        const int p = pix[k];
        const int y = p / W, x = p - y * W;
        NB w = none, nw = none, nn = none, ne = none;
        if (x > 0 && k > 0 && pix[k - 1] == p - 1) w = (NB)(k - 1);
        if (y > 0) {
            const int q0 = p - W - (x > 0 ? 1 : 0);
            int lo = 0, hi = k;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (pix[mid] < q0) lo = mid + 1;
                else hi = mid;
            }
            int j = lo;
            if (x > 0 && j < k && pix[j] == p - W - 1) nw = (NB)(j++);
            if (j < k && pix[j] == p - W) nn = (NB)(j++);
            if (x < W - 1 && j < k && pix[j] == p - W + 1) ne = (NB)j;
        }
        nb[k] = w;
        nb[nbs + k] = nw;
        nb[2 * nbs + k] = nn;
        nb[3 * nbs + k] = ne;
        lab[k] = 0;
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
    }
    int nc = 1, last = 0;
    for (int i = 1; i <= ls; i++) {
        const float ti = tthr[i];
        int anyl = 0;
        for (int k = tid; k < n; k += NT) {
            const bool in = m[k] >= i;
            par[k] = in ? k : -1;
            sz[k] = 0;
            anyl |= in;
        }
        if (!__syncthreads_or(anyl)) break;
        for (int k = tid; k < n; k += NT) {
            if (m[k] < i) continue;
            /* W, NW, N, NE; NW and NE are already joined through N when N is in the
               set (they are its W and E neighbours), NW also through W */
            const NB jw = nb[k], jnw = nb[nbs + k], jn = nb[2 * nbs + k], jne = nb[3 * nbs + k];
            const bool w = jw != none && m[jw] >= i, nn = jn != none && m[jn] >= i;
            if (w) uf_unite<G>(par, k, (int)jw);
            if (nn) {
                uf_unite<G>(par, k, (int)jn);
            } else {
                if (!w && jnw != none && m[jnw] >= i) uf_unite<G>(par, k, (int)jnw);
                if (jne != none && m[jne] >= i) uf_unite<G>(par, k, (int)jne);
            }
        }
        __syncthreads();
        for (int k = tid; k < n; k += NT)
            if (m[k] >= i) par[k] = uf_find<G>(par, k);
        __syncthreads();
        for (int k = tid; k < n; k += NT)
            if (m[k] >= i) atomicAdd(&sz[par[k]], 1);
        __syncthreads();
        /* kept components in raster order of their first pixel */
        int nkept = 0;
        for (int base = 0; base < n; base += NT) {
            const int k = base + tid;
            const bool isroot = k < n && m[k] >= i && ldw<G>(&par[k]) == k;
            const int cnt = isroot ? ldw<G>(&sz[k]) : 0;
            const bool kp = isroot && cnt >= a.debarea;
            int tot;
            const int pre = wg_prefix<NT>(kp, wsum, tot);
            if (isroot) {
                int kid = -1;
                if (kp) {
                    kid = nc + nkept + pre;
                    int parent = lab[k];
                    if (parent < 0) {
                        /* cannot happen when the level sets nest; reported, never written */
                        atomicOr(&a.ccount[4], 1);
                        parent = 0;
                    }
                    Comp cc;
                    cc.level = i; cc.parent = parent; cc.first = a.fpia[off + k]; cc.count = cnt;
                    cc.fdflux = 0.0f; cc.dthresh = ti; cc.nchild = 0; cc.nson = 0; cc.flags = 0;
                    C[kid] = cc;
                    if (atomicAdd(&C[parent].nchild, 1) == 1) s_multi = 1;
                }
                sz[k] = kid;
            }
            nkept += tot;
        }
        __syncthreads();
        for (int k = tid; k < n; k += NT) {
            const int l = m[k] >= i ? ldw<G>(&sz[ldw<G>(&par[k])]) : -1;
            lab[k] = l;
            a.lvl[(size_t)i * P0 + off + k] = l;
        }
        nc += nkept;
        last = i;
        if (tid == 0) levstart[i + 1] = nc;
        __syncthreads();
        /* significance of each new component whose parent has two or more children
           (one wavefront per component; only their significance matters) */
        for (int c = nc - nkept + wid; c < nc; c += NT / 64) {
            const int pc = C[c].parent;
            if (ldw<true>(&C[pc].nchild) < 2) continue;
            double ps = 0.0, pa = 0.0;
            for (int k = lane; k < n; k += 64)
                if (lab[k] == c) {
                    const double v = cdv[k];
                    ps += v;
                    pa += fabs(v);
                }
            ps = sxw_sum(ps);
            pa = sxw_sum(pa);
            const int cnt = C[c].count;
            const float b = C[c].dthresh * (float)cnt;
            int sig = deb_sig_bound(ps, pa, cnt, b, s_v0lo, s_v0hi);
            if (sig < 0) {
                /* too close to call: the exact sequential sums */
                const float f = wave_ordered_sum(cdv, n, [&](int k) { return lab[k] == c; });
                const float fr = wave_ordered_sum(cdv, n, [](int) { return true; });
                const float lhs = f - b;
                sig = (double)lhs > (double)fr * a.mincont;
                if (lane == 0) C[c].fdflux = f;
            }
            if (sig && lane == 0) {
                atomicOr(&C[c].flags, COMP_SIG);
                atomicAdd(&C[pc].nson, 1);
            }
        }
        if (nkept == 0) break;
    }
    __syncthreads();
    if (!s_multi) {
        /* no node has two children: the raw object is kept as it is */
        if (tid == 0) {
            C[0].flags = COMP_KEPT;
            a.rootnc[r] = nc;
            a.rootkept[r] = 1;
        }
        return;
    }
    for (int c = tid; c < nc; c += NT) atomicOr(&C[c].flags, COMP_OKA);
    __syncthreads();
    /* A(c) = nson(c) <= 1 and A(children), bottom-up one level at a time */
    for (int L = last; L >= 0; L--) {
        for (int c = levstart[L] + tid; c < levstart[L + 1]; c += NT) {
            int fl = ldw<true>(&C[c].flags);
            if (ldw<true>(&C[c].nson) > 1) {
                atomicAnd(&C[c].flags, ~COMP_OKA);
                fl &= ~COMP_OKA;
            }
            if (c > 0 && !(fl & COMP_OKA)) atomicAnd(&C[C[c].parent].flags, ~COMP_OKA);
        }
        __syncthreads();
    }
    for (int c = tid; c < nc; c += NT) {
        const int fl = ldw<true>(&C[c].flags);
        const bool okA = fl & COMP_OKA;
        const bool kept = c == 0 ? okA
                                 : (ldw<true>(&C[C[c].parent].nson) > 1 && okA && (fl & COMP_SIG));
        if (kept) {
            atomicOr(&C[c].flags, COMP_KEPT);
            atomicAdd(&s_nk, 1);
        }
    }
    __syncthreads();
    if (tid == 0) {
        a.rootnc[r] = nc;
        a.rootkept[r] = s_nk;
    }
}

/* LDS bytes per pixel of k_deb_tree_lds: cd, parent, label, size (float/int) + four
   16-bit neighbours + the level index */
#define DEB_LDS_BYTES_PER_PIX 25

template <int NT>
__global__ void __launch_bounds__(NT) k_deb_tree_lds(DebArgs a, int cls, int cap)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int r = a.clist[(size_t)cls * a.nraw + blockIdx.x];
    float *cdv = reinterpret_cast<float *>(smem);
    int *par = reinterpret_cast<int *>(smem + 4 * (size_t)cap);
    int *lab = par + cap;
    int *sz = lab + cap;
    unsigned short *nb = reinterpret_cast<unsigned short *>(sz + cap);
    unsigned char *m = reinterpret_cast<unsigned char *>(nb + 4 * (size_t)cap);
    /* the pixel list is only needed while the neighbours are found: it shares sz */
    deb_tree<NT, false, unsigned short>(a, r, cdv, m, nb, cap, par, lab, sz, sz, sz);
}

__global__ void __launch_bounds__(1024) k_deb_tree_glb(DebArgs a)
{
    const int r = a.clist[2 * (size_t)a.nraw + blockIdx.x];
    const int off = a.off[r];
    deb_tree<1024, true, int>(a, r, a.g_cdv + off, a.g_m + off, a.g_nb + off, a.P0,
                              a.g_par + off, a.g_lab + off, a.g_sz + off, nullptr,
                              a.fpia + off);
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
};

/* objects of each raw object, in component-table order (one wavefront per raw object) */
__global__ void __launch_bounds__(256) k_deb_objlist(ObjArgs A)
{
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= A.d.nraw) return;
    const int nc = A.d.rootnc[r], base = A.objbase[r];
    const Comp *C = A.d.comps + A.d.compoff[r];
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

__global__ void __launch_bounds__(256) k_deb_allocinit(ObjArgs A)
{
    const long g = (long)blockIdx.x * 256 + threadIdx.x;
    if (g < A.d.P0) A.alloc[A.d.fpia[g]] = -1;
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
    const int obj = blockIdx.x * 4 + wid;
    if (obj >= A.nobj) return;
    double *rows = s_rows[wid];
    float *frows = s_frows[wid];
    const DebArgs &d = A.d;
    const int r = A.objraw[obj], c = A.objcomp[obj];
    const Comp *C = d.comps + d.compoff[r];
    const int L = C[c].level;
    const int n = d.cnt[r], off = d.off[r], W = d.W, H = d.H;
    const int *lv = L > 0 ? d.lvl + (size_t)L * d.P0 + off : nullptr;
    /* pass 1: extent, peaks, count; rv is the raster-order float sum of cd */
    float peak = (float)-SXG_BIG, cpeak = (float)-SXG_BIG, racc = 0.0f;
    int xmin = 2 * SXG_MAXPICSIZE, ymin = 2 * SXG_MAXPICSIZE, xmax = 0, ymax = 0, cnt = 0;
    for (int base = 0; base < n; base += 64) {
        const int k = base + lane;
        const bool in = k < n && (!lv || lv[k] == c);
        float cval = 0.0f;
        if (in) {
            const int p = d.fpia[off + k];
            const int y = p / W, x = p - y * W;
            A.alloc[p] = obj;
            peak = fmaxf(peak, A.sub[p]);
            cval = d.cd[p];
            cpeak = fmaxf(cpeak, cval);
            xmin = min(xmin, x);
            xmax = max(xmax, x);
            ymin = min(ymin, y);
            ymax = max(ymax, y);
            cnt++;
        }
        const float tf[1] = {cval};
        sxw_fold<1>(frows, tf, in, racc);
    }
    const float rv = sxw_rl(racc, 0);
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
    /* pass 2: the moments, double sums of float terms in raster order */
    double acc = 0.0;
    int area2 = 0, dnpix = 0;
    for (int base = 0; base < n; base += 64) {
        const int k = base + lane;
        const bool in = k < n && (!lv || lv[k] == c);
        float val = 0.0f, tx = 0.0f, ty = 0.0f, txx = 0.0f, tyy = 0.0f, txy = 0.0f;
        if (in) {
            const int p = d.fpia[off + k];
            const int y = p / W - ymin, x = p - (p / W) * W - xmin;
            const float cval = d.cd[p];
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
        sxw_fold<6>(rows, t6, in, acc);
    }
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
__global__ void __launch_bounds__(256) k_deb_gather(ObjArgs A)
{
    const long g = (long)blockIdx.x * 256 + threadIdx.x;
    if (g >= A.d.P0) return;
    const int p = A.d.fpia[g];
    int a = A.alloc[p];
    if (a < 0) {
        const int r = A.d.rawkey[g];
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
    bool act = g < A.d.P0;
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

__global__ void __launch_bounds__(256) k_deb_zero(int *v, long n)
{
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    if (k < n) v[k] = 0;
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
                           (int)(S.weighted && S.wargs.rel), thr);
        sxg_kt_end(SXG_KT_DEB_THRESH);
    }
    SXG_HIP(hipGetLastError());
    S.deb_ready = true;
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
    int *alloc = sxg_take<int>("deb.alloc", S.N);
    auto *top1 = sxg_take<unsigned long long>("deb.top1", nr);
    auto *top2 = sxg_take<unsigned>("deb.top2", nr);
    int *lstar = sxg_take<int>("deb.lstar", nr);
    int *clist = sxg_take<int>("deb.clist", 3 * nr);
    int *ccount = sxg_take<int>("deb.ccount", 8);
    auto *dwork = sxg_take<unsigned long long>("deb.dwork", 2);
    if (!compoff || !dtot || !rootnc || !rootkept || !objbase || !alloc || !top1 || !top2 ||
        !lstar || !clist || !ccount || !dwork)
        return -1;
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
    if (!comps || !lvl) return -1;
    DebArgs a{};
    a.fpia = S.fpia; a.rawkey = S.fpia_keys; a.off = ro.off; a.cnt = ro.count; a.first = ro.first;
    a.cd = S.cd.as<float>(); a.lab = S.lab.as<int>();
    a.lvl = lvl; a.P0 = P0; a.thr = sxg_take<float>("deb.thr", nr * n_thresh);
    a.nthresh = n_thresh; a.debarea = deb_minarea; a.mincont = deblend_mincont;
    a.W = S.W; a.H = S.H; a.nraw = nraw; a.comps = comps; a.compoff = compoff;
    a.rootnc = rootnc; a.rootkept = rootkept;
    a.top1 = top1; a.top2 = top2; a.lstar = lstar; a.clist = clist; a.ccount = ccount;
    a.dwork = dwork;
    S.deb_pixels = S.deb_labels = 0;
    int nobj = 0;
    if (nraw > 0) {
        const int cap_small = (int)sxg_option(SXG_OPT_DEB_LDS_SMALL);
        const int cap_large = (int)sxg_option(SXG_OPT_DEB_LDS_LARGE);
        sxg_kt_begin(SXG_KT_DEB_LMAX);
        hipLaunchKernelGGL(k_deb_prep, dim3(sxg_div_up(nraw > 8 ? nraw : 8, 256)), dim3(256), 0,
                           st, a);
        hipLaunchKernelGGL(k_deb_lmax<0>, dim3(sxg_div_up(P0, 256)), dim3(256), 0, st, a);
        hipLaunchKernelGGL(k_deb_lmax<1>, dim3(sxg_div_up(P0, 256)), dim3(256), 0, st, a);
        hipLaunchKernelGGL(k_deb_classify, dim3(sxg_div_up(nraw, 256)), dim3(256), 0, st, a,
                           cap_small, cap_large);
        sxg_kt_end(SXG_KT_DEB_LMAX);
        int cc[4] = {0, 0, 0, 0};
        SXG_HIP_RET(hipMemcpyAsync(cc, ccount, sizeof(cc), hipMemcpyDeviceToHost, st), -1);
        SXG_HIP_RET(hipStreamSynchronize(st), -1);
        if (cc[2] > 0) {
            a.g_cdv = sxg_take<float>("deb.g_cdv", (size_t)P0);
            a.g_m = sxg_take<unsigned char>("deb.g_m", (size_t)P0);
            a.g_nb = sxg_take<int>("deb.g_nb", 4 * (size_t)P0);
            a.g_par = sxg_take<int>("deb.g_par", (size_t)P0);
            a.g_lab = sxg_take<int>("deb.g_lab", (size_t)P0);
            a.g_sz = sxg_take<int>("deb.g_sz", (size_t)P0);
            if (!a.g_cdv || !a.g_m || !a.g_nb || !a.g_par || !a.g_lab || !a.g_sz) return -1;
        }
        sxg_kt_begin(SXG_KT_DEB_LEVELS);
        if (cc[0] > 0) {
            const int cap = (cap_small + 63) & ~63;
            const size_t lds = (size_t)cap * DEB_LDS_BYTES_PER_PIX;
            SXG_HIP_RET(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_deb_tree_lds<256>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                        -1);
            hipLaunchKernelGGL(k_deb_tree_lds<256>, dim3(cc[0]), dim3(256), lds, st, a, 0, cap);
        }
        if (cc[1] > 0) {
            const int cap = (cc[3] + 63) & ~63;
            const size_t lds = (size_t)cap * DEB_LDS_BYTES_PER_PIX;
            SXG_HIP_RET(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_deb_tree_lds<1024>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                        -1);
            hipLaunchKernelGGL(k_deb_tree_lds<1024>, dim3(cc[1]), dim3(1024), lds, st, a, 1, cap);
        }
        if (cc[2] > 0)
            hipLaunchKernelGGL(k_deb_tree_glb, dim3(cc[2]), dim3(1024), 0, st, a);
        sxg_kt_end(SXG_KT_DEB_LEVELS);
        SXG_HIP_RET(hipGetLastError(), -1);
        sxg_scan_excl(rootkept, objbase, nraw, reinterpret_cast<int *>(dtot), st);
        SXG_HIP_RET(hipMemcpyAsync(&nobj, dtot, sizeof(int), hipMemcpyDeviceToHost, st), -1);
        SXG_HIP_RET(hipMemcpyAsync(cc, ccount + 4, sizeof(int), hipMemcpyDeviceToHost, st), -1);
        unsigned long long dw[2] = {0, 0};
        SXG_HIP_RET(hipMemcpyAsync(dw, dwork, sizeof(dw), hipMemcpyDeviceToHost, st), -1);
        SXG_HIP_RET(hipStreamSynchronize(st), -1);
        S.deb_pixels = (int64_t)dw[0];
        S.deb_labels = (int64_t)dw[1];
        if (cc[0] != 0) {
            sxg_set_error("parcel_out: inconsistent deblending levels (internal error)");
            return -1;
        }
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
        hipLaunchKernelGGL(k_deb_objlist, dim3(sxg_div_up(nraw, 4)), dim3(256), 0, st, A);
        hipLaunchKernelGGL(k_deb_allocinit, dim3(gp), dim3(256), 0, st, A);
        hipLaunchKernelGGL(k_deb_pre, dim3(sxg_div_up(nobj, 4)), dim3(256), 0, st, A);
        hipLaunchKernelGGL(k_deb_gather, dim3(gp), dim3(256), 0, st, A);
        sxg_kt_end(SXG_KT_DEB_OBJECTS);
        /* raw pixels regrouped by object (stable: raster order inside each object) */
        sxg_kt_begin(SXG_KT_OBJ_SORT);
        hipLaunchKernelGGL(k_deb_zero, dim3(sxg_div_up(nobj, 256)), dim3(256), 0, st, objcnt,
                           (long)nobj);
        hipLaunchKernelGGL(k_deb_objcount, dim3(gp), dim3(256), 0, st, A, objcnt);
        sxg_scan_excl(objcnt, objoff, nobj, objoff + nobj, st);
        int nbits = 1;
        while (nbits < 31 && ((nobj - 1) >> nbits) != 0) nbits++;
        const int where = sxg_radix_sort(k0, v0, k1, v1, (int)P0, nbits, hist, st);
        sxg_kt_end(SXG_KT_OBJ_SORT);
        S.opix = where ? v1 : v0;
        S.objoff = objoff;
    }
    sxg_stage_end(ST_DEBLEND);
    SXG_HIP_RET(hipGetLastError(), -1);
    SXG_HIP_RET(hipStreamSynchronize(st), -1);
    return nobj;
}

extern "C" void clear_deblend(int nthresh)
{
    (void)nthresh;
    sxg_state().deb_ready = false;
}
