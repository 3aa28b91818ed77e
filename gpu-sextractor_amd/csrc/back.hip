/*
 * back.hip -- background estimation and subtraction on gfx950 (back_estimate,
 * the reference's src/cuda/cudaback.cu:1212), bit-exact with the CPU code src/back.c.
 *
 * Kernels:
 *   k_meshstat   backstat (back.c:214-290).  The double sums of pix and pix^2 are
 *                sequential in raster order inside the mesh, as back.c does, so one lane
 *                runs each sum's chain; three wavefronts of the workgroup stream the mesh
 *                rows through a double-buffered LDS ring meanwhile.
 *   k_meshmode   one workgroup per mesh: histogram in LDS (back.c:297-341), then the
 *                clipped-mode iteration of backguess (back.c:347-426) on one wavefront.
 *                Inside the iteration every sum is an exact integer (< 2^53), so the
 *                sequential double sums equal prefix-sum differences (64-bit integer
 *                prefixes per 16 bins plus the bins of the two partial blocks); the
 *                two-pointer median walk is a merge of two non-decreasing sequences and
 *                its end point is found by a 64-way search on the prefix counts
 *                (merge-path), which reproduces the walk exactly.
 *   k_fb_*       bad-mesh interpolation and median filtering, one lane per mesh
 *                (back.c:434-530); global medians by radix selection, one workgroup
 *                per map (back.c:532-556); y second derivatives, one lane per mesh
 *                column with the recurrence in registers (back.c:739-780).
 *   k_rownodes   one lane per frame row: node interpolation along y and the x
 *                tridiagonal solve of subbackline (back.c:790-880); 64 rows per
 *                workgroup with the row arrays in LDS when they fit (k_rownodes_lds).
 *   k_subback    one lane per pixel: bicubic evaluation with the column tables of the
 *                cumulative dx recurrence (back.c:880-908), computed once on the host
 *                with the same float additions, and in-place subtraction.
 */
#include <algorithm>
#include <cmath>
#include <vector>

#include "sxg_internal.h"

#define QUANTIF_NSIGMA 5
#define QUANTIF_AMIN 4
#define BACK_MINGOODFRAC 0.5
#define BG_EPS (1e-4)

struct MeshStat {
    float mean, sigma, lcut, hcut, qzero, qscale;
    int nlevels, npix;
};

/* ---------------------------------------------------------------- backstat ---- */
/* statistics of one mesh from its two sequential passes */
struct MeshAcc {
    double mean = 0.0, sigma = 0.0;
    int npix = 0;
};

__device__ inline void meshstat_pass1_end(const MeshAcc &a, int bw, int bh, MeshStat &o, float &lcut,
                                          float &hcut, bool &bad)
{
    o.lcut = o.hcut = o.qzero = o.qscale = 0.0f;
    o.nlevels = o.npix = 0;
    bad = (float)a.npix < (float)(bw * bh * BACK_MINGOODFRAC);
    if (bad) {
        o.mean = o.sigma = (float)-SXG_BIG;
        return;
    }
    double sig;
    const double mean = a.mean / (double)a.npix;
    const double sigma = (sig = a.sigma / a.npix - mean * mean) > 0.0 ? sqrt(sig) : 0.0;
    lcut = o.lcut = (float)(mean - 2.0 * sigma);
    hcut = o.hcut = (float)(mean + 2.0 * sigma);
}

__device__ inline void meshstat_pass2_end(const MeshAcc &a, MeshStat &o)
{
    const double step = sqrt(2 / SXG_PI) * QUANTIF_NSIGMA / QUANTIF_AMIN;
    double sig;
    o.npix = a.npix;
    const double mean = a.mean / (double)a.npix;
    sig = a.sigma / a.npix - mean * mean;
    const double sigma = sig > 0.0 ? sqrt(sig) : 0.0;
    o.mean = (float)mean;
    o.sigma = (float)sigma;
    if ((o.nlevels = (int)(step * a.npix + 1)) > SXG_NMAXLEVELS) o.nlevels = SXG_NMAXLEVELS;
    o.qscale = (float)(sigma > 0.0 ? 2 * QUANTIF_NSIGMA * sigma / o.nlevels : 1.0);
    o.qzero = (float)(mean - QUANTIF_NSIGMA * sigma);
}

/*
 * One workgroup per band of up to MS_MPB meshes along x.  The sums of back.c are two
 * sequential double chains per mesh (pix and pix*pix in raster order); they cannot be
 * reordered (the squares carry up to 48 significant bits, so their double sum rounds),
 * but nothing else has to sit on the chain: wavefronts 1..3 stream the band's rows
 * through a double-buffered LDS ring and store, per pixel, t = (included ? pix : +0)
 * and t2 = (included ? pix*pix : -0) as doubles; in wavefront 0, lane m adds the t of
 * mesh m and lane MS_MPB + m its t2 (counting the non-negative ones), one double add per
 * pixel and lane.  Adding +0 to the first sum and -0 to the second is an exact identity
 * (both start at +0 and stay >= +0 in round-to-nearest).
 */
#define MS_MPB 16
#define MS_THREADS 256
/* loads of a loader thread in flight together */
#define MS_U 4
/* dynamic LDS of a workgroup, the ring: more rows per batch (a larger ring) measured no
   faster, and on a CU it leaves less room for other frames' kernels */
#define MS_LDS_BYTES (128 * 1024)

/* meshes whose rows do not fit the LDS ring: one lane per mesh reading global memory */
__global__ void __launch_bounds__(64) k_meshstat_direct(const float *__restrict__ img,
                                                        const float *__restrict__ wmap, int W,
                                                        int H, int bs, int nbx, int nmesh,
                                                        MeshStat *ms)
{
    const int mi = blockIdx.x * 64 + threadIdx.x;
    if (mi >= nmesh) return;
    const int x0 = (mi % nbx) * bs, y0 = (mi / nbx) * bs;
    const int bw = min(bs, W - x0), bh = min(bs, H - y0);
    MeshAcc acc;
    MeshStat o;
    float lcut = 0.0f, hcut = 0.0f;
    bool bad = false;
    for (int pass = 0; pass < 2 && !(pass == 1 && bad); pass++) {
        for (int y = 0; y < bh; y++) {
            const size_t ro = (size_t)(y0 + y) * W + x0;
            for (int x = 0; x < bw; x++) {
                const double pix = (wmap && !(wmap[ro + x] > 0.0f)) ? (float)-SXG_BIG : img[ro + x];
                if (pass == 0 ? pix > -SXG_BIG : (pix <= hcut && pix >= lcut)) {
                    acc.mean += pix;
                    acc.sigma += pix * pix;
                    acc.npix++;
                }
            }
        }
        if (pass == 0) {
            meshstat_pass1_end(acc, bw, bh, o, lcut, hcut, bad);
            acc = MeshAcc();
        } else {
            meshstat_pass2_end(acc, o);
        }
    }
    ms[mi] = o;
}
/* LDS segment of one mesh row in doubles: column c of the mesh sits at
   (c & 3) * ms_q(bs) + (c >> 2), so the loader lanes, which take four adjacent columns
   each, store to consecutive doubles; the segment length is odd, so the 16 chain lanes
   of an array, which read the same column of 16 meshes, start on distinct bank pairs */
__host__ __device__ inline int ms_q(int bs) { return (bs + 3) >> 2; }
__host__ __device__ inline int ms_seg(int bs) { return (4 * ms_q(bs)) | 1; }

__global__ void __launch_bounds__(MS_THREADS) k_meshstat(const float *__restrict__ img,
                                                         const float *__restrict__ wmap, int W,
                                                         int H, int bs, int nbx, int mpb, int R,
                                                         MeshStat *ms)
{
    extern __shared__ __align__(16) double ring[];
    __shared__ float s_lcut[MS_MPB], s_hcut[MS_MPB];
    __shared__ int s_cnt[MS_MPB];
    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    const int mx0 = blockIdx.x * mpb, my = blockIdx.y;
    const int nm = min(mpb, nbx - mx0);
    const int x0 = mx0 * bs, y0 = my * bs;
    const int bandw = min(W, x0 + nm * bs) - x0;
    const int bh = min(bs, H - y0);
    const int seg = ms_seg(bs), q = ms_q(bs);
    const int bufsz = R * mpb * seg;            /* doubles per buffer and array */
    const int nbatch = (bh + R - 1) / R;
    /* mesh of a band column: a shift when the mesh size is a power of two */
    const int bsh = (bs & (bs - 1)) == 0 ? __builtin_ctz(bs) : -1;
    /* the chain lane's mesh and array (0: t, 1: t2) */
    const int m = lane < mpb ? lane : lane - mpb;
    const int arr = lane < mpb ? 0 : 1;
    const bool chain = wid == 0 && lane < 2 * mpb && m < nm;
    const int bw = min(bs, W - (x0 + m * bs));
    MeshStat o;
    float lcut = 0.0f, hcut = 0.0f;
    bool bad = false;
    const bool vec4 = ((bs | W) & 3) == 0;
    /* wavefronts 1..3: rows [b*R, b*R + nr) of the band into buffer half h, and the
       number of included pixels of each mesh */
    auto load = [&](int pass, int b, int h) {
        double *bt = ring + (size_t)(2 * h) * bufsz, *bt2 = bt + bufsz;
        const int r0 = b * R, nr = min(R, bh - r0);
        const int lt = tid - 64, LT = MS_THREADS - 64;
        auto incl = [&](double pix, int mm) {
            return pass == 0 ? (pix > -SXG_BIG)
                             : ((pix <= (double)s_hcut[mm]) & (pix >= (double)s_lcut[mm]));
        };
        if (vec4) {
            /* all the batch's loads of a thread in flight before any store */
            const int q4 = bandw >> 2, tot = nr * q4;
            const int qsh = (q4 & (q4 - 1)) == 0 ? __builtin_ctz(q4) : -1;
            for (int e0 = lt; e0 < tot; e0 += MS_U * LT) {
                float4 v[MS_U];
#pragma unroll
                for (int u = 0; u < MS_U; u++) {
                    const int e = e0 + u * LT;
                    if (e < tot) {
                        const int r = qsh >= 0 ? e >> qsh : e / q4, c = 4 * (e - r * q4);
                        const size_t ofs = (size_t)(y0 + r0 + r) * W + x0 + c;
                        v[u] = *reinterpret_cast<const float4 *>(img + ofs);
                        if (wmap) {
                            /* pixels of non-positive weight are left out: -BIG */
                            const float4 wv = *reinterpret_cast<const float4 *>(wmap + ofs);
                            const float nb = (float)-SXG_BIG;
                            if (!(wv.x > 0.0f)) v[u].x = nb;
                            if (!(wv.y > 0.0f)) v[u].y = nb;
                            if (!(wv.z > 0.0f)) v[u].z = nb;
                            if (!(wv.w > 0.0f)) v[u].w = nb;
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < MS_U; u++) {
                    const int e = e0 + u * LT;
                    if (e < tot) {
                        const int r = qsh >= 0 ? e >> qsh : e / q4, c = 4 * (e - r * q4);
                        const int mm = bsh >= 0 ? c >> bsh : c / bs, off = c - mm * bs;
                        const double p0 = v[u].x, p1 = v[u].y, p2 = v[u].z, p3 = v[u].w;
                        const bool i0 = incl(p0, mm), i1 = incl(p1, mm), i2 = incl(p2, mm),
                                   i3 = incl(p3, mm);
                        /* off is a multiple of 4: its four columns at off/4 + j*q */
                        const int idx = (r * mpb + mm) * seg + (off >> 2);
                        double *d = bt + idx, *d2 = bt2 + idx;
                        d[0] = i0 ? p0 : 0.0;
                        d[q] = i1 ? p1 : 0.0;
                        d[2 * q] = i2 ? p2 : 0.0;
                        d[3 * q] = i3 ? p3 : 0.0;
                        d2[0] = i0 ? p0 * p0 : -0.0;
                        d2[q] = i1 ? p1 * p1 : -0.0;
                        d2[2 * q] = i2 ? p2 * p2 : -0.0;
                        d2[3 * q] = i3 ? p3 * p3 : -0.0;
                        const int k = (int)i0 + (int)i1 + (int)i2 + (int)i3;
                        if (k) atomicAdd(&s_cnt[mm], k);
                    }
                }
            }
        } else {
            const int tot = nr * bandw;
            for (int e0 = lt; e0 < tot; e0 += MS_U * LT) {
                float v[MS_U];
#pragma unroll
                for (int u = 0; u < MS_U; u++) {
                    const int e = e0 + u * LT;
                    if (e < tot) {
                        const int r = e / bandw, c = e - r * bandw;
                        const size_t ofs = (size_t)(y0 + r0 + r) * W + x0 + c;
                        v[u] = (wmap && !(wmap[ofs] > 0.0f)) ? (float)-SXG_BIG : img[ofs];
                    }
                }
#pragma unroll
                for (int u = 0; u < MS_U; u++) {
                    const int e = e0 + u * LT;
                    if (e < tot) {
                        const int r = e / bandw, c = e - r * bandw;
                        const int mm = bsh >= 0 ? c >> bsh : c / bs, off = c - mm * bs;
                        const double pix = v[u];
                        const bool in = incl(pix, mm);
                        const int idx = (r * mpb + mm) * seg + (off & 3) * q + (off >> 2);
                        bt[idx] = in ? pix : 0.0;
                        bt2[idx] = in ? pix * pix : -0.0;
                        if (in) atomicAdd(&s_cnt[mm], 1);
                    }
                }
            }
        }
    };
    for (int pass = 0; pass < 2; pass++) {
        double acc = 0.0;
        if (tid < MS_MPB) s_cnt[tid] = 0;
        __syncthreads();
        if (wid != 0) load(pass, 0, 0);
        __syncthreads();
        for (int b = 0; b < nbatch; b++) {
            if (wid != 0) {
                if (b + 1 < nbatch) load(pass, b + 1, (b + 1) & 1);
            } else if (chain) {
                /* the lane's chain: one double add per pixel in column order (column c
                   of a row at (c & 3) * q + (c >> 2)) */
                const double *src = ring + (size_t)(2 * (b & 1) + arr) * bufsz + m * seg;
                const int nr = min(R, bh - b * R), rstride = mpb * seg;
                if ((bw & 31) == 0 && ((nr * (bw >> 5)) & 1) == 0) {
                    /* an even number of chunks of 32 columns, in two register buffers
                       used in turn: the reads of the next chunk are in flight while this
                       one is added.  Every load is unconditional (the pointer wraps to
                       the batch's first chunk after its last), so no control-flow merge
                       makes the compiler wait for both buffers at once; no copy between
                       the buffers either, which would wait for the newer reads */
                    const int cpr = bw >> 5, K = nr * cpr;
                    const double *cp = src;
                    int cc = 0, rr = 0;
                    auto ld = [&](double (&d)[32]) {
#pragma unroll
                        for (int u = 0; u < 32; u++) d[u] = cp[(u & 3) * q + (u >> 2)];
                        const bool eor = ++cc == cpr;
                        cc = eor ? 0 : cc;
                        rr += eor ? 1 : 0;
                        const bool eob = rr == nr;
                        rr = eob ? 0 : rr;
                        cp = eob ? src : (eor ? cp + (rstride - (cpr - 1) * 8) : cp + 8);
                    };
                    double A[32], B[32];
                    ld(A);
                    for (int k = 0; k < K; k += 2) {
                        ld(B);
#pragma unroll
                        for (int u = 0; u < 32; u++) acc += A[u];
                        ld(A);
#pragma unroll
                        for (int u = 0; u < 32; u++) acc += B[u];
                    }
                } else {
                    for (int r = 0; r < nr; r++) {
                        const double *row = src + r * rstride;
                        for (int x = 0; x < bw; x++) acc += row[(x & 3) * q + (x >> 2)];
                    }
                }
            }
            __syncthreads();
        }
        /* mesh m: the pix sum from lane m, the pix^2 sum from lane mpb+m */
        const double sq = __shfl(acc, lane + mpb);
        if (wid == 0 && lane < nm) {
            MeshAcc a;
            a.mean = acc;
            a.sigma = sq;
            a.npix = s_cnt[lane];
            if (pass == 0) {
                meshstat_pass1_end(a, bw, bh, o, lcut, hcut, bad);
                s_lcut[lane] = lcut;
                s_hcut[lane] = hcut;
            } else if (!bad) {
                meshstat_pass2_end(a, o);
            }
        }
        __syncthreads();
    }
    if (wid == 0 && lane < nm) ms[my * nbx + mx0 + lane] = o;
}

/* --------------------------------------------------------------- backguess ---- */
#define MODE_THREADS 256

__device__ inline int d2i_sat(double v)
{
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (int)-2147483647 - 1;
    return (int)v;
}

__global__ void __launch_bounds__(MODE_THREADS) k_meshmode(const float *__restrict__ img,
                                                           const float *__restrict__ wmap, int W,
                                                           int H, int bs, int nbx,
                                                           const MeshStat *ms, float *back,
                                                           float *sigma)
{
    /* the histogram h, then in place its prefix counts P0[k] = sum_{j<k} h[j] for k in
       [0, 4096] (h[j] = P0[j + 1] - P0[j]); C1[t], C2[t]: the sums of h[j] j and h[j] j^2
       over the bins j < 16 t, exact in 64-bit integers */
    __shared__ int P0[SXG_NMAXLEVELS + 1];
    __shared__ long long C1[MODE_THREADS + 1], C2[MODE_THREADS + 1];
    __shared__ int sscan[MODE_THREADS / 64];
    __shared__ long long sscan1[MODE_THREADS / 64], sscan2[MODE_THREADS / 64];
    int *const h = P0;
    const int m = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const MeshStat st = ms[m];
    if (st.mean <= -SXG_BIG) {
        if (tid == 0) back[m] = sigma[m] = (float)-SXG_BIG;
        return;
    }
    const int mx = m % nbx, my = m / nbx;
    const int x0 = mx * bs, y0 = my * bs;
    const int bw = (x0 + bs <= W) ? bs : W - x0;
    const int bh = (y0 + bs <= H) ? bs : H - y0;
    const int nlev = st.nlevels;
    for (int i = tid; i < SXG_NMAXLEVELS; i += MODE_THREADS) h[i] = 0;
    __syncthreads();
    const float qscale = st.qscale;
    const float cste = (float)(0.499999 - st.qzero / qscale);
    const int npx = bw * bh;
    auto bin = [&](float v) {
        const float t = v / qscale + cste;
        if (t > -1.0f && t < (float)nlev) atomicAdd(&h[(int)t], 1);
    };
    if (((bw | W) & 3) == 0) {
        /* float4 reads, all of a thread's loads in flight before the binning */
        const int q4 = bw >> 2, nq = q4 * bh;
        for (int q0 = tid; q0 < nq; q0 += 4 * MODE_THREADS) {
            float4 v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int q = q0 + u * MODE_THREADS;
                if (q < nq) {
                    const int y = q / q4, x = 4 * (q - y * q4);
                    const size_t o = (size_t)(y0 + y) * W + x0 + x;
                    v[u] = *reinterpret_cast<const float4 *>(img + o);
                    if (wmap) {
                        const float4 wv = *reinterpret_cast<const float4 *>(wmap + o);
                        const float nb = (float)-SXG_BIG;
                        if (!(wv.x > 0.0f)) v[u].x = nb;
                        if (!(wv.y > 0.0f)) v[u].y = nb;
                        if (!(wv.z > 0.0f)) v[u].z = nb;
                        if (!(wv.w > 0.0f)) v[u].w = nb;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (q0 + u * MODE_THREADS < nq) {
                    bin(v[u].x);
                    bin(v[u].y);
                    bin(v[u].z);
                    bin(v[u].w);
                }
        }
    } else {
#pragma unroll 4
        for (int k = tid; k < npx; k += MODE_THREADS) {
            const int y = k / bw, x = k - y * bw;
            const size_t o = (size_t)(y0 + y) * W + x0 + x;
            bin((wmap && !(wmap[o] > 0.0f)) ? (float)-SXG_BIG : img[o]);
        }
    }
    __syncthreads();
    /* thread t owns bins [16t, 16t + 16): the exclusive prefixes of their counts and of
       their two moment sums over the workgroup */
    constexpr int PER = SXG_NMAXLEVELS / MODE_THREADS;
    const int b0 = tid * PER;
    int hc[PER];
    int csum = 0;
    long long q1 = 0, q2 = 0;
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int j = b0 + u;
        hc[u] = h[j];
        csum += hc[u];
        q1 += (long long)hc[u] * j;
        q2 += (long long)hc[u] * ((long long)j * j);
    }
    int x = csum;
    long long x1 = q1, x2 = q2;
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(x, d);
        const long long o1 = __shfl_up(x1, d), o2 = __shfl_up(x2, d);
        if (lane >= d) {
            x += o;
            x1 += o1;
            x2 += o2;
        }
    }
    if (lane == 63) {
        sscan[wid] = x;
        sscan1[wid] = x1;
        sscan2[wid] = x2;
    }
    __syncthreads();   /* every thread has read its bins of h and the wave totals are in */
    int before = 0;
    long long before1 = 0, before2 = 0;
    for (int w = 0; w < wid; w++) {
        before += sscan[w];
        before1 += sscan1[w];
        before2 += sscan2[w];
    }
    int pc = before + x - csum;
    C1[tid] = before1 + x1 - q1;
    C2[tid] = before2 + x2 - q2;
#pragma unroll
    for (int u = 0; u < PER; u++) {
        P0[b0 + u] = pc;
        pc += hc[u];
    }
    if (tid == MODE_THREADS - 1) {
        P0[SXG_NMAXLEVELS] = pc;
        C1[MODE_THREADS] = before1 + x1;
        C2[MODE_THREADS] = before2 + x2;
    }
    __syncthreads();
    /* The iteration of backguess runs on one wavefront (the others are done): each of
       its window sums is two coarse entries and at most 15 bins on each side, and the
       median bin a 64-way search on P0. */
    if (wid != 0) return;
    auto hget = [&](long i) -> long { return (i >= 0 && i < nlev) ? P0[i + 1] - P0[i] : 0; };
    /* exact sums of h[j] j and h[j] j^2 over j < k for the two ends of a window: lanes
       0..15 the bins of kA's 16-bin block below kA, lanes 16..31 those of kB's */
    auto wsums = [&](int kA, int kB, long long &s1, long long &s2) {
        const int k = lane < 16 ? kA : kB;
        const int j = (k & ~15) + (lane & 15);
        long long a1 = 0, a2 = 0;
        if (lane < 32 && j < k) {
            const long long c = P0[j + 1] - P0[j];
            a1 = c * j;
            a2 = c * ((long long)j * j);
        }
        for (int d = 8; d > 0; d >>= 1) {
            a1 += __shfl_xor(a1, d);
            a2 += __shfl_xor(a2, d);
        }
        const long long pA1 = C1[kA >> 4] + __shfl(a1, 0), pB1 = C1[kB >> 4] + __shfl(a1, 16);
        const long long pA2 = C2[kA >> 4] + __shfl(a2, 0), pB2 = C2[kB >> 4] + __shfl(a2, 16);
        s1 = pB1 - pA1;
        s2 = pB2 - pA2;
    };
    unsigned long lowsum, highsum, sum;
    double ftemp, mea, sig, sig1, med;
    int n, lcut, hcut, nlevelsm1;
    hcut = nlevelsm1 = nlev - 1;
    lcut = 0;
    sig = 10.0 * nlevelsm1;
    sig1 = 1.0;
    mea = med = st.mean;
    for (n = 100; n-- && (sig >= 0.1) && (fabs(sig / sig1 - 1.0) > BG_EPS);) {
        long lo, hi;
        sig1 = sig;
        const long nb = (long)hcut - lcut + 1;
        if (nb <= 0) {
            sum = 0;
            mea = sig = 0.0;
            lowsum = highsum = 0;
            lo = lcut;
            hi = hcut;
        } else {
            /* exact integer window sums (each below 2^53, so the sequential double sums
               of back.c equal them) */
            const int base = P0[lcut], top = P0[hcut + 1];
            const int target = base + top;
            long long s1, s2;
            wsums(lcut, hcut + 1, s1, s2);
            /* a = #{i in [0,nb): L(i) < H(nb-1-i)} with L(i) = P0[lcut+i]-P0[lcut] and
               H(nb-1-i) = P0[hcut+1]-P0[lcut+i+1], i.e. the first bin j >= lcut where the
               non-decreasing P0[j] + P0[j+1] reaches P0[lcut] + P0[hcut+1]: the range
               is cut into 64 chunks, the first chunk whose last bin reaches it holds
               that bin, and the search goes on inside it (two rounds for 4096 bins) */
            int first = 0x7fffffff;
            int rlo = lcut, rhi = hcut;
            while (true) {
                const int step = (rhi - rlo + 1 + 63) >> 6;
                const int jend = min(rlo + (lane + 1) * step - 1, rhi);
                const bool c = rlo + lane * step <= rhi && P0[jend] + P0[jend + 1] >= target;
                const unsigned long long bal = __ballot(c);
                if (!bal) break;
                const int ci = __builtin_ctzll(bal);
                if (step == 1) {
                    first = rlo + ci;
                    break;
                }
                rhi = min(rlo + (ci + 1) * step - 1, rhi);
                rlo = rlo + ci * step;
            }
            sum = (unsigned long)(top - base);
            mea = (double)s1;
            sig = (double)s2;
            const long a_lo = first == 0x7fffffff ? nb : (long)first - lcut;
            lo = lcut + a_lo;
            hi = lo - 1;
            lowsum = (unsigned long)(P0[lo] - base);
            highsum = (unsigned long)(top - P0[lo]);
        }
        if (hi >= 0) {
            long hl = hget(lo), hh = hget(hi);
            med = hi + 0.5 + ((double)highsum - lowsum) / (2.0 * (hl > hh ? hl : hh));
        } else
            med = 0.0;
        if (sum) {
            mea /= (double)sum;
            sig = sig / sum - mea * mea;
        }
        sig = sig > 0.0 ? sqrt(sig) : 0.0;
        lcut = (ftemp = med - 3.0 * sig) > 0.0 ? d2i_sat(ftemp > 0.0 ? ftemp + 0.5 : ftemp - 0.5) : 0;
        hcut = (ftemp = med + 3.0 * sig) < nlevelsm1
                   ? d2i_sat(ftemp > 0.0 ? ftemp + 0.5 : ftemp - 0.5)
                   : nlevelsm1;
    }
    if (threadIdx.x != 0) return;
    float mode;
    if (fabs(sig) > 0.0) {
        if (fabs(st.sigma / (sig * st.qscale) - 1) < 0.0)
            mode = (float)(st.qzero + mea * st.qscale);
        else if (fabs((mea - med) / sig) < 0.3)
            mode = (float)(st.qzero + (2.5 * med - 1.5 * mea) * st.qscale);
        else
            mode = (float)(st.qzero + med * st.qscale);
    } else
        mode = (float)(st.qzero + mea * st.qscale);
    back[m] = mode;
    sigma[m] = (float)(sig * st.qscale);
}

/* -------------------------------------------------------------- filterback ---- */
#define FB_THREADS 1024

/* median of a small array (sorted in place), hmedian's result (misc.c:30-72) */
__device__ inline float small_median(float *a, int n)
{
    for (int i = 1; i < n; i++) {
        float v = a[i];
        int j = i - 1;
        while (j >= 0 && a[j] > v) { a[j + 1] = a[j]; j--; }
        a[j + 1] = v;
    }
    if (n < 2) return a[0];
    return (n & 1) ? a[n / 2] : (float)((a[n / 2 - 1] + a[n / 2]) / 2.0);
}

/* bad meshes: average of the nearest valid meshes (back.c:461-490); one lane per mesh,
   into back2/sigma2 (good meshes copied) */
__global__ void __launch_bounds__(256) k_fb_interp(const float *back, const float *sigma,
                                                   float *back2, float *sigma2, int nx, int ny)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nx * ny) return;
    if (back[i] > -SXG_BIG) {
        back2[i] = back[i];
        sigma2[i] = sigma[i];
        return;
    }
    const int px = i % nx, py = i / nx;
    float d2min = (float)SXG_BIG, val = 0.0f, sval = 0.0f;
    int nmin = 0;
    for (int j = 0, y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++, j++)
            if (back[j] > -SXG_BIG) {
                float d2 = (float)(x - px) * (x - px) + (y - py) * (y - py);
                if (d2 < d2min) {
                    val = back[j];
                    sval = sigma[j];
                    nmin = 1;
                    d2min = d2;
                } else if (d2 == d2min) {
                    val += back[j];
                    sval += sigma[j];
                    nmin++;
                }
            }
    back2[i] = nmin ? val / nmin : 0.0f;
    sigma2[i] = nmin ? sval / nmin : 1.0f;
}

/* median filtering with a window clipped symmetrically at the edges (back.c:492-530);
   one lane per mesh, back2/sigma2 -> back/sigma */
__global__ void __launch_bounds__(256) k_fb_median(const float *back2, const float *sigma2,
                                                   float *back, float *sigma, int nx, int ny,
                                                   int nfx, int nfy, float fthresh)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nx * ny) return;
    const int npx = nfx / 2, npy = nfy / 2;
    const int px = i % nx, py = i / nx;
    int npx2 = nx - px - 1, npy2 = ny - py - 1;
    if (npx2 > npx) npx2 = npx;
    if (npx2 > px) npx2 = px;
    if (npy2 > npy) npy2 = npy;
    if (npy2 > py) npy2 = py;
    if (npx <= 1 && npy <= 1) {
        /* windows up to 3 x 3 in registers: the clipped window's values, +inf for the
           cells outside it, sorted by an odd-even transposition network (nine rounds
           sort nine values); its element k/2 is hmedian's for an odd count.  Equal
           values are then bitwise equal unless a zero or a NaN is present (+0 and -0,
           or the NaN orderings, where insertion order matters): those windows take
           the general path below. */
        float b9[9], s9[9];
        bool plain = true;
        int k = 0;
#pragma unroll
        for (int u = 0; u < 9; u++) {
            const int dy = u / 3 - 1, dx = u % 3 - 1;
            const bool in = dy >= -npy2 && dy <= npy2 && dx >= -npx2 && dx <= npx2;
            const int q = (py + (in ? dy : 0)) * nx + px + (in ? dx : 0);
            const float bv = back2[q], sv = sigma2[q];
            b9[u] = in ? bv : INFINITY;
            s9[u] = in ? sv : INFINITY;
            plain = plain && (!in || (bv != 0.0f && bv == bv && sv != 0.0f && sv == sv));
            k += in ? 1 : 0;
        }
        if (plain && (k & 1)) {
            auto sort9 = [](float (&a)[9]) {
#pragma unroll
                for (int r = 0; r < 9; r++)
#pragma unroll
                    for (int j = r & 1; j + 1 < 9; j += 2) {
                        const float lo = a[j + 1] < a[j] ? a[j + 1] : a[j];
                        const float hi = a[j + 1] < a[j] ? a[j] : a[j + 1];
                        a[j] = lo;
                        a[j + 1] = hi;
                    }
            };
            auto pick = [k](const float (&a)[9]) {
                float m = a[0];
#pragma unroll
                for (int u = 1; u < 9; u++)
                    if (u == k / 2) m = a[u];
                return m;
            };
            sort9(b9);
            const float med = pick(b9);
            const float b2i = back2[i], s2i = sigma2[i];
            if (fabs(med - b2i) >= fthresh) {
                sort9(s9);
                back[i] = med;
                sigma[i] = pick(s9);
            } else {
                back[i] = b2i;
                sigma[i] = s2i;
            }
            return;
        }
    }
    float bm[SXG_MAXBACKF * SXG_MAXBACKF], sm_[SXG_MAXBACKF * SXG_MAXBACKF];
    int k = 0;
    for (int dy = -npy2; dy <= npy2; dy++)
        for (int dx = -npx2; dx <= npx2; dx++) {
            bm[k] = back2[(py + dy) * nx + px + dx];
            sm_[k++] = sigma2[(py + dy) * nx + px + dx];
        }
    const float med = small_median(bm, k);
    if (fabs(med - back2[i]) >= fthresh) {
        back[i] = med;
        sigma[i] = small_median(sm_, k);
    } else {
        back[i] = back2[i];
        sigma[i] = sigma2[i];
    }
}

/* unsigned order-preserving key of a float */
__device__ inline unsigned fb_key(float v) { return (unsigned)f2ord(v) ^ 0x80000000u; }

/*
 * The q-th smallest (0-based) of the values of a[0..n) that pass the filter (all, or the
 * strictly positive ones), by radix selection on the order-preserving keys: four passes,
 * each an LDS histogram of the next 8 key bits among the values that match the prefix
 * found so far, then the bin that holds rank q.  One workgroup of FB_THREADS.
 */
__device__ float wg_select(const float *a, int n, int q, bool positive, int *hist, int *sel)
{
    const int tid = threadIdx.x, lane = tid & 63;
    unsigned prefix = 0, pmask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = tid; i < 256; i += FB_THREADS) hist[i] = 0;
        __syncthreads();
        for (int i = tid; i < n; i += FB_THREADS) {
            const float v = a[i];
            if (positive && !(v > 0.0f)) continue;
            const unsigned u = fb_key(v);
            if ((u & pmask) == prefix) atomicAdd(&hist[(u >> shift) & 255], 1);
        }
        __syncthreads();
        if (tid < 64) {
            /* four bins per lane; the lane whose running count passes q holds the bin */
            const int c0 = hist[4 * lane], c1 = hist[4 * lane + 1], c2 = hist[4 * lane + 2],
                      c3 = hist[4 * lane + 3];
            const int c = c0 + c1 + c2 + c3;
            int incl = c;
            for (int d = 1; d < 64; d <<= 1) {
                const int o = __shfl_up(incl, d);
                if (lane >= d) incl += o;
            }
            const unsigned long long over = __ballot(incl > q);
            const int L = over ? __builtin_ctzll(over) : 63;
            if (lane == L) {
                int before = incl - c, j = 0;
                const int cs[4] = {c0, c1, c2, c3};
                while (j < 3 && before + cs[j] <= q) before += cs[j++];
                sel[0] = 4 * lane + j;
                sel[1] = q - before;
            }
        }
        __syncthreads();
        prefix |= (unsigned)sel[0] << shift;
        pmask |= 255u << shift;
        q = sel[1];
        __syncthreads();
    }
    return ord2f((int)(prefix ^ 0x80000000u));
}

/* median of the cnt values of a[0..n) that pass the filter, by the even-count rule of
   hmedian: the mean of the two middle values */
__device__ float wg_median(const float *a, int n, int cnt, bool positive, int *hist, int *sel)
{
    if (cnt < 2) return wg_select(a, n, 0, positive, hist, sel);
    const float hi = wg_select(a, n, cnt / 2, positive, hist, sel);
    if (cnt & 1) return hi;
    const float lo = wg_select(a, n, cnt / 2 - 1, positive, hist, sel);
    return (float)((lo + hi) / 2.0);
}

/* global medians of the filtered maps (back.c:532-556): workgroup 0 the background,
   workgroup 1 the sigma (with the median of the strictly positive sigmas when the median
   is not positive, 1 when there is none) */
__global__ void __launch_bounds__(FB_THREADS) k_fb_global(const float *back, const float *sigma,
                                                          int np, float *out)
{
    __shared__ int hist[256];
    __shared__ int sel[2];
    __shared__ int s_npos;
    if (blockIdx.x == 0) {
        const float m = wg_median(back, np, np, false, hist, sel);
        if (threadIdx.x == 0) out[0] = m;
        return;
    }
    float bsig = wg_median(sigma, np, np, false, hist, sel);
    if (bsig <= 0.0f) {
        if (threadIdx.x == 0) s_npos = 0;
        __syncthreads();
        int c = 0;
        for (int i = threadIdx.x; i < np; i += FB_THREADS) c += sigma[i] > 0.0f;
        atomicAdd(&s_npos, c);
        __syncthreads();
        const int npos = s_npos;
        bsig = npos > 0 && npos < np ? wg_median(sigma, np, npos, true, hist, sel) : 1.0f;
    }
    if (threadIdx.x == 0) out[1] = bsig;
}

/* y second derivatives of back and sigma, one lane per mesh column (back.c:739-780);
   the forward recurrence runs in registers and keeps its per-row values (the
   tridiagonal factors and uu) for the backward pass in LDS when ny rows of 64 lanes fit
   (use_lds), else in uu_all (2 * ny floats per column) */
__global__ void __launch_bounds__(64) k_fb_yderiv(const float *back, const float *sigma,
                                                  float *dback, float *dsigma, int nx, int ny,
                                                  int use_lds, float *uu_all)
{
    extern __shared__ float yd[];
    const int c = blockIdx.x * 64 + threadIdx.x;
    const bool live = c < 2 * nx;
    const int x = c % nx;
    const float *map = c < nx ? back : sigma;
    float *dmap = c < nx ? dback : dsigma;
    if (!live) return;
    if (ny < 2) {
        dmap[x] = 0.0f;
        return;
    }
    /* element y of this column's two scratch rows */
    float *ta, *tb;
    size_t st;
    if (use_lds) {
        ta = yd + threadIdx.x;
        tb = yd + (size_t)ny * 64 + threadIdx.x;
        st = 64;
    } else {
        ta = uu_all + (size_t)c * 2 * ny;
        tb = ta + ny;
        st = 1;
    }
    float dprev = 0.0f, uprev = 0.0f;
    float m0 = map[x], m1 = map[nx + x];
    for (int y = 1; y < ny - 1; y++) {
        const float m2 = map[(y + 1) * nx + x];
        float temp = -1 / (dprev + 4);
        ta[y * st] = temp;
        dprev = temp;
        temp *= uprev - 6 * (m2 + m0 - 2 * m1);
        tb[y * st] = temp;
        uprev = temp;
        m0 = m1;
        m1 = m2;
    }
    dmap[x] = 0.0f;
    dmap[(ny - 1) * nx + x] = 0.0f;
    float dnext = 0.0f;
    for (int y = ny - 2; y >= 1; y--) {
        const float d = (float)((ta[y * st] * dnext + tb[y * st]) / 6.0);
        dmap[y * nx + x] = d;
        dnext = d;
    }
}

/* ------------------------------------------------------------ subtraction ---- */
/* per row: node[x] (y interpolation) and dnode[x] (x second derivatives) */
__global__ void __launch_bounds__(256) k_rownodes(const float *back, const float *dback, int nbx,
                                                  int nby, int bs, int H, float *nodes,
                                                  float *dnodes, float *utmp)
{
    const int y = blockIdx.x * 256 + threadIdx.x;
    if (y >= H) return;
    float *node = nodes + (size_t)y * nbx, *dnode = dnodes + (size_t)y * nbx;
    float *u = utmp + (size_t)y * nbx;
    if (nby > 1) {
        int yl;
        float dy = (float)((float)y / bs - 0.5);
        dy -= (yl = (int)dy);
        if (yl < 0) {
            yl = 0;
            dy = (float)(dy - 1.0);
        } else if (yl >= nby - 1) {
            yl = nby < 2 ? 0 : nby - 2;
            dy = (float)(dy + 1.0);
        }
        const float cdy = 1 - dy;
        const float dy3 = (dy * dy * dy - dy);
        const float cdy3 = (cdy * cdy * cdy - cdy);
        const float *blo = back + nbx * yl, *bhi = blo + nbx;
        const float *dblo = dback + nbx * yl, *dbhi = dblo + nbx;
        for (int x = 0; x < nbx; x++)
            node[x] = cdy * blo[x] + dy * bhi[x] + cdy3 * dblo[x] + dy3 * dbhi[x];
        if (nbx > 1) {
            float temp;
            dnode[0] = u[0] = 0.0f;
            for (int x = 1; x < nbx - 1; x++) {
                temp = -1 / (dnode[x - 1] + 4);
                dnode[x] = temp;
                temp *= u[x - 1] - 6 * (node[x + 1] + node[x - 1] - 2 * node[x]);
                u[x] = temp;
            }
            dnode[nbx - 1] = 0.0f;
            for (int x = nbx - 2; x >= 1; x--)
                dnode[x] = (float)((dnode[x] * dnode[x + 1] + u[x]) / 6.0);
        }
    } else {
        for (int x = 0; x < nbx; x++) {
            node[x] = back[x];
            dnode[x] = dback[x];
        }
    }
}

/* the same for RN_ROWS consecutive rows per workgroup (one lane each) with the row
   arrays in LDS, laid out [x][row] (no bank conflicts), and coalesced stores of the
   finished rows; used when 3 * nbx * RN_ROWS floats fit in LDS */
#define RN_ROWS 64
__global__ void __launch_bounds__(RN_ROWS) k_rownodes_lds(const float *back, const float *dback,
                                                          int nbx, int nby, int bs, int H,
                                                          float *nodes, float *dnodes)
{
    extern __shared__ float rn[];
    const int lane = threadIdx.x, y0 = blockIdx.x * RN_ROWS, y = y0 + lane;
    float *node = rn, *dnode = rn + nbx * RN_ROWS, *u = dnode + nbx * RN_ROWS;
#define RN(a, x) a[(x) * RN_ROWS + lane]
    if (y < H) {
        if (nby > 1) {
            int yl;
            float dy = (float)((float)y / bs - 0.5);
            dy -= (yl = (int)dy);
            if (yl < 0) {
                yl = 0;
                dy = (float)(dy - 1.0);
            } else if (yl >= nby - 1) {
                yl = nby < 2 ? 0 : nby - 2;
                dy = (float)(dy + 1.0);
            }
            const float cdy = 1 - dy;
            const float dy3 = (dy * dy * dy - dy);
            const float cdy3 = (cdy * cdy * cdy - cdy);
            const float *blo = back + nbx * yl, *bhi = blo + nbx;
            const float *dblo = dback + nbx * yl, *dbhi = dblo + nbx;
            for (int x = 0; x < nbx; x++)
                RN(node, x) = cdy * blo[x] + dy * bhi[x] + cdy3 * dblo[x] + dy3 * dbhi[x];
            if (nbx > 1) {
                float temp, dprev = 0.0f, uprev = 0.0f;
                RN(dnode, 0) = RN(u, 0) = 0.0f;
                for (int x = 1; x < nbx - 1; x++) {
                    temp = -1 / (dprev + 4);
                    RN(dnode, x) = dprev = temp;
                    temp *= uprev - 6 * (RN(node, x + 1) + RN(node, x - 1) - 2 * RN(node, x));
                    RN(u, x) = uprev = temp;
                }
                RN(dnode, nbx - 1) = 0.0f;
                float dnext = 0.0f;
                for (int x = nbx - 2; x >= 1; x--)
                    RN(dnode, x) = dnext = (float)((RN(dnode, x) * dnext + RN(u, x)) / 6.0);
            }
        } else {
            for (int x = 0; x < nbx; x++) {
                RN(node, x) = back[x];
                RN(dnode, x) = dback[x];
            }
        }
    }
#undef RN
    __syncthreads();
    /* rows y0.. are contiguous in the node tables */
    const int nr = min(RN_ROWS, H - y0), tot = nr * nbx;
    for (int e = lane; e < tot; e += RN_ROWS) {
        const int r = e / nbx, x = e - r * nbx;
        nodes[(size_t)y0 * nbx + e] = node[x * RN_ROWS + r];
        dnodes[(size_t)y0 * nbx + e] = dnode[x * RN_ROWS + r];
    }
}

/* colseg[x]: index of the left node, coldx[x]: dx of the cumulative recurrence.
   SUB: dst = src - the interpolated value (subbackline, back.c:790-930); else dst = the
   value (backrmsline on the sigma maps, back.c:938-1051).  One workgroup per 1024
   pixels of a row (blockIdx.y), four pixels per lane at a stride of 256. */
template <bool SUB>
__global__ void __launch_bounds__(256) k_subback(const float *src, float *dst, int W, int nbx,
                                                 const float *nodes, const float *dnodes,
                                                 const int *colseg, const float *coldx)
{
    const int y = blockIdx.y;
    const float *node = nodes + (size_t)y * nbx, *dnode = dnodes + (size_t)y * nbx;
    const size_t row = (size_t)y * W;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int x = blockIdx.x * 1024 + r * 256 + threadIdx.x;
        if (x >= W) break;
        const float bval = sxg_spline_at(node, dnode, colseg, coldx, nbx, x);
        if (SUB)
            dst[row + x] = src[row + x] - bval;
        else
            dst[row + x] = bval;
    }
}

/* dst = src - v, with v = *vdev when vdev is given (the global median on the device) */
__global__ void __launch_bounds__(256) k_subconst(const float *src, float *dst, long n, float v,
                                                  const float *vdev)
{
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    const float c = vdev ? *vdev : v;
    if (k < n) dst[k] = src[k] - c;
}

__global__ void k_setval(float *p, float v) { *p = v; }

/* The column tables of the x loop of subbackline (back.c:880-908): for each column, the
   index of its left node and the dx of the cumulative float recurrence dx += xstep,
   which starts at (xstep - 1)/2 and restarts at dx0 at the half of every mesh t in
   [1, nbx - 2] (the loop's pixel with i == bs/2, column t*bs + bs/2).  One lane per
   restart segment performs that segment's additions in the loop's order, so every value
   is the loop's own IEEE result. */
__global__ void __launch_bounds__(64) k_coltables(int W, int nbx, int bs, int *colseg,
                                                  float *coldx)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (nbx <= 1) {
        for (int j = t; j < W; j += gridDim.x * 64) {
            colseg[j] = 0;
            coldx[j] = 0.0f;
        }
        return;
    }
    const int cp = bs / 2;
    const int nseg = cp >= 1 && nbx > 2 ? nbx - 1 : 1;
    if (t >= nseg) return;
    const float xstep = (float)(1.0 / bs);
    float dx = t == 0 ? (xstep - 1) / 2 : ((bs + 1) % 2) * xstep / 2;
    const int j0 = t == 0 ? 0 : t * bs + cp;
    const int j1 = t + 1 < nseg ? min((t + 1) * bs + cp, W) : W;
    for (int j = j0; j < j1; j++) {
        colseg[j] = t;
        coldx[j] = dx;
        dx = dx + xstep;
    }
}

/* Byte layout of the device buffers of one background estimate of a W x H frame with
   nbx x nby meshes, each part 256-byte aligned: mesh statistics, work (4*n2 floats and
   the two global medians), the four mesh maps, the per-row spline tables (nodes, dnodes,
   utmp: H*nbx floats each) and the column tables (W ints, W floats). */
struct BackLayout {
    size_t ms, work, maps, rows, cols, total;
    int n2;
};
static inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
static BackLayout back_layout(int W, int H, int nbx, int nby)
{
    BackLayout L;
    const size_t nmesh = (size_t)nbx * nby;
    L.n2 = 1;
    while ((size_t)L.n2 < nmesh) L.n2 <<= 1;
    L.ms = 0;
    L.work = L.ms + al256(sizeof(MeshStat) * nmesh);
    L.maps = L.work + al256(sizeof(float) * (4 * (size_t)L.n2 + 64));
    L.rows = L.maps + al256(sizeof(float) * 4 * nmesh);
    L.cols = L.rows + al256(sizeof(float) * 3 * (size_t)H * nbx);
    L.total = L.cols + al256(sizeof(int) * (size_t)W) + al256(sizeof(float) * (size_t)W);
    return L;
}
static BackBufs back_bufs(char *base, const BackLayout &L, int W, int H, int nbx, int nby)
{
    BackBufs b;
    const size_t nmesh = (size_t)nbx * nby;
    b.ms = reinterpret_cast<MeshStat *>(base + L.ms);
    b.work = reinterpret_cast<float *>(base + L.work);
    b.outv = b.work + 4 * (size_t)L.n2;
    b.back = reinterpret_cast<float *>(base + L.maps);
    b.sigma = b.back + nmesh;
    b.dback = b.sigma + nmesh;
    b.dsigma = b.dback + nmesh;
    b.nodes = reinterpret_cast<float *>(base + L.rows);
    b.dnodes = b.nodes + (size_t)H * nbx;
    b.utmp = b.dnodes + (size_t)H * nbx;
    b.colseg = reinterpret_cast<int *>(base + L.cols);
    b.coldx = reinterpret_cast<float *>(base + L.cols + al256(sizeof(int) * (size_t)W));
    return b;
}

/* the per-row spline nodes of a mesh map (and its y second derivatives) into the row
   tables of b */
static bool spline_rows(const BackGeom &g, const BackBufs &b, const float *map,
                        const float *dmap, hipStream_t st)
{
    const size_t rnlds = 3 * (size_t)g.nbx * RN_ROWS * sizeof(float);
    if (rnlds <= 64 * 1024) {
        hipLaunchKernelGGL(k_rownodes_lds, dim3(sxg_div_up(g.H, RN_ROWS)), dim3(RN_ROWS), rnlds,
                           st, map, dmap, g.nbx, g.nby, g.bs, g.H, b.nodes, b.dnodes);
    } else {
        hipLaunchKernelGGL(k_rownodes, dim3(sxg_div_up(g.H, 256)), dim3(256), 0, st, map, dmap,
                           g.nbx, g.nby, g.bs, g.H, b.nodes, b.dnodes, b.utmp);
    }
    SXG_HIP_RET(hipGetLastError(), false);
    return true;
}

/* dst = src - the spline of the row tables (sub), or dst = the spline */
static bool spline_apply(const BackGeom &g, const BackBufs &b, const float *src, float *dst,
                         bool sub, hipStream_t st)
{
    const dim3 grid(sxg_div_up(g.W, 1024), g.H);
    if (sub)
        hipLaunchKernelGGL(k_subback<true>, grid, dim3(256), 0, st, src, dst, g.W, g.nbx,
                           b.nodes, b.dnodes, b.colseg, b.coldx);
    else
        hipLaunchKernelGGL(k_subback<false>, grid, dim3(256), 0, st, src, dst, g.W, g.nbx,
                           b.nodes, b.dnodes, b.colseg, b.coldx);
    SXG_HIP_RET(hipGetLastError(), false);
    return true;
}

/* The background of a W x H frame read from src and subtracted into dst (src == dst is
   allowed): mesh statistics and modes (pixels whose wmap value is not positive are left
   out when wmap is given), bad-mesh interpolation and median filtering, the global
   medians into b.outv[0..1] (device), the y derivatives, then the subtraction of the
   bicubic spline (back_type 0), of the global median b.outv[0] (1) or of back_value (2,
   which also becomes b.outv[0]).  Asynchronous on st; allocates nothing. */
bool sxg_back_run(const BackGeom &g, const BackBufs &b, const float *src, float *dst,
                  const float *wmap, float backfthresh, int back_type, float back_value,
                  bool timed, hipStream_t st)
{
    const int W = g.W, H = g.H, bs = g.bs, nbx = g.nbx, nby = g.nby, nmesh = nbx * nby;
    int n2 = 1;
    while (n2 < nmesh) n2 <<= 1;
    if (timed) sxg_kt_begin(SXG_KT_MESHSTAT);
    {
        /* meshes per workgroup and rows per batch: two buffers of R rows of mpb padded
           segments of t and t2 (doubles), each within 64 KB */
        const size_t segbytes = (size_t)ms_seg(bs) * 2 * sizeof(double);
        const int mpb = (int)std::min<size_t>(MS_MPB, (64 * 1024) / segbytes);
        if (mpb >= 1) {
            /* the LDS limit is an attribute of the function: raised once, to the most
               any launch asks for */
            static const bool attr = [] {
                return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_meshstat),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           MS_LDS_BYTES) == hipSuccess;
            }();
            if (!attr) {
                sxg_set_error("back_estimate: cannot raise the LDS limit of k_meshstat");
                return false;
            }
            const size_t rowbytes = mpb * segbytes;
            const long rows = sxg_option(SXG_OPT_MESH_ROWS);
            int R = (int)std::max<size_t>(1, std::min<size_t>(8, (64 * 1024) / rowbytes));
            if (rows > 0 && rows < R) R = (int)rows;
            const size_t lds = 2 * R * rowbytes;
            hipLaunchKernelGGL(k_meshstat, dim3(sxg_div_up(nbx, mpb), nby), dim3(MS_THREADS), lds,
                               st, src, wmap, W, H, bs, nbx, mpb, R, b.ms);
        } else {
            hipLaunchKernelGGL(k_meshstat_direct, dim3(sxg_div_up(nmesh, 64)), dim3(64), 0, st,
                               src, wmap, W, H, bs, nbx, nmesh, b.ms);
        }
    }
    if (timed) sxg_kt_end(SXG_KT_MESHSTAT);
    if (timed) sxg_kt_begin(SXG_KT_MESHMODE);
    hipLaunchKernelGGL(k_meshmode, dim3(nmesh), dim3(MODE_THREADS), 0, st, src, wmap, W, H, bs,
                       nbx, b.ms, b.back, b.sigma);
    if (timed) sxg_kt_end(SXG_KT_MESHMODE);
    if (timed) sxg_kt_begin(SXG_KT_FILTERBACK);
    {
        float *back2 = b.work, *sigma2 = b.work + n2;
        const int gm = sxg_div_up(nmesh, 256);
        hipLaunchKernelGGL(k_fb_interp, dim3(gm), dim3(256), 0, st, b.back, b.sigma, back2, sigma2,
                           nbx, nby);
        hipLaunchKernelGGL(k_fb_median, dim3(gm), dim3(256), 0, st, back2, sigma2, b.back, b.sigma,
                           nbx, nby, g.nfx, g.nfy, backfthresh);
        hipLaunchKernelGGL(k_fb_global, dim3(2), dim3(FB_THREADS), 0, st, b.back, b.sigma, nmesh,
                           b.outv);
        /* scratch of 2*nby floats per column: LDS when it fits, else back2/sigma2 and the
           sort buffer (4*n2 >= 2 * 2*nbx*nby floats) */
        const size_t ylds = 2 * (size_t)nby * 64 * sizeof(float);
        const int ylds_ok = ylds <= 128 * 1024;
        static const bool yattr = [] {
            return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fb_yderiv),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       128 * 1024) == hipSuccess;
        }();
        if (!yattr) {
            sxg_set_error("back_estimate: cannot raise the LDS limit of k_fb_yderiv");
            return false;
        }
        hipLaunchKernelGGL(k_fb_yderiv, dim3(sxg_div_up(2 * nbx, 64)), dim3(64),
                           ylds_ok ? ylds : 0, st, b.back, b.sigma, b.dback, b.dsigma, nbx, nby,
                           ylds_ok, back2);
    }
    if (timed) sxg_kt_end(SXG_KT_FILTERBACK);
    if (back_type == 1 || back_type == 2) {
        /* back_type 1 (BACK_TYPE MANUAL): the reference's device code subtracts the global
           median it has just computed (cudaback.cu:1336-1345), not BACK_VALUE; the caller
           (back.c:152-153) only overwrites its own backmean with BACK_VALUE afterwards.
           back_type 2 (an addition): stock SExtractor's MANUAL, the constant BACK_VALUE
           (back.c:148-149, :807-813), returned as backmean. */
        if (back_type == 2) hipLaunchKernelGGL(k_setval, dim3(1), dim3(1), 0, st, b.outv, back_value);
        hipLaunchKernelGGL(k_subconst, dim3(sxg_div_up((long)W * H, 256)), dim3(256), 0, st, src,
                           dst, (long)W * H, back_value, back_type == 1 ? b.outv : nullptr);
    } else {
        if (timed) sxg_kt_begin(SXG_KT_ROWNODES);
        hipLaunchKernelGGL(k_coltables, dim3(sxg_div_up(nbx, 64)), dim3(64), 0, st, W, nbx, bs,
                           b.colseg, b.coldx);
        if (!spline_rows(g, b, b.back, b.dback, st)) return false;
        if (timed) sxg_kt_end(SXG_KT_ROWNODES);
        /* dst nullptr: the caller subtracts the spline itself (from the row and column
           tables just built) */
        if (dst) {
            if (timed) sxg_kt_begin(SXG_KT_SUBBACK);
            if (!spline_apply(g, b, src, dst, true, st)) return false;
            if (timed) sxg_kt_end(SXG_KT_SUBBACK);
        }
    }
    SXG_HIP_RET(hipGetLastError(), false);
    return true;
}

/* the context's buffers for a background of geometry g (grow-only) */
static bool ctx_back_bufs(const BackGeom &g, BackBufs &b)
{
    SxgState &S = sxg_state();
    const BackLayout L = back_layout(g.W, g.H, g.nbx, g.nby);
    if (!S.backbuf.ensure(L.total)) return false;
    b = back_bufs(S.backbuf.as<char>(), L, g.W, g.H, g.nbx, g.nby);
    return true;
}

extern "C" int sxg_get_back_rms(float *out)
{
    SxgState &S = sxg_state();
    sxg_clear_error();
    /* the mesh maps of the last back_estimate stay until the next one (also after
       clear_device) */
    const int nm = S.nbx * S.nby;
    if (nm <= 0 || S.N == 0 || !S.backbuf.ptr || S.back_W != S.W || S.back_H != S.H) {
        sxg_set_error("sxg_get_back_rms: call after back_estimate");
        return -1;
    }
    BackBufs b;
    if (!ctx_back_bufs(S.bgeom, b)) return -1;
    float *rms = sxg_take<float>("back.rms", S.N);
    if (!rms || !spline_rows(S.bgeom, b, b.sigma, b.dsigma, S.stream) ||
        !spline_apply(S.bgeom, b, nullptr, rms, false, S.stream))
        return -1;
    SXG_HIP_RET(hipMemcpyAsync(out, rms, sizeof(float) * S.N, hipMemcpyDeviceToHost, S.stream),
                -1);
    SXG_HIP_RET(hipStreamSynchronize(S.stream), -1);
    return 0;
}

/* the geometry of a background of a W x H frame; false (error recorded) when invalid */
bool sxg_back_geom(int W, int H, int back_size, int nbackfx, int nbackfy, BackGeom &g)
{
    if (back_size < 1) {
        sxg_set_error("back_estimate: BACK_SIZE must be >= 1");
        return false;
    }
    g.W = W;
    g.H = H;
    g.bs = back_size;
    g.nbx = (W - 1) / back_size + 1;
    g.nby = (H - 1) / back_size + 1;
    /* field.c:139-140: a single mesh along an axis disables filtering along it */
    g.nfx = g.nbx > 1 ? nbackfx : 1;
    g.nfy = g.nby > 1 ? nbackfy : 1;
    if (g.nfx < 1 || g.nfy < 1 || g.nfx > SXG_MAXBACKF || g.nfy > SXG_MAXBACKF) {
        sxg_set_error("back_estimate: BACK_FILTERSIZE out of range (1..11)");
        return false;
    }
    return true;
}

/* back_estimate on the context's frame, reading the input frame from src (the context's
   own frame buffer for the reference entry point; a resident frame for the whole-frame
   runs, which saves the copy) and writing the subtracted frame into S.img */
void sxg_back_estimate_from(const float *src, int back_size, float backfthresh, float *backmean,
                            float *backsig, int nbackfx, int nbackfy, int back_type)
{
    SxgState &S = sxg_state();
    if (!S.have_frame) {
        sxg_set_error("back_estimate: no frame (call init_device first)");
        return;
    }
    BackGeom g;
    if (!sxg_back_geom(S.W, S.H, back_size, nbackfx, nbackfy, g)) return;
    S.bs = g.bs;
    S.nbx = g.nbx;
    S.nby = g.nby;
    S.back_W = S.back_H = 0;    /* set once the maps exist for this size */
    BackBufs b;
    if (!ctx_back_bufs(g, b)) return;
    S.bgeom = g;
    S.back_W = S.W;
    S.back_H = S.H;
    float *img = S.img.as<float>();
    WeightArgs wa;
    if (!sxg_weight_args(wa)) return;
    sxg_stage_begin(ST_BACK);
    /* pixels with a non-positive mask value leave the mesh statistics; a whole-frame run
       may leave the spline's subtraction to its fused filtering kernel */
    const bool defer = S.defer_sub && back_type == 0;
    S.defer_sub = false;
    S.sub_src = defer ? src : nullptr;
    if (!sxg_back_run(g, b, src, defer ? nullptr : img, wa.mask, backfthresh, back_type,
                      S.back_value, true, S.stream))
        return;
    if (!S.plan) {
        float hv[2];
        SXG_HIP(hipMemcpyAsync(hv, b.outv, sizeof(hv), hipMemcpyDeviceToHost, S.stream));
        SXG_HIP(hipStreamSynchronize(S.stream));
        S.backmean = hv[0];
        S.backsig = hv[1];
    }
    /* a planned run leaves the medians on the device (b.outv): the thresholds are
       derived from them there */
    if (S.weighted && S.wtype == SXG_WEIGHT_BACKGROUND) {
        /* WEIGHT_TYPE BACKGROUND: the variance map of this frame is the square of the
           interpolated background rms */
        if (!S.weight.ensure(sizeof(float) * S.N) ||
            !spline_rows(g, b, b.sigma, b.dsigma, S.stream) ||
            !spline_apply(g, b, nullptr, S.weight.as<float>(), false, S.stream))
            return;
        S.ww = S.W;
        S.wh = S.H;
        S.wrms_frame = true;
    }
    sxg_stage_end(ST_BACK);
    SXG_HIP(hipGetLastError());
    if (backmean) *backmean = S.backmean;
    if (backsig) *backsig = S.backsig;
}

extern "C" void back_estimate(int back_size, float backfthresh, float *backmean, float *backsig,
                              int nbackfx, int nbackfy, int back_type)
{
    SxgState &S = sxg_state();
    sxg_back_estimate_from(S.img.as<float>(), back_size, backfthresh, backmean, backsig, nbackfx,
                           nbackfy, back_type);
}

extern "C" int sxg_get_back_maps(float *back, float *sigma, float *dback, float *dsigma)
{
    SxgState &S = sxg_state();
    const int nm = S.nbx * S.nby;
    if (nm <= 0 || !S.backbuf.ptr) return -1;
    BackBufs b;
    if (!ctx_back_bufs(S.bgeom, b)) return -1;
    const float *src[4] = {b.back, b.sigma, b.dback, b.dsigma};
    float *dst[4] = {back, sigma, dback, dsigma};
    for (int i = 0; i < 4; i++)
        if (dst[i])
            SXG_HIP_RET(hipMemcpyAsync(dst[i], src[i], sizeof(float) * nm, hipMemcpyDeviceToHost,
                                       S.stream), -1);
    SXG_HIP_RET(hipStreamSynchronize(S.stream), -1);
    return 0;
}

const float *sxg_back_map()
{
    SxgState &S = sxg_state();
    if (!S.backbuf.ptr) return nullptr;
    const BackLayout L = back_layout(S.bgeom.W, S.bgeom.H, S.bgeom.nbx, S.bgeom.nby);
    return reinterpret_cast<const float *>(S.backbuf.as<char>() + L.maps);
}

/* the BackBufs block of geometry g: its size in bytes, and its parts at base */
size_t sxg_back_bytes(const BackGeom &g) { return back_layout(g.W, g.H, g.nbx, g.nby).total; }
BackBufs sxg_back_bufs_at(char *base, const BackGeom &g)
{
    return back_bufs(base, back_layout(g.W, g.H, g.nbx, g.nby), g.W, g.H, g.nbx, g.nby);
}
