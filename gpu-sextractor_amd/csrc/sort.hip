/*
 * sort.hip -- stable LSD radix sort of (key, value) int32 pairs, written for 64-lane
 * wavefronts.  Used to group the detected pixels by object while keeping each object's
 * pixels in raster order (the reference uses CUDPP's radix sort for this,
 * src/cuda/cudadetection.cu:720-732).
 *
 * One pass per 8-bit digit:
 *   k_rs_hist     per tile of 4096 elements, digit histogram in LDS;
 *   scan          exclusive scan of the digit-major (digit, tile) counts;
 *   k_rs_scatter  per tile (its pairs read into registers first), 16 rounds of 256
 *                 elements in index order: lanes with equal digits find each other with
 *                 8 ballots (one per digit bit), their rank is the popcount of the lower
 *                 lanes, and wave totals are combined in LDS; a running per-digit
 *                 counter keeps the scatter stable.
 */
#include "sxg_internal.h"

#define RS_THREADS 256
#define RS_ROUNDS 16
#define RS_TILE (RS_THREADS * RS_ROUNDS)
#define RS_BINS 256

__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const int *__restrict__ keys, int n,
                                                        int shift, int ntiles, int *hist,
                                                        const int *nd)
{
    __shared__ int h[RS_BINS];
    n = sxg_dn(nd, n);
    h[threadIdx.x] = 0;
    __syncthreads();
    const long base = (long)blockIdx.x * RS_TILE;
    /* every key of the thread read before the first is counted */
    int kr[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const long e = base + (long)r * RS_THREADS + threadIdx.x;
        kr[r] = e < n ? keys[e] : -1;
    }
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++)
        if (base + (long)r * RS_THREADS + threadIdx.x < n)
            atomicAdd(&h[((unsigned)kr[r] >> shift) & 255], 1);
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const int *__restrict__ kin,
                                                           const int *__restrict__ vin, int n,
                                                           int shift, int ntiles,
                                                           const int *__restrict__ offs,
                                                           int *kout, int *vout, const int *nd)
{
    __shared__ int cnt[RS_BINS];
    n = sxg_dn(nd, n);
    __shared__ int wc[RS_THREADS / 64][RS_BINS];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    cnt[t] = offs[t * ntiles + blockIdx.x];
    for (int w = 0; w < RS_THREADS / 64; w++) wc[w][t] = 0;
    __syncthreads();
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const long base = (long)blockIdx.x * RS_TILE;
    /* the thread's pairs of all rounds read before the first round: one memory round
       trip per tile instead of one per round */
    int kr[RS_ROUNDS], vr[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const long e = base + (long)r * RS_THREADS + t;
        kr[r] = e < n ? kin[e] : 0;
        vr[r] = e < n ? vin[e] : 0;
    }
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const long e = base + (long)r * RS_THREADS + t;
        const bool valid = e < n;
        const int key = kr[r], val = vr[r];
        const int d = ((unsigned)key >> shift) & 255;
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            unsigned long long bb = __ballot((d >> b) & 1);
            m &= ((d >> b) & 1) ? bb : ~bb;
        }
        const int rank = __popcll(m & below);
        if (valid && rank == 0) wc[wave][d] = __popcll(m);
        __syncthreads();
        if (valid) {
            int pos = cnt[d] + rank;
            for (int w = 0; w < wave; w++) pos += wc[w][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        int add = 0;
        for (int w = 0; w < RS_THREADS / 64; w++) {
            add += wc[w][t];
            wc[w][t] = 0;
        }
        cnt[t] += add;
        __syncthreads();
    }
}

/* Sorts n pairs by the low `nbits` bits of the keys (keys must be >= 0).  The result is
   in (k0, v0) when the number of passes is even, otherwise in (k1, v1); the function
   returns which (0 or 1).  hist must hold 256 * ceil(n/4096) + 64 ints.  With nd the
   pairs are the first *nd (<= n): the tiles past them count nothing, which leaves every
   offset as it is. */
int sxg_radix_sort(int *k0, int *v0, int *k1, int *v1, int n, int nbits, int *hist,
                   hipStream_t s, const int *nd)
{
    if (n <= 0) return 0;
    const int ntiles = sxg_div_up(n, RS_TILE);
    int cur = 0;
    for (int shift = 0; shift < nbits; shift += 8) {
        int *ki = cur ? k1 : k0, *vi = cur ? v1 : v0;
        int *ko = cur ? k0 : k1, *vo = cur ? v0 : v1;
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_THREADS), 0, s, ki, n, shift, ntiles,
                           hist, nd);
        sxg_scan_excl(hist, hist, RS_BINS * ntiles, nullptr, s);
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_THREADS), 0, s, ki, vi, n, shift,
                           ntiles, hist, ko, vo, nd);
        cur ^= 1;
    }
    return cur;
}
