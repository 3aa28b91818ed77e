/*
 * sxg_wave.h -- 64-lane wavefront helpers shared by the per-object kernels.
 *
 * The reference accumulates every per-object sum in one sequential loop over the
 * object's pixels in raster order.  A wavefront reproduces those sums bit for bit by
 * letting each lane compute the terms of one pixel (in parallel) and then adding the
 * terms of the member lanes in lane order (the raster order of a chunk of 64 list
 * entries) with readlane: the additions are the only sequential part.
 */
#pragma once
#include <hip/hip_runtime.h>

__device__ inline unsigned long long sxw_lanes_below(int lane)
{
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

/* value of v in lane j (j uniform across the wavefront) */
__device__ inline float sxw_rl(float v, int j)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ inline double sxw_rl(double v, int j)
{
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)u, j);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), j);
    return __hiloint2double(hi, lo);
}

template <class Op> __device__ inline float sxw_dpp_allf(float v, Op op);
struct SxwMaxF { __device__ float operator()(float a, float b) const { return fmaxf(a, b); } };
struct SxwMinF { __device__ float operator()(float a, float b) const { return fminf(a, b); } };
/* float extrema over the full wavefront (every lane active), without LDS */
__device__ inline float sxw_max(float v) { return sxw_dpp_allf(v, SxwMaxF{}); }
__device__ inline float sxw_minf(float v) { return sxw_dpp_allf(v, SxwMinF{}); }
template <class Op> __device__ inline int sxw_dpp_all(int v, Op op);
struct SxwMax { __device__ int operator()(int a, int b) const { return max(a, b); } };
struct SxwMin { __device__ int operator()(int a, int b) const { return min(a, b); } };
struct SxwAdd { __device__ int operator()(int a, int b) const { return a + b; } };
/* int reductions over the full wavefront (every lane active), without LDS */
__device__ inline int sxw_max(int v) { return sxw_dpp_all(v, SxwMax{}); }
__device__ inline int sxw_min(int v) { return sxw_dpp_all(v, SxwMin{}); }
__device__ inline int sxw_sum(int v) { return sxw_dpp_all(v, SxwAdd{}); }
__device__ inline double sxw_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

/* op over the 64 lanes of an int, the result in every lane, without LDS: DPP moves
   inside each row of 16 lanes (quad permutations for the partners at distance 1 and 2,
   then the half-row and row mirrors, each of which pairs a lane with one of the group it
   has not seen yet) and one readlane per row.  Every lane of the wavefront must be
   active (the readlanes take rows 0..3 from lanes 0, 16, 32 and 48). */
template <class Op> __device__ inline int sxw_dpp_all(int v, Op op)
{
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));    /* 1,0,3,2 */
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));    /* 2,3,0,1 */
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));   /* half mirror */
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));   /* mirror */
    const int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    const int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return op(op(a, b), op(c, d));
}
template <class Op> __device__ inline float sxw_dpp_allf(float v, Op op)
{
    auto mv = [](float x, int ctrl) {
        const int i = __float_as_int(x);
        int r;
        switch (ctrl) {
        case 0: r = __builtin_amdgcn_update_dpp(i, i, 0xB1, 0xF, 0xF, false); break;
        case 1: r = __builtin_amdgcn_update_dpp(i, i, 0x4E, 0xF, 0xF, false); break;
        case 2: r = __builtin_amdgcn_update_dpp(i, i, 0x141, 0xF, 0xF, false); break;
        default: r = __builtin_amdgcn_update_dpp(i, i, 0x140, 0xF, 0xF, false); break;
        }
        return __int_as_float(r);
    };
    v = op(v, mv(v, 0));
    v = op(v, mv(v, 1));
    v = op(v, mv(v, 2));
    v = op(v, mv(v, 3));
    const float a = sxw_rl(v, 0), b = sxw_rl(v, 16), c = sxw_rl(v, 32), d = sxw_rl(v, 48);
    return op(op(a, b), op(c, d));
}

/* keep the five largest values seen by this lane, t[0] >= ... >= t[4] (-inf = empty) */
struct SxwTop5 {
    float t[5];
    __device__ void init()
    {
#pragma unroll
        for (int i = 0; i < 5; i++) t[i] = -INFINITY;
    }
    __device__ void add(float v)
    {
#pragma unroll
        for (int i = 0; i < 5; i++) {
            const float hi = fmaxf(t[i], v), lo = fminf(t[i], v);
            t[i] = hi;
            v = lo;
        }
    }
    /* the fifth largest value over the wavefront (needs >= 5 values in total) */
    __device__ float fifth()
    {
        const int lane = threadIdx.x & 63;
        float res = 0.0f;
#pragma unroll
        for (int round = 0; round < 5; round++) {
            const float m = sxw_max(t[0]);
            const unsigned long long b = __ballot(t[0] == m);
            if (lane == __builtin_ctzll(b)) {
#pragma unroll
                for (int i = 0; i < 4; i++) t[i] = t[i + 1];
                t[4] = -INFINITY;
            }
            res = m;
        }
        return res;
    }
};

/*
 * Transposed ordered folds.  The NS terms of each member lane of a chunk are written to
 * LDS compacted in lane order (row s holds the terms of sum s); then lane s adds row s
 * in order into its own accumulator.  NS sums advance together at one LDS read and one
 * add per member, and each sum is still the sequential sum in raster order.
 * rows points to this wavefront's scratch of NS rows of SXW_ROW elements.
 */
#define SXW_ROW 66   /* row stride: rows start on different LDS banks */

template <int NS, class T>
__device__ inline void sxw_fold(T *rows, const T (&t)[NS], bool member, T &acc)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long b = __ballot(member);
    const int cnt = __popcll(b);
    if (member) {
        const int pos = __popcll(b & sxw_lanes_below(lane));
#pragma unroll
        for (int s = 0; s < NS; s++) rows[s * SXW_ROW + pos] = t[s];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < NS) {
        const T *row = rows + lane * SXW_ROW;
        T a = acc;
#pragma unroll 16
        for (int j = 0; j < cnt; j++) a += row[j];
        acc = a;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/*
 * The same with ND double sums and NF float sums advancing together: lane s < ND adds
 * double row s, lane ND + s adds float row s.  Every chain lane performs one double and
 * one float add per member (the add it does not own reads a valid row and is ignored),
 * so the chains of both types interleave in one instruction stream.  drows: ND rows,
 * frows: NF rows (at least one row each).
 */
template <int ND, int NF>
__device__ inline void sxw_fold2(double *drows, float *frows, const double (&td)[ND],
                                 const float (&tf)[NF], bool member, double &dacc, float &facc)
{
    static_assert(ND + NF <= 64, "one lane per sum");
    const int lane = threadIdx.x & 63;
    const unsigned long long b = __ballot(member);
    const int cnt = __popcll(b);
    if (member) {
        const int pos = __popcll(b & sxw_lanes_below(lane));
#pragma unroll
        for (int s = 0; s < ND; s++) drows[s * SXW_ROW + pos] = td[s];
#pragma unroll
        for (int s = 0; s < NF; s++) frows[s * SXW_ROW + pos] = tf[s];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < ND + NF) {
        const double *dr = drows + (lane < ND ? lane : 0) * SXW_ROW;
        const float *fr = frows + (lane >= ND ? lane - ND : 0) * SXW_ROW;
        double da = dacc;
        float fa = facc;
#pragma unroll 16
        for (int j = 0; j < cnt; j++) {
            da += dr[j];
            fa += fr[j];
        }
        dacc = da;
        facc = fa;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

