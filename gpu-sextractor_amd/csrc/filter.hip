/*
 * filter.hip -- detection filtering (cudaFilter, the reference's
 * src/cuda/cudafilter.cu:150), bit-exact with the CPU convolve of src/filter.c:39-94.
 *
 * Definition: out(x,y) = sum over mask rows i then columns j (row-major) of
 * mask[i][j] * in(x+j-w/2, y+i-h/2), accumulated in FP32 from +0 with every product and
 * every addition rounded separately (no FMA); taps outside the frame contribute nothing.
 * Here they are zero-padded: m*0 is a signed zero and the accumulator can never be -0
 * (it starts at +0 and x + (-x) rounds to +0), so adding it leaves the sum unchanged.
 *
 * Layout: a workgroup of 256 lanes produces a 64 x 32 output tile.  The input tile with
 * its halo is staged in LDS once (coalesced row loads, one global read per input pixel
 * plus halo); each lane owns one column and 8 consecutive rows.  Sweeping the input
 * rows once, input row t feeds output row r through mask row i = t - r, so every
 * output still sees its taps in mask row-major order while each LDS row is read once
 * per lane for 8 outputs.  Bound: VALU (2 K^2 FP32 operations per pixel, no FMA).
 */
#include <cstring>

#include "sxg_internal.h"

#define CV_TX 64
#define CV_TY 32
#define CV_R 8                       /* output rows per lane */
#define CV_THREADS (CV_TX * CV_TY / CV_R)

/* The tile of this workgroup.  Workgroups go to the eight XCDs round robin by linear
   index, so the linear index is mapped to tiles such that each XCD takes one run of
   consecutive tiles in raster order: neighbouring tiles, which read each other's halo
   columns and rows, then share them in that XCD's L2 instead of each XCD fetching them
   from HBM.  A bijection on the tiles (the first n % 8 XCDs take one tile more). */
__device__ inline void cv_tile(int &X0, int &Y0)
{
    const int gx = gridDim.x, n = gx * gridDim.y;
    const int b = blockIdx.y * gx + blockIdx.x;
    const int q = n >> 3, r = n & 7, xcd = b & 7, slot = b >> 3;
    const int t = xcd * q + min(xcd, r) + slot;
    X0 = (t % gx) * CV_TX;
    Y0 = (t / gx) * CV_TY;
}

template <int K>
__global__ void __launch_bounds__(CV_THREADS) k_convolve(const float *__restrict__ in,
                                                         float *__restrict__ out, int W, int H,
                                                         const float *__restrict__ mask)
{
    constexpr int HW = K / 2;
    constexpr int LW = CV_TX + K - 1, LH = CV_TY + K - 1;
    __shared__ float tile[LH][LW + 1];
    int X0, Y0;
    cv_tile(X0, Y0);
    for (int k = threadIdx.x; k < LW * LH; k += CV_THREADS) {
        const int ly = k / LW, lx = k - ly * LW;
        const int gx = X0 + lx - HW, gy = Y0 + ly - HW;
        tile[ly][lx] = (gx >= 0 && gx < W && gy >= 0 && gy < H) ? in[(size_t)gy * W + gx] : 0.0f;
    }
    __syncthreads();
    const int tx = threadIdx.x % CV_TX, ty = threadIdx.x / CV_TX;
    const int r0 = ty * CV_R;
    float acc[CV_R];
#pragma unroll
    for (int r = 0; r < CV_R; r++) acc[r] = 0.0f;
#pragma unroll
    for (int t = 0; t < CV_R + K - 1; t++) {
        float v[K];
#pragma unroll
        for (int j = 0; j < K; j++) v[j] = tile[r0 + t][tx + j];
#pragma unroll
        for (int r = 0; r < CV_R; r++) {
            const int i = t - r;
            if (i >= 0 && i < K) {
#pragma unroll
                for (int j = 0; j < K; j++) {
                    float prod = mask[i * K + j] * v[j];
                    acc[r] = acc[r] + prod;
                }
            }
        }
    }
    const int x = X0 + tx;
    if (x < W) {
#pragma unroll
        for (int r = 0; r < CV_R; r++) {
            const int y = Y0 + r0 + r;
            if (y < H) out[(size_t)y * W + x] = acc[r];
        }
    }
}

/* generic square masks up to SXG_MAXMASK (runtime width), same order */
__global__ void __launch_bounds__(CV_THREADS) k_convolve_any(const float *__restrict__ in,
                                                             float *__restrict__ out, int W,
                                                             int H, const float *__restrict__ mask,
                                                             int K)
{
    extern __shared__ float tile[];
    const int HW = K / 2, LW = CV_TX + K - 1, LH = CV_TY + K - 1;
    int X0, Y0;
    cv_tile(X0, Y0);
    for (int k = threadIdx.x; k < LW * LH; k += CV_THREADS) {
        const int ly = k / LW, lx = k - ly * LW;
        const int gx = X0 + lx - HW, gy = Y0 + ly - HW;
        tile[ly * LW + lx] = (gx >= 0 && gx < W && gy >= 0 && gy < H) ? in[(size_t)gy * W + gx] : 0.0f;
    }
    __syncthreads();
    const int tx = threadIdx.x % CV_TX, ty = threadIdx.x / CV_TX;
    const int r0 = ty * CV_R;
    float acc[CV_R];
    for (int r = 0; r < CV_R; r++) acc[r] = 0.0f;
    for (int t = 0; t < CV_R + K - 1; t++) {
        const float *row = tile + (r0 + t) * LW + tx;
#pragma unroll
        for (int r = 0; r < CV_R; r++) {
            const int i = t - r;
            if (i >= 0 && i < K)
                for (int j = 0; j < K; j++) {
                    float prod = mask[i * K + j] * row[j];
                    acc[r] = acc[r] + prod;
                }
        }
    }
    const int x = X0 + tx;
    if (x < W)
        for (int r = 0; r < CV_R; r++) {
            const int y = Y0 + r0 + r;
            if (y < H) out[(size_t)y * W + x] = acc[r];
        }
}

/* conv(in -> out) with a square K x K mask already on the device */
void sxg_convolve(const float *in, float *out, int W, int H, const float *dmask, int K,
                  hipStream_t s)
{
    dim3 grid(sxg_div_up(W, CV_TX), sxg_div_up(H, CV_TY)), block(CV_THREADS);
    switch (K) {
    case 3: hipLaunchKernelGGL(k_convolve<3>, grid, block, 0, s, in, out, W, H, dmask); break;
    case 5: hipLaunchKernelGGL(k_convolve<5>, grid, block, 0, s, in, out, W, H, dmask); break;
    case 7: hipLaunchKernelGGL(k_convolve<7>, grid, block, 0, s, in, out, W, H, dmask); break;
    case 9: hipLaunchKernelGGL(k_convolve<9>, grid, block, 0, s, in, out, W, H, dmask); break;
    case 11: hipLaunchKernelGGL(k_convolve<11>, grid, block, 0, s, in, out, W, H, dmask); break;
    default: {
        size_t lds = sizeof(float) * (CV_TX + K - 1) * (CV_TY + K - 1);
        hipLaunchKernelGGL(k_convolve_any, grid, block, lds, s, in, out, W, H, dmask, K);
    }
    }
}

extern "C" void cudaFilter(float *h_mask, unsigned int mask_width)
{
    SxgState &S = sxg_state();
    if (!S.have_frame) {
        sxg_set_error("cudaFilter: no frame");
        return;
    }
    const int K = (int)mask_width;
    if (K < 1 || K > SXG_MAXMASK || !h_mask) {
        sxg_set_error("cudaFilter: mask width must be 1..31");
        return;
    }
    if (!S.dmask.ensure(sizeof(float) * SXG_MAXMASK * SXG_MAXMASK)) return;
    float *dmask = S.dmask.as<float>();
    /* the mask stays on the device while the caller passes the same one; a new one is
       copied synchronously, so the caller's array is free again on return */
    const size_t mb = sizeof(float) * K * K;
    if (S.mask_k != K || S.mask_buf != S.dmask.ptr || memcmp(S.mask_host, h_mask, mb) != 0) {
        if (S.capturing) {
            sxg_set_error("cudaFilter: a new mask while a graph is captured");
            return;
        }
        SXG_HIP(hipStreamSynchronize(S.stream));
        SXG_HIP(hipMemcpy(dmask, h_mask, mb, hipMemcpyHostToDevice));
        memcpy(S.mask_host, h_mask, mb);
        S.mask_k = K;
        S.mask_buf = S.dmask.ptr;
    }
    sxg_stage_begin(ST_FILTER);
    sxg_kt_begin(SXG_KT_CONV);
    sxg_convolve(S.img.as<float>(), S.cd.as<float>(), S.W, S.H, dmask, K, S.stream);
    sxg_kt_end(SXG_KT_CONV);
    sxg_stage_end(ST_FILTER);
    S.filtered = true;
    SXG_HIP(hipGetLastError());
    sxg_stage_wait();
}
