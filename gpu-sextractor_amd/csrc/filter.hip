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
#include <string>

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

/* the K x K mask on the device: it stays there while the caller passes the same one; a
   new one is copied synchronously, so the caller's array is free again on return */
static const float *upload_mask(const float *h_mask, int K, const char *who)
{
    SxgState &S = sxg_state();
    if (!S.dmask.ensure(sizeof(float) * SXG_MAXMASK * SXG_MAXMASK)) return nullptr;
    float *dmask = S.dmask.as<float>();
    const size_t mb = sizeof(float) * K * K;
    if (S.mask_k != K || S.mask_buf != S.dmask.ptr || memcmp(S.mask_host, h_mask, mb) != 0) {
        if (S.capturing) {
            sxg_set_error(std::string(who) + ": a new mask while a graph is captured");
            return nullptr;
        }
        SXG_HIP_RET(hipStreamSynchronize(S.stream), nullptr);
        SXG_HIP_RET(hipMemcpy(dmask, h_mask, mb, hipMemcpyHostToDevice), nullptr);
        memcpy(S.mask_host, h_mask, mb);
        S.mask_k = K;
        S.mask_buf = S.dmask.ptr;
    }
    return dmask;
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
    const float *dmask = upload_mask(h_mask, K, "cudaFilter");
    if (!dmask) return;
    sxg_stage_begin(ST_FILTER);
    sxg_kt_begin(SXG_KT_CONV);
    sxg_convolve(S.img.as<float>(), S.cd.as<float>(), S.W, S.H, dmask, K, S.stream);
    sxg_kt_end(SXG_KT_CONV);
    sxg_stage_end(ST_FILTER);
    S.filtered = true;
    SXG_HIP(hipGetLastError());
    sxg_stage_wait();
}

/* ------------------------------------------------- fused subtraction + filter ---- */
/* Whole-frame runs: one pass over the input frame does the background subtraction
   (k_subback's spline, the same float expression), the convolution (k_convolve's tiles
   and tap order) and the detection parents (k_det_init's ballot over aligned 64-pixel
   windows: with W a multiple of 64 these are the tile rows a wavefront holds).  The
   tile loader subtracts the spline from every input pixel it stages (the halo too, with
   the same value the neighbouring tile computes) and writes the tile's own pixels to
   sub; the lanes then convolve, write cd, compare with the detection threshold and link
   each detected pixel to the first pixel of its run inside its window, or write -1.
   One read of the frame and three plane writes instead of three kernels' six passes. */
struct SplineArgs {
    const float *nodes, *dnodes;   /* per frame row: nbx nodes and second derivatives */
    const int *colseg;
    const float *coldx;
    int nbx;
};

template <int K>
__global__ void __launch_bounds__(CV_THREADS) k_subconv_det(const float *__restrict__ src,
                                                            float *__restrict__ sub,
                                                            float *__restrict__ out, int *par,
                                                            int W, int H,
                                                            const float *__restrict__ mask,
                                                            SplineArgs sp, float t, WeightArgs wa,
                                                            const float *td)
{
    constexpr int HW = K / 2;
    constexpr int LW = CV_TX + K - 1, LH = CV_TY + K - 1;
    __shared__ float tile[LH][LW + 1];
    if (td) t = *td;
    int X0, Y0;
    cv_tile(X0, Y0);
    for (int k = threadIdx.x; k < LW * LH; k += CV_THREADS) {
        const int ly = k / LW, lx = k - ly * LW;
        const int gx = X0 + lx - HW, gy = Y0 + ly - HW;
        float v = 0.0f;
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const size_t o = (size_t)gy * W + gx;
            const float bval = sxg_spline_at(sp.nodes + (size_t)gy * sp.nbx,
                                             sp.dnodes + (size_t)gy * sp.nbx, sp.colseg,
                                             sp.coldx, sp.nbx, gx);
            v = src[o] - bval;
            if (lx >= HW && lx < HW + CV_TX && ly >= HW && ly < HW + CV_TY) sub[o] = v;
        }
        tile[ly][lx] = v;
    }
    __syncthreads();
    const int tx = threadIdx.x % CV_TX, ty = threadIdx.x / CV_TX;
    const int r0 = ty * CV_R;
    float acc[CV_R];
#pragma unroll
    for (int r = 0; r < CV_R; r++) acc[r] = 0.0f;
#pragma unroll
    for (int tt = 0; tt < CV_R + K - 1; tt++) {
        float v[K];
#pragma unroll
        for (int j = 0; j < K; j++) v[j] = tile[r0 + tt][tx + j];
#pragma unroll
        for (int r = 0; r < CV_R; r++) {
            const int i = tt - r;
            if (i >= 0 && i < K) {
#pragma unroll
                for (int j = 0; j < K; j++) {
                    float prod = mask[i * K + j] * v[j];
                    acc[r] = acc[r] + prod;
                }
            }
        }
    }
    const int x = X0 + tx, lane = tx;   /* X0 is a multiple of 64: lane = x mod 64 */
    const unsigned long long below = ~0ull >> (63 - lane);
#pragma unroll
    for (int r = 0; r < CV_R; r++) {
        const int y = Y0 + r0 + r;
        const bool live = x < W && y < H;
        const long p = (long)y * W + x;
        const bool fg = live && acc[r] >= sxg_det_thresh(wa, t, p);
        const unsigned long long F = __ballot(fg);
        const bool start = fg && (lane == 0 || !((F >> (lane - 1)) & 1ull));
        const unsigned long long S = __ballot(start);
        if (live) {
            out[p] = acc[r];
            par[p] = fg ? (int)(p - lane + (63 - __builtin_clzll(S & below))) : -1;
        }
    }
}

bool sxg_subconv_detect(const float *h_mask, int K, float dthresh)
{
    SxgState &S = sxg_state();
    if (!S.have_frame || !S.sub_src || !h_mask || (S.W & 63) != 0 ||
        !(K == 3 || K == 5 || K == 7 || K == 9 || K == 11))
        return false;
    const float *dmask = upload_mask(h_mask, K, "sxg_subconv_detect");
    if (!dmask) return false;
    WeightArgs wa;
    if (!sxg_weight_args(wa)) return false;
    const BackBufs b = sxg_back_bufs_at(S.backbuf.as<char>(), S.bgeom);
    const SplineArgs sp{b.nodes, b.dnodes, b.colseg, b.coldx, S.bgeom.nbx};
    const float *td = S.plan ? &S.pdev->dthresh : nullptr;
    const dim3 grid(sxg_div_up(S.W, CV_TX), sxg_div_up(S.H, CV_TY)), block(CV_THREADS);
    float *sub = S.img.as<float>(), *cd = S.cd.as<float>();
    int *par = S.aux.as<int>();
    const float *src = S.sub_src;
    S.sub_src = nullptr;
    sxg_stage_begin(ST_FILTER);
    sxg_kt_begin(SXG_KT_CONV);
    switch (K) {
    case 3: hipLaunchKernelGGL(k_subconv_det<3>, grid, block, 0, S.stream, src, sub, cd, par, S.W, S.H, dmask, sp, dthresh, wa, td); break;
    case 5: hipLaunchKernelGGL(k_subconv_det<5>, grid, block, 0, S.stream, src, sub, cd, par, S.W, S.H, dmask, sp, dthresh, wa, td); break;
    case 7: hipLaunchKernelGGL(k_subconv_det<7>, grid, block, 0, S.stream, src, sub, cd, par, S.W, S.H, dmask, sp, dthresh, wa, td); break;
    case 9: hipLaunchKernelGGL(k_subconv_det<9>, grid, block, 0, S.stream, src, sub, cd, par, S.W, S.H, dmask, sp, dthresh, wa, td); break;
    default: hipLaunchKernelGGL(k_subconv_det<11>, grid, block, 0, S.stream, src, sub, cd, par, S.W, S.H, dmask, sp, dthresh, wa, td); break;
    }
    sxg_kt_end(SXG_KT_CONV);
    sxg_stage_end(ST_FILTER);
    SXG_HIP_RET(hipGetLastError(), false);
    S.filtered = true;
    S.det_init_done = true;
    return true;
}
