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
 * plus halo, all of a lane's loads in flight together: cv_load); each lane owns one
 * column and 8 consecutive rows.  Sweeping the input
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

/* the background spline of a frame: per frame row, nbx nodes and their second
   derivatives; per column, its left node and dx (the cumulative recurrence of
   subbackline, back.c:880-908) */
struct SplineArgs {
    const float *nodes, *dnodes;   /* per frame row: nbx nodes and second derivatives */
    const int *colseg;
    const float *coldx;
    int nbx;
};

/* what a staged tile pixel holds: the frame value (TS_COPY), or the frame value minus the
   background spline at that pixel (TS_SPLINE: rows of nbx > 1 nodes; TS_NODE0: one node
   per row), the expression of sxg_spline_at */
enum TileSrc { TS_COPY = 0, TS_SPLINE, TS_NODE0 };

/*
 * Stage the (CV_TY + K - 1) x (CV_TX + K - 1) input tile of the workgroup at (X0, Y0) in
 * LDS, 0 outside the frame.  Every global load of a lane is issued before the first is
 * used, so a workgroup waits for about two memory round trips instead of one or more per
 * staged pixel: the columns 0..63 of the tile go by rows, one row per wavefront and step,
 * with a lane's column tables read once for all its rows; the right strip (columns
 * 64..63 + K - 1) is a flat list over the workgroup.  The addresses are clamped into the
 * frame, so that the loads need no branch, and the values outside it are replaced by 0
 * afterwards.
 */
template <int K, int MODE>
__device__ __forceinline__ void cv_load(float (*tile)[CV_TX + K], const float *__restrict__ src,
                                        const SplineArgs &sp, int W, int H, int X0, int Y0)
{
    constexpr int HW = K / 2;
    constexpr int LH = CV_TY + K - 1;
    constexpr int NW = CV_THREADS / 64;
    constexpr int NRW = (LH + NW - 1) / NW;              /* rows per wavefront */
    constexpr int NS = (K - 1) * LH;                     /* right-strip pixels */
    constexpr int NSU = (NS + CV_THREADS - 1) / CV_THREADS;
    constexpr bool SPL = MODE == TS_SPLINE;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nbx = sp.nbx;
    /* the lane's column of the main part and its spline segment */
    const int gxm = X0 + lane - HW;
    const bool cm = gxm >= 0 && gxm < W;
    const int gxmc = min(max(gxm, 0), W - 1);
    int sm = 0;
    float dxm = 0.0f;
    if (SPL) {
        sm = sp.colseg[gxmc];
        dxm = sp.coldx[gxmc];
    }
    /* the lane's right-strip pixels */
    int sly[NSU > 0 ? NSU : 1], slx[NSU > 0 ? NSU : 1], sgx[NSU > 0 ? NSU : 1],
        ss[NSU > 0 ? NSU : 1];
    float sdx[NSU > 0 ? NSU : 1];
    bool sin_[NSU > 0 ? NSU : 1];
#pragma unroll
    for (int u = 0; u < NSU; u++) {
        const int e = threadIdx.x + u * CV_THREADS;
        const int ly = e / (K - 1), lx = CV_TX + (e - ly * (K - 1));
        const int gx = X0 + lx - HW, gy = Y0 + ly - HW;
        sly[u] = ly;
        slx[u] = lx;
        sin_[u] = e < NS && gx >= 0 && gx < W && gy >= 0 && gy < H;
        sgx[u] = min(max(gx, 0), W - 1);
        ss[u] = 0;
        sdx[u] = 0.0f;
        if (SPL) {
            ss[u] = sp.colseg[sgx[u]];
            sdx[u] = sp.coldx[sgx[u]];
        }
    }
    /* every load: frame values, then the spline nodes of their rows */
    float mv[NRW], mn0[NRW], mn1[NRW], md0[NRW], md1[NRW];
#pragma unroll
    for (int u = 0; u < NRW; u++) {
        const int gy = min(max(Y0 + wid + NW * u - HW, 0), H - 1);
        mv[u] = src[(size_t)gy * W + gxmc];
        const float *nr = sp.nodes + (size_t)gy * nbx, *dr = sp.dnodes + (size_t)gy * nbx;
        if (SPL) {
            mn0[u] = nr[sm];
            mn1[u] = nr[sm + 1];
            md0[u] = dr[sm];
            md1[u] = dr[sm + 1];
        } else if (MODE == TS_NODE0) {
            mn0[u] = nr[0];
        }
    }
    float sv[NSU > 0 ? NSU : 1], sn0[NSU > 0 ? NSU : 1], sn1[NSU > 0 ? NSU : 1],
        sd0[NSU > 0 ? NSU : 1], sd1[NSU > 0 ? NSU : 1];
#pragma unroll
    for (int u = 0; u < NSU; u++) {
        const int gy = min(max(Y0 + sly[u] - HW, 0), H - 1);
        sv[u] = src[(size_t)gy * W + sgx[u]];
        const float *nr = sp.nodes + (size_t)gy * nbx, *dr = sp.dnodes + (size_t)gy * nbx;
        if (SPL) {
            sn0[u] = nr[ss[u]];
            sn1[u] = nr[ss[u] + 1];
            sd0[u] = dr[ss[u]];
            sd1[u] = dr[ss[u] + 1];
        } else if (MODE == TS_NODE0) {
            sn0[u] = nr[0];
        }
    }
    /* the staged values (sxg_spline_at's expression, term for term) */
    auto back = [](float dx, float n0, float n1, float d0, float d1) {
        const float cdx = 1 - dx;
        return cdx * (n0 + (cdx * cdx - 1) * d0) + dx * (n1 + (dx * dx - 1) * d1);
    };
#pragma unroll
    for (int u = 0; u < NRW; u++) {
        const int ly = wid + NW * u;
        if (ly < LH) {
            const int gy = Y0 + ly - HW;
            float v = mv[u];
            if (SPL) v = v - back(dxm, mn0[u], mn1[u], md0[u], md1[u]);
            else if (MODE == TS_NODE0) v = v - mn0[u];
            tile[ly][lane] = cm && gy >= 0 && gy < H ? v : 0.0f;
        }
    }
#pragma unroll
    for (int u = 0; u < NSU; u++) {
        if (threadIdx.x + u * CV_THREADS < NS) {
            float v = sv[u];
            if (SPL) v = v - back(sdx[u], sn0[u], sn1[u], sd0[u], sd1[u]);
            else if (MODE == TS_NODE0) v = v - sn0[u];
            tile[sly[u]][slx[u]] = sin_[u] ? v : 0.0f;
        }
    }
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
    cv_load<K, TS_COPY>(tile, in, SplineArgs{}, W, H, X0, Y0);
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
   the same value the neighbouring tile computes); the lanes then convolve, write the
   tile's own subtracted pixels to sub and the filtered ones to cd, compare with the
   detection threshold and link each detected pixel to the first pixel of its run inside
   its window, or write -1.
   One read of the frame and three plane writes instead of three kernels' six passes. */
template <int K>
__global__ void __launch_bounds__(CV_THREADS) k_subconv_det(const float *__restrict__ src,
                                                            float *__restrict__ sub,
                                                            float *__restrict__ out, int *par,
                                                            unsigned long long *fgm, int W, int H,
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
    if (sp.nbx > 1) cv_load<K, TS_SPLINE>(tile, src, sp, W, H, X0, Y0);
    else cv_load<K, TS_NODE0>(tile, src, sp, W, H, X0, Y0);
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
            /* the window's detection mask (W is a multiple of 64: a tile row of a
               wavefront is one window) */
            if (lane == 0) fgm[p >> 6] = F;
            /* the subtracted pixel, as staged */
            sub[p] = tile[HW + r0 + r][HW + tx];
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
    unsigned long long *fgm =
        sxg_take<unsigned long long>("det.fgmask", sxg_div_up((long)S.N, 64));
    if (!fgm) return false;
    const float *src = S.sub_src;
    S.sub_src = nullptr;
    sxg_stage_begin(ST_FILTER);
    sxg_kt_begin(SXG_KT_CONV);
    switch (K) {
    case 3: hipLaunchKernelGGL(k_subconv_det<3>, grid, block, 0, S.stream, src, sub, cd, par, fgm, S.W, S.H, dmask, sp, dthresh, wa, td); break;
    case 5: hipLaunchKernelGGL(k_subconv_det<5>, grid, block, 0, S.stream, src, sub, cd, par, fgm, S.W, S.H, dmask, sp, dthresh, wa, td); break;
    case 7: hipLaunchKernelGGL(k_subconv_det<7>, grid, block, 0, S.stream, src, sub, cd, par, fgm, S.W, S.H, dmask, sp, dthresh, wa, td); break;
    case 9: hipLaunchKernelGGL(k_subconv_det<9>, grid, block, 0, S.stream, src, sub, cd, par, fgm, S.W, S.H, dmask, sp, dthresh, wa, td); break;
    default: hipLaunchKernelGGL(k_subconv_det<11>, grid, block, 0, S.stream, src, sub, cd, par, fgm, S.W, S.H, dmask, sp, dthresh, wa, td); break;
    }
    sxg_kt_end(SXG_KT_CONV);
    sxg_stage_end(ST_FILTER);
    SXG_HIP_RET(hipGetLastError(), false);
    S.filtered = true;
    S.det_init_done = true;
    return true;
}
