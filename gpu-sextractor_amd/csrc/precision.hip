/*
 * precision.hip -- optional re-evaluation of the graded catalogue columns (FLUX_ISO,
 * X/Y_IMAGE, X2/Y2/XY_IMAGE, A/B/THETA_IMAGE, CXX/CYY/CXY_IMAGE) after CLEAN, in a
 * precision mode chosen by SXG_OPT_PRECISION:
 *
 *   0  the reference's arithmetic (default; nothing runs here): float products of cd and
 *      the pixel offsets summed in double in raster order, the non-central formula
 *      x2 = Sxx/S - xm^2 (cudaanalyse.cu:1770-1898, :2048-2125);
 *   1  FP64 throughout, two passes: the barycentre, then the central moments about it,
 *      each a parallel (tree) reduction of double terms;
 *   2  compensated FP32: the flux and the first and second moment sums (pixel offsets
 *      taken from the object's first pixel) are Kahan sums in FP32 per lane, merged
 *      across lanes with the same error-free transformation; FP64 only for the
 *      central-moment step (x2 - xm^2, y2 - ym^2, xy - xm*ym) and the ellipse.
 *
 * Definitions are the reference's: X, Y is the cd-weighted barycentre of the object's
 * level pixels (its kept component, the FULL preanalysis); the second moments are taken
 * about it over the final pixels (the robust / MERGED pass for deblended objects; the
 * same pixels for the others); the singular-matrix rule adds 1/12 to the variances
 * (:1847-1853); FLUX_ISO is the sum of the background-subtracted values.  The columns
 * used by gatherup and CLEAN, and therefore the segmentation map, are the reference's
 * in every mode: this runs after CLEAN and rewrites catalogue columns only.
 *
 * One wavefront per object; the level pixels of a deblended object are the raw pixels
 * whose kept component (kcomp, written by the deblending) is the object's.
 */
#include "sxg_internal.h"
#include "sxg_objects.h"
#include "sxg_wave.h"

struct PrecArgs {
    const int *opix, *objoff;        /* final pixels of each object */
    const int *fpia, *rawoff, *rawcnt;   /* raw pixel lists */
    const int *kcomp;                /* kept component of each raw pixel, -1 if none */
    const int *objcomp;              /* component of each object in its raw object */
    const float *sub, *cd;
    int W, nobj;
    const int *nd;                   /* planned run: the device's object count */
    ObjCols o;
};

/* Kahan sum in one lane */
struct Kahan {
    float s = 0.0f, c = 0.0f;
    __device__ void add(float v)
    {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

/* merge lane sums: a Kahan sum over the 64 partial (value, correction) pairs in lane
   order, done by every lane identically */
__device__ inline float kahan_merge(const Kahan &k)
{
    Kahan m;
    for (int j = 0; j < 64; j++) {
        m.add(__shfl(k.s, j));
        m.add(-__shfl(k.c, j));
    }
    return m.s - m.c;
}

__device__ inline void ellipse_store(ObjCols &o, int obj, double xm, double ym, double x2,
                                     double y2, double xy)
{
    double temp2 = x2 * y2 - xy * xy;
    if (temp2 < 0.00694) {
        x2 += 0.0833333;
        y2 += 0.0833333;
        temp2 = x2 * y2 - xy * xy;
    }
    const double d = x2 - y2;
    const double theta = fabs(d) > 0.0 ? atan2(2.0 * xy, d) / 2.0 : SXG_PI / 4.0;
    const double temp = sqrt(0.25 * d * d + xy * xy);
    const double pm = 0.5 * (x2 + y2);
    o.mx[obj] = xm;
    o.my[obj] = ym;
    o.mx2[obj] = x2;
    o.my2[obj] = y2;
    o.mxy[obj] = xy;
    o.a[obj] = (float)sqrt(pm + temp);
    o.b[obj] = (float)sqrt(pm - temp);
    o.theta[obj] = (float)(theta * 180.0 / SXG_PI);
    o.cxx[obj] = (float)(y2 / temp2);
    o.cyy[obj] = (float)(x2 / temp2);
    o.cxy[obj] = (float)(-2 * xy / temp2);
}

template <int MODE> __global__ void __launch_bounds__(256) k_precision(PrecArgs A)
{
    const int lane = threadIdx.x & 63;
    const int obj = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (obj >= sxg_dn(A.nd, A.nobj)) return;
    ObjCols o = A.o;
    const int W = A.W;
    const int L = o.level[obj];
    const int r = o.root[obj];
    const int c = A.objcomp[obj];
    const int off = A.rawoff[r], n = A.rawcnt[r];
    const int *lv = L > 0 ? A.kcomp + off : nullptr;
    const int g0 = A.objoff[obj], nf = A.objoff[obj + 1] - g0;
    const int *F = A.opix + g0;
    /* reference point for the offsets: the object's first final pixel */
    const int x0 = F[0] % W, y0 = F[0] / W;
    double xm, ym, x2, y2, xy, flux;
    if (MODE == 1) {
        double s = 0.0, sx = 0.0, sy = 0.0;
        for (int k = lane; k < (lv ? n : nf); k += 64) {
            int p;
            if (lv) {
                if (lv[k] != c) continue;
                p = A.fpia[off + k];
            } else {
                p = F[k];
            }
            const double w = A.cd[p];
            s += w;
            sx += w * (double)(p % W - x0);
            sy += w * (double)(p / W - y0);
        }
        s = sxw_sum(s);
        sx = sxw_sum(sx);
        sy = sxw_sum(sy);
        const double xb = sx / s, yb = sy / s;
        double t = 0.0, sxx = 0.0, syy = 0.0, sxy = 0.0, fl = 0.0;
        for (int k = lane; k < nf; k += 64) {
            const int p = F[k];
            const double w = A.cd[p], dx = (double)(p % W - x0) - xb, dy = (double)(p / W - y0) - yb;
            t += w;
            sxx += w * dx * dx;
            syy += w * dy * dy;
            sxy += w * dx * dy;
            fl += (double)A.sub[p];
        }
        t = sxw_sum(t);
        x2 = sxw_sum(sxx) / t;
        y2 = sxw_sum(syy) / t;
        xy = sxw_sum(sxy) / t;
        flux = sxw_sum(fl);
        xm = xb + x0;
        ym = yb + y0;
    } else {
        Kahan s, sx, sy;
        for (int k = lane; k < (lv ? n : nf); k += 64) {
            int p;
            if (lv) {
                if (lv[k] != c) continue;
                p = A.fpia[off + k];
            } else {
                p = F[k];
            }
            const float w = A.cd[p];
            s.add(w);
            sx.add(w * (float)(p % W - x0));
            sy.add(w * (float)(p / W - y0));
        }
        const double S = kahan_merge(s), SX = kahan_merge(sx), SY = kahan_merge(sy);
        const double xb = SX / S, yb = SY / S;
        Kahan t, sxx, syy, sxy, fl, tx, ty;
        for (int k = lane; k < nf; k += 64) {
            const int p = F[k];
            const float w = A.cd[p], dx = (float)(p % W - x0), dy = (float)(p / W - y0);
            t.add(w);
            tx.add(w * dx);
            ty.add(w * dy);
            sxx.add(w * dx * dx);
            syy.add(w * dy * dy);
            sxy.add(w * dx * dy);
            fl.add(A.sub[p]);
        }
        const double T = kahan_merge(t), TX = kahan_merge(tx), TY = kahan_merge(ty);
        const double SXX = kahan_merge(sxx), SYY = kahan_merge(syy), SXY = kahan_merge(sxy);
        flux = kahan_merge(fl);
        /* the central-moment step in FP64, about the level barycentre (xb, yb):
           E[(x-xb)^2] = Sxx/T - 2 xb Tx/T + xb^2 */
        const double ex = TX / T, ey = TY / T;
        x2 = SXX / T - 2.0 * xb * ex + xb * xb;
        y2 = SYY / T - 2.0 * yb * ey + yb * yb;
        xy = SXY / T - xb * ey - yb * ex + xb * yb;
        xm = xb + x0;
        ym = yb + y0;
    }
    if (lane != 0) return;
    ellipse_store(o, obj, xm, ym, x2, y2, xy);
    o.flux[obj] = (float)flux;
}

/* rewrite the graded columns of the current frame's objects in precision mode `mode`
   (1 or 2); called after CLEAN */
bool sxg_precision_columns(int mode)
{
    SxgState &S = sxg_state();
    if (mode == 0 || S.nobj == 0 || !S.opix || !S.objoff) return true;
    RawObj ro;
    if (!sxg_raw_alloc(S.nraw, ro)) return false;
    PrecArgs A;
    A.opix = S.opix;
    A.objoff = S.objoff;
    A.fpia = S.fpia;
    A.rawoff = ro.off;
    A.rawcnt = ro.count;
    A.kcomp = sxg_take<int>("deb.kcomp", (size_t)(S.npix_raw > 0 ? S.npix_raw : 1));
    A.objcomp = sxg_take<int>("obj.comp", S.nobj);
    A.sub = S.img.as<float>();
    A.cd = S.cd.as<float>();
    A.W = S.W;
    A.nobj = S.nobj;
    A.nd = S.plan ? &S.pdev->nobj : nullptr;
    if (!sxg_obj_view(A.o) || !A.kcomp || !A.objcomp) return false;
    const dim3 grid(sxg_div_up(S.nobj, 4)), block(256);
    if (mode == 1)
        hipLaunchKernelGGL(k_precision<1>, grid, block, 0, S.stream, A);
    else
        hipLaunchKernelGGL(k_precision<2>, grid, block, 0, S.stream, A);
    SXG_HIP_RET(hipGetLastError(), false);
    return true;
}
