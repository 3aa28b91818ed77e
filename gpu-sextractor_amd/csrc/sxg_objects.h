/*
 * sxg_objects.h -- deblending-tree nodes, per-object columns and the moment/ellipse
 * derivation shared by the FULL and ROBUST preanalysis passes
 * (the reference's src/cuda/cudaanalyse.cu:1837-1903 and :2048-2125).
 */
#pragma once
#include "sxg_internal.h"

#define COMP_KEPT 1
#define COMP_SIG 2
#define COMP_OKA 4

/* deblending-tree node as the object kernels see it, stored in the raw object's region
   of the table in the reference's order (level, then raster order of the first pixel):
   level, first (raster index of its first pixel), flags & COMP_KEPT */
struct Comp {
    int level, first, flags;
};

/* per-object columns (structure of arrays, object order) */
struct ObjCols {
    int *level, *root, *firstpix, *flag, *fdnpix, *dnpix, *npix, *xmin, *xmax, *ymin, *ymax,
        *peakx, *peaky, *iso;
    double *mx, *my, *mx2, *my2, *mxy, *poserr_mx2, *poserr_my2, *poserr_mxy;
    float *a, *b, *theta, *cxx, *cyy, *cxy, *abcor, *fdflux, *dflux, *flux, *fluxerr, *fdpeak,
        *dpeak, *peak, *bkg, *thresh, *mthresh, *fwhm, *dthresh;
    unsigned *master;
};
bool sxg_obj_alloc(int nobj, ObjCols &o);
bool sxg_obj_view(ObjCols &o);   /* the current frame's columns */

struct Moments {
    double mx, my, mx2, my2, mxy;
    float a, b, theta, cxx, cyy, cxy, abcor;
};

/* xm = mx/rv ...; the MERGED branch re-centres on the previous barycentre
   (cudaanalyse.cu:2059-2070); singular moment matrices get 1/12 added to the variances
   (:1847-1853); a, b, theta, cxx, cyy, cxy as in :1857-1883; abcor as in :1885-1894
   with the canonical log. */
__device__ inline void sxg_moments(double mx, double my, double mx2, double my2, double mxy,
                                   float rv, bool merged, double prev_mx, double prev_my,
                                   int xmin, int ymin, int area2, int dnpix, double dthresh,
                                   double dthresh2, Moments &m)
{
    double xm = mx / rv, ym = my / rv, xm2, ym2, xym, temp, temp2, theta, pmx2, pmy2;
    if (merged) {
        const double xn = prev_mx - xmin, yn = prev_my - ymin;
        xm2 = mx2 / rv + xn * xn - 2 * xm * xn;
        ym2 = my2 / rv + yn * yn - 2 * ym * yn;
        xym = mxy / rv + xn * yn - xm * yn - xn * ym;
        xm = xn;
        ym = yn;
    } else {
        xm2 = mx2 / rv - xm * xm;
        ym2 = my2 / rv - ym * ym;
        xym = mxy / rv - xm * ym;
    }
    if ((temp2 = xm2 * ym2 - xym * xym) < 0.00694) {
        xm2 += 0.0833333;
        ym2 += 0.0833333;
        temp2 = xm2 * ym2 - xym * xym;
    }
    if ((fabs(temp = xm2 - ym2)) > 0.0)
        theta = atan2(2.0 * xym, temp) / 2.0;
    else
        theta = SXG_PI / 4.0;
    temp = sqrt(0.25 * temp * temp + xym * xym);
    pmy2 = pmx2 = 0.5 * (xm2 + ym2);
    pmx2 += temp;
    pmy2 -= temp;
    m.mx = xm + xmin;
    m.my = ym + ymin;
    m.mx2 = xm2;
    m.my2 = ym2;
    m.mxy = xym;
    m.a = (float)sqrt(pmx2);
    m.b = (float)sqrt(pmy2);
    m.theta = (float)(theta * 180.0 / SXG_PI);
    m.cxx = (float)(ym2 / temp2);
    m.cyy = (float)(xm2 / temp2);
    m.cxy = (float)(-2 * xym / temp2);
    const double darea = (double)area2 - dnpix;
    const double t1t2 = dthresh / dthresh2;
    if (t1t2 > 0.0) {
        m.abcor = (float)((darea < 0.0 ? darea : -1.0)
                          / (2 * SXG_PI * sx_log(t1t2 < 1.0 ? t1t2 : 0.99) * m.a * m.b));
        if (m.abcor > 1.0) m.abcor = 1.0f;
    } else
        m.abcor = 1.0f;
}

__device__ inline void sxg_store_moments(ObjCols &o, int i, const Moments &m)
{
    o.mx[i] = m.mx; o.my[i] = m.my; o.mx2[i] = m.mx2; o.my2[i] = m.my2; o.mxy[i] = m.mxy;
    o.a[i] = m.a; o.b[i] = m.b; o.theta[i] = m.theta;
    o.cxx[i] = m.cxx; o.cyy[i] = m.cyy; o.cxy[i] = m.cxy; o.abcor[i] = m.abcor;
}
