/*
 * analyse.hip -- robust preanalysis and isophotal analysis (analyse_gpu, the reference's
 * src/cuda/cudaanalyse.cu:3171) over each object's final pixels.
 *
 * Semantics: preanalyse_robust (cudaanalyse.cu:1906-2126) for objects of level > 0 with
 * the detection threshold and the previous barycentre; examineiso (:2176-2581): the
 * background at the rounded centroid (bilinear in the filtered mesh map, :2128-2172),
 * FLUX_ISO and its variance, ISOAREA above the analysis threshold, the peak and its
 * position, the position-error moments, the 5th brightest cd - dthresh (mthresh, with
 * the reference's EXT_MINAREA = 5, cudatypes.h:62), the 8 isophotal areas and the FWHM
 * log fit; the SATUR flag of endobject (:2873-2874).
 *
 * Layout: one 64-lane wavefront per raw object; lane j owns object j of that raw object
 * and walks the raw pixel list in raster order, taking the pixels whose object index is
 * its own, so every sum is the sequential raster-order sum of the reference's loops.
 */
#include "sxg_internal.h"
#include "sxg_objects.h"

#define AW 64
#define HEAPN 5

struct AnaArgs {
    const int *fpia, *off, *cnt;
    const int *objbase, *rootkept, *alloc;
    const float *sub, *cd, *back;
    int W, H, nbx, nby, backw, backh;
    float backsig, thresh, dthresh;
    double gain, satur_level;
    int errflag, peakflag, isoflag, fwhmflag, cleanflag;
    ObjCols o;
};

/* back.c:685-732 */
__device__ inline float back_at(const float *b, int x, int y, int nx, int ny, int backw, int backh)
{
    int xl, yl, pos;
    double dx, dy, cdx;
    float b0, b1, b2, b3;
    dx = (double)x / backw - 0.5;
    dy = (double)y / backh - 0.5;
    dx -= (xl = (int)dx);
    dy -= (yl = (int)dy);
    if (xl < 0) { xl = 0; dx -= 1.0; }
    else if (xl >= nx - 1) { xl = nx < 2 ? 0 : nx - 2; dx += 1.0; }
    if (yl < 0) { yl = 0; dy -= 1.0; }
    else if (yl >= ny - 1) { yl = ny < 2 ? 0 : ny - 2; dy += 1.0; }
    pos = yl * nx + xl;
    cdx = 1 - dx;
    b += pos;
    b0 = *b;
    b1 = nx < 2 ? b0 : *(++b);
    b2 = ny < 2 ? *b : *(b += nx);
    b3 = nx < 2 ? *b : *(--b);
    return (float)((1 - dy) * (cdx * b0 + dx * b1) + dy * (dx * b2 + cdx * b3));
}

__device__ void analyse_object(const AnaArgs &A, int off, int n, int obj)
{
    const int W = A.W, H = A.H;
    ObjCols o = A.o;
    /* ---- pass 1: extent, peaks, float flux of the final pixels ---- */
    float peak = (float)-SXG_BIG, cpeak = (float)-SXG_BIG, val, cval, rv = 0.0f;
    int xmin = 2 * SXG_MAXPICSIZE, ymin = 2 * SXG_MAXPICSIZE, xmax = 0, ymax = 0, npx = 0;
    for (int k = 0; k < n; k++) {
        const int p = A.fpia[off + k];
        if (A.alloc[p] != obj) continue;
        const int y = p / W, x = p - y * W;
        if (peak < (val = A.sub[p])) peak = val;
        if (cpeak < (cval = A.cd[p])) cpeak = cval;
        rv += cval;
        if (xmin > x) xmin = x;
        if (xmax < x) xmax = x;
        if (ymin > y) ymin = y;
        if (ymax < y) ymax = y;
        npx++;
    }
    o.fdnpix[obj] = npx;
    int flag = o.flag[obj];
    if (o.level[obj] > 0) {
        /* ---- preanalyse_robust ---- */
        if (xmin == 0 || xmax == W - 1 || ymin == 0 || ymax == H - 1) flag |= SXG_OBJ_TRUNC;
        o.fdflux[obj] = rv;
        o.fdpeak[obj] = cpeak;
        o.dpeak[obj] = peak;
        o.xmin[obj] = xmin; o.xmax[obj] = xmax; o.ymin[obj] = ymin; o.ymax[obj] = ymax;
        const double dthresh = (double)A.dthresh;
        const double dthresh2 = (dthresh + peak) / 2.0;
        double mx = 0.0, my = 0.0, tv = 0.0, mx2 = 0.0, my2 = 0.0, mxy = 0.0;
        int area2 = 0, dnpix = 0;
        for (int k = 0; k < n; k++) {
            const int p = A.fpia[off + k];
            if (A.alloc[p] != obj) continue;
            const int y = p / W - ymin, x = p - (p / W) * W - xmin;
            cval = A.cd[p];
            tv += (val = A.sub[p]);
            if (val > dthresh) dnpix++;
            if (val > dthresh2) area2++;
            mx += cval * x;
            my += cval * y;
            mx2 += cval * x * x;
            my2 += cval * y * y;
            mxy += cval * x * y;
        }
        Moments m;
        sxg_moments(mx, my, mx2, my2, mxy, rv, (flag & SXG_OBJ_MERGED) != 0, o.mx[obj], o.my[obj],
                    xmin, ymin, area2, dnpix, dthresh, dthresh2, m);
        sxg_store_moments(o, obj, m);
        o.dnpix[obj] = dnpix;
        o.dflux[obj] = (float)tv;
    }
    o.dthresh[obj] = A.dthresh;   /* the detection threshold from here on */
    const double mxo = o.mx[obj], myo = o.my[obj];
    const float fdflux = o.fdflux[obj];
    /* ---- examineiso ---- */
    const float bkg = back_at(A.back, (int)(mxo + 0.5), (int)(myo + 0.5), A.nbx, A.nby, A.backw, A.backh);
    const float dbkg = 0.0f;
    double tv = 0.0, sigtv = 0.0, esum = 0.0, emx2 = 0.0, emy2 = 0.0, emxy = 0.0, err;
    const double var = (double)(A.backsig * A.backsig), dbacknoise2 = var;
    const double xm = A.errflag ? mxo : 0.0, ym = A.errflag ? myo : 0.0;
    float pk = (float)-SXG_BIG, cdpeak = (float)-SXG_BIG, pix, cdpix;
    int area = 0, peakx = 0, peaky = 0;
    int cleanflag = A.cleanflag && npx >= HEAPN;
    float top[HEAPN];
    int nfill = 0;
    for (int k = 0; k < n; k++) {
        const int p = A.fpia[off + k];
        if (A.alloc[p] != obj) continue;
        double x = p % W, y = p / W;
        pix = A.sub[p];
        if (pix > pk) pk = pix;
        cdpix = A.cd[p];
        if (A.peakflag && cdpix > cdpeak) {
            cdpeak = cdpix;
            peakx = (int)x + 1;
            peaky = (int)y + 1;
        }
        sigtv += var;
        if (pix > A.thresh) area++;
        tv += pix;
        if (A.errflag) {
            err = dbacknoise2;
            if (A.gain > 0.0 && cdpix > 0.0) err += cdpix / A.gain;
            x = x - xm;
            y = y - ym;
            esum += err;
            emx2 += err * x * x;
            emy2 += err * y * y;
            emxy += err * x * y;
        }
        if (cleanflag) {
            const float tpix = cdpix - A.dthresh;
            if (nfill < HEAPN) {
                int i = nfill++;
                while (i > 0 && top[i - 1] > tpix) { top[i] = top[i - 1]; i--; }
                top[i] = tpix;
            } else if (tpix > top[0]) {
                int i = 0;
                while (i + 1 < HEAPN && top[i + 1] < tpix) { top[i] = top[i + 1]; i++; }
                top[i] = tpix;
            }
        }
    }
    o.mthresh[obj] = cleanflag ? top[0] : 0.0f;
    if (A.errflag) {
        const double flux2 = (float)(fdflux * fdflux);
        o.poserr_mx2[obj] = emx2 / flux2;
        o.poserr_my2[obj] = emy2 / flux2;
        o.poserr_mxy[obj] = emxy / flux2;
    }
    tv -= npx * dbkg;
    if (A.gain > 0.0 && tv > 0.0) sigtv += tv / A.gain;
    o.npix[obj] = area;
    o.flux[obj] = (float)tv;
    o.fluxerr[obj] = (float)sigtv;
    const float thr_o = A.thresh - dbkg, pk_o = pk - dbkg;
    o.thresh[obj] = thr_o;
    o.peak[obj] = pk_o;
    o.peakx[obj] = peakx;
    o.peaky[obj] = peaky;
    o.bkg[obj] = bkg;
    /* ---- isophotal areas and FWHM ---- */
    float threshs[SXG_NISO];
    if (pk_o > 0.0 && thr_o > 0.0)
        for (int i = 0; i < SXG_NISO; i++)
            threshs[i] = (float)(thr_o * sx_pow((double)pk_o / thr_o, (double)i / SXG_NISO));
    else
        for (int i = 0; i < SXG_NISO; i++) threshs[i] = 0.0f;
    int iso[SXG_NISO] = {0, 0, 0, 0, 0, 0, 0, 0};
    float thresh0 = (float)(pk_o / 5.0);
    if (thresh0 < thr_o) thresh0 = thr_o;
    const double sat = (double)(A.satur_level - dbkg);
    double s = 0.0, sx = 0.0, sy = 0.0, sxx = 0.0, sxy = 0.0;
    for (int k = 0; k < n; k++) {
        const int p = A.fpia[off + k];
        if (A.alloc[p] != obj) continue;
        const float v = A.sub[p];
        for (int i = 0; i < SXG_NISO && v > threshs[i]; i++) iso[i]++;
        const double dpix = v - (double)dbkg;
        if (thresh0 > 0.0 && dpix > thresh0 && dpix < sat) {
            const double dx = (p % W) - mxo, dy = (p / W) - myo;
            const double lpix = sx_log(dpix);
            const double inverr2 = dpix * dpix;
            s += inverr2;
            const double d2 = dx * dx + dy * dy;
            sx += d2 * inverr2;
            sxx += d2 * d2 * inverr2;
            sy += lpix * inverr2;
            sxy += lpix * d2 * inverr2;
        }
    }
    if (A.isoflag)
        for (int i = 0; i < SXG_NISO; i++) o.iso[(size_t)obj * SXG_NISO + i] = iso[i];
    float fwhm = 0.0f;
    if (A.fwhmflag && thresh0 > 0.0) {
        const double d = s * sxx - sx * sx;
        if (fabs(d) > 0.0) {
            double bb = -(s * sxy - sx * sy) / d, bmax;
            if (bb < (bmax = 1 / (13 * o.a[obj] * o.b[obj]))) bb = bmax;
            fwhm = (float)(1.6651 / sqrt(bb));
            if (fwhm > 0.0) fwhm -= 1 / (4 * fwhm);
        }
    }
    o.fwhm[obj] = fwhm;
    if (pk_o + bkg >= A.satur_level) flag |= SXG_OBJ_SATUR;
    o.flag[obj] = flag;
}

__global__ void __launch_bounds__(AW) k_analyse(AnaArgs A)
{
    const int r = blockIdx.x;
    const int n = A.cnt[r], off = A.off[r], base = A.objbase[r], nk = A.rootkept[r];
    for (int j = threadIdx.x; j < nk; j += AW) analyse_object(A, off, n, base + j);
}

extern "C" void analyse_gpu(double flagobj_poserr_mx2, int flagobj_peakx, int flagobj_iso_0,
                            float flagobj_fwhm, float backsig, double gain, double ngamma,
                            float thresh, float dthresh, double satur_level, int plistexist_var,
                            int plistoff_var, int prefs_weightgain_flag, int prefs_ext_minarea,
                            int prefs_clean_flag, int detect_type, int nbackx, int nbacky,
                            int backw, int backh, int width, int numobj)
{
    SxgState &S = sxg_state();
    (void)ngamma; (void)plistexist_var; (void)plistoff_var; (void)prefs_weightgain_flag;
    (void)prefs_ext_minarea;
    if (!S.have_frame) {
        sxg_set_error("analyse_gpu: no frame");
        return;
    }
    if (numobj != S.nobj || width != S.W) {
        sxg_set_error("analyse_gpu: numobj/width differ from parcel_out");
        return;
    }
    if (detect_type != 0) {
        sxg_set_error("analyse_gpu: only DETECT_TYPE CCD is supported");
        return;
    }
    if (nbackx != S.nbx || nbacky != S.nby || backw != S.bs || backh != S.bs) {
        sxg_set_error("analyse_gpu: background geometry differs from back_estimate");
        return;
    }
    if (S.nraw == 0 || numobj == 0) return;
    RawObj ro;
    if (!sxg_raw_alloc(S.nraw, ro)) return;
    AnaArgs A;
    A.fpia = S.fpia; A.off = ro.off; A.cnt = ro.count;
    A.objbase = sxg_take<int>("deb.objbase", S.nraw);
    A.rootkept = sxg_take<int>("deb.rootkept", S.nraw);
    A.alloc = sxg_take<int>("deb.alloc", S.N);
    A.sub = S.img.as<float>(); A.cd = S.cd.as<float>(); A.back = S.maps.as<float>();
    A.W = S.W; A.H = S.H; A.nbx = nbackx; A.nby = nbacky; A.backw = backw; A.backh = backh;
    A.backsig = backsig; A.thresh = thresh; A.dthresh = dthresh;
    A.gain = gain; A.satur_level = satur_level;
    /* the FLAG() convention: a boolean in the first byte of the argument */
    A.errflag = *reinterpret_cast<const char *>(&flagobj_poserr_mx2) != 0;
    A.peakflag = *reinterpret_cast<const char *>(&flagobj_peakx) != 0;
    A.isoflag = *reinterpret_cast<const char *>(&flagobj_iso_0) != 0;
    A.fwhmflag = *reinterpret_cast<const char *>(&flagobj_fwhm) != 0;
    A.cleanflag = prefs_clean_flag != 0;
    if (!sxg_obj_view(A.o)) return;
    S.clean_enabled = prefs_clean_flag != 0;
    sxg_stage_begin(ST_ANALYSE);
    sxg_kt_begin(SXG_KT_ANALYSE);
    hipLaunchKernelGGL(k_analyse, dim3(S.nraw), dim3(AW), 0, S.stream, A);
    sxg_kt_end(SXG_KT_ANALYSE);
    sxg_stage_end(ST_ANALYSE);
    SXG_HIP(hipGetLastError());
    SXG_HIP(hipStreamSynchronize(S.stream));
}
