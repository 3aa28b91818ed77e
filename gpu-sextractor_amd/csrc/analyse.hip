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
 * Layout: one 64-lane wavefront per object over its contiguous, raster-ordered range of
 * the object pixel list (opix/objoff, built after deblending).  Lanes compute the
 * per-pixel terms of a chunk of 64 pixels in parallel; extrema, counts and the
 * first-maximum position are wavefront reductions; the floating-point sums are folded in
 * raster order by transposed folds (sxg_wave.h: lane s adds the terms of sum s), so each
 * sum is the reference's sequential sum bit for bit.
 */
#include "sxg_internal.h"
#include "sxg_objects.h"
#include "sxg_wave.h"

#define HEAPN 5

struct AnaArgs {
    const int *opix, *objoff;
    const float *sub, *cd, *back;
    int W, H, nbx, nby, backw, backh, nobj;
    float backsig, thresh, dthresh;
    double gain, satur_level;
    int errflag, peakflag, isoflag, fwhmflag, cleanflag;
    int gainflag;           /* WEIGHT_GAIN with a variance map (cudaanalyse.cu:2281) */
    ObjCols o;
    WeightArgs wa;          /* MAP_WEIGHT: per-pixel variance and thresholds */
    const float *rawt0;     /* detection threshold of each raw object */
    const int *order;       /* objects in launch order: the large ones first */
    const PlanDev *pd;      /* planned run: nobj, backsig, thresh and dthresh on the device */
};

/* objects of at least ANA_BIG pixels are analysed first: their sequential sums are the
   longest, so they start at once and the small objects fill the machine around them */
#define ANA_BIG 256

__global__ void __launch_bounds__(256) k_obj_order(const int *objoff, int nobj, int *cnt,
                                                   int *order, const PlanDev *pd)
{
    const int obj = blockIdx.x * 256 + threadIdx.x;
    if (pd) nobj = pd->nobj;
    if (obj >= nobj) return;
    if (objoff[obj + 1] - objoff[obj] >= ANA_BIG)
        order[atomicAdd(&cnt[0], 1)] = obj;
    else
        order[nobj - 1 - atomicAdd(&cnt[1], 1)] = obj;
}

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

__global__ void __launch_bounds__(256) k_analyse(AnaArgs A)
{
    __shared__ double s_rows[4][11 * SXW_ROW];
    __shared__ float s_frows[4][SXW_ROW];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int slot = blockIdx.x * 4 + wid;
    if (slot >= (A.pd ? A.pd->nobj : A.nobj)) return;
    const int obj = A.order[slot];
    /* the frame's noise and thresholds (from the device in a planned run) */
    const float backsig = A.pd ? A.pd->backsig : A.backsig;
    const float athresh = A.pd ? A.pd->thresh : A.thresh;
    const float adthresh = A.pd ? A.pd->dthresh : A.dthresh;
    double *rows = s_rows[wid];
    float *frows = s_frows[wid];
    const int W = A.W, H = A.H;
    ObjCols o = A.o;
    const int g0 = A.objoff[obj], npx = A.objoff[obj + 1] - g0;
    const int *P = A.opix + g0;
    int flag = o.flag[obj];
    /* the object's detection threshold: with per-pixel thresholds, the lowest among the
       pixels of its raw object */
    const bool perpix = A.wa.w && A.wa.rel;
    const float odth = perpix ? A.rawt0[o.root[obj]] : adthresh;
    double mxo, myo;
    float fdflux, oa, ob;
    if (o.level[obj] > 0) {
        /* ---- preanalyse_robust over the final pixels ----
           extent and peaks first (reductions), then one fold pass for the float flux and
           the six double moment sums (independent chains, one lane each) */
        float peak = (float)-SXG_BIG, cpeak = (float)-SXG_BIG;
        int xmin = 2 * SXG_MAXPICSIZE, ymin = 2 * SXG_MAXPICSIZE, xmax = 0, ymax = 0;
#pragma unroll 4
        for (int k = lane; k < npx; k += 64) {
            const int p = P[k];
            const int y = p / W, x = p - y * W;
            peak = fmaxf(peak, A.sub[p]);
            cpeak = fmaxf(cpeak, A.cd[p]);
            xmin = min(xmin, x);
            xmax = max(xmax, x);
            ymin = min(ymin, y);
            ymax = max(ymax, y);
        }
        peak = sxw_max(peak);
        cpeak = sxw_max(cpeak);
        xmin = sxw_min(xmin);
        xmax = sxw_max(xmax);
        ymin = sxw_min(ymin);
        ymax = sxw_max(ymax);
        if (xmin == 0 || xmax == W - 1 || ymin == 0 || ymax == H - 1) flag |= SXG_OBJ_TRUNC;
        const double dthresh = (double)odth;
        const double dthresh2 = (dthresh + peak) / 2.0;
        double acc = 0.0;
        float racc = 0.0f;
        int area2 = 0, dnpix = 0;
        /* the next chunk's pixels are read while this chunk is folded */
        int np_ = lane < npx ? P[lane] : 0;
        float nsub = lane < npx ? A.sub[np_] : 0.0f, ncd = lane < npx ? A.cd[np_] : 0.0f;
        for (int base = 0; base < npx; base += 64) {
            const int k = base + lane;
            const int pc = np_;
            const float csub = nsub, ccd = ncd;
            if (k + 64 < npx) {
                np_ = P[k + 64];
                nsub = A.sub[np_];
                ncd = A.cd[np_];
            }
            float val = 0.0f, cval = 0.0f, tx = 0.0f, ty = 0.0f, txx = 0.0f, tyy = 0.0f,
                  txy = 0.0f;
            if (k < npx) {
                const int p = pc;
                const int y = p / W - ymin, x = p - (p / W) * W - xmin;
                cval = ccd;
                val = csub;
                if (val > dthresh) dnpix++;
                if (val > dthresh2) area2++;
                tx = cval * x;
                ty = cval * y;
                txx = cval * x * x;
                tyy = cval * y * y;
                txy = cval * x * y;
            }
            const double t6[6] = {val, tx, ty, txx, tyy, txy};
            const float tf[1] = {cval};
            sxw_fold2<6, 1>(rows, frows, t6, tf, k < npx, acc, racc);
        }
        const float rv = sxw_rl(racc, 6);
        const double tv = sxw_rl(acc, 0), mx = sxw_rl(acc, 1), my = sxw_rl(acc, 2);
        const double mx2 = sxw_rl(acc, 3), my2 = sxw_rl(acc, 4), mxy = sxw_rl(acc, 5);
        area2 = sxw_sum(area2);
        dnpix = sxw_sum(dnpix);
        Moments m;
        sxg_moments(mx, my, mx2, my2, mxy, rv, (flag & SXG_OBJ_MERGED) != 0, o.mx[obj], o.my[obj],
                    xmin, ymin, area2, dnpix, dthresh, dthresh2, m);
        mxo = m.mx;
        myo = m.my;
        fdflux = rv;
        oa = m.a;
        ob = m.b;
        if (lane == 0) {
            o.fdflux[obj] = rv;
            o.fdpeak[obj] = cpeak;
            o.dpeak[obj] = peak;
            o.xmin[obj] = xmin; o.xmax[obj] = xmax; o.ymin[obj] = ymin; o.ymax[obj] = ymax;
            sxg_store_moments(o, obj, m);
            o.dnpix[obj] = dnpix;
            o.dflux[obj] = (float)tv;
        }
    } else {
        mxo = o.mx[obj];
        myo = o.my[obj];
        fdflux = o.fdflux[obj];
        oa = o.a[obj];
        ob = o.b[obj];
    }
    /* ---- examineiso ----
       reductions first (peak, its position, the analysis threshold, ISOAREA, mthresh),
       then one fold pass for the six isophotal sums and the five FWHM sums (a pixel
       outside the FWHM selection adds +0, an identity: those sums never hold -0) */
    const float bkg = back_at(A.back, (int)(mxo + 0.5), (int)(myo + 0.5), A.nbx, A.nby, A.backw,
                              A.backh);
    const float dbkg = 0.0f;
    const double var = (double)(backsig * backsig), dbacknoise2 = var;
    const double xm = A.errflag ? mxo : 0.0, ym = A.errflag ? myo : 0.0;
    float pk = (float)-SXG_BIG, cdpeak = (float)-SXG_BIG, othr = (float)SXG_BIG;
    int area = 0, peakp = 0x7fffffff;
    const bool cleanflag = A.cleanflag && npx >= HEAPN;
    SxwTop5 top;
    top.init();
#pragma unroll 4
    for (int k = lane; k < npx; k += 64) {
        const int p = P[k];
        const float pix = A.sub[p];
        pk = fmaxf(pk, pix);
        const float cdpix = A.cd[p];
        if (A.peakflag && cdpix > cdpeak) {
            cdpeak = cdpix;
            peakp = p;
        }
        const float varp = A.wa.w ? sxg_var(A.wa, p) : 0.0f;
        const float athr = perpix ? sxg_pix_thresh(A.wa.kana, varp) : athresh;
        othr = fminf(othr, athr);
        if (pix > athr) area++;
        if (cleanflag) top.add(cdpix - odth);
    }
    pk = sxw_max(pk);
    othr = sxw_minf(othr);
    area = sxw_sum(area);
    int peakx = 0, peaky = 0;
    if (A.peakflag) {
        const float cmax = sxw_max(cdpeak);
        const int pp = sxw_min(cdpeak == cmax ? peakp : 0x7fffffff);
        peakx = pp % W + 1;
        peaky = pp / W + 1;
    }
    const float mthresh = cleanflag ? top.fifth() : 0.0f;
    const float thr_o = othr - dbkg, pk_o = pk - dbkg;
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
    double eacc = 0.0;
    /* the next chunk's pixels are read while this chunk is folded */
    int np_ = lane < npx ? P[lane] : 0;
    float nsub = lane < npx ? A.sub[np_] : 0.0f, ncd = lane < npx ? A.cd[np_] : 0.0f;
    for (int base = 0; base < npx; base += 64) {
        const int k = base + lane;
        const int pc = np_;
        const float csub = nsub, ccd = ncd;
        if (k + 64 < npx) {
            np_ = P[k + 64];
            nsub = A.sub[np_];
            ncd = A.cd[np_];
        }
        float pix = 0.0f;
        double err = 0.0, ex2 = 0.0, ey2 = 0.0, exy = 0.0, var2 = 0.0;
        double ts = 0.0, tsx = 0.0, tsxx = 0.0, tsy = 0.0, tsxy = 0.0;
        if (k < npx) {
            const int p = pc;
            double x = p % W, y = p / W;
            pix = csub;
            const float cdpix = ccd;
            /* per-pixel variance with a weight map */
            const float varp = A.wa.w ? sxg_var(A.wa, p) : 0.0f;
            var2 = A.wa.w ? (double)varp : var;
            err = A.wa.w ? var2 : dbacknoise2;
            /* var is the background noise squared */
            if (A.gainflag && pix > 0.0 && A.gain > 0.0) var2 += pix / A.gain * varp / var;
            if (A.gain > 0.0 && cdpix > 0.0) err += cdpix / A.gain;
            x = x - xm;
            y = y - ym;
            ex2 = err * x * x;
            ey2 = err * y * y;
            exy = err * x * y;
            bool go = true;
#pragma unroll
            for (int i = 0; i < SXG_NISO; i++) {
                go = go && pix > threshs[i];
                iso[i] += go;
            }
            const double dpix = pix - (double)dbkg;
            if (thresh0 > 0.0 && dpix > thresh0 && dpix < sat) {
                const double dx = (p % W) - mxo, dy = (p / W) - myo;
                const double lpix = sx_log(dpix);
                const double inverr2 = dpix * dpix;
                const double d2 = dx * dx + dy * dy;
                ts = inverr2;
                tsx = d2 * inverr2;
                tsxx = d2 * d2 * inverr2;
                tsy = lpix * inverr2;
                tsxy = lpix * d2 * inverr2;
            }
        }
        /* tv, sigtv, esum, emx2, emy2, emxy (the error sums only with errflag), then the
           FWHM fit sums */
        const double t11[11] = {pix, var2, err, ex2, ey2, exy, ts, tsx, tsxx, tsy, tsxy};
        sxw_fold<11, double>(rows, t11, k < npx, eacc);
    }
    double tv = sxw_rl(eacc, 0), sigtv = sxw_rl(eacc, 1);
    const double emx2 = sxw_rl(eacc, 3), emy2 = sxw_rl(eacc, 4), emxy = sxw_rl(eacc, 5);
    double pmx2 = 0.0, pmy2 = 0.0, pmxy = 0.0;
    if (A.errflag) {
        const double flux2 = (float)(fdflux * fdflux);
        pmx2 = emx2 / flux2;
        pmy2 = emy2 / flux2;
        pmxy = emxy / flux2;
    }
    tv -= npx * dbkg;
    if (!A.gainflag && A.gain > 0.0 && tv > 0.0) sigtv += tv / A.gain;
    const double s = sxw_rl(eacc, 6), sx = sxw_rl(eacc, 7), sxx = sxw_rl(eacc, 8);
    const double sy = sxw_rl(eacc, 9), sxy = sxw_rl(eacc, 10);
#pragma unroll
    for (int i = 0; i < SXG_NISO; i++) iso[i] = sxw_sum(iso[i]);
    float fwhm = 0.0f;
    if (A.fwhmflag && thresh0 > 0.0) {
        const double d = s * sxx - sx * sx;
        if (fabs(d) > 0.0) {
            double bb = -(s * sxy - sx * sy) / d, bmax;
            if (bb < (bmax = 1 / (13 * oa * ob))) bb = bmax;
            fwhm = (float)(1.6651 / sqrt(bb));
            if (fwhm > 0.0) fwhm -= 1 / (4 * fwhm);
        }
    }
    if (pk_o + bkg >= A.satur_level) flag |= SXG_OBJ_SATUR;
    if (lane != 0) return;
    o.fdnpix[obj] = npx;
    o.dthresh[obj] = odth;   /* the detection threshold from here on */
    o.mthresh[obj] = mthresh;
    if (A.errflag) {
        o.poserr_mx2[obj] = pmx2;
        o.poserr_my2[obj] = pmy2;
        o.poserr_mxy[obj] = pmxy;
    }
    o.npix[obj] = area;
    o.flux[obj] = (float)tv;
    o.fluxerr[obj] = (float)sigtv;
    o.thresh[obj] = thr_o;
    o.peak[obj] = pk_o;
    o.peakx[obj] = peakx;
    o.peaky[obj] = peaky;
    o.bkg[obj] = bkg;
    if (A.isoflag)
        for (int i = 0; i < SXG_NISO; i++) o.iso[(size_t)obj * SXG_NISO + i] = iso[i];
    o.fwhm[obj] = fwhm;
    o.flag[obj] = flag;
}

__global__ void k_ana_zero2(int *p)
{
    if (threadIdx.x < 2) p[threadIdx.x] = 0;
}

extern "C" void analyse_gpu(double flagobj_poserr_mx2, int flagobj_peakx, int flagobj_iso_0,
                            float flagobj_fwhm, float backsig, double gain, double ngamma,
                            float thresh, float dthresh, double satur_level, int plistexist_var,
                            int plistoff_var, int prefs_weightgain_flag, int prefs_ext_minarea,
                            int prefs_clean_flag, int detect_type, int nbackx, int nbacky,
                            int backw, int backh, int width, int numobj)
{
    SxgState &S = sxg_state();
    (void)ngamma; (void)plistexist_var; (void)plistoff_var; 
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
    if (!S.opix || !S.objoff) {
        sxg_set_error("analyse_gpu: no object pixel list (run parcel_out first)");
        return;
    }
    AnaArgs A;
    A.opix = S.opix; A.objoff = S.objoff; A.nobj = numobj;
    A.sub = S.img.as<float>(); A.cd = S.cd.as<float>(); A.back = sxg_back_map();
    A.W = S.W; A.H = S.H; A.nbx = nbackx; A.nby = nbacky; A.backw = backw; A.backh = backh;
    A.backsig = backsig; A.thresh = thresh; A.dthresh = dthresh;
    A.gain = gain; A.satur_level = satur_level;
    /* the FLAG() convention: a boolean in the first byte of the argument */
    A.errflag = *reinterpret_cast<const char *>(&flagobj_poserr_mx2) != 0;
    A.peakflag = *reinterpret_cast<const char *>(&flagobj_peakx) != 0;
    A.isoflag = *reinterpret_cast<const char *>(&flagobj_iso_0) != 0;
    A.fwhmflag = *reinterpret_cast<const char *>(&flagobj_fwhm) != 0;
    A.cleanflag = prefs_clean_flag != 0;
    if (!sxg_weight_args(A.wa)) return;
    A.gainflag = A.wa.w && prefs_weightgain_flag;
    RawObj ro;
    if (!sxg_raw_alloc(S.nraw, ro)) return;
    A.rawt0 = ro.t0;
    if (!sxg_obj_view(A.o)) return;
    S.clean_enabled = prefs_clean_flag != 0;
    int *order = sxg_take<int>("ana.order", (size_t)numobj + 2);
    if (!order) return;
    A.order = order;
    /* a planned run: numobj is a capacity, the count and thresholds are on the device */
    A.pd = S.plan ? S.pdev : nullptr;
    sxg_stage_begin(ST_ANALYSE);
    /* the two counters of k_obj_order (a fill kernel: no memset node in a graph) */
    hipLaunchKernelGGL(k_ana_zero2, dim3(1), dim3(64), 0, S.stream, order + numobj);
    hipLaunchKernelGGL(k_obj_order, dim3(sxg_div_up(numobj, 256)), dim3(256), 0, S.stream,
                       S.objoff, numobj, order + numobj, order, A.pd);
    sxg_kt_begin(SXG_KT_ANALYSE);
    hipLaunchKernelGGL(k_analyse, dim3(sxg_div_up(numobj, 4)), dim3(256), 0, S.stream, A);
    sxg_kt_end(SXG_KT_ANALYSE);
    sxg_stage_end(ST_ANALYSE);
    /* no host output: the next entry points run on the same stream */
    SXG_HIP(hipGetLastError());
    sxg_stage_wait();
}
