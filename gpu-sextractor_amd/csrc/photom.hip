/*
 * photom.hip -- catalogue photometry and rows (endobject_gpu, the reference's
 * src/cuda/cudaanalyse.cu:3344), run after CLEAN.
 *
 * Semantics (cudaanalyse.cu):
 *   - labels for the photometry (updateAllocateLabel, :3320-3342): each pixel carries the
 *     object it was allocated to, -1, or INT_MAX for the pixels of objects that CLEAN
 *     removed (label > 0 only, as in the reference) and for detected pixels of no object
 *     (cd >= dthresh);
 *   - ISOCOR (:62-98); APER in one circular aperture with 5x5 oversampling of the pixels
 *     that straddle its edge (:104-287, NAPER 1); AUTO: the Kron factor from the first
 *     moment inside 6 isophotal radii, then the flux inside the Kron ellipse (:293-590);
 *     pixels of other objects replaced by their mirror through the centroid when
 *     MASK_TYPE is CORRECT, else zero; the OBJ_APERT_PB and OBJ_CROWDED flags;
 *   - CLASS_STAR: the star/galaxy network on the isophotal areas, the peak and the
 *     seeing (:599-636, :2903-2921); ELONGATION, ELLIPTICITY, FLUXERR_ISO, MAG_ISO,
 *     MAGERR_ISO (:2778-2792, :638-653); the SATUR flag (:2873-2874);
 *   - one catalogue row per object in the outobjstruct layout (:3049-3095).
 * Expressions keep the reference's operand types and order (float where the CUDA code
 * computes in float: sqrt and log10 of a float are the float functions there).  The
 * transcendental functions are the canonical ones of sx_cmath.h.
 *
 * Layout: one 64-lane wavefront per object, objects with a large bounding box first.
 * The aperture boxes are walked in raster order 64 pixels at a time: the lanes compute
 * the per-pixel terms (oversampled area, masking and mirror, PHOTO-mode exponential), and
 * the sums are transposed ordered folds (sxg_wave.h), so every sum is the reference's
 * sequential raster-order sum bit for bit.  Counts are ballots.  The network runs one
 * neuron per lane.
 */
#include <climits>
#include <cstring>

#include "sxg_internal.h"
#include "sxg_objects.h"
#include "sxg_wave.h"

#define PH_CROWDED 0x0001
#define PH_APERT_PB 0x0010
#define PH_OVERSAMP 5                 /* APER_OVERSAMP, src/define.h:172 */
#define PH_KRON_NSIG (3 * 2.0)        /* KRON_NSIG = 3*MARGIN_SCALE, src/define.h:173 */
#define PH_CROWD 0.1                  /* CROWD_THRESHOLD, src/define.h:177 */
#define PH_BIG_BOX 256                /* objects with a larger bounding box go first */

struct PhotCols {
    float *flux_iso, *fluxerr_iso, *mag_iso, *magerr_iso, *flux_isocor, *fluxerr_isocor,
        *flux_aper, *fluxerr_aper, *mag_aper, *magerr_aper, *kron, *flux_auto, *fluxerr_auto,
        *mag_auto, *magerr_auto, *elong, *ellip, *sprob;
};

struct PhotArgs {
    const float *sub;          /* background-subtracted frame */
    const int *lab;            /* photometry labels */
    const int *order;          /* objects in launch order */
    int W, H, nobj;
    ObjCols o;
    PhotCols p;
    float backsig, thresh;
    double apert, autoparam0, autoparam1, autoaper0, autoaper1, seeing, pixscale, zp, ngamma,
        gain, satur;
    int corr, photo, do_aper, do_auto, do_class;
    const sxg_brain *brain;    /* device copy, or nullptr */
    int *npix;                 /* per object: aperture pixels summed */
};

/* ------------------------------------------------------------------ labels ---- */
/* four pixels per lane with 16-byte accesses; the n % 4 tail pixels by the first lanes */
__global__ void __launch_bounds__(256) k_phot_lab_init(const float *__restrict__ cd, int *lab,
                                                       long n, float dthresh)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    const long n4 = n >> 2;
    if (q < n4) {
        const float4 v = reinterpret_cast<const float4 *>(cd)[q];
        int4 o;
        o.x = v.x >= dthresh ? INT_MAX : -1;
        o.y = v.y >= dthresh ? INT_MAX : -1;
        o.z = v.z >= dthresh ? INT_MAX : -1;
        o.w = v.w >= dthresh ? INT_MAX : -1;
        reinterpret_cast<int4 *>(lab)[q] = o;
    }
    if (q < (n & 3)) {
        const long p = (n4 << 2) + q;
        lab[p] = cd[p] >= dthresh ? INT_MAX : -1;
    }
}

__global__ void __launch_bounds__(256) k_phot_lab_scatter(const int *__restrict__ fpia, long P0,
                                                          const int *__restrict__ alloc,
                                                          const unsigned *__restrict__ master,
                                                          int *lab)
{
    const long g = (long)blockIdx.x * 256 + threadIdx.x;
    if (g >= P0) return;
    const int p = fpia[g];
    const int t = alloc[p];
    lab[p] = (t > 0 && master[t] != 0xffffffffu) ? INT_MAX : t;
}

__global__ void __launch_bounds__(256) k_phot_order(ObjCols o, int nobj, int *cnt, int *order)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nobj) return;
    const long box = (long)(o.xmax[i] - o.xmin[i] + 1) * (o.ymax[i] - o.ymin[i] + 1);
    if (box >= PH_BIG_BOX)
        order[atomicAdd(&cnt[0], 1)] = i;
    else
        order[nobj - 1 - atomicAdd(&cnt[1], 1)] = i;
}

/* ---------------------------------------------------------------- per object ---- */
/* the value used for pixel (x, y) of object obj's aperture: its own, or, when the pixel
   belongs to another object, the mirror pixel through the centroid if that one is free
   or the object's own (MASK_TYPE CORRECT), else 0 (cudaanalyse.cu:222-243, :523-545) */
__device__ inline double ph_value(const PhotArgs &A, int obj, double mx, double my, int x,
                                  int y, double pix, bool &masked)
{
    int t = A.lab[(long)y * A.W + x];
    masked = t >= 0 && t != obj;
    if (!masked) return pix;
    int x2, y2;
    if (A.corr && (x2 = (int)(2 * mx + 0.49999 - x)) >= 0 && x2 < A.W &&
        (y2 = (int)(2 * my + 0.49999 - y)) >= 0 && y2 < A.H &&
        ((t = A.lab[(long)y2 * A.W + x2]) == obj || t == -1))
        return A.sub[(long)y2 * A.W + x2];
    return 0.0;
}

__device__ inline int ph_rint(double x) { return (int)floor(x + 0.5); }   /* define.h:262 */

__device__ inline void ph_box(double mx, double my, float dxlim, float dylim, int W, int H,
                              int &flag, int &xmin, int &xmax, int &ymin, int &ymax)
{
    if ((xmin = ph_rint(mx - dxlim)) < 0) { xmin = 0; flag |= PH_APERT_PB; }
    if ((xmax = ph_rint(mx + dxlim) + 1) > W) { xmax = W; flag |= PH_APERT_PB; }
    if ((ymin = ph_rint(my - dylim)) < 0) { ymin = 0; flag |= PH_APERT_PB; }
    if ((ymax = ph_rint(my + dylim) + 1) > H) { ymax = H; flag |= PH_APERT_PB; }
}

/* the ellipse (or circle) of an AUTO pass: coefficients, half extents and limit
   (cudaanalyse.cu:354-372, :440-460) */
__device__ inline void ph_ellipse(const PhotArgs &A, int obj, bool use_ellipse, double circ_diam,
                                  float &cx2, float &cy2, float &cxy, float &dxlim, float &dylim)
{
    if (use_ellipse) {
        cx2 = A.o.cxx[obj];
        cy2 = A.o.cyy[obj];
        cxy = A.o.cxy[obj];
        dxlim = (float)(cx2 - cxy * cxy / (4.0 * cy2));
        dylim = (float)(cy2 - cxy * cxy / (4.0 * cx2));
    } else {
        cx2 = cy2 = 1.0f;
        cxy = 0.0f;
        dxlim = dylim = (float)(circ_diam / 2.0);
    }
}

__global__ void __launch_bounds__(256) k_endobject(PhotArgs A)
{
    __shared__ double s_rows[4][3 * SXW_ROW];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gw = blockIdx.x * 4 + wid;
    if (gw >= A.nobj) return;            /* whole wavefronts only: no workgroup barrier */
    const int obj = A.order[gw];
    double *rows = s_rows[wid];
    const ObjCols &o = A.o;
    const int W = A.W, H = A.H;
    const double mx = o.mx[obj], my = o.my[obj];
    const float a = o.a[obj], b = o.b[obj];
    int flag = o.flag[obj];
    double var = A.backsig * A.backsig;   /* float product, as the reference forms it */
    int visited = 0;                      /* aperture pixels summed (wave-uniform) */

    /* ---- APER (cudaanalyse.cu:104-287) */
    float flux_aper = 0.0f, fluxerr_aper = 0.0f, mag_aper = 0.0f, magerr_aper = 0.0f;
    if (A.do_aper) {
        const float raper = (float)(A.apert / 2.0);
        const float raper2 = raper * raper;
        const float rintlim = (float)(raper - 0.75);
        const float rintlim2 = (rintlim > 0.0) ? rintlim * rintlim : 0.0f;
        const float rextlim2 = (float)((raper + 0.75) * (raper + 0.75));
        const float scale = (float)(1.0 / PH_OVERSAMP);
        const float scale2 = scale * scale;
        const float offset = (float)(0.5 * (scale - 1.0));
        int xmin = (int)(mx - raper + 0.499999), xmax = (int)(mx + raper + 1.499999);
        int ymin = (int)(my - raper + 0.499999), ymax = (int)(my + raper + 1.499999);
        if (xmin < 0) { xmin = 0; flag |= PH_APERT_PB; }
        if (xmax > W) { xmax = W; flag |= PH_APERT_PB; }
        if (ymin < 0) { ymin = 0; flag |= PH_APERT_PB; }
        if (ymax > H) { ymax = H; flag |= PH_APERT_PB; }
        const int bw = xmax - xmin, bh = ymax - ymin;
        const int nb = (bw > 0 && bh > 0) ? bw * bh : 0;
        double acc = 0.0;   /* lane 0: area, 1: tv, 2: sigtv */
        for (int k0 = 0; k0 < nb; k0 += 64) {
            const int k = k0 + lane;
            bool member = false;
            double t[3] = {0.0, 0.0, 0.0};
            if (k < nb) {
                const int y = ymin + k / bw, x = xmin + k % bw;
                float dx = (float)(x - mx), dy = (float)(y - my);
                const float r2 = dx * dx + dy * dy;
                if (r2 < rextlim2) {
                    member = true;
                    float locarea;
                    if (r2 > rintlim2) {
                        dx += offset;
                        dy += offset;
                        locarea = 0.0f;
                        for (int sy = PH_OVERSAMP; sy--; dy += scale) {
                            float dx1 = dx;
                            const float dy2 = dy * dy;
                            for (int sx = PH_OVERSAMP; sx--; dx1 += scale)
                                if (dx1 * dx1 + dy2 < raper2) locarea += scale2;
                        }
                    } else {
                        locarea = 1.0f;
                    }
                    bool masked;
                    double pix = ph_value(A, obj, mx, my, x, y, (double)A.sub[(long)y * W + x],
                                          masked);
                    double sig;
                    if (A.photo) {
                        pix = sx_exp(pix / A.ngamma);
                        sig = var * locarea * pix * pix;
                    } else {
                        sig = var * locarea;
                    }
                    t[0] = locarea;
                    t[1] = locarea * pix;
                    t[2] = sig;
                }
            }
            visited += __popcll(__ballot(member));
            sxw_fold<3, double>(rows, t, member, acc);
        }
        const double area = sxw_rl(acc, 0);
        double tv = sxw_rl(acc, 1), sigtv = sxw_rl(acc, 2);
        if (A.photo) {
            tv = A.ngamma * (tv - area * sx_exp(0.0 / A.ngamma));   /* dbkg = 0 (:2244) */
            sigtv /= A.ngamma * A.ngamma;
        } else {
            tv -= area * 0.0;
            if (A.gain > 0.0 && tv > 0.0) sigtv += tv / A.gain;
        }
        flux_aper = (float)tv;
        fluxerr_aper = (float)sqrt(sigtv);
        mag_aper = tv > 0.0 ? (float)(-2.5 * sx_log10(tv) + A.zp) : 99.0f;
        magerr_aper = tv > 0.0 ? (float)(1.086 * sqrt(sigtv) / tv) : 99.0f;
    }

    /* ---- AUTO (cudaanalyse.cu:293-590) */
    float kf = 0.0f, flux_auto = 0.0f, fluxerr_auto = 0.0f, mag_auto = 0.0f, magerr_auto = 0.0f;
    if (A.do_auto) {
        float cx2, cy2, cxy, klim2, dxlim, dylim;
        int xmin, xmax, ymin, ymax;
        const bool ell1 = PH_KRON_NSIG * sqrtf(a * b) > A.autoaper0 / 2.0;
        ph_ellipse(A, obj, ell1, A.autoaper0, cx2, cy2, cxy, dxlim, dylim);
        if (ell1) {
            dxlim = dxlim > 0.0 ? (float)(PH_KRON_NSIG / sqrtf(dxlim)) : 0.0f;
            dylim = dylim > 0.0 ? (float)(PH_KRON_NSIG / sqrtf(dylim)) : 0.0f;
            klim2 = (float)(PH_KRON_NSIG * PH_KRON_NSIG);
        } else {
            klim2 = dxlim * dxlim;
        }
        ph_box(mx, my, dxlim, dylim, W, H, flag, xmin, xmax, ymin, ymax);
        int bw = xmax - xmin, bh = ymax - ymin;
        int nb = (bw > 0 && bh > 0) ? bw * bh : 0;
        int area = 0, areab = 0;
        double acc = 0.0;   /* lane 0: r1, 1: v1 */
        for (int k0 = 0; k0 < nb; k0 += 64) {
            const int k = k0 + lane;
            bool inside = false, good = false;
            double t[2] = {0.0, 0.0};
            if (k < nb) {
                const int y = ymin + k / bw, x = xmin + k % bw;
                const float dx = (float)(x - mx), dy = (float)(y - my);
                const float r2 = cx2 * dx * dx + cy2 * dy * dy + cxy * dx * dy;
                if (r2 <= klim2) {
                    inside = true;
                    const float pix = A.sub[(long)y * W + x];
                    const int tl = A.lab[(long)y * W + x];
                    good = (tl == obj || tl == -1) && pix > -SXG_BIG;
                    if (good) {
                        t[0] = sqrtf(r2) * pix;
                        t[1] = pix;
                    }
                }
            }
            const unsigned long long bi = __ballot(inside), bg = __ballot(good);
            area += __popcll(bg);
            areab += __popcll(bi & ~bg);
            visited += __popcll(bi);
            sxw_fold<2, double>(rows, t, good, acc);
        }
        const double r1 = sxw_rl(acc, 0), v1 = sxw_rl(acc, 1);
        area += areab;
        double tv = 0.0, sigtv = 0.0;
        if (area) {
            if (r1 > 0.0 && v1 > 0.0) {
                kf = (float)(A.autoparam0 * r1 / v1);
                if (kf < A.autoparam1) kf = (float)A.autoparam1;
            } else {
                kf = (float)A.autoparam1;
            }
            if ((float)areab / area > PH_CROWD) flag |= PH_CROWDED;
            const bool ell2 = kf * sqrtf(a * b) > A.autoaper1 / 2.0;
            ph_ellipse(A, obj, ell2, A.autoaper1, cx2, cy2, cxy, dxlim, dylim);
            if (ell2) {
                dxlim = dxlim > 0.0 ? kf / sqrtf(dxlim) : 0.0f;
                dylim = dylim > 0.0 ? kf / sqrtf(dylim) : 0.0f;
                klim2 = kf * kf;
            } else {
                klim2 = dxlim * dxlim;
                kf = 0.0f;
            }
            ph_box(mx, my, dxlim, dylim, W, H, flag, xmin, xmax, ymin, ymax);
            bw = xmax - xmin;
            bh = ymax - ymin;
            nb = (bw > 0 && bh > 0) ? bw * bh : 0;
            area = areab = 0;
            acc = 0.0;      /* lane 0: tv, 1: sigtv */
            for (int k0 = 0; k0 < nb; k0 += 64) {
                const int k = k0 + lane;
                bool inside = false, masked = false;
                double t[2] = {0.0, 0.0};
                if (k < nb) {
                    const int y = ymin + k / bw, x = xmin + k % bw;
                    const float dx = (float)(x - mx), dy = (float)(y - my);
                    if (cx2 * dx * dx + cy2 * dy * dy + cxy * dx * dy <= klim2) {
                        inside = true;
                        float pix = (float)ph_value(A, obj, mx, my, x, y,
                                                    (double)A.sub[(long)y * W + x], masked);
                        double sig;
                        if (A.photo) {
                            pix = (float)sx_exp(pix / A.ngamma);
                            sig = var * pix * pix;
                        } else {
                            sig = var;
                        }
                        t[0] = pix;
                        t[1] = sig;
                    }
                }
                const unsigned long long bi = __ballot(inside), bm = __ballot(inside && masked);
                area += __popcll(bi);
                areab += __popcll(bm);
                visited += __popcll(bi);
                sxw_fold<2, double>(rows, t, inside, acc);
            }
            tv = sxw_rl(acc, 0);
            sigtv = sxw_rl(acc, 1);
            if ((float)areab > PH_CROWD * area) flag |= PH_CROWDED;
            if (A.photo) {
                tv = A.ngamma * (tv - area * sx_exp(0.0 / A.ngamma));
                sigtv /= A.ngamma * A.ngamma;
            } else {
                tv -= area * 0.0;
                if (A.gain > 0.0 && tv > 0.0) sigtv += tv / A.gain;
            }
        }
        flux_auto = (float)tv;
        fluxerr_auto = (float)sqrt(sigtv);
        mag_auto = flux_auto > 0.0 ? (float)(-2.5 * sx_log10f(flux_auto) + A.zp) : 99.0f;
        magerr_auto = flux_auto > 0.0 ? (float)(1.086 * fluxerr_auto / flux_auto) : 99.0f;
        if (tv <= 0.0) kf = 0.0f;
    }

    /* ---- CLASS_STAR (cudaanalyse.cu:2903-2921, :599-636): lane i holds input i, then
       lane j neuron j of each layer */
    float sprob = 0.0f;
    if (A.do_class && A.brain) {
        const sxg_brain *br = A.brain;
        const double fwhm = A.seeing;
        double fac2 = fwhm / A.pixscale;
        fac2 *= fac2;
        const int *iso = o.iso + (size_t)obj * SXG_NISO;
        double in = 0.0;
        if (lane == 0) {
            in = sx_log10(iso[0] ? iso[0] / fac2 : 0.01);
        } else if (lane == 1) {
            const float pk = o.peak[obj];
            in = A.thresh > 0.0 ? sx_log10(pk > 0.0 ? (double)(pk / A.thresh) : 0.1) : -1.0;
        } else if (lane < 9) {
            in = sx_log10(iso[lane - 1] ? iso[lane - 1] / fac2 : 0.01);
        } else if (lane == 9) {
            in = sx_log10(fwhm);
        }
        double cur = lane < br->nn[0] ? in * br->inscale[lane] + br->inbias[lane] : 0.0;
        const int last = br->layersnb - 1;
        for (int l = 0; l < last; l++) {
            const bool mine = lane < br->nn[l + 1];
            double s = mine ? br->b[l][lane] : 0.0;
            for (int i = 0; i < br->nn[l]; i++) {
                const double ni = sxw_rl(cur, i);
                if (mine) s += br->w[l][i][lane] * ni;
            }
            cur = mine ? 1.0 / (1.0 + sx_exp(-s)) : 0.0;
        }
        sprob = (float)((sxw_rl(cur, 0) - br->outbias[0]) / br->outscale[0]);
    }

    if (lane == 0) {
        const PhotCols &p = A.p;
        p.elong[obj] = a / b;
        p.ellip[obj] = 1 - b / a;
        const float fl = o.flux[obj], fe = o.fluxerr[obj];
        const float fei = sqrtf(fe);
        p.flux_iso[obj] = fl;
        p.fluxerr_iso[obj] = fei;
        /* ISOCOR (cudaanalyse.cu:62-98) */
        const unsigned fdnpix = (unsigned)o.fdnpix[obj];
        double ati = (fl > 0.0) ? (double)((float)fdnpix * o.dthresh[obj] / fl) : 0.0;
        if (ati > 1.0)
            ati = 1.0;
        else if (ati < 0.0)
            ati = 0.0;
        p.flux_isocor[obj] = (float)(fl / (1.0 - 0.196099 * ati - 0.751208 * ati * ati));
        if (fl > 0.0) {
            const double sigtv = (double)(fe / (fl * fl));
            double dati = fdnpix ? ati * sqrt(sigtv + 1.0 / fdnpix) : 0.0;
            dati = 0.196099 * dati + 0.751208 * 2 * ati * dati;
            p.fluxerr_isocor[obj] = (float)(sqrt(sigtv + dati * dati) * fl);
        } else {
            p.fluxerr_isocor[obj] = fei;
        }
        p.flux_aper[obj] = flux_aper;
        p.fluxerr_aper[obj] = fluxerr_aper;
        p.mag_aper[obj] = mag_aper;
        p.magerr_aper[obj] = magerr_aper;
        p.kron[obj] = kf;
        p.flux_auto[obj] = flux_auto;
        p.fluxerr_auto[obj] = fluxerr_auto;
        p.mag_auto[obj] = mag_auto;
        p.magerr_auto[obj] = magerr_auto;
        if (o.peak[obj] + o.bkg[obj] >= A.satur) flag |= SXG_OBJ_SATUR;
        p.sprob[obj] = sprob;
        p.mag_iso[obj] = fl > 0.0 ? (float)(-2.5 * sx_log10f(fl) + A.zp) : 99.0f;
        p.magerr_iso[obj] = fl > 0.0 ? (float)(1.086 * fei / fl) : 99.0f;
        o.flag[obj] = flag;
        A.npix[obj] = visited;
    }
}

/* the total of the per-object aperture pixel counts (one workgroup) */
__global__ void __launch_bounds__(1024) k_phot_sum(const int *v, int n, unsigned long long *out)
{
    __shared__ unsigned long long part[16];
    unsigned long long s = 0;
    for (int i = threadIdx.x; i < n; i += 1024) s += (unsigned)v[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < 16; w++) t += part[w];
        *out = t;
    }
}

/* catalogue rows, the outobjstruct fields endobject_kernel fills (cudaanalyse.cu:3049-3095) */
__global__ void __launch_bounds__(256) k_phot_rows(PhotArgs A, const double *alpha,
                                                   const double *delta, sxg_outobj *rows)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= A.nobj) return;
    const ObjCols &o = A.o;
    const PhotCols &p = A.p;
    sxg_outobj r = {};
    r.number = i + 1;
    r.alpha2000 = alpha[i];
    r.delta2000 = delta[i];
    r.sposx = (float)(o.mx[i] + 1.0);
    r.sposy = (float)(o.my[i] + 1.0);
    r.flux_iso = p.flux_iso[i];
    r.fluxerr_iso = p.fluxerr_iso[i];
    r.mag_iso = p.mag_iso[i];
    r.magerr_iso = p.magerr_iso[i];
    r.flux_isocor = p.flux_isocor[i];
    r.fluxerr_isocor = p.fluxerr_isocor[i];
    r.flux_aper[0] = p.flux_aper[i];
    r.fluxerr_aper[0] = p.fluxerr_aper[i];
    r.mag_aper[0] = p.mag_aper[i];
    r.magerr_aper[0] = p.magerr_aper[i];
    r.flux_auto = p.flux_auto[i];
    r.fluxerr_auto = p.fluxerr_auto[i];
    r.mag_auto = p.mag_auto[i];
    r.magerr_auto = p.magerr_auto[i];
    r.theta = o.theta[i];
    r.poserr_mx2 = o.poserr_mx2[i];
    r.poserr_my2 = o.poserr_my2[i];
    r.poserr_mxy = o.poserr_mxy[i];
    r.flag = (short)o.flag[i];
    r.elong = p.elong[i];
    r.ellip = p.ellip[i];
    r.sprob = p.sprob[i];
    r.peak = o.peak[i];
    r.kronfactor = p.kron[i];
    r.bkg = o.bkg[i];
    r.fwhm = o.fwhm[i];
    rows[i] = r;
}

/* ------------------------------------------------------------------- host ---- */
static bool phot_cols(int nobj, PhotCols &p)
{
    const size_t n = (size_t)(nobj > 0 ? nobj : 1);
#define PF(f) p.f = sxg_take<float>("phot." #f, n)
    PF(flux_iso); PF(fluxerr_iso); PF(mag_iso); PF(magerr_iso); PF(flux_isocor);
    PF(fluxerr_isocor); PF(flux_aper); PF(fluxerr_aper); PF(mag_aper); PF(magerr_aper);
    PF(kron); PF(flux_auto); PF(fluxerr_auto); PF(mag_auto); PF(magerr_auto); PF(elong);
    PF(ellip); PF(sprob);
#undef PF
    const float *all[] = {p.flux_iso, p.fluxerr_iso, p.mag_iso, p.magerr_iso, p.flux_isocor,
                          p.fluxerr_isocor, p.flux_aper, p.fluxerr_aper, p.mag_aper,
                          p.magerr_aper, p.kron, p.flux_auto, p.fluxerr_auto, p.mag_auto,
                          p.magerr_auto, p.elong, p.ellip, p.sprob};
    for (const float *q : all)
        if (!q) return false;
    return true;
}

static bool brain_ok(const sxg_brain *b)
{
    if (b->layersnb < 2 || b->layersnb > 3 || b->nn[0] != 10) return false;
    for (int l = 0; l < b->layersnb; l++)
        if (b->nn[l] < 1 || b->nn[l] > 10) return false;
    return b->outscale[0] != 0.0;
}

extern "C" int sxg_set_photometry(const sxg_photom_params *p)
{
    sxg_clear_error();
    if (!p || !(p->apert > 0.0) || !(p->autoparam[0] >= 0.0) || !(p->autoparam[1] >= 0.0) ||
        !(p->autoaper[0] >= 0.0) || !(p->autoaper[1] >= 0.0) || !(p->seeing_fwhm > 0.0)) {
        sxg_set_error("sxg_set_photometry: invalid parameters");
        return -1;
    }
    SxgState &S = sxg_state();
    S.phot = *p;
    S.mag_zeropoint = p->mag_zeropoint;
    return 0;
}

/* Byte offsets of the fields endobject_gpu reads from the caller's prefstruct and
   obj2struct, as the reference's build lays them out on x86-64 (src/prefs.h,
   src/types.h, NAPER 1): the ones its endobject kernels read (cudaanalyse.cu:145-151,
   :349-584, :2905, :3363, :3491-3507), and the enum values they are compared with
   (DETECT_TYPE PHOTO, MASK_TYPE CORRECT).  sxg_reference_layout() returns them, so a test
   compiled against the reference's headers can check them. */
enum RefLayout : long {
    RL_PREFS_NAPER = 2096, RL_PREFS_APERT = 1840, RL_PREFS_AUTOPARAM = 2120,
    RL_PREFS_AUTOAPER = 2168, RL_PREFS_SEEING = 2232, RL_PREFS_MAGZP = 2192,
    RL_PREFS_MASKTYPE = 7152, RL_PREFS_DETECTTYPE = 1696, RL_OBJ2_FLUX_APER = 88,
    RL_OBJ2_FLUX_AUTO = 36, RL_OBJ2_SPROB = 528, RL_DETECT_PHOTO = 1, RL_MASK_CORRECT = 2,
    RL_PREFS_SIZE = 7696, RL_OBJ2_SIZE = 576
};

extern "C" int sxg_reference_layout(long *out, int n)
{
    static const long v[] = {RL_PREFS_SIZE, RL_PREFS_NAPER, RL_PREFS_APERT, RL_PREFS_AUTOPARAM,
                             RL_PREFS_AUTOAPER, RL_PREFS_SEEING, RL_PREFS_MAGZP,
                             RL_PREFS_MASKTYPE, RL_PREFS_DETECTTYPE, RL_OBJ2_SIZE,
                             RL_OBJ2_FLUX_APER, RL_OBJ2_FLUX_AUTO, RL_OBJ2_SPROB,
                             RL_DETECT_PHOTO, RL_MASK_CORRECT};
    const int nv = (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; out && i < n && i < nv; i++) out[i] = v[i];
    /* then the wcsstruct offsets (wcs.hip) */
    return nv + sxg_wcs_layout(out && n > nv ? out + nv : nullptr, n > nv ? n - nv : 0);
}

/* the photometry settings of the reference's prefs (and the groups its obj2 flags ask for:
   the first byte of each flag field, define.h:283); false (error recorded) when its NAPER
   is not this build's (the reference aborts, cudaanalyse.cu:3363-3367) */
static bool prefs_settings(const void *prefs, const void *flagobj2, sxg_photom_params &pp)
{
    if (prefs) {
        const char *pb = static_cast<const char *>(prefs);
        int naper = 0, mask = 0, detect = 0;
        memcpy(&naper, pb + RL_PREFS_NAPER, sizeof(int));
        if (naper != 1) {
            sxg_set_error("endobject_gpu: prefs.naper is " + std::to_string(naper) +
                          ", this library measures one aperture (NAPER 1)");
            return false;
        }
        memcpy(&pp.apert, pb + RL_PREFS_APERT, sizeof(double));
        memcpy(pp.autoparam, pb + RL_PREFS_AUTOPARAM, 2 * sizeof(double));
        memcpy(pp.autoaper, pb + RL_PREFS_AUTOAPER, 2 * sizeof(double));
        memcpy(&pp.seeing_fwhm, pb + RL_PREFS_SEEING, sizeof(double));
        memcpy(&pp.mag_zeropoint, pb + RL_PREFS_MAGZP, sizeof(double));
        memcpy(&mask, pb + RL_PREFS_MASKTYPE, sizeof(int));
        memcpy(&detect, pb + RL_PREFS_DETECTTYPE, sizeof(int));
        pp.mask_correct = mask == RL_MASK_CORRECT;
        pp.photo = detect == RL_DETECT_PHOTO;
    }
    if (flagobj2) {
        const char *ob = static_cast<const char *>(flagobj2);
        pp.do_aper = ob[RL_OBJ2_FLUX_APER] != 0;
        pp.do_auto = ob[RL_OBJ2_FLUX_AUTO] != 0;
        pp.do_class = ob[RL_OBJ2_SPROB] != 0;
    }
    return true;
}

/* the photometry of the current frame's objects into the "phot.*" columns and the flags;
   brain may be nullptr (CLASS_STAR 0) */
static bool run_photometry(const sxg_photom_params &pp, const sxg_brain *brain, float backsig,
                           float thresh, double satur_level, double pixscale, double ngamma,
                           double gain)
{
    SxgState &S = sxg_state();
    hipStream_t st = S.stream;
    const int n = S.nobj;
    PhotArgs A;
    A.W = S.W;
    A.H = S.H;
    A.nobj = n;
    if (!sxg_obj_view(A.o) || !phot_cols(n, A.p)) return false;
    if (pp.photo && !(ngamma > 0.0)) {
        sxg_set_error("endobject_gpu: DETECT_TYPE PHOTO needs ngamma > 0");
        return false;
    }
    if (pp.do_class && brain && !(pixscale > 0.0)) {
        sxg_set_error("endobject_gpu: CLASS_STAR needs a pixel scale > 0");
        return false;
    }
    int *lab = sxg_take<int>("phot.lab", S.N);
    int *order = sxg_take<int>("phot.order", (size_t)n + 2);
    A.npix = sxg_take<int>("phot.npix", (size_t)n);
    if (!lab || !order || !A.npix) return false;
    A.brain = nullptr;
    if (brain && pp.do_class) {
        if (!brain_ok(brain)) {
            sxg_set_error("endobject_gpu: the classifier network has an unsupported shape");
            return false;
        }
        sxg_brain *db = sxg_take<sxg_brain>("phot.brain", 1);
        if (!db) return false;
        SXG_HIP_RET(hipMemcpyAsync(db, brain, sizeof(sxg_brain), hipMemcpyHostToDevice, st),
                    false);
        A.brain = db;
    }
    const long N = (long)S.N;
    sxg_kt_begin(SXG_KT_PHOT);
    hipLaunchKernelGGL(k_phot_lab_init, dim3(sxg_div_up((N >> 2) + 4, 256)), dim3(256), 0, st,
                       S.cd.as<float>(), lab, N, S.dthresh);
    if (S.npix_raw > 0)
        hipLaunchKernelGGL(k_phot_lab_scatter, dim3(sxg_div_up((long)S.npix_raw, 256)),
                           dim3(256), 0, st, S.fpia, (long)S.npix_raw,
                           sxg_take<int>("deb.alloc", S.N), A.o.master, lab);
    SXG_HIP_RET(hipMemsetAsync(order + n, 0, 2 * sizeof(int), st), false);
    hipLaunchKernelGGL(k_phot_order, dim3(sxg_div_up(n, 256)), dim3(256), 0, st, A.o, n,
                       order + n, order);
    A.sub = S.img.as<float>();
    A.lab = lab;
    A.order = order;
    A.backsig = backsig;
    A.thresh = thresh;
    A.apert = pp.apert;
    A.autoparam0 = pp.autoparam[0];
    A.autoparam1 = pp.autoparam[1];
    A.autoaper0 = pp.autoaper[0];
    A.autoaper1 = pp.autoaper[1];
    A.seeing = pp.seeing_fwhm;
    A.pixscale = pixscale;
    A.zp = pp.mag_zeropoint;
    A.ngamma = ngamma;
    A.gain = gain;
    A.satur = satur_level;
    A.corr = pp.mask_correct;
    A.photo = pp.photo;
    A.do_aper = pp.do_aper;
    A.do_auto = pp.do_auto;
    A.do_class = pp.do_class;
    hipLaunchKernelGGL(k_endobject, dim3(sxg_div_up(n, 4)), dim3(256), 0, st, A);
    sxg_kt_end(SXG_KT_PHOT);
    SXG_HIP_RET(hipGetLastError(), false);
    return true;
}

extern "C" void endobject_gpu(void *prefs, void *brain, void *flagobj2, void *objlist, void *wcs,
                              int fymin, int fymax, int width, int height, float backsig,
                              float thresh, double satur_level, double pixscale, double ngamma,
                              double gain, int numobj)
{
    SxgState &S = sxg_state();
    sxg_clear_error();
    if (!S.have_frame || !S.cleaned || numobj != S.nobj || width != S.W || height != S.H) {
        sxg_set_error("endobject_gpu: call after run_clean, with parcel_out's numobj and the "
                      "frame's width and height");
        return;
    }
    if (fymin != 0 || fymax != height) {
        sxg_set_error("endobject_gpu: only whole frames are supported (fymin 0, fymax height)");
        return;
    }
    S.phot_ready = false;
    S.phot_pixels = 0;
    if (numobj == 0) {
        S.phot_ready = true;
        return;
    }
    /* the reference front-end's prefs and obj2 flags when given, else the settings of
       sxg_set_photometry (MAG_ZEROPOINT: sxg_set_mag_zeropoint) */
    sxg_photom_params pp = S.phot;
    pp.mag_zeropoint = S.mag_zeropoint;
    if (!prefs_settings(prefs, flagobj2, pp)) return;
    if (!run_photometry(pp, static_cast<const sxg_brain *>(brain), backsig, thresh, satur_level,
                        pixscale, ngamma, gain) ||
        !sxg_astrometry(wcs))
        return;
    unsigned long long visited = 0;
    unsigned long long *vsum = sxg_take<unsigned long long>("phot.npix_sum", 1);
    if (!vsum) return;
    hipLaunchKernelGGL(k_phot_sum, dim3(1), dim3(1024), 0, S.stream,
                       sxg_take<int>("phot.npix", (size_t)numobj), numobj, vsum);
    SXG_HIP(hipMemcpyAsync(&visited, vsum, sizeof(visited), hipMemcpyDeviceToHost, S.stream));
    if (objlist) {
        PhotArgs A;
        A.nobj = numobj;
        if (!sxg_obj_view(A.o) || !phot_cols(numobj, A.p)) return;
        sxg_outobj *drows = sxg_take<sxg_outobj>("phot.rows", (size_t)numobj);
        if (!drows) return;
        hipLaunchKernelGGL(k_phot_rows, dim3(sxg_div_up(numobj, 256)), dim3(256), 0, S.stream, A,
                           (const double *)sxg_take<double>("wcs.alpha", (size_t)numobj),
                           (const double *)sxg_take<double>("wcs.delta", (size_t)numobj),
                           drows);
        SXG_HIP(hipMemcpyAsync(objlist, drows, sizeof(sxg_outobj) * numobj,
                               hipMemcpyDeviceToHost, S.stream));
    }
    SXG_HIP(hipGetLastError());
    SXG_HIP(hipStreamSynchronize(S.stream));
    S.phot_pixels = (int64_t)visited;
    S.phot_ready = true;
}

extern "C" int sxg_get_photometry(int column, float *out, size_t count)
{
    SxgState &S = sxg_state();
    PhotCols p;
    if (!S.phot_ready) {
        sxg_set_error("sxg_get_photometry: call after endobject_gpu");
        return -1;
    }
    if (column < 0 || column >= SXG_NPHOT || !phot_cols(S.nobj, p)) {
        sxg_set_error("sxg_get_photometry: bad column");
        return -1;
    }
    if (count < (size_t)S.nobj) {
        sxg_set_error("sxg_get_photometry: output buffer too small");
        return -1;
    }
    if (S.nobj == 0) return 0;
    const float *src[SXG_NPHOT] = {p.flux_iso, p.fluxerr_iso, p.mag_iso, p.magerr_iso,
                                   p.flux_isocor, p.fluxerr_isocor, p.flux_aper,
                                   p.fluxerr_aper, p.mag_aper, p.magerr_aper, p.kron,
                                   p.flux_auto, p.fluxerr_auto, p.mag_auto, p.magerr_auto,
                                   p.elong, p.ellip, p.sprob};
    SXG_HIP_RET(hipMemcpyAsync(out, src[column], sizeof(float) * S.nobj, hipMemcpyDeviceToHost,
                               S.stream), -1);
    SXG_HIP_RET(hipStreamSynchronize(S.stream), -1);
    return 0;
}
