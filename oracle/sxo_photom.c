/*
 * sxo_photom.c -- CPU oracle of the catalogue photometry (TEST INFRASTRUCTURE ONLY).
 *
 * Restates, object by object and in plain sequential C, what the reference's
 * endobject_gpu (src/cuda/cudaanalyse.cu:3344) computes after CLEAN:
 *   - the photometry label map of updateAllocateLabel (cudaanalyse.cu:3320-3342);
 *   - computeisocorflux (:62-98), computeaperflux (:104-287, one aperture: NAPER 1,
 *     src/define.h:50), computeautoflux (:293-590);
 *   - the star/galaxy network neurresp and its sigmoid (:599-636), fed as in :2903-2921;
 *   - ELONGATION / ELLIPTICITY (:2778-2786), FLUXERR_ISO (:2791-2792), the SATUR flag
 *     (:2873-2874) and computemags for MAG_ISO (:638-653).
 * Arithmetic types follow the CUDA source as compiled (C++ overloads: sqrt and log10 of a
 * float are the float functions): each expression below keeps the reference's operand
 * types and order.  The transcendental functions are the canonical ones of sx_cmath.h
 * (shared with the HIP kernels); log10 is sx_log(x) times 1/ln(10).
 * Every floating-point sum is sequential in raster order over the aperture box, as the
 * reference's loops run.
 */
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sxo.h"
#include "../include/sx_cmath.h"

#define PH_OBJ_CROWDED 0x0001
#define PH_OBJ_SATUR 0x0004
#define PH_OBJ_APERT_PB 0x0010
#define PH_APER_OVERSAMP 5                 /* src/define.h:172 */
#define PH_KRON_NSIG (3 * 2.0)             /* src/define.h:173, MARGIN_SCALE 2.0 */
#define PH_CROWD_THRESHOLD 0.1             /* src/define.h:177 */

double sxo_log10(double x) { return sx_log10(x); }

static int ph_rint(double x) { return (int)floor(x + 0.5); }   /* src/define.h:262 */

/* cudaanalyse.cu:3320-3342: pixels of objects removed by CLEAN (label > 0, master set)
   and detected pixels of no object (cd >= dthresh, label -1) become INT_MAX */
void sxo_photom_labels(const sxo_result *r, int *lab)
{
    const long N = (long)r->width * r->height;
    for (long p = 0; p < N; p++) {
        int t = r->alloc[p];
        if (t > 0 && r->master[t] != 0xffffffffu) t = INT_MAX;
        if (r->cd[p] >= r->dthresh && t == -1) t = INT_MAX;
        lab[p] = t;
    }
}

/* cudaanalyse.cu:62-98 */
static void ph_isocor(const sxo_result *r, int o, float *flux_isocor, float *fluxerr_isocor)
{
    const float flux = r->flux[o], fluxerr = r->fluxerr[o];
    const unsigned fdnpix = (unsigned)r->fdnpix[o];
    double ati = (flux > 0.0) ? (double)((float)fdnpix * r->dthresh_obj[o] / flux) : 0.0;
    if (ati > 1.0)
        ati = 1.0;
    else if (ati < 0.0)
        ati = 0.0;
    *flux_isocor = (float)(flux / (1.0 - 0.196099 * ati - 0.751208 * ati * ati));
    if (flux > 0.0) {
        const double sigtv = (double)(fluxerr / (flux * flux));
        double dati = fdnpix ? ati * sqrt(sigtv + 1.0 / fdnpix) : 0.0;
        dati = 0.196099 * dati + 0.751208 * 2 * ati * dati;
        *fluxerr_isocor = (float)(sqrt(sigtv + dati * dati) * flux);
    } else {
        *fluxerr_isocor = sqrtf(fluxerr);
    }
}

/* the replacement value of a pixel that belongs to another object (MASK_TYPE CORRECT:
   its mirror through the centroid when that one is free or the object's own, else 0);
   cudaanalyse.cu:222-243 (APER) and :523-545 (AUTO) */
static double ph_masked(const sxo_result *r, const int *lab, int o, int x, int y, int corr,
                        double pix)
{
    const int W = r->width, H = r->height;
    int t = lab[(long)y * W + x];
    if (t >= 0 && t != o) {
        int x2, y2;
        if (corr && (x2 = (int)(2 * r->mx[o] + 0.49999 - x)) >= 0 && x2 < W &&
            (y2 = (int)(2 * r->my[o] + 0.49999 - y)) >= 0 && y2 < H &&
            ((t = lab[(long)y2 * W + x2]) == o || t == -1))
            return r->sub[(long)y2 * W + x2];
        return 0.0;
    }
    return pix;
}

/* cudaanalyse.cu:104-287, aperture i = 0 */
static void ph_aper(const sxo_result *r, const int *lab, int o, const sxo_photom_config *pc,
                    int *flag, float *flux, float *fluxerr, float *mag, float *magerr)
{
    const int W = r->width, H = r->height;
    const double mx = r->mx[o], my = r->my[o];
    const int pflag = pc->photo, corr = pc->mask_correct;
    double tv, sigtv, area, pix, var;
    var = pc->backsig * pc->backsig;
    const float raper = (float)(pc->apert / 2.0);
    const float raper2 = raper * raper;
    const float rintlim = (float)(raper - 0.75);
    const float rintlim2 = (rintlim > 0.0) ? rintlim * rintlim : 0.0f;
    const float rextlim2 = (float)((raper + 0.75) * (raper + 0.75));
    tv = sigtv = area = 0.0;
    const float scalex = (float)(1.0 / PH_APER_OVERSAMP), scaley = scalex;
    const float scale2 = scalex * scaley;
    const float offsetx = (float)(0.5 * (scalex - 1.0)), offsety = (float)(0.5 * (scaley - 1.0));
    int xmin = (int)(mx - raper + 0.499999), xmax = (int)(mx + raper + 1.499999);
    int ymin = (int)(my - raper + 0.499999), ymax = (int)(my + raper + 1.499999);
    if (xmin < 0) { xmin = 0; *flag |= PH_OBJ_APERT_PB; }
    if (xmax > W) { xmax = W; *flag |= PH_OBJ_APERT_PB; }
    if (ymin < 0) { ymin = 0; *flag |= PH_OBJ_APERT_PB; }
    if (ymax > H) { ymax = H; *flag |= PH_OBJ_APERT_PB; }
    for (int y = ymin; y < ymax; y++)
        for (int x = xmin; x < xmax; x++) {
            float dx = (float)(x - mx), dy = (float)(y - my);
            const float r2 = dx * dx + dy * dy;
            if (!(r2 < rextlim2)) continue;
            float locarea;
            if (r2 > rintlim2) {
                dx += offsetx;
                dy += offsety;
                locarea = 0.0f;
                for (int sy = PH_APER_OVERSAMP; sy--; dy += scaley) {
                    float dx1 = dx;
                    const float dy2 = dy * dy;
                    for (int sx = PH_APER_OVERSAMP; sx--; dx1 += scalex)
                        if (dx1 * dx1 + dy2 < raper2) locarea += scale2;
                }
            } else {
                locarea = 1.0f;
            }
            area += locarea;
            pix = ph_masked(r, lab, o, x, y, corr, (double)r->sub[(long)y * W + x]);
            if (pflag) {
                pix = sx_exp(pix / pc->ngamma);
                sigtv += var * locarea * pix * pix;
            } else {
                sigtv += var * locarea;
            }
            tv += locarea * pix;
        }
    if (pflag) {
        tv = pc->ngamma * (tv - area * sx_exp(0.0 / pc->ngamma));
        sigtv /= pc->ngamma * pc->ngamma;
    } else {
        tv -= area * 0.0;   /* dbkg is 0 (cudaanalyse.cu:2244) */
        if (pc->gain > 0.0 && tv > 0.0) sigtv += tv / pc->gain;
    }
    *flux = (float)tv;
    *fluxerr = (float)sqrt(sigtv);
    *mag = tv > 0.0 ? (float)(-2.5 * sx_log10(tv) + pc->mag_zeropoint) : 99.0f;
    *magerr = tv > 0.0 ? (float)(1.086 * sqrt(sigtv) / tv) : 99.0f;
}

/* the search / integration box of computeautoflux (cudaanalyse.cu:354-397, :452-494) */
static void ph_box(double mx, double my, float dxlim, float dylim, int W, int H, int *flag,
                   int *xmin, int *xmax, int *ymin, int *ymax)
{
    if ((*xmin = ph_rint(mx - dxlim)) < 0) { *xmin = 0; *flag |= PH_OBJ_APERT_PB; }
    if ((*xmax = ph_rint(mx + dxlim) + 1) > W) { *xmax = W; *flag |= PH_OBJ_APERT_PB; }
    if ((*ymin = ph_rint(my - dylim)) < 0) { *ymin = 0; *flag |= PH_OBJ_APERT_PB; }
    if ((*ymax = ph_rint(my + dylim) + 1) > H) { *ymax = H; *flag |= PH_OBJ_APERT_PB; }
}

/* cudaanalyse.cu:293-590 */
static void ph_auto(const sxo_result *r, const int *lab, int o, const sxo_photom_config *pc,
                    int *flag, float *kron, float *flux, float *fluxerr, float *mag,
                    float *magerr)
{
    const int W = r->width, H = r->height;
    const double mx = r->mx[o], my = r->my[o];
    const float a = r->a[o], b = r->b[o];
    const int pflag = pc->photo, corr = pc->mask_correct;
    double sigtv, tv, r1, v1, var;
    float cx2, cy2, cxy, klim2, dxlim, dylim;
    int area, areab, xmin, xmax, ymin, ymax;
    var = pc->backsig * pc->backsig;
    if (PH_KRON_NSIG * sqrtf(a * b) > pc->autoaper[0] / 2.0) {
        cx2 = r->cxx[o];
        cy2 = r->cyy[o];
        cxy = r->cxy[o];
        dxlim = (float)(cx2 - cxy * cxy / (4.0 * cy2));
        dxlim = dxlim > 0.0 ? (float)(PH_KRON_NSIG / sqrtf(dxlim)) : 0.0f;
        dylim = (float)(cy2 - cxy * cxy / (4.0 * cx2));
        dylim = dylim > 0.0 ? (float)(PH_KRON_NSIG / sqrtf(dylim)) : 0.0f;
        klim2 = (float)(PH_KRON_NSIG * PH_KRON_NSIG);
    } else {
        cx2 = cy2 = 1.0f;
        cxy = 0.0f;
        dxlim = dylim = (float)(pc->autoaper[0] / 2.0);
        klim2 = dxlim * dxlim;
    }
    ph_box(mx, my, dxlim, dylim, W, H, flag, &xmin, &xmax, &ymin, &ymax);
    v1 = r1 = 0.0;
    area = areab = 0;
    for (int y = ymin; y < ymax; y++)
        for (int x = xmin; x < xmax; x++) {
            const float dx = (float)(x - mx), dy = (float)(y - my);
            const float r2 = cx2 * dx * dx + cy2 * dy * dy + cxy * dx * dy;
            if (r2 <= klim2) {
                const float pix = r->sub[(long)y * W + x];
                const int t = lab[(long)y * W + x];
                if ((t == o || t == -1) && pix > -SXO_BIG) {
                    area++;
                    r1 += sqrtf(r2) * pix;
                    v1 += pix;
                } else {
                    areab++;
                }
            }
        }
    area += areab;
    float kf = 0.0f;
    if (area) {
        if (r1 > 0.0 && v1 > 0.0) {
            kf = (float)(pc->autoparam[0] * r1 / v1);
            if (kf < pc->autoparam[1]) kf = (float)pc->autoparam[1];
        } else {
            kf = (float)pc->autoparam[1];
        }
        if ((float)areab / area > PH_CROWD_THRESHOLD) *flag |= PH_OBJ_CROWDED;
        if (kf * sqrtf(a * b) > pc->autoaper[1] / 2.0) {
            cx2 = r->cxx[o];
            cy2 = r->cyy[o];
            cxy = r->cxy[o];
            dxlim = (float)(cx2 - cxy * cxy / (4.0 * cy2));
            dxlim = dxlim > 0.0 ? kf / sqrtf(dxlim) : 0.0f;
            dylim = (float)(cy2 - cxy * cxy / (4.0 * cx2));
            dylim = dylim > 0.0 ? kf / sqrtf(dylim) : 0.0f;
            klim2 = kf * kf;
        } else {
            cx2 = cy2 = 1.0f;
            cxy = 0.0f;
            dxlim = dylim = (float)(pc->autoaper[1] / 2.0);
            klim2 = dxlim * dxlim;
            kf = 0.0f;
        }
        ph_box(mx, my, dxlim, dylim, W, H, flag, &xmin, &xmax, &ymin, &ymax);
        area = areab = 0;
        tv = sigtv = 0.0;
        for (int y = ymin; y < ymax; y++)
            for (int x = xmin; x < xmax; x++) {
                const float dx = (float)(x - mx), dy = (float)(y - my);
                if (cx2 * dx * dx + cy2 * dy * dy + cxy * dx * dy <= klim2) {
                    area++;
                    float pix = r->sub[(long)y * W + x];
                    const int t = lab[(long)y * W + x];
                    if (t >= 0 && t != o) {
                        areab++;
                        pix = (float)ph_masked(r, lab, o, x, y, corr, pix);
                    }
                    if (pflag) {
                        pix = (float)sx_exp(pix / pc->ngamma);
                        sigtv += var * pix * pix;
                    } else {
                        sigtv += var;
                    }
                    tv += pix;
                }
            }
        if ((float)areab > PH_CROWD_THRESHOLD * area) *flag |= PH_OBJ_CROWDED;
        if (pflag) {
            tv = pc->ngamma * (tv - area * sx_exp(0.0 / pc->ngamma));
            sigtv /= pc->ngamma * pc->ngamma;
        } else {
            tv -= area * 0.0;
            if (pc->gain > 0.0 && tv > 0.0) sigtv += tv / pc->gain;
        }
    } else {
        tv = sigtv = 0.0;
    }
    *flux = (float)tv;
    *fluxerr = (float)sqrt(sigtv);
    *mag = *flux > 0.0 ? (float)(-2.5 * sx_log10f(*flux) + pc->mag_zeropoint) : 99.0f;
    *magerr = *flux > 0.0 ? (float)(1.086 * *fluxerr / *flux) : 99.0f;
    if (tv <= 0.0) kf = 0.0f;
    *kron = kf;
}

/* cudaanalyse.cu:599-636 */
double sxo_neurresp(const sxo_brain *br, const double *input)
{
    double n[3][10];
    const int last = br->layersnb - 1;
    for (int i = 0; i < br->nn[0]; i++) n[0][i] = input[i] * br->inscale[i] + br->inbias[i];
    for (int l = 0; l < last; l++)
        for (int j = 0; j < br->nn[l + 1]; j++) {
            double s = br->b[l][j];
            for (int i = 0; i < br->nn[l]; i++) s += br->w[l][i][j] * n[l][i];
            n[l + 1][j] = 1.0 / (1.0 + sx_exp(-s));
        }
    return (n[last][0] - br->outbias[0]) / br->outscale[0];
}

/* cudaanalyse.cu:2903-2921 */
static float ph_class_star(const sxo_result *r, int o, const sxo_photom_config *pc)
{
    double input[10];
    const double fwhm = pc->seeing_fwhm;
    double fac2 = fwhm / pc->pixscale;
    fac2 *= fac2;
    const int *iso = r->iso + 8 * (long)o;
    int j = 0;
    input[j] = sx_log10(iso[0] ? iso[0] / fac2 : 0.01);
    input[++j] = pc->thresh > 0.0
                     ? sx_log10(r->peak[o] > 0.0 ? (double)(r->peak[o] / pc->thresh) : 0.1)
                     : -1.0;
    for (int i = 1; i < 8; i++) input[++j] = sx_log10(iso[i] ? iso[i] / fac2 : 0.01);
    input[++j] = sx_log10(fwhm);
    return (float)sxo_neurresp(pc->brain, input);
}

int sxo_endobject(const sxo_result *r, const sxo_photom_config *pc, sxo_photom *out)
{
    const int n = r->nobj;
    const long N = (long)r->width * r->height;
    int *lab = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
    if (!lab) return -1;
    sxo_photom_labels(r, lab);
    for (int o = 0; o < n; o++) {
        int flag = r->flag[o];
        const float a = r->a[o], b = r->b[o];
        out->elong[o] = a / b;
        out->ellip[o] = 1 - b / a;
        out->flux_iso[o] = r->flux[o];
        out->fluxerr_iso[o] = sqrtf(r->fluxerr[o]);
        ph_isocor(r, o, &out->flux_isocor[o], &out->fluxerr_isocor[o]);
        if (pc->do_aper)
            ph_aper(r, lab, o, pc, &flag, &out->flux_aper[o], &out->fluxerr_aper[o],
                    &out->mag_aper[o], &out->magerr_aper[o]);
        if (pc->do_auto)
            ph_auto(r, lab, o, pc, &flag, &out->kronfactor[o], &out->flux_auto[o],
                    &out->fluxerr_auto[o], &out->mag_auto[o], &out->magerr_auto[o]);
        if (r->peak[o] + r->bkg[o] >= pc->satur_level) flag |= PH_OBJ_SATUR;
        out->sprob[o] = (pc->brain && pc->do_class) ? ph_class_star(r, o, pc) : 0.0f;
        const float fi = out->flux_iso[o];
        out->mag_iso[o] = fi > 0.0 ? (float)(-2.5 * sx_log10f(fi) + pc->mag_zeropoint) : 99.0f;
        out->magerr_iso[o] = fi > 0.0 ? (float)(1.086 * out->fluxerr_iso[o] / fi) : 99.0f;
        out->flag[o] = flag;
    }
    free(lab);
    return n;
}
