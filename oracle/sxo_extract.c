/*
 * sxo_extract.c -- oracle restatement of detection, deblending, analysis and cleaning.
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the CPU-baseline leg of
 * the benchmark, never by the product.
 *
 * The reference's CPU code for these stages was removed (src/analyse.c is empty,
 * src/extract.c and src/refine.c hold no functions), so this file restates the
 * reference's CUDA algorithms sequentially:
 *   detection  src/cuda/cudadetection.cu:36-64 (cd >= dthresh), :156-270 (8-connected
 *              labels converge to the minimum raster index), :677-882 (objects sorted by
 *              label, count >= DETECT_MINAREA kept, numbered in label order), :899-1090
 *              (fdpeak, fdflux per object);
 *   deblending src/cuda/cudadeblend.cu:199-217 (sub-thresholds), :223-380 and :579-727
 *              (per-level labelling of the previous level's kept pixels), :748-1175
 *              (count >= deb_minarea kept, parent, fdflux), :422-483 and :1425-1615
 *              (tree pruning, kept objects ordered by root, level, label);
 *   analysis   src/cuda/cudaanalyse.cu:1686-1904 (preanalyse_full on level pixels),
 *              cudadeblend.cu:1262-1360 (gatherup), :1876-1935 (final pixel lists),
 *              cudaanalyse.cu:1906-2126 (preanalyse_robust), :2128-2581 (_back and
 *              examineiso);
 *   cleaning   src/cuda/cudaclean.cu:10-29 (amp, alpha), :120-174 (pair test),
 *              :260-386 (window of 16*ceil(MEMORY_OBJSTACK/16) objects, victims dropped).
 * Canonical choices where the reference order is set by a library or is undefined:
 *   - every per-object float or double sum runs sequentially over the object's pixels
 *     in raster order (the reference's per-object loops, cudaanalyse.cu:1770-1835, do
 *     this; its CUDPP segmented-scan sums, cudadetection.cu:950 and cudadeblend.cu:1015,
 *     are defined here the same way);
 *   - pow/log on decision paths use the canonical functions of include/sx_cmath.h.
 * Deliberate deviations from reference bugs (each documented in the design notes):
 *   - gatherup falls back to the root's first object when no distance is finite (the
 *     reference falls back to global object 0, cudadeblend.cu:1300) and never reads
 *     past the root's object segment (:1306-1311);
 *   - CLEAN N skips cleaning (the reference cleans regardless of the flag).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "sxo.h"
#include "../include/sx_cmath.h"

#define SXO_PI 3.1415926535898     /* src/define.h:56 */
#define OBJ_MERGED 0x0002
#define OBJ_SATUR 0x0004
#define OBJ_TRUNC 0x0008
#define NISO 8
#define MAXDEBAREA 3
#define CLEAN_ZONE 10.0
#define MAXPICSIZE 1048576

double sxo_log(double x) { return sx_log(x); }
double sxo_exp(double x) { return sx_exp(x); }
double sxo_pow(double x, double y) { return sx_pow(x, y); }

void sxo_config_defaults(sxo_config *c)
{
    memset(c, 0, sizeof(*c));
    c->back_size = 64;
    c->back_fsize = 3;
    c->back_fthresh = 0.0;
    c->detect_thresh = 1.5;
    c->analysis_thresh = 1.5;
    c->ext_minarea = 5;
    c->deblend_nthresh = 32;
    c->deblend_mincont = 0.005;
    c->clean = 1;
    c->clean_param = 1.0;
    c->clean_stacksize = 3000;
    c->satur_level = 50000.0;
    c->gain = 0.0;
}

/* ----------------------------------------------------------- union-find ---- */
static int uf_find(int *par, int x)
{
    while (par[x] != x) {
        par[x] = par[par[x]];
        x = par[x];
    }
    return x;
}

static void uf_union(int *par, int a, int b)
{
    a = uf_find(par, a);
    b = uf_find(par, b);
    if (a == b) return;
    if (a < b) par[b] = a; else par[a] = b;   /* the root is the smallest index */
}

/* ------------------------------------------------------- per-object work ---- */
typedef struct {
    int level, parent, first, count;
    float fdflux, dthresh;
    int nson, okA, kept, sig;
} comp_t;

/* preanalyse over one pixel list (cudaanalyse.cu:1686-1904 FULL, :1906-2126 ROBUST) */
typedef struct {
    int flag, xmin, xmax, ymin, ymax, dnpix;
    float fdflux, fdpeak, dpeak, dflux, a, b, theta, cxx, cyy, cxy, abcor;
    double mx, my, mx2, my2, mxy;
} pre_t;

static void preanalyse(const int *pix, int n, const float *sub, const float *cd, int W, int H,
                       double dthresh, int flag_in, int merged_prev, double prev_mx,
                       double prev_my, pre_t *o)
{
    float peak, cpeak, val, cval, rv;
    double dthresh2, t1t2, darea, mx, my, mx2, my2, mxy, tv, xm, ym, xm2, ym2, xym, temp,
        temp2, theta, pmx2, pmy2;
    int x, y, xmin, xmax, ymin, ymax, area2, dnpix, flag = flag_in;

    dnpix = 0;
    rv = 0.0f;
    peak = cpeak = (float)-SXO_BIG;
    ymin = xmin = 2 * MAXPICSIZE;
    ymax = xmax = 0;
    for (int i = 0; i < n; i++) {
        int p = pix[i];
        x = p % W;
        y = p / W;
        if (peak < (val = sub[p])) peak = val;
        if (cpeak < (cval = cd[p])) cpeak = cval;
        rv += cval;
        if (xmin > x) xmin = x;
        if (xmax < x) xmax = x;
        if (ymin > y) ymin = y;
        if (ymax < y) ymax = y;
    }
    if (xmin == 0 || xmax == W - 1 || ymin == 0 || ymax == H - 1) flag |= OBJ_TRUNC;
    o->fdflux = rv;
    o->fdpeak = cpeak;
    o->dpeak = peak;
    o->xmin = xmin; o->xmax = xmax; o->ymin = ymin; o->ymax = ymax;

    mx = my = tv = 0.0;
    mx2 = my2 = mxy = 0.0;
    dthresh2 = (dthresh + peak) / 2.0;
    area2 = 0;
    for (int i = 0; i < n; i++) {
        int p = pix[i];
        x = p % W - xmin;
        y = p / W - ymin;
        cval = cd[p];
        tv += (val = sub[p]);
        if (val > dthresh) dnpix++;
        if (val > dthresh2) area2++;
        mx += cval * x;
        my += cval * y;
        mx2 += cval * x * x;
        my2 += cval * y * y;
        mxy += cval * x * y;
    }
    xm = mx / rv;
    ym = my / rv;
    if (merged_prev) {
        double xn = prev_mx - xmin, yn = prev_my - ymin;
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
        theta = SXO_PI / 4.0;
    temp = sqrt(0.25 * temp * temp + xym * xym);
    pmy2 = pmx2 = 0.5 * (xm2 + ym2);
    pmx2 += temp;
    pmy2 -= temp;
    o->dnpix = dnpix;
    o->dflux = (float)tv;
    o->mx = xm + xmin;
    o->my = ym + ymin;
    o->mx2 = xm2;
    o->my2 = ym2;
    o->mxy = xym;
    o->a = (float)sqrt(pmx2);
    o->b = (float)sqrt(pmy2);
    o->theta = (float)(theta * 180.0 / SXO_PI);
    o->cxx = (float)(ym2 / temp2);
    o->cyy = (float)(xm2 / temp2);
    o->cxy = (float)(-2 * xym / temp2);
    darea = (double)area2 - dnpix;
    t1t2 = dthresh / dthresh2;
    if (t1t2 > 0.0) {
        o->abcor = (float)((darea < 0.0 ? darea : -1.0)
                           / (2 * SXO_PI * sx_log(t1t2 < 1.0 ? t1t2 : 0.99) * o->a * o->b));
        if (o->abcor > 1.0) o->abcor = 1.0f;
    } else
        o->abcor = 1.0f;
    o->flag = flag;
}

/* back.c:685-732 / cudaanalyse.cu:2128-2172: bilinear mesh value */
float sxo_back_at(const float *b, int x, int y, int nx, int ny, int backw, int backh)
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

/* k-th largest of a float heap of size h (cudaanalyse.cu:2386-2421 with a sorted buffer) */
static void top_insert(float *top, int h, int *nfill, float v)
{
    if (*nfill < h) {
        int i = (*nfill)++;
        while (i > 0 && top[i - 1] > v) { top[i] = top[i - 1]; i--; }
        top[i] = v;
    } else if (v > top[0]) {
        int i = 0;
        while (i + 1 < h && top[i + 1] < v) { top[i] = top[i + 1]; i++; }
        top[i] = v;
    }
}

/* -------------------------------------------------------------- the run ---- */
/* monotonic wall clock in milliseconds (the CPU baseline's per-stage times) */
static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

sxo_result *sxo_run(const float *img, int W, int H, const sxo_config *cfg)
{
    sxo_result *r = (sxo_result *)calloc(1, sizeof(sxo_result));
    if (!r) return NULL;
    double t_mark = now_ms();
#define STAGE_END(k)                                                                       \
    do {                                                                                   \
        const double t_now = now_ms();                                                     \
        r->stage_ms[k] = t_now - t_mark;                                                   \
        t_mark = t_now;                                                                    \
    } while (0)
    const size_t N = (size_t)W * H;
    const int bw = cfg->back_size < W ? cfg->back_size : W;
    const int bh = cfg->back_size < H ? cfg->back_size : H;
    if (bw != bh || bw <= 0) { free(r); return NULL; }   /* back.c:99-103 */
    const int bs = bw;
    const int nbx = (W - 1) / bs + 1, nby = (H - 1) / bs + 1;
    r->width = W; r->height = H; r->nbx = nbx; r->nby = nby;
    r->back = (float *)malloc(sizeof(float) * nbx * nby);
    r->sigma = (float *)malloc(sizeof(float) * nbx * nby);
    r->dback = (float *)malloc(sizeof(float) * nbx * nby);
    r->dsigma = (float *)malloc(sizeof(float) * nbx * nby);
    r->sub = (float *)malloc(sizeof(float) * N);
    r->cd = (float *)malloc(sizeof(float) * N);
    r->alloc = (int *)malloc(sizeof(int) * N);
    r->seg = (int *)calloc(N, sizeof(int));

    /* ---- weight map: variance and mesh mask (weight_to_var, weight.c:93-126, with
       sigfac = 1, weight.c:80).  MAP_WEIGHT: pixels of non-positive weight stay out of the
       mesh statistics; MAP_RMS / MAP_VAR: pixels whose variance is BIG or more.  WEIGHT_TYPE
       BACKGROUND has no map: the variance is the square of the interpolated background rms
       (back.c:938-1051), built after the background below. ---- */
    const int wtype = cfg->weight_type;
    const float *wmap = wtype == 3 ? NULL : cfg->weight;
    float *masked = NULL;
    if (wmap) {
        masked = (float *)malloc(sizeof(float) * N);
        r->var = (float *)malloc(sizeof(float) * N);
        for (size_t p = 0; p < N; p++) {
            const float v = wmap[p];
            float var;
            int good;
            if (wtype == 1) {
                var = v < SXO_BIG ? v * v : v;
                good = var < (float)SXO_BIG;
            } else if (wtype == 2) {
                var = v;
                good = var < (float)SXO_BIG;
            } else {
                good = v > 0.0f;
                var = good ? 1.0f / v : (float)SXO_BIG;
            }
            masked[p] = good ? img[p] : (float)-SXO_BIG;
            r->var[p] = var;
        }
    }
    /* ---- background and thresholds (back.c:43-208) ---- */
    sxo_background(masked ? masked : img, W, H, bs, cfg->back_fsize, cfg->back_fsize,
                   cfg->back_fthresh, cfg->back_type, cfg->back_value, r->sub, r->back,
                   r->sigma, r->dback, r->dsigma, &r->backmean, &r->backsig);
    if (masked) {
        /* the maps come from the masked frame; the subtraction applies to the frame */
        memcpy(r->sub, img, sizeof(float) * N);
        if (cfg->back_type == 1 || cfg->back_type == 2) {
            /* the constant sxo_background subtracted (the median of the masked maps for
               back_type 1, BACK_VALUE for 2) */
            const float bval = cfg->back_type == 1 ? r->backmean : (float)cfg->back_value;
            for (size_t k = 0; k < N; k++) r->sub[k] -= bval;
        } else {
            for (int y = 0; y < H; y++)
                sxo_subbackline(r->back, r->dback, nbx, nby, bs, W, y, r->sub + (size_t)y * W, NULL);
        }
        free(masked);
    }
    /* back.c:152-153: with BACK_TYPE MANUAL the reported backmean is BACK_VALUE */
    if (cfg->back_type == 1) r->backmean = (float)cfg->back_value;
    if (wtype == 3) {
        r->var = (float *)malloc(sizeof(float) * N);
        for (int y = 0; y < H; y++)
            sxo_backrmsline(r->sigma, r->dsigma, nbx, nby, bs, W, y, r->var + (size_t)y * W);
        for (size_t p = 0; p < N; p++) {
            const float v = r->var[p];
            r->var[p] = v < SXO_BIG ? v * v : v;
        }
    }
    const int haswt = r->var != NULL;
    /* WEIGHT_GAIN with a variance map (cudaanalyse.cu:2281) */
    const int gainflag = haswt && cfg->weight_gain;
    if (cfg->thresh_type == 1) {
        r->dthresh = (float)cfg->detect_thresh;
        r->thresh = (float)cfg->analysis_thresh;
    } else {
        r->dthresh = (float)(cfg->detect_thresh * r->backsig);
        r->thresh = (float)(cfg->analysis_thresh * r->backsig);
    }
    const float dthresh = r->dthresh, thresh = r->thresh;
    /* per-pixel thresholds: with a weight map, relative thresholds are in units of the
       local noise sqrt(var) */
    const int wrel = haswt && cfg->thresh_type != 1;
#define DTH(p) (wrel ? (float)(cfg->detect_thresh * sqrt((double)r->var[p])) : dthresh)
#define ATH(p) (wrel ? (float)(cfg->analysis_thresh * sqrt((double)r->var[p])) : thresh)

    STAGE_END(0);
    /* ---- filtering (filter.c:39-94) ---- */
    if (cfg->mask)
        sxo_convolve(r->sub, W, H, cfg->mask, cfg->mask_w, cfg->mask_h, r->cd);
    else
        memcpy(r->cd, r->sub, sizeof(float) * N);
    const float *cd = r->cd, *sub = r->sub;

    STAGE_END(1);
    /* ---- detection: 8-connected labels, root = smallest raster index ---- */
    int *par = (int *)malloc(sizeof(int) * N);
    long nabove = 0;
    for (size_t p = 0; p < N; p++) {
        if (cd[p] >= DTH(p)) {
            par[p] = (int)p;
            nabove++;
        } else
            par[p] = -1;
    }
    r->npix_above = nabove;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int p = y * W + x;
            if (par[p] < 0) continue;
            if (x > 0 && par[p - 1] >= 0) uf_union(par, p, p - 1);
            if (y > 0) {
                if (x > 0 && par[p - W - 1] >= 0) uf_union(par, p, p - W - 1);
                if (par[p - W] >= 0) uf_union(par, p, p - W);
                if (x < W - 1 && par[p - W + 1] >= 0) uf_union(par, p, p - W + 1);
            }
        }
    int *cnt = (int *)calloc(N, sizeof(int));
    for (size_t p = 0; p < N; p++)
        if (par[p] >= 0) cnt[uf_find(par, (int)p)]++;
    /* raw objects: components with >= DETECT_MINAREA pixels, in label order */
    int *rawid = (int *)malloc(sizeof(int) * N);
    int nraw = 0;
    for (size_t p = 0; p < N; p++) {
        rawid[p] = -1;
        if (par[p] == (int)p && cnt[p] >= cfg->ext_minarea) rawid[p] = nraw++;
    }
    r->nraw = nraw;
    r->raw_first = (int *)malloc(sizeof(int) * (nraw + 1));
    r->raw_count = (int *)malloc(sizeof(int) * (nraw + 1));
    r->raw_fdflux = (float *)malloc(sizeof(float) * (nraw + 1));
    r->raw_fdpeak = (float *)malloc(sizeof(float) * (nraw + 1));
    int *rawoff = (int *)malloc(sizeof(int) * (nraw + 1));
    {
        int k = 0;
        long off = 0;
        for (size_t p = 0; p < N; p++)
            if (rawid[p] >= 0) {
                r->raw_first[k] = (int)p;
                r->raw_count[k] = cnt[p];
                rawoff[k] = (int)off;
                off += cnt[p];
                k++;
            }
        rawoff[nraw] = (int)off;
        r->npix_raw = off;
    }
    /* FPIA0: pixels grouped by raw object, raster order inside (cudadetection.cu:677-882) */
    int *fpia = (int *)malloc(sizeof(int) * (r->npix_raw + 1));
    int *pixobj = (int *)malloc(sizeof(int) * N);   /* raw object of each pixel, -1 */
    {
        int *fill = (int *)calloc(nraw + 1, sizeof(int));
        for (size_t p = 0; p < N; p++) {
            pixobj[p] = -1;
            if (par[p] < 0) continue;
            int root = uf_find(par, (int)p);
            int k = rawid[root];
            if (k < 0) continue;
            pixobj[p] = k;
            fpia[rawoff[k] + fill[k]++] = (int)p;
        }
        free(fill);
    }
    free(cnt);
    /* fdpeak, fdflux per raw object (cudadetection.cu:899-1090; sequential raster sum);
       the object's base threshold is the lowest detection threshold among its pixels
       (the global one without a weight map) */
    float *rawt0 = (float *)malloc(sizeof(float) * (nraw + 1));
    for (int k = 0; k < nraw; k++) {
        float pk = (float)-SXO_BIG, fl = 0.0f, t0 = dthresh;
        for (int i = rawoff[k]; i < rawoff[k + 1]; i++) {
            float v = cd[fpia[i]];
            if (v > pk) pk = v;
            fl += v;
            if (wrel) {
                const float t = DTH(fpia[i]);
                if (i == rawoff[k] || t < t0) t0 = t;
            }
        }
        r->raw_fdpeak[k] = pk;
        r->raw_fdflux[k] = fl;
        rawt0[k] = t0;
    }

    STAGE_END(2);
    /* ---- deblending, preanalyse_full and gatherup, one root at a time ---- */
    const int nthr = cfg->deblend_nthresh;
    const int debarea = cfg->ext_minarea < MAXDEBAREA ? cfg->ext_minarea : MAXDEBAREA;
    int cap = nraw * 2 + 16, nobj = 0;
    pre_t *pre = (pre_t *)malloc(sizeof(pre_t) * cap);
    int *o_level = (int *)malloc(sizeof(int) * cap), *o_root = (int *)malloc(sizeof(int) * cap);
    int *o_first = (int *)malloc(sizeof(int) * cap);
    float *o_dthresh = (float *)malloc(sizeof(float) * cap);
    for (size_t p = 0; p < N; p++) r->alloc[p] = -1;
    float *thr = (float *)malloc(sizeof(float) * (nthr > 1 ? nthr : 1));
    float *thr_eff = (float *)malloc(sizeof(float) * (nthr > 1 ? nthr : 1));

    int gcap = 0;
    int *grid = NULL, *gid = NULL;
    int *loc = NULL, *lvl = NULL, *tmp = NULL;
    int locap = 0;
    comp_t *comps = NULL;
    int ccap = 0;

    for (int k = 0; k < nraw; k++) {
        const int n = rawoff[k + 1] - rawoff[k];
        const int *P = fpia + rawoff[k];
        int x0 = W, x1 = -1, y0 = P[0] / W, y1 = P[n - 1] / W;
        for (int i = 0; i < n; i++) {
            int x = P[i] % W;
            if (x < x0) x0 = x;
            if (x > x1) x1 = x;
        }
        const int gw = x1 - x0 + 1, gh = y1 - y0 + 1;
        if ((long)gw * gh > gcap) {
            gcap = gw * gh;
            grid = (int *)realloc(grid, sizeof(int) * gcap);
            gid = (int *)realloc(gid, sizeof(int) * gcap);
        }
        if (n * (nthr > 0 ? nthr : 1) > locap || n > locap) {
            locap = n * (nthr > 0 ? nthr : 1) + n;
            loc = (int *)realloc(loc, sizeof(int) * locap);
            lvl = (int *)realloc(lvl, sizeof(int) * locap);
            tmp = (int *)realloc(tmp, sizeof(int) * locap);
        }
        for (int i = 0; i < n; i++) loc[i] = (P[i] / W - y0) * gw + (P[i] % W - x0);
        for (int i = 0; i < gw * gh; i++) grid[i] = -1;

        /* sub-thresholds (cudadeblend.cu:199-217) */
        const double t0 = (double)rawt0[k], pk = (double)r->raw_fdpeak[k];
        float run = rawt0[k];
        for (int i = 1; i < nthr; i++) {
            thr[i] = (float)(t0 * sx_pow(pk / t0, (double)i / nthr));
            /* a pixel enters level i only through the kept pixels of level i-1, so the
               effective membership threshold is the running maximum */
            run = thr[i] > run ? thr[i] : run;
            thr_eff[i] = run;
        }
        /* component list: level 0 is the root itself */
        int nc = 0;
        if (ccap < 64) { ccap = 64; comps = (comp_t *)realloc(comps, sizeof(comp_t) * ccap); }
        comps[0].level = 0; comps[0].parent = -1; comps[0].first = P[0]; comps[0].count = n;
        comps[0].fdflux = r->raw_fdflux[k]; comps[0].dthresh = rawt0[k];
        nc = 1;
        /* lvl[l*n + i]: component of pixel i at level l (-1 if none) */
        for (int i = 0; i < n; i++) lvl[i] = 0;
        int nlev = 1;
        for (int l = 1; l < nthr; l++) {
            const int *prev = lvl + (size_t)(l - 1) * n;
            int *cur = lvl + (size_t)l * n;
            int any = 0;
            for (int i = 0; i < n; i++) {
                cur[i] = -1;
                if (prev[i] >= 0 && cd[P[i]] >= thr[l]) {
                    grid[loc[i]] = loc[i];
                    any = 1;
                } else
                    grid[loc[i]] = -1;
            }
            if (!any) break;
            for (int i = 0; i < n; i++) {
                int g = loc[i];
                if (grid[g] < 0) continue;
                int gx = g % gw, gy = g / gw;
                if (gx > 0 && grid[g - 1] >= 0) uf_union(grid, g, g - 1);
                if (gy > 0) {
                    if (gx > 0 && grid[g - gw - 1] >= 0) uf_union(grid, g, g - gw - 1);
                    if (grid[g - gw] >= 0) uf_union(grid, g, g - gw);
                    if (gx < gw - 1 && grid[g - gw + 1] >= 0) uf_union(grid, g, g - gw + 1);
                }
            }
            /* components in label order, counts and raster-order float flux */
            int base = nc;
            for (int i = 0; i < n; i++) {
                int g = loc[i];
                if (grid[g] < 0) continue;
                int rt = uf_find(grid, g);
                int c;
                if (rt == g) {
                    c = nc++;
                    if (nc > ccap) { ccap *= 2; comps = (comp_t *)realloc(comps, sizeof(comp_t) * ccap); }
                    gid[g] = c;
                    comps[c].level = l;
                    comps[c].parent = prev[i];
                    comps[c].first = P[i];
                    comps[c].count = 0;
                    comps[c].fdflux = 0.0f;
                    comps[c].dthresh = thr[l];
                } else
                    c = gid[rt];
                comps[c].count++;
                comps[c].fdflux += cd[P[i]];
                tmp[i] = c;
            }
            /* drop components below deb_minarea (cudadeblend.cu:748-958) */
            int *remap = (int *)malloc(sizeof(int) * (nc - base + 1));
            int w = base;
            for (int c = base; c < nc; c++) {
                if (comps[c].count >= debarea) {
                    remap[c - base] = w;
                    comps[w++] = comps[c];
                } else
                    remap[c - base] = -1;
            }
            for (int i = 0; i < n; i++) {
                int g = loc[i];
                if (grid[g] < 0) continue;
                cur[i] = remap[tmp[i] - base];
            }
            free(remap);
            nc = w;
            nlev = l + 1;
            if (w == base) break;
        }
        (void)thr_eff;

        /* tree pruning (cudadeblend.cu:422-483): bottom-up */
        for (int c = 0; c < nc; c++) { comps[c].nson = 0; comps[c].okA = 1; comps[c].kept = 0; comps[c].sig = 0; }
        const double value0 = (double)r->raw_fdflux[k] * cfg->deblend_mincont;
        for (int c = nc - 1; c >= 1; c--) {
            float lhs = comps[c].fdflux - comps[c].dthresh * (float)comps[c].count;
            comps[c].sig = (double)lhs > value0;
        }
        for (int c = nc - 1; c >= 1; c--) comps[comps[c].parent].nson += comps[c].sig;
        for (int c = nc - 1; c >= 0; c--) {
            if (comps[c].nson > 1) comps[c].okA = 0;
            if (c > 0 && !comps[c].okA) comps[comps[c].parent].okA = 0;
        }
        for (int c = 0; c < nc; c++) {
            if (c == 0)
                comps[c].kept = comps[c].okA;
            else
                comps[c].kept = comps[comps[c].parent].nson > 1 && comps[c].okA && comps[c].sig;
        }

        /* kept objects, in (level, label) order within the root */
        int objbase = nobj;
        for (int c = 0; c < nc; c++) {
            if (!comps[c].kept) continue;
            if (nobj >= cap) {
                cap *= 2;
                pre = (pre_t *)realloc(pre, sizeof(pre_t) * cap);
                o_level = (int *)realloc(o_level, sizeof(int) * cap);
                o_root = (int *)realloc(o_root, sizeof(int) * cap);
                o_first = (int *)realloc(o_first, sizeof(int) * cap);
                o_dthresh = (float *)realloc(o_dthresh, sizeof(float) * cap);
            }
            /* level pixels of c in raster order */
            const int l = comps[c].level;
            const int *cl = lvl + (size_t)l * n;
            int m = 0;
            for (int i = 0; i < n; i++)
                if (cl[i] == c) tmp[m++] = P[i];
            preanalyse(tmp, m, sub, cd, W, H, (double)comps[c].dthresh,
                       l > 0 ? OBJ_MERGED : 0, 0, 0.0, 0.0, &pre[nobj]);
            for (int i = 0; i < m; i++) r->alloc[tmp[i]] = nobj;
            o_level[nobj] = l;
            o_root[nobj] = k;
            o_first[nobj] = comps[c].first;
            o_dthresh[nobj] = comps[c].dthresh;
            nobj++;
        }
        (void)nlev;
        /* gatherup (cudadeblend.cu:1262-1360) */
        for (int i = 0; i < n; i++) {
            int p = P[i];
            if (r->alloc[p] >= 0) continue;
            int x = p % W, y = p / W;
            float distmin = (float)1e+31, dist, dx, dy;
            int iclst = objbase;
            for (int o = objbase; o < nobj; o++) {
                dx = (float)(x - pre[o].mx);
                dy = (float)(y - pre[o].my);
                dist = (float)(0.5 * (pre[o].cxx * dx * dx + pre[o].cyy * dy * dy + pre[o].cxy * dx * dy)
                               / pre[o].abcor);
                if (dist < distmin) {
                    distmin = dist;
                    iclst = o;
                }
            }
            r->alloc[p] = -2 - iclst;   /* gathered: resolved below, after every root */
        }
    }
    free(grid); free(gid); free(loc); free(lvl); free(tmp); free(comps); free(thr); free(thr_eff);
    /* level pixels: assigned by a kept component (gathered ones are still negative) */
    unsigned char *islev = (unsigned char *)calloc(N, 1);
    for (size_t p = 0; p < N; p++) islev[p] = r->alloc[p] >= 0;
    for (size_t p = 0; p < N; p++)
        if (r->alloc[p] <= -2) r->alloc[p] = -2 - r->alloc[p];

    /* ---- final pixel lists per object, raster order (cudadeblend.cu:1876-1935) ---- */
    r->nobj = nobj;
    r->objpix_start = (int *)calloc(nobj + 1, sizeof(int));
    r->objpix = (int *)malloc(sizeof(int) * (r->npix_raw + 1));
    for (size_t p = 0; p < N; p++)
        if (r->alloc[p] >= 0) r->objpix_start[r->alloc[p] + 1]++;
    for (int o = 0; o < nobj; o++) r->objpix_start[o + 1] += r->objpix_start[o];
    {
        int *fill = (int *)calloc(nobj + 1, sizeof(int));
        for (size_t p = 0; p < N; p++) {
            int o = r->alloc[p];
            if (o >= 0) r->objpix[r->objpix_start[o] + fill[o]++] = (int)p;
        }
        free(fill);
    }

    r->levpix_start = (int *)calloc(nobj + 1, sizeof(int));
    r->levpix = (int *)malloc(sizeof(int) * (r->npix_raw + 1));
    {
        int m = 0;
        for (int o = 0; o < nobj; o++) {
            r->levpix_start[o] = m;
            for (int i = r->objpix_start[o]; i < r->objpix_start[o + 1]; i++)
                if (islev[r->objpix[i]]) r->levpix[m++] = r->objpix[i];
        }
        r->levpix_start[nobj] = m;
    }
    free(islev);

    STAGE_END(3);
    /* ---- object arrays ---- */
#define ALLOC_I(f) r->f = (int *)calloc(nobj + 1, sizeof(int))
#define ALLOC_F(f) r->f = (float *)calloc(nobj + 1, sizeof(float))
#define ALLOC_D(f) r->f = (double *)calloc(nobj + 1, sizeof(double))
    ALLOC_I(level); ALLOC_I(root); ALLOC_I(firstpix); ALLOC_I(flag); ALLOC_I(fdnpix); ALLOC_I(dnpix);
    ALLOC_I(npix); ALLOC_I(xmin); ALLOC_I(xmax); ALLOC_I(ymin); ALLOC_I(ymax); ALLOC_I(peakx);
    ALLOC_I(peaky);
    r->iso = (int *)calloc((size_t)(nobj + 1) * NISO, sizeof(int));
    ALLOC_D(mx); ALLOC_D(my); ALLOC_D(mx2); ALLOC_D(my2); ALLOC_D(mxy); ALLOC_D(poserr_mx2);
    ALLOC_D(poserr_my2); ALLOC_D(poserr_mxy);
    ALLOC_F(a); ALLOC_F(b); ALLOC_F(theta); ALLOC_F(cxx); ALLOC_F(cyy); ALLOC_F(cxy); ALLOC_F(abcor);
    ALLOC_F(fdflux); ALLOC_F(dflux); ALLOC_F(flux); ALLOC_F(fluxerr); ALLOC_F(fdpeak); ALLOC_F(dpeak);
    ALLOC_F(peak); ALLOC_F(bkg); ALLOC_F(thresh_obj); ALLOC_F(mthresh); ALLOC_F(fwhm);
    ALLOC_F(dthresh_obj);
    r->master = (uint32_t *)malloc(sizeof(uint32_t) * (nobj + 1));

    const double backnoise2 = (double)(r->backsig * r->backsig);
    /* the mthresh order statistic uses the reference's EXT_MINAREA = 5 (cudatypes.h:62,
       cudaanalyse.cu:2322-2335), whatever DETECT_MINAREA is */
    const int hsize = 5;
    float *top = (float *)malloc(sizeof(float) * hsize);
    for (int o = 0; o < nobj; o++) {
        const int *pl = r->objpix + r->objpix_start[o];
        const int n = r->objpix_start[o + 1] - r->objpix_start[o];
        pre_t q = pre[o];
        /* the object's detection threshold: its raw object's base threshold (the global
           one without a weight map) */
        const float odth = rawt0[o_root[o]];
        /* preanalyse_robust for level > 0 (cudaanalyse.cu:1906-2126; dthresh reset to the
           detection threshold, cudadeblend.cu:1402-1413) */
        if (o_level[o] > 0)
            preanalyse(pl, n, sub, cd, W, H, (double)odth, pre[o].flag, 1, pre[o].mx,
                       pre[o].my, &q);
        r->level[o] = o_level[o];
        r->root[o] = o_root[o];
        r->firstpix[o] = o_first[o];
        r->fdnpix[o] = n;
        r->dnpix[o] = q.dnpix;
        r->xmin[o] = q.xmin; r->xmax[o] = q.xmax; r->ymin[o] = q.ymin; r->ymax[o] = q.ymax;
        r->mx[o] = q.mx; r->my[o] = q.my; r->mx2[o] = q.mx2; r->my2[o] = q.my2; r->mxy[o] = q.mxy;
        r->a[o] = q.a; r->b[o] = q.b; r->theta[o] = q.theta;
        r->cxx[o] = q.cxx; r->cyy[o] = q.cyy; r->cxy[o] = q.cxy; r->abcor[o] = q.abcor;
        r->fdflux[o] = q.fdflux; r->dflux[o] = q.dflux; r->fdpeak[o] = q.fdpeak; r->dpeak[o] = q.dpeak;
        r->dthresh_obj[o] = odth;

        /* examineiso (cudaanalyse.cu:2176-2581), CCD mode, GLOBAL background */
        float bkg = sxo_back_at(r->back, (int)(q.mx + 0.5), (int)(q.my + 0.5), nbx, nby, bs, bs);
        float dbkg = 0.0f;
        double tv = 0.0, sigtv = 0.0, esum = 0.0, emx2 = 0.0, emy2 = 0.0, emxy = 0.0, err, var2;
        double xm = q.mx, ym = q.my, x, y;
        const double var = backnoise2, dbacknoise2 = backnoise2;
        float peak = (float)-SXO_BIG, cdpeak = (float)-SXO_BIG, pix, cdpix;
        int area = 0, peakx = 0, peaky = 0, nfill = 0;
        float othr = thresh;   /* the object's analysis threshold: lowest among its pixels */
        int cleanflag = cfg->clean;
        if (n < hsize) cleanflag = 0;
        for (int i = 0; i < n; i++) {
            int p = pl[i];
            x = p % W;
            y = p / W;
            pix = sub[p];
            if (pix > peak) peak = pix;
            cdpix = cd[p];
            if (cdpix > cdpeak) {
                cdpeak = cdpix;
                peakx = (int)x + 1;
                peaky = (int)y + 1;
            }
            /* per-pixel variance with a weight map, else the background's */
            var2 = haswt ? (double)r->var[p] : var;
            /* cudaanalyse.cu:2354-2355; var is the background noise squared */
            if (gainflag && pix > 0.0 && cfg->gain > 0.0) var2 += pix / cfg->gain * r->var[p] / var;
            sigtv += var2;
            const float athr = ATH(p);
            if (i == 0 || athr < othr) othr = athr;
            if (pix > athr) area++;
            tv += pix;
            err = haswt ? (double)r->var[p] : dbacknoise2;
            if (cfg->gain > 0.0 && cdpix > 0.0) err += cdpix / cfg->gain;
            x = x - xm;
            y = y - ym;
            esum += err;
            emx2 += err * x * x;
            emy2 += err * y * y;
            emxy += err * x * y;
            if (cleanflag) top_insert(top, hsize, &nfill, (float)(cdpix - odth));
        }
        r->mthresh[o] = cleanflag ? top[0] : 0.0f;
        {
            double flux2 = (float)(q.fdflux * q.fdflux);
            emx2 /= flux2;
            emy2 /= flux2;
            emxy /= flux2;
            r->poserr_mx2[o] = emx2;
            r->poserr_my2[o] = emy2;
            r->poserr_mxy[o] = emxy;
        }
        tv -= n * dbkg;
        if (!gainflag && cfg->gain > 0.0 && tv > 0.0) sigtv += tv / cfg->gain;
        r->npix[o] = area;
        r->flux[o] = (float)tv;
        r->fluxerr[o] = (float)sigtv;
        r->thresh_obj[o] = othr - dbkg;
        r->peak[o] = peak - dbkg;
        r->peakx[o] = peakx;
        r->peaky[o] = peaky;
        r->bkg[o] = bkg;
        {
            float threshs[NISO];
            float tho = r->thresh_obj[o], pko = r->peak[o];
            if (pko > 0.0 && tho > 0.0)
                for (int i = 0; i < NISO; i++)
                    threshs[i] = (float)(tho * sx_pow((double)pko / tho, (double)i / NISO));
            else
                for (int i = 0; i < NISO; i++) threshs[i] = 0.0f;
            int *iso = r->iso + (size_t)o * NISO;
            for (int j = 0; j < n; j++) {
                float v = sub[pl[j]];
                for (int i = 0; i < NISO && v > threshs[i]; i++) iso[i]++;
            }
        }
        {
            float thresh0 = (float)(r->peak[o] / 5.0);
            if (thresh0 < r->thresh_obj[o]) thresh0 = r->thresh_obj[o];
            if (thresh0 > 0.0) {
                double mxo = q.mx, myo = q.my, s, sx, sy, sxx, sxy, dx, dy, d2, lpix, dpix, bb,
                       inverr2, sat, d, bmax;
                sat = (double)(cfg->satur_level - dbkg);
                s = sx = sy = sxx = sxy = 0.0;
                for (int j = 0; j < n; j++) {
                    int p = pl[j];
                    dpix = sub[p] - (double)dbkg;
                    if (dpix > thresh0 && dpix < sat) {
                        dx = (p % W) - mxo;
                        dy = (p / W) - myo;
                        lpix = sx_log(dpix);
                        inverr2 = dpix * dpix;
                        s += inverr2;
                        d2 = dx * dx + dy * dy;
                        sx += d2 * inverr2;
                        sxx += d2 * d2 * inverr2;
                        sy += lpix * inverr2;
                        sxy += lpix * d2 * inverr2;
                    }
                }
                d = s * sxx - sx * sx;
                if (fabs(d) > 0.0) {
                    bb = -(s * sxy - sx * sy) / d;
                    if (bb < (bmax = 1 / (13 * q.a * q.b))) bb = bmax;
                    r->fwhm[o] = (float)(1.6651 / sqrt(bb));
                    if (r->fwhm[o] > 0.0) r->fwhm[o] -= 1 / (4 * r->fwhm[o]);
                } else
                    r->fwhm[o] = 0.0f;
            } else
                r->fwhm[o] = 0.0f;
        }
        /* endobject: saturation flag (cudaanalyse.cu:2873-2874) */
        int flag = q.flag;
        if (r->peak[o] + r->bkg[o] >= cfg->satur_level) flag |= OBJ_SATUR;
        r->flag[o] = flag;
    }
    free(top);

    STAGE_END(4);
    /* ---- CLEAN (cudaclean.cu:10-29, :120-174) ---- */
    for (int o = 0; o < nobj; o++) r->master[o] = 0xffffffffu;
    if (cfg->clean && nobj > 1) {
        double *amp = (double *)malloc(sizeof(double) * nobj), *alpha = (double *)malloc(sizeof(double) * nobj);
        for (int o = 0; o < nobj; o++) {
            float unitarea = (float)(SXO_PI * r->a[o] * r->b[o]);
            double ampin = amp[o] = r->fdflux[o] / (2 * unitarea * r->abcor[o]);
            alpha[o] = (sx_pow(ampin / r->dthresh_obj[o], 1.0 / cfg->clean_param) - 1) * unitarea
                       / r->fdnpix[o];
        }
        const int win = ((cfg->clean_stacksize - 1) / 16 + 1) * 16;
        for (int i = 0; i < nobj; i++)
            for (int kk = 0; kk < win; kk++) {
                int j = i + kk + 1;
                if (j >= nobj) break;
                float dx = (float)(r->mx[i] - r->mx[j]);
                float dy = (float)(r->my[i] - r->my[j]);
                float rlim = r->a[i] + r->a[j];
                double val;
                rlim *= rlim;
                if (dx * dx + dy * dy < rlim * CLEAN_ZONE * CLEAN_ZONE) {
                    if (r->fdflux[i] < r->fdflux[j]) {
                        val = 1 + alpha[j] * (r->cxx[j] * dx * dx + r->cyy[j] * dy * dy + r->cxy[j] * dx * dy);
                        if (val > 1.0 && ((float)(val < 1e10 ? amp[j] * (1.0 / val) : 0.0) > r->mthresh[i]))
                            if ((uint32_t)j < r->master[i]) r->master[i] = (uint32_t)j;
                    } else {
                        val = 1 + alpha[i] * (r->cxx[i] * dx * dx + r->cyy[i] * dy * dy + r->cxy[i] * dx * dy);
                        if (val > 1.0 && ((float)(val < 1e10 ? amp[i] * (1.0 / val) : 0.0) > r->mthresh[j]))
                            if ((uint32_t)i < r->master[j]) r->master[j] = (uint32_t)i;
                    }
                }
            }
        free(amp);
        free(alpha);
    }

    STAGE_END(5);
    /* ---- segmentation map: survivors numbered 1.. in object order ---- */
    {
        int *num = (int *)malloc(sizeof(int) * (nobj + 1));
        int ns = 0;
        for (int o = 0; o < nobj; o++) num[o] = (r->master[o] == 0xffffffffu) ? ++ns : 0;
        r->nsurv = ns;
        for (size_t p = 0; p < N; p++) {
            int o = r->alloc[p];
            r->seg[p] = o >= 0 ? num[o] : 0;
        }
        free(num);
    }

    free(pre); free(o_level); free(o_root); free(o_first); free(o_dthresh); free(rawt0);
    free(par); free(rawid); free(rawoff); free(fpia); free(pixobj);
    STAGE_END(6);
#undef STAGE_END
    return r;
}

void sxo_free(sxo_result *r)
{
    if (!r) return;
    void *ptrs[] = {r->levpix_start, r->levpix, r->back, r->sigma, r->dback, r->dsigma, r->sub,
                    r->cd, r->var, r->raw_first,
                    r->raw_count, r->raw_fdflux, r->raw_fdpeak, r->level, r->root, r->firstpix,
                    r->flag, r->fdnpix, r->dnpix, r->npix, r->xmin, r->xmax, r->ymin, r->ymax,
                    r->peakx, r->peaky, r->iso, r->mx, r->my, r->mx2, r->my2, r->mxy,
                    r->poserr_mx2, r->poserr_my2, r->poserr_mxy, r->a, r->b, r->theta, r->cxx,
                    r->cyy, r->cxy, r->abcor, r->fdflux, r->dflux, r->flux, r->fluxerr, r->fdpeak,
                    r->dpeak, r->peak, r->bkg, r->thresh_obj, r->mthresh, r->fwhm, r->dthresh_obj,
                    r->master, r->alloc, r->seg, r->objpix_start, r->objpix};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(r);
}

int sxo_measure_fp64(const sxo_result *r, double *out, int ncol)
{
    if (ncol != 12 || !r->levpix) return -1;
    const int W = r->width;
    for (int o = 0; o < r->nobj; o++) {
        typedef long double LD;
        const int *F = r->objpix + r->objpix_start[o];
        const int nf = r->objpix_start[o + 1] - r->objpix_start[o];
        const int *L = r->levpix + r->levpix_start[o];
        const int nl = r->levpix_start[o + 1] - r->levpix_start[o];
        /* barycentre of the level pixels, weights cd */
        LD s = 0, sx = 0, sy = 0;
        for (int i = 0; i < nl; i++) {
            const LD c = r->cd[L[i]], x = L[i] % W, y = L[i] / W;
            s += c;
            sx += c * x;
            sy += c * y;
        }
        const LD xb = sx / s, yb = sy / s;
        /* central moments about it over the final pixels */
        LD t = 0, x2 = 0, y2 = 0, xy = 0, flux = 0;
        for (int i = 0; i < nf; i++) {
            const LD c = r->cd[F[i]], dx = (LD)(F[i] % W) - xb, dy = (LD)(F[i] / W) - yb;
            t += c;
            x2 += c * dx * dx;
            y2 += c * dy * dy;
            xy += c * dx * dy;
            flux += r->sub[F[i]];
        }
        x2 /= t;
        y2 /= t;
        xy /= t;
        LD temp2 = x2 * y2 - xy * xy;
        if (temp2 < 0.00694L) {
            x2 += 0.0833333L;
            y2 += 0.0833333L;
            temp2 = x2 * y2 - xy * xy;
        }
        const LD d = x2 - y2;
        const LD theta = fabsl(d) > 0 ? atan2l(2 * xy, d) / 2 : 3.14159265358979323846L / 4;
        const LD temp = sqrtl(0.25L * d * d + xy * xy);
        const LD pm = 0.5L * (x2 + y2);
        double *q = out + 12 * (size_t)o;
        q[0] = (double)flux;
        q[1] = (double)xb;
        q[2] = (double)yb;
        q[3] = (double)x2;
        q[4] = (double)y2;
        q[5] = (double)xy;
        q[6] = (double)sqrtl(pm + temp);
        q[7] = (double)sqrtl(pm - temp);
        q[8] = (double)(theta * 180 / 3.14159265358979323846L);
        q[9] = (double)(y2 / temp2);
        q[10] = (double)(x2 / temp2);
        q[11] = (double)(-2 * xy / temp2);
    }
    return r->nobj;
}
