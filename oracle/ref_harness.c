/*
 * ref_harness.c -- drives the reference's own CPU background and filter code, compiled
 * from the reference tree (src/back.c, src/filter.c, src/misc.c, src/globals.c,
 * src/fits/fitsmisc.c), to produce golden vectors for the oracle.
 * TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ by `make -C oracle ref` and loaded
 * only by tests/golden/make_golden.py.
 *
 * The entry points below set up the reference data structures (picstruct, backstruct,
 * filterstruct) for a whole frame held in memory and call the reference functions
 * unchanged: backstat/backhisto/backguess (back.c:214-426) once per mesh row, exactly as
 * stock makeback does when a mesh row fits BACK_BUFSIZE; filterback (back.c:434-560);
 * makebackspline (back.c:739-780); subbackline (back.c:790-921); convolve (filter.c:39)
 * and getconv (filter.c:101); hmedian (misc.c:30); and the world coordinates of
 * computeastrom (astrom.c:97-120): create_wcs/init_wcs (fitswcs.c:99-252), raw_to_wcs
 * (fitswcs.c:1251) with the reference's WCSLIB (src/wcs/) and precess (fitswcs.c:1718).
 */
#include "config.h"
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "define.h"
#include "globals.h"
#include "prefs.h"
#include "fits/fitscat.h"
#include "back.h"
#include "field.h"
#include "filter.h"
#include "fitswcs.h"

/* per-mesh statistics: mean sigma lcut hcut qzero qscale nlevels npix, then mode/sigma */
int ref_makeback(const float *img, int W, int H, int bs, float *stats, float *back, float *sigma)
{
    const int nbx = (W - 1) / bs + 1, nby = (H - 1) / bs + 1;
    backstruct *bm = (backstruct *)calloc(nbx, sizeof(backstruct));
    if (!bm) return -1;
    for (int my = 0; my < nby; my++) {
        int h = (my * bs + bs <= H) ? bs : H - my * bs;
        PIXTYPE *buf = (PIXTYPE *)(img + (size_t)my * bs * W);
        backstat(bm, buf, (size_t)W * h, nbx, W, bs);
        for (int mx = 0; mx < nbx; mx++) {
            bm[mx].histo = (LONG *)calloc(bm[mx].nlevels > 0 ? bm[mx].nlevels : 1, sizeof(LONG));
        }
        backhisto(bm, buf, (size_t)W * h, nbx, W, bs);
        for (int mx = 0; mx < nbx; mx++) {
            float *st = stats + 8 * (size_t)(my * nbx + mx);
            st[0] = bm[mx].mean; st[1] = bm[mx].sigma; st[2] = bm[mx].lcut; st[3] = bm[mx].hcut;
            st[4] = bm[mx].qzero; st[5] = bm[mx].qscale; st[6] = (float)bm[mx].nlevels;
            st[7] = (float)bm[mx].npix;
            backguess(&bm[mx], &back[my * nbx + mx], &sigma[my * nbx + mx]);
            free(bm[mx].histo);
            memset(&bm[mx], 0, sizeof(backstruct));
        }
    }
    free(bm);
    return 0;
}

static void setup_field(picstruct *f, int nbx, int nby, int nfx, int nfy)
{
    memset(f, 0, sizeof(*f));
    f->nbackx = nbx; f->nbacky = nby; f->nback = nbx * nby;
    f->nbackfx = nfx; f->nbackfy = nfy;
    f->flags = DETECT_FIELD;
}

int ref_filterback(float *back, float *sigma, int nbx, int nby, int nfx, int nfy, double fthresh,
                   float *backmean, float *backsig)
{
    picstruct f;
    setup_field(&f, nbx, nby, nfx, nfy);
    f.back = back;
    f.sigma = sigma;
    prefs.backfthresh = fthresh;
    filterback(&f);
    *backmean = f.backmean;
    *backsig = f.backsig;
    return 0;
}

int ref_makebackspline(const float *map, int nbx, int nby, float *dmap)
{
    picstruct f;
    setup_field(&f, nbx, nby, 1, 1);
    float *d = makebackspline(&f, (float *)map);
    memcpy(dmap, d, sizeof(float) * nbx * nby);
    free(d);
    return 0;
}

/* subtract the interpolated background from every row of img (in place) */
int ref_subback(const float *back, const float *dback, int nbx, int nby, int bs, int W, int H,
                float *img)
{
    picstruct f;
    setup_field(&f, nbx, nby, 1, 1);
    f.width = W; f.height = H;
    f.backw = f.backh = bs;
    f.back = (float *)back;
    f.dback = (float *)dback;
    f.back_type = BACK_RELATIVE;
    f.backline = (PIXTYPE *)malloc(sizeof(PIXTYPE) * W);
    for (int y = 0; y < H; y++) subbackline(&f, y, img + (size_t)y * W);
    free(f.backline);
    return 0;
}

/* the interpolated background rms of every row (back.c:938-1051) */
int ref_backrms(const float *sigma, const float *dsigma, int nbx, int nby, int bs, int W, int H,
                float *out)
{
    picstruct f;
    setup_field(&f, nbx, nby, 1, 1);
    f.width = W; f.height = H;
    f.backw = f.backh = bs;
    f.sigma = (float *)sigma;
    f.dsigma = (float *)dsigma;
    for (int y = 0; y < H; y++) backrmsline(&f, y, out + (size_t)y * W);
    return 0;
}

int ref_convolve(const float *img, int W, int H, const float *mask, int mw, int mh, float *out)
{
    picstruct f;
    filterstruct filt;
    memset(&f, 0, sizeof(f));
    memset(&filt, 0, sizeof(filt));
    filt.conv = (float *)mask;
    filt.convw = mw; filt.convh = mh; filt.nconv = mw * mh;
    thefilter = &filt;
    f.width = W; f.height = H;
    f.strip = (PIXTYPE *)img;
    f.stripheight = H;
    f.ymin = 0; f.ymax = H;
    for (int y = 0; y < H; y++) {
        f.y = y;
        convolve(&f, out + (size_t)y * W);
    }
    thefilter = NULL;
    return 0;
}

int ref_getconv(const char *path, float *mask, int maxn, int *mw, int *mh)
{
    filterstruct filt;
    memset(&filt, 0, sizeof(filt));
    thefilter = &filt;
    if (getconv((char *)path) != RETURN_OK) { thefilter = NULL; return -1; }
    int n = filt.convw * filt.convh;
    if (n > maxn) { free(filt.conv); thefilter = NULL; return -2; }
    memcpy(mask, filt.conv, sizeof(float) * n);
    *mw = filt.convw; *mh = filt.convh;
    free(filt.conv);
    thefilter = NULL;
    return 0;
}

float ref_hmedian(float *ra, int n) { return hmedian(ra, n); }

/* a wcsstruct built by the reference's create_wcs from header values, with the CD matrix,
   LONPOLE/LATPOLE and EQUINOX set and init_wcs run again (as read_wcs leaves it) */
void *ref_wcs_new(const char *ctype1, const char *ctype2, const double *crval,
                  const double *crpix, const double *cd, double longpole, double latpole,
                  double equinox)
{
    char c1[9] = {0}, c2[9] = {0};
    char *ct[2] = {c1, c2};
    int naxisn[2] = {4096, 4096};
    strncpy(c1, ctype1, 8);
    strncpy(c2, ctype2, 8);
    wcsstruct *w = create_wcs(ct, (double *)crval, (double *)crpix, NULL, naxisn, 2);
    if (!w) return NULL;
    for (int k = 0; k < 4; k++) w->cd[k] = cd[k];
    w->longpole = longpole;
    w->latpole = latpole;
    w->epoch = w->equinox = equinox;
    init_wcs(w);
    return w;
}

/* astrom.c:97-120: world coordinates (mxw, myw) and J2000 (alpha2000, delta2000) of n
   1-based pixel positions; -1 when a position has no mapping */
int ref_wcs_world(void *wv, int n, const double *pix, double *world, double *radec)
{
    wcsstruct *wcs = (wcsstruct *)wv;
    const int lng = wcs->lng, lat = wcs->lat;
    for (int i = 0; i < n; i++) {
        double raw[2] = {pix[2 * i], pix[2 * i + 1]}, wp[2];
        if (raw_to_wcs(wcs, raw, wp)) return -1;
        world[2 * i] = wp[0];
        world[2 * i + 1] = wp[1];
        if (fabs(wcs->equinox - 2000.0) > 0.003)
            precess(wcs->equinox, wp[lng < lat ? 0 : 1], wp[lng < lat ? 1 : 0], 2000.0,
                    &radec[2 * i], &radec[2 * i + 1]);
        else {
            radec[2 * i] = lng < lat ? wp[0] : wp[1];
            radec[2 * i + 1] = lng < lat ? wp[1] : wp[0];
        }
    }
    return 0;
}

/* offsets of the wcsstruct fields a GPU astrometry reads (types.h:547-592) */
int ref_wcs_layout(long *out, int n)
{
    const long v[] = {(long)sizeof(wcsstruct), (long)offsetof(wcsstruct, naxis),
                      (long)offsetof(wcsstruct, ctype), (long)offsetof(wcsstruct, crval),
                      (long)offsetof(wcsstruct, crpix), (long)offsetof(wcsstruct, cd),
                      (long)offsetof(wcsstruct, longpole), (long)offsetof(wcsstruct, latpole),
                      (long)offsetof(wcsstruct, lat), (long)offsetof(wcsstruct, lng),
                      (long)offsetof(wcsstruct, equinox)};
    const int nv = (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n && i < nv; i++) out[i] = v[i];
    return nv;
}
