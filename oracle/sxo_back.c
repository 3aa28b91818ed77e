/*
 * sxo_back.c -- oracle restatement of the reference CPU background code.
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the CPU-baseline leg
 * of the benchmark, never by the product.
 *
 * Follows src/back.c (backstat :214-290, backhisto :297-341, backguess :347-426,
 * filterback :434-560, makebackspline :739-780, subbackline :790-921) and
 * src/misc.c:30-72 (hmedian), with whole meshes as the reference GPU reads them
 * (src/cuda/cudaback.cu:1212-1390): the mesh grid is nbx = (W-1)/bs+1 by
 * nby = (H-1)/bs+1, and the last column/row of meshes is truncated at the frame edge.
 *
 * Deliberate, documented deviations:
 *   - backguess uses fabs(sig) (stock SExtractor and cudaback.cu:413) where this fork's
 *     back.c:406 wrote the integer abs();
 *   - histogram bins are range-checked in float before the int conversion, which gives
 *     the reference's results for every finite value and defines NaN/overflow;
 *   - out-of-range lcut/hcut conversions saturate (the GPU conversion rule) instead of
 *     being undefined, and histogram reads outside [0, nlevels) read 0.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sxo.h"

#define BACK_MINGOODFRAC 0.5
#define QUANTIF_NSIGMA 5
#define QUANTIF_NMAXLEVELS 4096
#define QUANTIF_AMIN 4
#define SXO_PI 3.1415926535898     /* src/define.h:56 */
#define SXO_EPS (1e-4)             /* src/back.c:345 */

/* saturating double -> int conversion (defined for NaN and out-of-range values) */
static int d2i_sat(double v)
{
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (int)-2147483647 - 1;
    return (int)v;
}

/* back.c:214-290 for one mesh; pixels are visited in raster order inside the mesh */
void sxo_backstat_mesh(const float *img, int W, int x0, int y0, int bw, int bh, sxo_mesh *bm)
{
    double pix, sig, mean, sigma;
    const double step = sqrt(2 / SXO_PI) * QUANTIF_NSIGMA / QUANTIF_AMIN;
    float lcut, hcut;
    int npix = 0;

    mean = sigma = 0.0;
    for (int y = 0; y < bh; y++) {
        const float *r = img + (size_t)(y0 + y) * W + x0;
        for (int x = 0; x < bw; x++)
            if ((pix = r[x]) > -SXO_BIG) {
                mean += pix;
                sigma += pix * pix;
                npix++;
            }
    }
    memset(bm, 0, sizeof(*bm));
    if ((float)npix < (float)(bw * bh * BACK_MINGOODFRAC)) {
        bm->mean = bm->sigma = (float)-SXO_BIG;
        return;
    }
    mean /= (double)npix;
    sigma = (sig = sigma / npix - mean * mean) > 0.0 ? sqrt(sig) : 0.0;
    lcut = bm->lcut = (float)(mean - 2.0 * sigma);
    hcut = bm->hcut = (float)(mean + 2.0 * sigma);

    mean = sigma = 0.0;
    npix = 0;
    for (int y = 0; y < bh; y++) {
        const float *r = img + (size_t)(y0 + y) * W + x0;
        for (int x = 0; x < bw; x++) {
            pix = r[x];
            if (pix <= hcut && pix >= lcut) {
                mean += pix;
                sigma += pix * pix;
                npix++;
            }
        }
    }
    bm->npix = npix;
    mean /= (double)npix;
    sig = sigma / npix - mean * mean;
    sigma = sig > 0.0 ? sqrt(sig) : 0.0;
    bm->mean = (float)mean;
    bm->sigma = (float)sigma;
    if ((bm->nlevels = (int)(step * npix + 1)) > QUANTIF_NMAXLEVELS)
        bm->nlevels = QUANTIF_NMAXLEVELS;
    bm->qscale = (float)(sigma > 0.0 ? 2 * QUANTIF_NSIGMA * sigma / bm->nlevels : 1.0);
    bm->qzero = (float)(mean - QUANTIF_NSIGMA * sigma);
}

/* back.c:297-341 for one mesh; histo must hold nlevels zeroed ints */
void sxo_backhisto_mesh(const float *img, int W, int x0, int y0, int bw, int bh,
                        const sxo_mesh *bm, int *histo)
{
    if (bm->mean <= -SXO_BIG) return;
    const int nlevels = bm->nlevels;
    const float qscale = bm->qscale;
    const float cste = (float)(0.499999 - bm->qzero / qscale);
    for (int y = 0; y < bh; y++) {
        const float *r = img + (size_t)(y0 + y) * W + x0;
        for (int x = 0; x < bw; x++) {
            float t = r[x] / qscale + cste;
            if (t > -1.0f && t < (float)nlevels) histo[(int)t]++;
        }
    }
}

static long hget(const int *h, int n, long i) { return (i >= 0 && i < n) ? h[i] : 0; }

/* back.c:347-426 (mode of the clipped histogram) */
float sxo_backguess(const sxo_mesh *bkg, const int *histo, float *mean, float *sigma)
{
    unsigned long lowsum, highsum, sum;
    double ftemp, mea, sig, sig1, med, dpix;
    int i, n, lcut, hcut, nlevelsm1, pix;
    const int nlev = bkg->nlevels;

    if (bkg->mean <= -SXO_BIG) {
        *mean = *sigma = (float)-SXO_BIG;
        return (float)-SXO_BIG;
    }
    hcut = nlevelsm1 = nlev - 1;
    lcut = 0;
    sig = 10.0 * nlevelsm1;
    sig1 = 1.0;
    mea = med = bkg->mean;
    for (n = 100; n-- && (sig >= 0.1) && (fabs(sig / sig1 - 1.0) > SXO_EPS);) {
        long lo, hi, t;
        sig1 = sig;
        sum = 0;
        mea = sig = 0.0;
        lowsum = highsum = 0;
        t = lo = lcut;
        hi = hcut;
        for (i = lcut; i <= hcut; i++) {
            if (lowsum < highsum)
                lowsum += hget(histo, nlev, lo++);
            else
                highsum += hget(histo, nlev, hi--);
            sum += (pix = (int)hget(histo, nlev, t++));
            mea += (dpix = (double)pix * i);
            sig += dpix * i;
        }
        if (hi >= 0) {
            long hl = hget(histo, nlev, lo), hh = hget(histo, nlev, hi);
            med = hi + 0.5 + ((double)highsum - lowsum) / (2.0 * (hl > hh ? hl : hh));
        } else
            med = 0.0;
        if (sum) {
            mea /= (double)sum;
            sig = sig / sum - mea * mea;
        }
        sig = sig > 0.0 ? sqrt(sig) : 0.0;
        lcut = (ftemp = med - 3.0 * sig) > 0.0 ? d2i_sat(ftemp > 0.0 ? ftemp + 0.5 : ftemp - 0.5) : 0;
        hcut = (ftemp = med + 3.0 * sig) < nlevelsm1
                   ? d2i_sat(ftemp > 0.0 ? ftemp + 0.5 : ftemp - 0.5)
                   : nlevelsm1;
    }
    if (fabs(sig) > 0.0) {
        if (fabs(bkg->sigma / (sig * bkg->qscale) - 1) < 0.0)
            *mean = (float)(bkg->qzero + mea * bkg->qscale);
        else if (fabs((mea - med) / sig) < 0.3)
            *mean = (float)(bkg->qzero + (2.5 * med - 1.5 * mea) * bkg->qscale);
        else
            *mean = (float)(bkg->qzero + med * bkg->qscale);
    } else
        *mean = (float)(bkg->qzero + mea * bkg->qscale);
    *sigma = (float)(sig * bkg->qscale);
    return *mean;
}

/* misc.c:30-72: Numerical Recipes heapsort median; sorts ra[] in place */
float sxo_hmedian(float *ra, int n)
{
    int l, j, ir, i;
    float rra;
    if (n < 2) return *ra;
    ra--;
    for (l = ((ir = n) >> 1) + 1;;) {
        if (l > 1)
            rra = ra[--l];
        else {
            rra = ra[ir];
            ra[ir] = ra[1];
            if (--ir == 1) {
                ra[1] = rra;
                return n & 1 ? ra[n / 2 + 1] : (float)((ra[n / 2] + ra[n / 2 + 1]) / 2.0);
            }
        }
        for (j = (i = l) << 1; j <= ir;) {
            if (j < ir && ra[j] < ra[j + 1]) ++j;
            if (rra < ra[j]) {
                ra[i] = ra[j];
                j += (i = j);
            } else
                j = ir + 1;
        }
        ra[i] = rra;
    }
}

/* Mesh statistics, histograms and mode for every mesh (back.c:43-208 with whole meshes,
   cudaback.cu:1278-1287); back/sigma get the unfiltered mode and sigma maps. */
int sxo_makeback(const float *img, int W, int H, int bs, int nfx, int nfy, double fthresh,
                 float *back, float *sigma, float *backmean, float *backsig)
{
    (void)nfx; (void)nfy; (void)fthresh; (void)backmean; (void)backsig;
    const int nbx = (W - 1) / bs + 1, nby = (H - 1) / bs + 1;
    int *histo = (int *)malloc(QUANTIF_NMAXLEVELS * sizeof(int));
    if (!histo) return -1;
    for (int my = 0; my < nby; my++)
        for (int mx = 0; mx < nbx; mx++) {
            int x0 = mx * bs, y0 = my * bs;
            int bw = (x0 + bs <= W) ? bs : W - x0;
            int bh = (y0 + bs <= H) ? bs : H - y0;
            sxo_mesh m;
            sxo_backstat_mesh(img, W, x0, y0, bw, bh, &m);
            memset(histo, 0, QUANTIF_NMAXLEVELS * sizeof(int));
            sxo_backhisto_mesh(img, W, x0, y0, bw, bh, &m, histo);
            sxo_backguess(&m, histo, &back[my * nbx + mx], &sigma[my * nbx + mx]);
        }
    free(histo);
    return 0;
}

/* back.c:434-560: bad-mesh interpolation, median filtering, global medians */
void sxo_filterback(float *back, float *sigma, int nx, int ny, int nfx, int nfy,
                    double backfthresh, float *backmean, float *backsig)
{
    float *back2, *sigma2, *bmask, *smask, *sigmat, d2, d2min, fthresh, med, val, sval;
    int i, j, px, py, np, npx, npx2, npy, npy2, dpx, dpy, x, y, nmin;

    fthresh = (float)backfthresh;
    np = nx * ny;
    npx = nfx / 2;
    npy = nfy / 2;
    npy *= nx;
    bmask = (float *)malloc(sizeof(float) * (2 * npx + 1) * (2 * (nfy / 2) + 1));
    smask = (float *)malloc(sizeof(float) * (2 * npx + 1) * (2 * (nfy / 2) + 1));
    back2 = (float *)malloc(sizeof(float) * np);
    sigma2 = (float *)malloc(sizeof(float) * np);
    val = sval = 0.0;

    for (i = 0, py = 0; py < ny; py++)
        for (px = 0; px < nx; px++, i++)
            if ((back2[i] = back[i]) <= -SXO_BIG) {
                d2min = (float)SXO_BIG;
                nmin = 0;
                for (j = 0, y = 0; y < ny; y++)
                    for (x = 0; x < nx; x++, j++)
                        if (back[j] > -SXO_BIG) {
                            d2 = (float)(x - px) * (x - px) + (y - py) * (y - py);
                            if (d2 < d2min) {
                                val = back[j];
                                sval = sigma[j];
                                nmin = 1;
                                d2min = d2;
                            } else if (d2 == d2min) {
                                val += back[j];
                                sval += sigma[j];
                                nmin++;
                            }
                        }
                back2[i] = nmin ? val / nmin : 0.0f;
                sigma[i] = nmin ? sval / nmin : 1.0f;
            }
    memcpy(back, back2, (size_t)np * sizeof(float));

    for (py = 0; py < np; py += nx) {
        npy2 = np - py - nx;
        if (npy2 > npy) npy2 = npy;
        if (npy2 > py) npy2 = py;
        for (px = 0; px < nx; px++) {
            npx2 = nx - px - 1;
            if (npx2 > npx) npx2 = npx;
            if (npx2 > px) npx2 = px;
            i = 0;
            for (dpy = -npy2; dpy <= npy2; dpy += nx) {
                y = py + dpy;
                for (dpx = -npx2; dpx <= npx2; dpx++) {
                    x = px + dpx;
                    bmask[i] = back[x + y];
                    smask[i++] = sigma[x + y];
                }
            }
            if (fabs((med = sxo_hmedian(bmask, i)) - back[px + py]) >= fthresh) {
                back2[px + py] = med;
                sigma2[px + py] = sxo_hmedian(smask, i);
            } else {
                back2[px + py] = back[px + py];
                sigma2[px + py] = sigma[px + py];
            }
        }
    }
    free(bmask);
    free(smask);
    memcpy(back, back2, (size_t)np * sizeof(float));
    *backmean = sxo_hmedian(back2, np);
    free(back2);
    memcpy(sigma, sigma2, (size_t)np * sizeof(float));
    *backsig = sxo_hmedian(sigma2, np);
    if (*backsig <= 0.0) {
        sigmat = sigma2 + np;
        for (i = np; i-- && *(--sigmat) > 0.0;)
            ;
        if (i >= 0 && i < (np - 1))
            *backsig = sxo_hmedian(sigmat + 1, np - 1 - i);
        else
            *backsig = 1.0;
    }
    free(sigma2);
}

/* back.c:739-780: natural cubic-spline second derivatives along y */
void sxo_makebackspline(const float *map, int nbx, int nby, float *dmap)
{
    int x, y, nbym1 = nby - 1;
    float *u = (float *)malloc(sizeof(float) * (nbym1 > 0 ? nbym1 : 1)), temp;
    for (x = 0; x < nbx; x++) {
        if (nby > 1) {
            dmap[x] = u[0] = 0.0f;
            for (y = 1; y < nbym1; y++) {
                temp = -1 / (dmap[(y - 1) * nbx + x] + 4);
                dmap[y * nbx + x] = temp;
                temp *= u[y - 1] - 6 * (map[(y + 1) * nbx + x] + map[(y - 1) * nbx + x]
                                        - 2 * map[y * nbx + x]);
                u[y] = temp;
            }
            dmap[nbym1 * nbx + x] = 0.0f;
            for (y = nby - 2; y >= 1; y--)
                dmap[y * nbx + x] = (float)((dmap[y * nbx + x] * dmap[(y + 1) * nbx + x] + u[y]) / 6.0);
        } else
            dmap[x] = 0.0f;
    }
    free(u);
}

/* back.c:790-921 for one row (AUTO background); backline may be NULL */
void sxo_backrmsline(const float *sigma, const float *dsigma, int nbx, int nby, int bs, int W,
                     int y, float *out)
{
    float *scratch = (float *)calloc((size_t)W, sizeof(float));
    sxo_subbackline(sigma, dsigma, nbx, nby, bs, W, y, scratch, out);
    free(scratch);
}

void sxo_subbackline(const float *back, const float *dback, int nbx, int nby, int bs,
                     int W, int y, float *line, float *backline)
{
    int i, j, x, yl, nbxm1 = nbx - 1, changepoint;
    float dx, dx0, dy, dy3, cdx, cdy, cdy3, temp, xstep;
    const float *blo, *bhi, *dblo, *dbhi;
    float *node, *dnode, *u;
    float bval;

    node = (float *)malloc(sizeof(float) * nbx);
    dnode = (float *)malloc(sizeof(float) * nbx);
    u = (float *)malloc(sizeof(float) * (nbx > 1 ? nbx - 1 : 1));
    if (nby > 1) {
        dy = (float)((float)y / bs - 0.5);
        dy -= (yl = (int)dy);
        if (yl < 0) {
            yl = 0;
            dy = (float)(dy - 1.0);
        } else if (yl >= nby - 1) {
            yl = nby < 2 ? 0 : nby - 2;
            dy = (float)(dy + 1.0);
        }
        cdy = 1 - dy;
        dy3 = (dy * dy * dy - dy);
        cdy3 = (cdy * cdy * cdy - cdy);
        blo = back + nbx * yl;
        bhi = blo + nbx;
        dblo = dback + nbx * yl;
        dbhi = dblo + nbx;
        for (x = 0; x < nbx; x++)
            node[x] = cdy * blo[x] + dy * bhi[x] + cdy3 * dblo[x] + dy3 * dbhi[x];
        if (nbx > 1) {
            dnode[0] = u[0] = 0.0f;
            for (x = 1; x < nbxm1; x++) {
                temp = -1 / (dnode[x - 1] + 4);
                dnode[x] = temp;
                temp *= u[x - 1] - 6 * (node[x + 1] + node[x - 1] - 2 * node[x]);
                u[x] = temp;
            }
            dnode[nbxm1] = 0.0f;
            for (x = nbx - 2; x >= 1; x--)
                dnode[x] = (float)((dnode[x] * dnode[x + 1] + u[x]) / 6.0);
        }
    } else {
        memcpy(node, back, sizeof(float) * nbx);
        memcpy(dnode, dback, sizeof(float) * nbx);
    }

    if (nbx > 1) {
        const int nx = bs;
        xstep = (float)(1.0 / nx);
        changepoint = nx / 2;
        dx = (xstep - 1) / 2;
        dx0 = ((nx + 1) % 2) * xstep / 2;
        blo = node;
        bhi = node + 1;
        dblo = dnode;
        dbhi = dnode + 1;
        for (x = i = 0, j = 0; j < W; j++, i++, dx += xstep) {
            if (i == changepoint && x > 0 && x < nbxm1) {
                blo++; bhi++; dblo++; dbhi++;
                dx = dx0;
            }
            cdx = 1 - dx;
            bval = cdx * (*blo + (cdx * cdx - 1) * *dblo) + dx * (*bhi + (dx * dx - 1) * *dbhi);
            if (backline) backline[j] = bval;
            line[j] -= bval;
            if (i == nx) {
                x++;
                i = 0;
            }
        }
    } else
        for (j = 0; j < W; j++) {
            bval = node[0];
            if (backline) backline[j] = bval;
            line[j] -= bval;
        }
    free(node);
    free(dnode);
    free(u);
}

/* Full background stage: maps, filtering, splines, subtraction (makeback, back.c:43-208) */
int sxo_background(const float *img, int W, int H, int bs, int nfx, int nfy, double fthresh,
                   int back_type, double back_value, float *sub, float *back, float *sigma,
                   float *dback, float *dsigma, float *backmean, float *backsig)
{
    const int nbx = (W - 1) / bs + 1, nby = (H - 1) / bs + 1;
    /* field.c:119-140: the filter size collapses to 1 along a single-mesh axis */
    const int fx = nbx > 1 ? nfx : 1, fy = nby > 1 ? nfy : 1;
    if (sxo_makeback(img, W, H, bs, fx, fy, fthresh, back, sigma, backmean, backsig)) return -1;
    sxo_filterback(back, sigma, nbx, nby, fx, fy, fthresh, backmean, backsig);
    sxo_makebackspline(back, nbx, nby, dback);
    sxo_makebackspline(sigma, nbx, nby, dsigma);
    memcpy(sub, img, sizeof(float) * (size_t)W * H);
    if (back_type == 1 || back_type == 2) {
        /* back_type 1, BACK_TYPE MANUAL as the reference's device code runs it: the
           constant subtracted is the global median just computed (cudaback.cu:1336-1345);
           *backmean keeps that median, as back_estimate returns it; the caller reports
           BACK_VALUE as backmean afterwards (back.c:152-153).
           back_type 2, stock SExtractor's MANUAL: the constant BACK_VALUE itself
           (back.c:148-149, :807-813), which becomes backmean. */
        const float bval = back_type == 1 ? *backmean : (float)back_value;
        for (size_t k = 0; k < (size_t)W * H; k++) sub[k] -= bval;
        if (back_type == 2) *backmean = bval;
    } else {
        for (int y = 0; y < H; y++)
            sxo_subbackline(back, dback, nbx, nby, bs, W, y, sub + (size_t)y * W, NULL);
    }
    return 0;
}
