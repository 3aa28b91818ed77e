/*
 * synth.c -- deterministic synthetic astronomical frames for the detection path.
 *
 * The frame model: a sky of `bkg` ADU with a linear gradient of +-grad*bkg along x,
 * optional large-scale Gaussian structure, Gaussian read noise of `sigma` ADU, and a
 * catalogue of point sources (Moffat profiles) and galaxies (elliptical Sersic
 * profiles) whose fluxes follow dN/dF ~ F^-2 between fmin and fmax.  A fraction of
 * the sources is placed in close pairs or triples (2-6 px) to force deblending.
 *
 * Random numbers: the source catalogue comes from xoshiro256** (Blackman & Vigna)
 * seeded through SplitMix64; pixel noise is counter based (SplitMix64 of the pixel
 * index mixed with the seed) so that rows can be rendered by independent threads and
 * the frame stays bit-identical whatever the thread count.  Each pixel accumulates
 * the sources in catalogue order, which fixes the float summation order.
 *
 * Sersic b_n follows Ciotti & Bertin (1999), A&A 352, 447, eq. 18 (asymptotic series).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "synth.h"

#define SX_PI 3.14159265358979323846

/* ---------------------------------------------------------------- RNG ---- */
static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

typedef struct { uint64_t s[4]; } xoshiro;

static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t xo_next(xoshiro *r)
{
    uint64_t *s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return result;
}

static void xo_seed(xoshiro *r, uint64_t seed)
{
    uint64_t sm = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&sm);
}

/* uniform in [0,1) with 53 random bits */
static double xo_uniform(xoshiro *r) { return (double)(xo_next(r) >> 11) * 0x1.0p-53; }

/* counter-based standard normal for pixel `idx` (Box-Muller on two hashed uniforms) */
static double pixel_normal(uint64_t seed, uint64_t idx)
{
    uint64_t s = seed * 0xd1342543de82ef95ULL ^ (idx * 0x9e3779b97f4a7c15ULL);
    double u1 = ((double)(splitmix64(&s) >> 11) + 0.5) * 0x1.0p-53;   /* (0,1) */
    double u2 = (double)(splitmix64(&s) >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * SX_PI * u2);
}

/* ------------------------------------------------------------ sources ---- */
typedef struct {
    double x, y, flux;
    int kind;          /* 0 Moffat, 1 Sersic */
    double p0, p1;     /* Moffat: alpha, beta      Sersic: r_e, n */
    double q, pa;      /* Sersic axis ratio and position angle */
    double amp;        /* central (Moffat) or effective (Sersic) intensity */
    double bn;         /* Sersic b_n */
    int rad;           /* stamp half size in pixels */
} source_t;

static double sersic_bn(double n)
{
    return 2.0 * n - 1.0 / 3.0 + 4.0 / (405.0 * n) + 46.0 / (25515.0 * n * n)
           + 131.0 / (1148175.0 * n * n * n);
}

static double profile(const source_t *s, double dx, double dy)
{
    if (s->kind == 0) {
        double r2 = (dx * dx + dy * dy) / (s->p0 * s->p0);
        return s->amp * pow(1.0 + r2, -s->p1);
    } else {
        double c = cos(s->pa), sn = sin(s->pa);
        double xp = dx * c + dy * sn, yp = -dx * sn + dy * c;
        double r = sqrt(xp * xp + (yp / s->q) * (yp / s->q));
        return s->amp * exp(-s->bn * (pow(r / s->p0, 1.0 / s->p1) - 1.0));
    }
}

static void finish_source(source_t *s, double eps, int maxrad)
{
    double R;
    if (s->kind == 0) {
        double a = s->p0, b = s->p1;
        s->amp = s->flux * (b - 1.0) / (SX_PI * a * a);
        R = a * sqrt(fmax(pow(s->amp / eps, 1.0 / b) - 1.0, 0.0));
    } else {
        double re = s->p0, n = s->p1;
        s->bn = sersic_bn(n);
        /* total flux of an elliptical Sersic profile */
        double tot = re * re * 2.0 * SX_PI * n * s->q * exp(s->bn) * tgamma(2.0 * n)
                     / pow(s->bn, 2.0 * n);
        s->amp = s->flux / tot;
        double l = log(fmax(s->amp / eps, 1.0));
        R = re * pow(1.0 + l / s->bn, n);
    }
    int r = (int)ceil(R);
    if (r < 2) r = 2;
    if (r > maxrad) r = maxrad;
    s->rad = r;
}

static void draw_source(xoshiro *rng, const sxs_params *p, source_t *s, int star)
{
    double u = xo_uniform(rng);
    s->flux = 1.0 / (1.0 / p->fmin - u * (1.0 / p->fmin - 1.0 / p->fmax));
    if (star) {
        s->kind = 0;
        s->p1 = p->moffat_beta;
        s->p0 = p->fwhm / (2.0 * sqrt(pow(2.0, 1.0 / s->p1) - 1.0));
        s->q = 1.0; s->pa = 0.0;
    } else {
        s->kind = 1;
        s->p1 = (xo_uniform(rng) < 0.5) ? 1.0 : 4.0;
        s->p0 = p->re_min + (p->re_max - p->re_min) * xo_uniform(rng);
        s->q = p->q_min + (1.0 - p->q_min) * xo_uniform(rng);
        s->pa = SX_PI * xo_uniform(rng);
    }
}

/* ------------------------------------------------------------- render ---- */
int sxs_render(const sxs_params *p, float *img, double *src_out, int src_out_cap)
{
    const int W = p->width, H = p->height;
    if (W <= 0 || H <= 0 || !img) return -1;
    const int nsrc = p->nsrc > 0 ? p->nsrc : 0;
    source_t *src = (source_t *)calloc((size_t)(nsrc > 0 ? nsrc : 1), sizeof(source_t));
    if (!src) return -2;

    xoshiro rng;
    xo_seed(&rng, p->seed);

    /* catalogue: anchors at uniform positions, companions of close groups at 2-6 px
       (from the anchor for pairs and triples; from a random earlier member for the
       grp_min..grp_max groups, which nests them) */
    int k = 0;
    const int big = p->grp_min > 0 && p->grp_max >= p->grp_min;
    while (k < nsrc) {
        int group = 1;
        if (big) {
            if (p->group_frac > 0.0 && xo_uniform(&rng) < p->group_frac)
                group = p->grp_min + (int)(xo_uniform(&rng) * (p->grp_max - p->grp_min + 1));
        } else if (p->group_frac > 0.0 && xo_uniform(&rng) < p->group_frac / 2.5) {
            group = (xo_uniform(&rng) < 0.5) ? 2 : 3;
        }
        double ax = W * xo_uniform(&rng), ay = H * xo_uniform(&rng);
        const int k0 = k;
        for (int g = 0; g < group && k < nsrc; g++, k++) {
            source_t *s = &src[k];
            int star = xo_uniform(&rng) < p->star_frac;
            draw_source(&rng, p, s, star);
            if (g == 0) {
                s->x = ax; s->y = ay;
            } else {
                const source_t *ref = big ? &src[k0 + (int)(xo_uniform(&rng) * g)] : &src[k0];
                double sep = 2.0 + 4.0 * xo_uniform(&rng);
                double ang = 2.0 * SX_PI * xo_uniform(&rng);
                s->x = ref->x + sep * cos(ang);
                s->y = ref->y + sep * sin(ang);
            }
        }
    }
    /* bright saturated stars (high-dynamic-range frames) */
    int nsat = p->nsat;
    source_t *sat = NULL;
    if (nsat > 0) {
        sat = (source_t *)calloc((size_t)nsat, sizeof(source_t));
        if (!sat) { free(src); return -2; }
        for (int i = 0; i < nsat; i++) {
            source_t *s = &sat[i];
            s->kind = 0;
            s->p1 = p->sat_beta;
            s->p0 = p->fwhm / (2.0 * sqrt(pow(2.0, 1.0 / s->p1) - 1.0));
            s->x = W * xo_uniform(&rng); s->y = H * xo_uniform(&rng);
            s->amp = p->sat_peak;
            s->flux = s->amp * SX_PI * s->p0 * s->p0 / (s->p1 - 1.0);
            s->q = 1.0; s->pa = 0.0;
        }
    }
    for (int i = 0; i < nsrc; i++) finish_source(&src[i], p->eps, p->maxrad);
    for (int i = 0; i < nsat; i++) {
        double keep = sat[i].amp;
        finish_source(&sat[i], p->eps, p->maxrad);
        sat[i].amp = keep;
    }

    if (src_out && src_out_cap > 0) {
        int n = nsrc < src_out_cap ? nsrc : src_out_cap;
        for (int i = 0; i < n; i++) {
            src_out[4 * i + 0] = src[i].x;
            src_out[4 * i + 1] = src[i].y;
            src_out[4 * i + 2] = src[i].flux;
            src_out[4 * i + 3] = src[i].kind;
        }
    }

    /* row bands rendered independently; each pixel sums sources in catalogue order */
#pragma omp parallel for schedule(dynamic, 8)
    for (int y = 0; y < H; y++) {
        float *row = img + (size_t)y * W;
        double gx = (W > 1) ? 2.0 / (W - 1) : 0.0;
        for (int x = 0; x < W; x++) {
            double v = p->bkg * (1.0 + p->grad * (x * gx - 1.0));
            for (int b = 0; b < p->nblob; b++) {
                double dx = x - p->blob_x[b] * W, dy = y - p->blob_y[b] * H;
                v += p->blob_amp * exp(-0.5 * (dx * dx + dy * dy) / (p->blob_sigma * p->blob_sigma));
            }
            row[x] = (float)v;
        }
        for (int pass = 0; pass < 2; pass++) {
            const source_t *list = pass ? sat : src;
            int n = pass ? nsat : nsrc;
            for (int i = 0; i < n; i++) {
                const source_t *s = &list[i];
                if (y < s->y - s->rad - 1 || y > s->y + s->rad + 1) continue;
                int x0 = (int)floor(s->x) - s->rad, x1 = (int)floor(s->x) + s->rad + 1;
                if (x0 < 0) x0 = 0;
                if (x1 > W - 1) x1 = W - 1;
                for (int x = x0; x <= x1; x++) {
                    double dx = x - s->x, dy = y - s->y;
                    double v;
                    if (fabs(dx) <= 3.0 && fabs(dy) <= 3.0) {
                        v = 0.0;   /* 4x4 sub-pixel sampling near the centre */
                        for (int sy = 0; sy < 4; sy++)
                            for (int sx = 0; sx < 4; sx++)
                                v += profile(s, dx + (sx - 1.5) * 0.25, dy + (sy - 1.5) * 0.25);
                        v *= 1.0 / 16.0;
                    } else {
                        v = profile(s, dx, dy);
                    }
                    row[x] = (float)(row[x] + v);
                }
            }
        }
        for (int x = 0; x < W; x++) {
            double v = row[x] + p->sigma * pixel_normal(p->seed, (uint64_t)y * (uint64_t)W + x);
            if (p->clip_level > 0.0 && v > p->clip_level) v = p->clip_level;
            row[x] = (float)v;
        }
    }
    free(src);
    free(sat);
    return 0;
}

/* weight map: inverse variance 1/sigma^2 with a smooth +-vign vignetting pattern and a
   fraction of zeroed (bad) columns chosen from the seed */
int sxs_weight(const sxs_params *p, double vign, double badcol_frac, float *w)
{
    const int W = p->width, H = p->height;
    if (W <= 0 || H <= 0 || !w) return -1;
    xoshiro rng;
    xo_seed(&rng, p->seed ^ 0x5bd1e995ULL);
    unsigned char *bad = (unsigned char *)calloc((size_t)W, 1);
    if (!bad) return -2;
    int nbad = (int)(badcol_frac * W + 0.5);
    for (int i = 0; i < nbad; i++) bad[(int)(W * xo_uniform(&rng)) % W] = 1;
    double w0 = 1.0 / (p->sigma * p->sigma);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++) {
        double fy = (double)y / (H > 1 ? H - 1 : 1);
        for (int x = 0; x < W; x++) {
            double fx = (double)x / (W > 1 ? W - 1 : 1);
            double v = 1.0 + vign * cos(SX_PI * (fx - 0.5)) * cos(SX_PI * (fy - 0.5)) * 2.0 - vign;
            w[(size_t)y * W + x] = bad[x] ? 0.0f : (float)(w0 * v);
        }
    }
    free(bad);
    return 0;
}

/* 64-bit FNV-1a of a float buffer, for logging frame identity */
uint64_t sxs_hash(const float *img, size_t n)
{
    const unsigned char *b = (const unsigned char *)img;
    uint64_t h = 1469598103934665603ULL;
    for (size_t i = 0; i < n * sizeof(float); i++) { h ^= b[i]; h *= 1099511628211ULL; }
    return h;
}

void sxs_defaults(sxs_params *p)
{
    memset(p, 0, sizeof(*p));
    p->width = 1024; p->height = 1024; p->nsrc = 3000;
    p->star_frac = 0.7; p->fmin = 1e2; p->fmax = 1e5;
    p->bkg = 1000.0; p->grad = 0.05; p->sigma = 10.0;
    p->moffat_beta = 3.5; p->fwhm = 3.0;
    p->re_min = 2.0; p->re_max = 8.0; p->q_min = 0.3;
    p->eps = 1e-3; p->maxrad = 64;
    p->seed = 1000;
    p->sat_beta = 2.5; p->sat_peak = 1e6;
}
