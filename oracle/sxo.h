/*
 * sxo.h -- CPU oracle of the SExtractor detection path (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the parity checker for the HIP implementation.  Only the tests, the
 * smoke check and the benchmark's CPU-baseline leg may load it; the product never
 * links or calls it.  It restates, in plain sequential C:
 *   - the CPU background code of the reference (src/back.c, src/misc.c:30 hmedian),
 *   - the CPU convolution (src/filter.c:39-186),
 *   - the algorithms of the reference's CUDA detection path for the parts whose CPU
 *     code was removed (src/cuda/cudadetection.cu, cudadeblend.cu, cudaanalyse.cu,
 *     cudaclean.cu), each function citing the file:line it follows.
 * Summation orders that the reference leaves to a library (CUDPP segmented scans) are
 * fixed here as "sequential over the object's pixels in raster order", the order the
 * reference's own per-object loops use (cudaanalyse.cu:1770-1835).
 */
#ifndef SXO_H
#define SXO_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SXO_BIG 1e+30

typedef struct {
    int back_size;           /* BACK_SIZE (square meshes) */
    int back_fsize;          /* BACK_FILTERSIZE */
    double back_fthresh;     /* BACK_FILTTHRESH */
    int back_type;           /* 0 AUTO; 1 MANUAL as the reference's device code runs it
                                (subtract the computed global median, report BACK_VALUE);
                                2 stock MANUAL (subtract BACK_VALUE) */
    double back_value;       /* BACK_VALUE */
    int thresh_type;         /* 0 RELATIVE, 1 ABSOLUTE */
    double detect_thresh;    /* DETECT_THRESH */
    double analysis_thresh;  /* ANALYSIS_THRESH */
    int ext_minarea;         /* DETECT_MINAREA */
    int deblend_nthresh;     /* DEBLEND_NTHRESH */
    double deblend_mincont;  /* DEBLEND_MINCONT */
    int clean;               /* CLEAN */
    double clean_param;      /* CLEAN_PARAM */
    int clean_stacksize;     /* MEMORY_OBJSTACK */
    double satur_level;      /* SATUR_LEVEL */
    double gain;             /* GAIN */
    const float *mask;       /* filter mask (row-major), NULL = no filtering */
    int mask_w, mask_h;
    const float *weight;     /* weight map (MAP_WEIGHT, MAP_RMS or MAP_VAR) or NULL */
    double weight_thresh;    /* unused unless weight != NULL */
    int weight_type;         /* 0 MAP_WEIGHT, 1 MAP_RMS, 2 MAP_VAR, 3 BACKGROUND (no map) */
    int weight_gain;         /* WEIGHT_GAIN */
} sxo_config;

/* stages timed by sxo_run: background, filtering, detection, deblending (with gatherup
   and the final pixel lists), analysis, CLEAN, segmentation map */
#define SXO_NSTAGE 7
typedef struct {
    int width, height, nbx, nby;
    float backmean, backsig, dthresh, thresh;
    float *back, *sigma, *dback, *dsigma;   /* filtered mesh maps and y derivatives */
    float *sub;                             /* background-subtracted frame */
    float *cd;                              /* filtered frame */
    float *var;                             /* per-pixel variance (weight maps) or NULL */
    long npix_above;                        /* P_all */
    long npix_raw;                          /* P0: pixels of raw objects */
    int nraw;                               /* R */
    int *raw_first, *raw_count;             /* raw objects: first raster pixel, size */
    float *raw_fdflux, *raw_fdpeak;
    int nobj;                               /* M: deblended objects */
    int *level, *root, *firstpix, *flag, *fdnpix, *dnpix, *npix;
    int *xmin, *xmax, *ymin, *ymax, *peakx, *peaky, *iso;   /* iso: 8 per object */
    double *mx, *my, *mx2, *my2, *mxy, *poserr_mx2, *poserr_my2, *poserr_mxy;
    float *a, *b, *theta, *cxx, *cyy, *cxy, *abcor, *fdflux, *dflux, *flux, *fluxerr;
    float *fdpeak, *dpeak, *peak, *bkg, *thresh_obj, *mthresh, *fwhm, *dthresh_obj;
    uint32_t *master;                       /* 0xffffffff: kept by CLEAN */
    int *alloc;                             /* per pixel: object index or -1 */
    int *seg;                               /* per pixel: catalogue NUMBER or 0 */
    int nsurv;
    int *objpix_start, *objpix;             /* final pixel lists (raster order) */
    int *levpix_start, *levpix;             /* each object's level pixels (its kept
                                               component before gatherup), raster order */
    double stage_ms[SXO_NSTAGE];            /* wall time of each stage of sxo_run */
} sxo_result;

/* ---------------- background (src/back.c) ---------------- */
typedef struct {
    float mean, sigma, lcut, hcut, qzero, qscale;
    int nlevels, npix;
} sxo_mesh;

void sxo_backstat_mesh(const float *img, int W, int x0, int y0, int bw, int bh, sxo_mesh *m);
void sxo_backhisto_mesh(const float *img, int W, int x0, int y0, int bw, int bh,
                        const sxo_mesh *m, int *histo);
float sxo_backguess(const sxo_mesh *m, const int *histo, float *mean, float *sigma);
float sxo_hmedian(float *ra, int n);
int sxo_makeback(const float *img, int W, int H, int bs, int nfx, int nfy, double fthresh,
                 float *back, float *sigma, float *backmean, float *backsig);
void sxo_filterback(float *back, float *sigma, int nbx, int nby, int nfx, int nfy,
                    double fthresh, float *backmean, float *backsig);
void sxo_makebackspline(const float *map, int nbx, int nby, float *dmap);
/* back.c:938-1051: the same spline as subbackline over the sigma maps, its values written
   into out (width W) */
void sxo_backrmsline(const float *sigma, const float *dsigma, int nbx, int nby, int bs, int W,
                     int y, float *out);
void sxo_subbackline(const float *back, const float *dback, int nbx, int nby, int bs,
                     int W, int y, float *line, float *backline);
int sxo_background(const float *img, int W, int H, int bs, int nfx, int nfy, double fthresh,
                   int back_type, double back_value, float *sub, float *back, float *sigma,
                   float *dback, float *dsigma, float *backmean, float *backsig);

/* ---------------- filtering (src/filter.c) ---------------- */
int sxo_getconv(const char *filename, float *mask, int maxn, int *mw, int *mh);
void sxo_convolve(const float *in, int W, int H, const float *mask, int mw, int mh, float *out);

/* ---------------- whole path ---------------- */
void sxo_config_defaults(sxo_config *c);
sxo_result *sxo_run(const float *img, int W, int H, const sxo_config *cfg);

/* world coordinates (sxo_wcs.c): a celestial WCS with a zenithal projection */
typedef struct {
    int lng, lat;             /* the axes of right ascension and declination */
    int proj;                 /* 0 TAN, 1 SIN, 2 ARC, 3 STG, 4 ZEA */
    double crval[2], crpix[2], cd[4];
    double phip;              /* native longitude of the celestial pole (LONPOLE) */
    double equinox;
} sxo_wcs;
/* 0, or -1 for a WCS this restatement does not cover; longpole 999: the default */
int sxo_wcs_setup(sxo_wcs *w, const char *ctype1, const char *ctype2, const double *crval,
                  const double *crpix, const double *cd, double longpole, double equinox);
/* pixel (px, py), 1-based, to world[lng, lat] (degrees, the equinox of the WCS) and
   radec[2] (J2000); 0, or -1 outside the projection's domain */
int sxo_wcs_world(const sxo_wcs *w, double px, double py, double *world, double *radec);
void sxo_precess(double yin, double ain, double din, double yout, double *aout, double *dout);
/* the bilinear value of a nx x ny mesh map b at pixel (x, y), meshes of backw x backh
   (back.c:685-732, cudaanalyse.cu:2128-2172): the BACKGROUND of an object's centroid */
float sxo_back_at(const float *b, int x, int y, int nx, int ny, int backw, int backh);
void sxo_free(sxo_result *r);
/* all-double re-measurement of the catalogue columns over the same final pixel lists */
/* All-extended-precision (long double) evaluation of the graded catalogue quantities on
   the result's own pixel sets: for each object, out[12*o + k] with k = 0 FLUX_ISO, 1 X,
   2 Y (0-based), 3 X2, 4 Y2, 5 XY, 6 A, 7 B, 8 THETA (degrees), 9 CXX, 10 CYY, 11 CXY.
   X, Y: the barycentre of the object's level pixels; second moments about it over the
   final pixels (the reference's FULL and robust/MERGED definitions); the singularity rule
   of the reference applied.  Returns the number of objects. */
int sxo_measure_fp64(const sxo_result *r, double *out, int ncol);

/* ---------------- catalogue photometry (endobject_gpu) ---------------- */
/* the classifier network, layout of the reference's brainstruct (src/types.h:930-943,
   LAYERS 3, NEURONS 10) without its scratch arrays */
typedef struct {
    int layersnb;
    int nn[3];
    double inbias[10], inscale[10], outbias[10], outscale[10];
    double w[2][10][10];
    double b[2][10];
} sxo_brain;

typedef struct {
    double apert;            /* PHOT_APERTURES: the one aperture's diameter (NAPER 1) */
    double autoparam[2];     /* PHOT_AUTOPARAMS: Kron factor, minimum radius */
    double autoaper[2];      /* PHOT_AUTOAPERS: estimation, measurement minimum diameters */
    double seeing_fwhm;      /* SEEING_FWHM */
    double pixscale;         /* field pixel scale */
    double mag_zeropoint;    /* MAG_ZEROPOINT */
    double ngamma, gain, satur_level;
    float backsig, thresh;   /* field backsig and analysis threshold */
    int mask_correct;        /* MASK_TYPE CORRECT */
    int photo;               /* DETECT_TYPE PHOTO */
    int do_aper, do_auto, do_class;
    const sxo_brain *brain;
} sxo_photom_config;

typedef struct {             /* caller-allocated arrays of nobj entries */
    float *flux_iso, *fluxerr_iso, *mag_iso, *magerr_iso;
    float *flux_isocor, *fluxerr_isocor;
    float *flux_aper, *fluxerr_aper, *mag_aper, *magerr_aper;
    float *kronfactor, *flux_auto, *fluxerr_auto, *mag_auto, *magerr_auto;
    float *elong, *ellip, *sprob;
    int *flag;
} sxo_photom;

/* per-pixel photometry labels (object index, -1, or INT_MAX), N entries */
void sxo_photom_labels(const sxo_result *r, int *lab);
double sxo_neurresp(const sxo_brain *br, const double *input);
/* fills out for every object of r; returns nobj or < 0 */
int sxo_endobject(const sxo_result *r, const sxo_photom_config *pc, sxo_photom *out);

/* canonical math (shared definition with the HIP kernels) */
double sxo_log(double x);
double sxo_exp(double x);
double sxo_pow(double x, double y);
double sxo_log10(double x);

#ifdef __cplusplus
}
#endif
#endif
