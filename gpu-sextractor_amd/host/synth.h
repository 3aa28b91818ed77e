/* synth.h -- deterministic synthetic frames (see synth.c for the model). */
#ifndef SXS_SYNTH_H
#define SXS_SYNTH_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SXS_MAXBLOB 8

typedef struct {
    int width, height;
    int nsrc;                 /* catalogue sources (stars + galaxies) */
    double star_frac;         /* fraction of Moffat stars */
    double fmin, fmax;        /* flux range, dN/dF ~ F^-2 */
    double bkg, grad;         /* sky level and +-fractional x gradient */
    double sigma;             /* Gaussian noise rms */
    double moffat_beta, fwhm; /* star profile */
    double re_min, re_max;    /* Sersic effective radius range */
    double q_min;             /* minimum axis ratio */
    double group_frac;        /* fraction of sources in close pairs/triples */
    double eps;               /* stamp truncation level (ADU) */
    int maxrad;               /* maximum stamp half size */
    uint64_t seed;
    int nblob;                /* large-scale Gaussian structure */
    double blob_x[SXS_MAXBLOB], blob_y[SXS_MAXBLOB]; /* fractional centres */
    double blob_sigma, blob_amp;
    int nsat;                 /* extra saturated Moffat stars */
    double sat_beta, sat_peak;
    double clip_level;        /* >0: clip pixel values (detector saturation) */
    int grp_min, grp_max;     /* >0: groups of grp_min..grp_max members (group_frac of the
                                 anchors), each companion 2-6 px from an earlier member of
                                 its group (nested clumps); 0: pairs and triples */
} sxs_params;

void sxs_defaults(sxs_params *p);
/* Render into img[height*width]; optionally return (x, y, flux, kind) per source. */
int sxs_render(const sxs_params *p, float *img, double *src_out, int src_out_cap);
/* Inverse-variance weight map with vignetting and zeroed columns. */
int sxs_weight(const sxs_params *p, double vign, double badcol_frac, float *w);
uint64_t sxs_hash(const float *img, size_t n);

#ifdef __cplusplus
}
#endif
#endif
