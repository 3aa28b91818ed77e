/*
 * sexgpu.h -- C-ABI of the MI355X SExtractor detection path (libsexgpu.so).
 *
 * The first group of entry points has exactly the names, argument lists and calling
 * order of the reference's CUDA library (the src/cuda sources, called from src/makeit.c:189,
 * src/back.c:104 and src/scan.c:78-146), so that the reference front-end can link this
 * library instead.  Semantics follow the reference:
 *   - a stateful singleton: width, height and every device buffer live in the library;
 *     one frame at a time, calls in the order listed; not re-entrant, not thread-safe;
 *   - every call returns synchronously;
 *   - the library owns device memory; init_device copies the caller's frame in and
 *     clear_device overwrites the caller's buffer with the background-subtracted frame;
 *   - run_clean fills the caller's masterIndex[numobj] (0xffffffff = kept).
 * The reference exits on any error; this library records the error instead: the void
 * entry points keep their signatures, sxg_last_error() returns the message (NULL when
 * the last call succeeded), and the int-returning ones return a negative value.
 */
#ifndef SEXGPU_H
#define SEXGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference entry points ---------------------------------------------------- */

/* src/cuda/cudainit.cu:16 -- copy the width x height float frame to the device. */
void init_device(int width, int height, float *img);

/* src/cuda/cudaback.cu:1212 -- background mesh statistics, mode, median filter,
   splines, global medians (written to *backmean, *backsig) and in-place subtraction.
   back_type 1 (MANUAL) subtracts the global median *backmean, as the reference's device
   code does (cudaback.cu:1336-1345; its caller back.c:152-153 then reports BACK_VALUE).
   back_type 2 (an addition) is stock SExtractor's MANUAL: it subtracts the value set by
   sxg_set_back_value() and returns it as *backmean. */
void back_estimate(int back_size, float backfthresh, float *backmean, float *backsig,
                   int nbackfx, int nbackfy, int back_type);

/* src/cuda/cudadetection.cu:492 */
void init_detection(void);

/* src/cuda/cudafilter.cu:150 -- convolve the subtracted frame with a square
   mask_width x mask_width mask (any odd or even width up to 31, row-major). */
void cudaFilter(float *mask, unsigned int mask_width);

/* src/cuda/cudadetection.cu:1107 -- threshold (cd >= dthresh), 8-connected labelling,
   DETECT_MINAREA pruning; returns the number of raw objects and of their pixels. */
void run_detection(float dthresh, float thresh, int ext_minarea, size_t *numValidObj,
                   size_t *numValidPix);

/* src/cuda/cudadeblend.cu:508 -- per-object sub-thresholds. */
void init_deblend(float basethresh, int nthresh, size_t numValidObj, size_t numValidPix);

/* src/cuda/cudadeblend.cu:579 -- multi-threshold deblending, tree pruning,
   preanalysis of the kept objects and gather-up; returns the number of objects. */
int parcel_out(float basethresh, int n_thresh, int deb_minarea, size_t numValidObj,
               size_t numValidPix, double deblend_mincont);

/* src/cuda/cudadeblend.cu:1937 */
void clear_deblend(int nthresh);

/* src/cuda/cudaanalyse.cu:3171 -- robust preanalysis and isophotal analysis. The
   flagobj_* arguments carry a boolean in their first byte (src/define.h:283). */
void analyse_gpu(double flagobj_poserr_mx2, int flagobj_peakx, int flagobj_iso_0,
                 float flagobj_fwhm, float backsig, double gain, double ngamma, float thresh,
                 float dthresh, double satur_level, int plistexist_var, int plistoff_var,
                 int prefs_weightgain_flag, int prefs_ext_minarea, int prefs_clean_flag,
                 int detect_type, int nbackx, int nbacky, int backw, int backh, int width,
                 int numobj);

/* src/cuda/cudaclean.cu:260 -- masterIndex[i] = index of the object that makes object
   i spurious, or 0xffffffff.  (The reference header cudaclean.h:4 declares an int
   return; the definition is void, as here.) */
void run_clean(int *masterIndex, double clean_param, int clean_stacksize, int numobj);

/* src/cuda/cudaanalyse.cu:3344 -- catalogue photometry (ISOCOR, APER, AUTO, CLASS_STAR,
   magnitudes, ELONGATION/ELLIPTICITY, the APERT_PB/CROWDED/SATUR flags) and one row per
   object (outobjstruct layout of src/types.h:417-455, mirrored by sxg_outobj below; objlist
   may be NULL to keep the results on the device).  prefs is the reference's prefstruct:
   PHOT_APERTURES, PHOT_AUTOPARAMS, PHOT_AUTOAPERS, SEEING_FWHM, MAG_ZEROPOINT, MASK_TYPE
   and DETECT_TYPE are read from it at the offsets of the reference's layout
   (sxg_reference_layout; a naper other than 1 is an error, as the reference aborts);
   flagobj2 (the reference's obj2struct of flags) switches the APER, AUTO and CLASS_STAR
   groups by the first byte of its flux_aper, flux_auto and sprob fields.  With prefs NULL
   the settings of sxg_set_photometry apply, with flagobj2 NULL every group is computed.
   brain is read as sxg_brain (NULL: no CLASS_STAR).  wcs is the reference's wcsstruct:
   with a celestial pair of a zenithal projection (TAN, SIN, ARC, STG, ZEA) the rows get
   ALPHA_J2000/DELTA_J2000 (see sxg_set_wcs, which takes precedence; the reference's GPU
   code writes zeros, cudaanalyse.cu:3033-3047).  Whole frames only: fymin 0, fymax
   height. */
void endobject_gpu(void *prefs, void *brain, void *flagobj2, void *objlist, void *wcs,
                   int fymin, int fymax, int width, int height, float backsig, float thresh,
                   double satur_level, double pixscale, double ngamma, double gain,
                   int numobj);

/* src/cuda/cudadetection.cu:1092 */
void clear_detection(void);

/* src/cuda/cudainit.cu:160 -- copy the background-subtracted frame into img. */
void clear_device(float *img);

/* ---- additions (no counterpart in the reference) -------------------------------- */

/* Message of the last failed call, or NULL. */
const char *sxg_last_error(void);

/* BACK_VALUE: subtracted by back_type 2, reported as backmean by back_type 1 in the
   whole-frame runs (the reference reads prefs.back_val on the host, back.c:152-153). */
void sxg_set_back_value(float value);

/* MAG_ZEROPOINT used by endobject_gpu (the reference reads prefs.mag_zeropoint). */
void sxg_set_mag_zeropoint(double zp);

/* The star/galaxy network, in the layout of the reference's brainstruct
   (src/types.h:930-943, LAYERS 3, NEURONS 10): endobject_gpu reads its brain argument as
   this structure, so the reference front-end's own brain pointer can be passed.  ni, no
   and n are scratch in the reference and are not read. */
typedef struct __attribute__((aligned(16))) {
    int layersnb;
    int nn[3];
    double inbias[10], inscale[10], outbias[10], outscale[10];
    double ni[10], no[10];
    double n[3][10];
    double w[2][10][10];
    double b[2][10];
} sxg_brain;

/* Photometry settings of endobject_gpu.  The reference reads them from its prefs
   structure (src/prefs.h): PHOT_APERTURES (one aperture, NAPER 1), PHOT_AUTOPARAMS,
   PHOT_AUTOAPERS, SEEING_FWHM, MAG_ZEROPOINT, MASK_TYPE, DETECT_TYPE; the do_* fields
   stand for the catalogue parameters that request them (the obj2 flags).  Defaults:
   the reference's default.sex (5, {2.5, 3.5}, {0, 0}, 1.2, 0, CORRECT, CCD, all on). */
typedef struct {
    double apert;           /* aperture diameter in pixels */
    double autoparam[2];    /* Kron factor, minimum radius */
    double autoaper[2];     /* minimum diameters: estimation, measurement */
    double seeing_fwhm;     /* arcsec */
    double mag_zeropoint;
    int mask_correct;       /* 1: MASK_TYPE CORRECT (mirror), 0: other objects' pixels = 0 */
    int photo;              /* 1: DETECT_TYPE PHOTO */
    int do_aper, do_auto, do_class;
} sxg_photom_params;
/* Returns 0, or -1 for invalid settings. */
int sxg_set_photometry(const sxg_photom_params *p);

/* The byte offsets endobject_gpu assumes in the reference's structures (NAPER 1, NAXIS 2,
   x86-64): sizeof(prefstruct), prefs.naper, .apert, .autoparam, .autoaper, .seeing_fwhm,
   .mag_zeropoint, .mask_type, .detect_type, sizeof(obj2struct), obj2.flux_aper, .flux_auto,
   .sprob, the values of PHOTO and MASK_CORRECT, then sizeof(wcsstruct), wcs.naxis, .ctype,
   .crval, .crpix, .cd, .longpole, .latpole, .lat, .lng, .equinox.  Writes min(n, count)
   values to out and returns the count. */
int sxg_reference_layout(long *out, int n);

/* World coordinates (no counterpart in the reference's GPU path, which writes zeros;
   its CPU semantics: astrom.c:86-140, raw_to_wcs fitswcs.c:1251, precess fitswcs.c:1718):
   the WCS of the frames processed next, from the FITS header keywords.  ctype: CTYPE1/2,
   a right ascension / declination pair in either order with a zenithal projection (TAN,
   SIN, ARC, STG, ZEA); cd: CD1_1, CD1_2, CD2_1, CD2_2; longpole 999: the default (180);
   positions are precessed to J2000 from an equinox other than 2000 (IAU 1976).  NULL
   removes it.  Returns 0, or -1 for a WCS that is not covered. */
typedef struct {
    char ctype[2][9];
    double crval[2], crpix[2], cd[4];
    double longpole, latpole, equinox;
} sxg_wcs;
int sxg_set_wcs(const sxg_wcs *wcs);
/* After endobject_gpu: X_WORLD, Y_WORLD (the axes' order, the WCS's equinox) and
   ALPHA_J2000, DELTA_J2000 in degrees, object order (any pointer may be NULL); zeros
   without a WCS, 1e31 outside the projection's domain.  Returns 0 or < 0. */
int sxg_get_world(double *xw, double *yw, double *alpha2000, double *delta2000, size_t count);

/* Photometry columns of the last endobject_gpu call (float, object order). */
enum sxg_phot_column {
    SXG_PHOT_FLUX_ISO = 0, SXG_PHOT_FLUXERR_ISO, SXG_PHOT_MAG_ISO, SXG_PHOT_MAGERR_ISO,
    SXG_PHOT_FLUX_ISOCOR, SXG_PHOT_FLUXERR_ISOCOR, SXG_PHOT_FLUX_APER, SXG_PHOT_FLUXERR_APER,
    SXG_PHOT_MAG_APER, SXG_PHOT_MAGERR_APER, SXG_PHOT_KRON_RADIUS, SXG_PHOT_FLUX_AUTO,
    SXG_PHOT_FLUXERR_AUTO, SXG_PHOT_MAG_AUTO, SXG_PHOT_MAGERR_AUTO, SXG_PHOT_ELONGATION,
    SXG_PHOT_ELLIPTICITY, SXG_PHOT_CLASS_STAR, SXG_NPHOT
};
/* out[count >= numobj]; returns 0 or < 0 on error. */
int sxg_get_photometry(int column, float *out, size_t count);

/* Segmentation check-image: out[width*height] = catalogue NUMBER (1-based, survivors of
   CLEAN in object order) on each object's final pixels, 0 elsewhere.  Call after
   run_clean (before clear_detection). Returns the number of survivors or < 0. */
int sxg_get_segmap(int32_t *out);

/* Catalogue row of the isophotal analysis, one per survivor of CLEAN in NUMBER order:
   the columns of the reference's analysis (cudaanalyse.cu:1906-2581, :2715-2716). */
typedef struct {
    int32_t number;             /* NUMBER, 1-based: the object's segmentation value */
    int32_t flags;              /* FLAGS: MERGED 2 | SATUR 4 | TRUNC 8 */
    int32_t isoarea;            /* ISOAREA_IMAGE: pixels above the analysis threshold */
    int32_t fdnpix;             /* pixels of the object */
    double x, y;                /* X_IMAGE, Y_IMAGE (1-based) */
    double x2, y2, xy;          /* X2_IMAGE, Y2_IMAGE, XY_IMAGE */
    float a, b, theta;          /* A_IMAGE, B_IMAGE, THETA_IMAGE (degrees) */
    float cxx, cyy, cxy;        /* CXX_IMAGE, CYY_IMAGE, CXY_IMAGE */
    float flux_iso, fluxerr_iso;   /* FLUX_ISO, FLUXERR_ISO */
    float peak, bkg;            /* FLUX_MAX, BACKGROUND */
} sxg_isorow;
/* The rows of the last frame into out[0..min(survivors, cap)), with one device-to-host
   copy: after run_clean, until the next frame.  Returns the number of survivors or < 0. */
int sxg_get_catalog(sxg_isorow *out, int cap);

/* Per-object columns (object order, numobj entries).  Returns 0 or < 0 on error. */
enum sxg_column {
    SXG_COL_LEVEL = 0,     /* int32  deblending level of the object */
    SXG_COL_ROOT,          /* int32  raw (level-0) object it comes from */
    SXG_COL_FIRSTPIX,      /* int32  raster index of its first level pixel */
    SXG_COL_FLAG,          /* int32  MERGED 2 | SATUR 4 | TRUNC 8 */
    SXG_COL_FDNPIX,        /* int32  final pixel count */
    SXG_COL_DNPIX,         /* int32  pixels above the detection threshold */
    SXG_COL_NPIX,          /* int32  ISOAREA_IMAGE (pixels above the analysis thresh) */
    SXG_COL_XMIN, SXG_COL_XMAX, SXG_COL_YMIN, SXG_COL_YMAX,   /* int32, 0-based */
    SXG_COL_PEAKX, SXG_COL_PEAKY,                            /* int32, 1-based */
    SXG_COL_ISO,           /* int32[8] isophotal areas */
    SXG_COL_MX, SXG_COL_MY, SXG_COL_MX2, SXG_COL_MY2, SXG_COL_MXY,   /* double */
    SXG_COL_POSERR_MX2, SXG_COL_POSERR_MY2, SXG_COL_POSERR_MXY,      /* double */
    SXG_COL_A, SXG_COL_B, SXG_COL_THETA, SXG_COL_CXX, SXG_COL_CYY, SXG_COL_CXY,  /* float */
    SXG_COL_ABCOR, SXG_COL_FDFLUX, SXG_COL_DFLUX, SXG_COL_FLUX, SXG_COL_FLUXERR,
    SXG_COL_FDPEAK, SXG_COL_DPEAK, SXG_COL_PEAK, SXG_COL_BKG, SXG_COL_THRESH,
    SXG_COL_MTHRESH, SXG_COL_FWHM,                                   /* float */
    SXG_COL_MASTER,        /* uint32 CLEAN master (0xffffffff = kept) */
    SXG_NCOL
};
int sxg_get_column(int column, void *out, size_t bytes);

/* Weight maps beyond MAP_WEIGHT (weight.c:35-126): the type of the map given to
   sxg_set_weight_map.  MAP_RMS: var = rms^2 (rms < BIG); MAP_VAR: var as given;
   BACKGROUND: no map, the variance is the square of the frame's interpolated background
   rms (backrmsline, back.c:938-1051).  Pixels whose variance is BIG or more (weight <= 0
   for MAP_WEIGHT) stay out of the mesh statistics; with RELATIVE thresholds the detection
   and analysis thresholds are per pixel (DETECT_THRESH * sqrt(var)).  weight_gain:
   WEIGHT_GAIN Y (the per-pixel flux variance is scaled by var/backsig^2, cudaanalyse.cu:
   2281, :2354-2355, :2464-2465).  map = NULL with a type other than BACKGROUND removes
   the weighting. */
enum sxg_weight_type { SXG_WEIGHT_MAP_WEIGHT = 0, SXG_WEIGHT_MAP_RMS, SXG_WEIGHT_MAP_VAR,
                       SXG_WEIGHT_BACKGROUND };
int sxg_set_weight_map(const float *map, int width, int height, int type, int thresh_type,
                       double detect_thresh, double analysis_thresh, int weight_gain);
/* The interpolated background rms of the last back_estimate, width*height floats (the
   BACKGROUND_RMS check-image, backrmsline: back.c:938-1051).  Returns 0 or < 0. */
int sxg_get_back_rms(float *out);

/* Frame-level values: backmean, backsig, dthresh, thresh, and counts
   (pixels above threshold, raw objects, raw-object pixels, objects, survivors). */
typedef struct {
    float backmean, backsig, dthresh, thresh;
    int64_t npix_above, npix_raw;
    int32_t nraw, nobj, nsurv, nbx, nby;
    /* deblending work: pixels of the raw objects with two or more local maxima (the only
       ones that get a component tree), and the sum over them of pixels x L* (the
       pixel-levels a level-by-level construction would visit) */
    int64_t deb_pixels, deb_labels;
    /* photometry work of the last endobject_gpu call: the aperture pixels summed (inside
       the APER outer radius, inside both AUTO ellipses) */
    int64_t phot_pixels;
    /* pixels of the multi-peak raw objects in each deblending size class (0: 256-lane LDS
       trees, 1 and 2: 1024-lane LDS trees, 3: global-memory trees) */
    int64_t deb_class_pixels[4];
} sxg_frame_info;
int sxg_get_frame_info(sxg_frame_info *info);

/* Background maps after filtering (nbx*nby floats each): back, sigma, dback, dsigma. */
int sxg_get_back_maps(float *back, float *sigma, float *dback, float *dsigma);

/* Filtered (detection) frame, width*height floats. */
int sxg_get_filtered(float *out);

/* Per-pixel object index (deblended objects, before CLEAN) or -1: width*height int32. */
int sxg_get_alloc_map(int32_t *out);

/* Device pointers of the resident frame (for callers that keep data on the GPU):
   background-subtracted frame and filtered frame.  Valid until clear_device. */
void *sxg_device_frame(int which);

/* Timing of the last frame's stages in milliseconds (HIP events): background, filter,
   detection, deblending, analysis, clean.  Returns the number of stages written. */
int sxg_stage_times(float *ms, int n);

/* Release every device buffer the library holds (the workspace is grow-only and kept
   across frames otherwise). */
void sxg_release(void);

/* Configuration of a whole-frame run (the .sex keys the reference front-end passes). */
typedef struct {
    int back_size, back_fsize;
    float back_fthresh;
    int back_type;              /* 0 AUTO, 1 MANUAL (reference), 2 MANUAL (stock) */
    float back_value;
    int thresh_type;            /* 0 RELATIVE, 1 ABSOLUTE */
    double detect_thresh, analysis_thresh;
    int ext_minarea, deblend_nthresh;
    double deblend_mincont;
    int clean;
    double clean_param;
    int clean_stacksize;
    double satur_level, gain;
    const float *mask;          /* square filter mask (host memory) or NULL */
    int mask_width;
} sxg_params;

/* Keep a copy of a frame in device memory (one H2D transfer). */
int sxg_set_resident_frame(const float *host, int width, int height);
/* Run the whole path on the resident frame, in the reference's scan.c order, leaving
   the catalogue columns and the segmentation map on the device.  Returns the number
   of deblended objects or < 0. */
int sxg_process_resident(const sxg_params *p, sxg_frame_info *info);
/* Wait for the library's stream. */
void sxg_sync(void);

/* MAP_WEIGHT weight map for the frames processed next on this context (no counterpart
   in the reference GPU path; the CPU semantics of weight.c:91-126 with sigfac = 1 as
   weight.c:80 fixes it): var = 1/w for w > 0, else BIG; pixels of non-positive weight
   leave the mesh statistics; with relative thresholds (thresh_type 0) the detection and
   analysis thresholds of each pixel are detect_thresh*sqrt(var) and
   analysis_thresh*sqrt(var), a raw object's deblending base is the lowest threshold among
   its pixels; FLUXERR_ISO and the position errors use the per-pixel variance.  w == NULL
   removes the map.  The map must have the size of the frame.  Returns 0 or < 0. */
int sxg_set_weight(const float *w, int width, int height, int thresh_type,
                   double detect_thresh, double analysis_thresh);

/* ---- contexts and batches (no counterpart in the reference) -------------------
   A context holds everything the entry points above keep between calls (frame buffers,
   workspace, stream).  The calling thread uses the context it bound last, or the
   process-wide default context: several host threads, each bound to its own context,
   can run the whole path on different frames at the same time.  Errors are per
   thread. */
typedef struct sxg_ctx sxg_ctx;
sxg_ctx *sxg_ctx_create(void);
void sxg_ctx_destroy(sxg_ctx *ctx);
/* Bind ctx to the calling thread (NULL: the default context).  Returns 0. */
int sxg_ctx_bind(sxg_ctx *ctx);

#define SXG_MAX_WORKERS 16
/* Copy n frames into device memory once (replacing any previous batch). */
int sxg_batch_set_frames(int n, const float *const *host, const int *width, const int *height);
/* Run the whole path on every resident frame with nworkers concurrent contexts, one
   host thread and one stream each.  info[i] and frame_ms[i] (host wall time of frame i
   from its start to its segmentation map) may be NULL.  The segmentation map of each
   frame stays on the device.  Returns the number of frames or < 0. */
int sxg_batch_process(const sxg_params *p, int nworkers, sxg_frame_info *info,
                      double *frame_ms);
/* The batched path end to end, from host memory: frame i (width[i] x height[i] floats at
   frames[i]) is copied in by the worker that takes it, on that worker's stream, the whole
   path runs, and the segmentation map is copied to segs[i] (width*height int32) and the
   catalogue rows to rows[i] (up to row_cap[i] rows; rows or rows[i] NULL: none).  With
   pinned host buffers (hipHostMalloc) the copies of one worker overlap the kernels of the
   others.  info[i] and frame_ms[i] (host wall time of frame i from its copy in to its
   last copy out) may be NULL.  Returns the number of frames or < 0. */
int sxg_batch_process_host(const sxg_params *p, int nworkers, int n, const float *const *frames,
                           const int *width, const int *height, int32_t *const *segs,
                           sxg_isorow *const *rows, const int *row_cap, sxg_frame_info *info,
                           double *frame_ms);
/* The resident batch as one HIP graph: a branch per worker context (frames assigned
   longest first to the least loaded branch), every frame a count-planned run whose
   launches are sized by capacities from an eager calibration run of that frame (counts
   plus the plan margin, sxg_batch_set_plan_margin), with no host synchronisation and no
   allocation inside.  The graph is built on the first call, and again when the
   parameters, nworkers or the frames change or a worker's buffers were reallocated; each
   call replays it once.  A frame whose counts exceed its plan (e.g. new pixels loaded
   with sxg_batch_update_frame) is run again eagerly on its context.  The results equal
   sxg_batch_process's.  info[i] as there; frame_ms[i]: device time from frame i's first
   kernel to its last (its branch's clock, wall_clock64).  Returns the number of frames
   run again eagerly, or < 0. */
int sxg_batch_process_graph(const sxg_params *p, int nworkers, sxg_frame_info *info,
                            double *frame_ms);
/* Plan margin of the graph batch in percent (default 25; 0..1000). */
int sxg_batch_set_plan_margin(int percent);
/* Graph builds, launches and eager reruns so far; returns 1 when a graph is built. */
int sxg_batch_graph_stats(long *builds, long *launches, long *reruns);
/* Replace the pixels of resident frame i (same size; the graph stays valid). */
int sxg_batch_update_frame(int frame, const float *host);
/* Autotuner of the graph batch (no counterpart in the reference): the number of graph
   branches (1, 2, 4, 8, 16 up to max_branches) and the deblending size classes
   (SXG_OPT_DEB_LDS_TINY/SMALL/LARGE) that replay the resident batch fastest, by
   coordinate descent, each candidate timed over `replays` graph replays.  Choices are
   kept in a table keyed by the batch's mean frame size, detected-pixel density and frame
   count (log2 buckets): a batch with a known key takes the stored choice at once
   (candidates = 0).  The chosen classes are set as the options; pass `branches` to
   sxg_batch_process_graph.  Every candidate's results (each frame's counts and a hash of
   its segmentation map) must equal the first one's, else an error is returned.
   Returns 0 or < 0. */
typedef struct {
    int branches;
    long deb_lds_tiny, deb_lds_small, deb_lds_large;
    double ms_per_batch;    /* the chosen setting's replay time when it was tuned */
    int candidates;         /* settings timed by this call */
    int size_bucket, density_bucket, count_bucket;   /* the table key */
} sxg_tune_choice;
int sxg_batch_autotune(const sxg_params *p, int max_branches, int replays, sxg_tune_choice *out);
/* Empty the autotuner's table. */
void sxg_autotune_forget(void);
/* Segmentation map of frame i of the last sxg_batch_process. */
int sxg_batch_segmap(int frame, int32_t *out);
/* Free the batch and its worker contexts. */
void sxg_batch_release(void);

/* Per-kernel HIP-event timers for the default context (off by default; enabling resets
   them). */
enum sxg_kernel_id {
    SXG_KT_COPY = 0, SXG_KT_MESHSTAT, SXG_KT_MESHMODE, SXG_KT_FILTERBACK, SXG_KT_ROWNODES,
    SXG_KT_SUBBACK, SXG_KT_CONV, SXG_KT_DET_INIT, SXG_KT_DET_FG, SXG_KT_DET_UNION,
    SXG_KT_DET_FLATTEN, SXG_KT_DET_STAT, SXG_KT_DET_SORT, SXG_KT_DEB_THRESH,
    SXG_KT_DEB_LEVELS, SXG_KT_DEB_OBJECTS, SXG_KT_ANALYSE, SXG_KT_CLEAN, SXG_KT_SEG,
    SXG_KT_DEB_LMAX, SXG_KT_OBJ_SORT, SXG_KT_PHOT,
    /* the deblending tree kernels one by one, each timed on the stream it runs on: size
       classes 0, 1, 2 (k_deb_tree_lds), the one-workgroup global-memory trees
       (k_deb_tree_glb) and the level-parallel chain (k_deb_glb_prep/levels/tree) */
    SXG_KT_DEB_TREE0, SXG_KT_DEB_TREE1, SXG_KT_DEB_TREE2, SXG_KT_DEB_TREE_GLB, SXG_KT_DEB_LVPAR,
    SXG_NKT
};
void sxg_kernel_timing(int enable);
int sxg_kernel_times(int id, double *total_ms, int *launches);

/* Options (no counterpart in the reference).  These never change a result: they only
   choose which kernel variant handles which work.
     SXG_OPT_DEB_LDS_TINY   raw objects (with two or more local maxima) up to this many
                            pixels are deblended by 256-lane workgroups in LDS (default 512)
     SXG_OPT_DEB_LDS_SMALL  up to this many by 1024-lane workgroups in LDS (default 1024)
     SXG_OPT_DEB_LDS_LARGE  up to this many by 1024-lane workgroups in LDS (default 6144,
                            at most 6400); larger ones use global scratch memory
     SXG_OPT_DEB_TABLE      deblending-tree nodes an LDS workgroup can hold (0: 192, 512
                            and 1024 for the three classes; at most 2048, and what the
                            class's pixels leave of the LDS); a raw object whose tree may
                            need more is built in global memory
     SXG_OPT_SIDE_STREAMS   1 (default): the deblending size classes of a frame run side by
                            side on streams of their own; 0: all on the context's stream
     SXG_OPT_DEB_GLB_LEVELS the raw objects built in global memory with at least this many
                            pixels (default 12288; at most 2^30) get the components of all
                            their levels from one workgroup per level, in parallel; the
                            others, and all of them at 0, one growing union-find per tree,
                            level after level
     SXG_OPT_DEB_ORDER_FULL the global-memory objects are launched largest first when there
                            are at most this many of them (default 4096); beyond, only the
                            level-parallel ones are ordered, the others follow in any order
     SXG_OPT_MESH_ROWS      frame rows per batch of the mesh-statistics kernel's LDS ring
                            (0, default: as many as fit in 64 KB per buffer, at most 8; else
                            1..8): fewer rows hold less LDS while its chains run, which
                            leaves room on each CU for other frames' kernels
     SXG_OPT_FUSED_FILTER   1 (default): the whole-frame runs (sxg_process_resident and the
                            batches) subtract the background spline, convolve and write the
                            detection parents in one kernel for frames of a width that is a
                            multiple of 64 and masks of 3 to 11; 0: three kernels, as the
                            reference's entry points run them
   This one changes the catalogue (never the segmentation map, the object list or the
   CLEAN decisions):
     SXG_OPT_PRECISION      arithmetic of FLUX_ISO, X/Y, X2/Y2/XY, A/B/THETA, CXX/CYY/CXY:
                            0 the reference's (default, bit-exact), 1 two-pass FP64,
                            2 compensated FP32 sums with FP64 central moments */
enum sxg_option_id { SXG_OPT_DEB_LDS_SMALL = 0, SXG_OPT_DEB_LDS_LARGE, SXG_OPT_PRECISION,
                     SXG_OPT_DEB_LDS_TINY, SXG_OPT_DEB_TABLE, SXG_OPT_SIDE_STREAMS,
                     SXG_OPT_DEB_GLB_LEVELS, SXG_OPT_DEB_ORDER_FULL, SXG_OPT_MESH_ROWS,
                     SXG_OPT_FUSED_FILTER, SXG_NOPT };
/* Returns 0, or -1 for an unknown option or a value out of range. */
int sxg_set_option(int id, long value);
long sxg_get_option(int id);

/* ---- allocation-free entry point ----------------------------------------------
   No counterpart in the reference.  For callers that cannot allocate device memory
   (code under HIP stream capture, tight serving loops, harnesses that forbid
   hipMalloc): back_estimate's kernels on a frame the caller holds in device memory.
   src: width x height floats (device), dst: the background-subtracted frame (device;
   src == dst allowed), on stream (a hipStream_t; NULL: the null stream).  back_type 0
   AUTO (spline), 1 MANUAL as the reference's device code (the global median), 2 MANUAL
   with back_value.  backmean, backsig (host, may be NULL) receive the global medians
   (back_value for type 2); the call waits for its stream.  No weight map.
   Every intermediate table lives in one device array of SXG_NOALLOC_SCRATCH_BYTES per
   GPU, held until the process ends: 8192 x 8192 with BACK_SIZE 64 needs 13.7 MB; a
   frame whose tables need more is rejected (-1, sxg_last_error says how much).  Calls
   that overlap on one GPU share that array and race.  Allocates nothing: no hipMalloc,
   no hipFree, no context workspace.  Returns 0 or -1. */
#define SXG_NOALLOC_SCRATCH_BYTES (16 * 1024 * 1024)
int sxg_back_estimate_noalloc(const float *src, float *dst, int width, int height,
                              int back_size, float backfthresh, int nbackfx, int nbackfy,
                              int back_type, float back_value, float *backmean,
                              float *backsig, void *stream);

/* Catalogue row, layout of the reference's outobjstruct (src/types.h:417-455, NAPER 1). */
typedef struct __attribute__((aligned(16))) {
    int number;
    double alpha2000, delta2000;
    float sposx, sposy;
    float flux_iso, fluxerr_iso, mag_iso, magerr_iso;
    float flux_isocor, fluxerr_isocor;
    float flux_aper[1], fluxerr_aper[1], mag_aper[1], magerr_aper[1];
    float flux_auto, fluxerr_auto, mag_auto, magerr_auto;
    float theta;
    double poserr_mx2, poserr_my2, poserr_mxy;
    short flag;
    float elong, ellip, sprob, peak, kronfactor, bkg, fwhm;
} sxg_outobj;

#ifdef __cplusplus
}
#endif
#endif
