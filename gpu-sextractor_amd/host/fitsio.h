/* fitsio.h -- minimal FITS image reader/writer for the front-end. */
#ifndef SXF_FITSIO_H
#define SXF_FITSIO_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
/* Read the primary image as float: BITPIX 8/16/32/-32/-64 with BSCALE/BZERO; for
   floating-point data NaN/Inf pixels become -1e30 (src/fits/fitsbody.c:195-466). Returns
   a malloc'd buffer or NULL; *err receives a message on failure. */
float *sxf_read_image(const char *path, int *width, int *height, double *gain,
                      double *satur, const char **err);
/* The celestial WCS keywords of the primary header, as read_wcs reads them
   (src/fitswcs.c:320-470): CTYPE1/2, CRVAL (0), CRPIX (1), the CD matrix from CDi_j, else
   PC00i00j times CDELTi, else CDELTi and CROTA2; LONPOLE/LATPOLE (999 when absent);
   EQUINOX (else EPOCH, else 2000).  Returns 1 when CTYPE1 and CTYPE2 are present, 0 when
   not, -1 when the file cannot be read. */
int sxf_read_wcs(const char *path, char ctype[2][9], double crval[2], double crpix[2],
                 double cd[4], double *longpole, double *latpole, double *equinox);
/* Write a 2-D image: bitpix -32 (float data) or 32 (int32 data). */
int sxf_write_image(const char *path, const void *data, int width, int height, int bitpix);
#ifdef __cplusplus
}
#endif
#endif
