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
/* Write a 2-D image: bitpix -32 (float data) or 32 (int32 data). */
int sxf_write_image(const char *path, const void *data, int width, int height, int bitpix);
#ifdef __cplusplus
}
#endif
#endif
