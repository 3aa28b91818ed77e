/*
 * fitsio.c -- minimal FITS image I/O (primary HDU, 2880-byte blocks, big-endian data).
 * Reading follows the conversions of the reference's FITS library
 * (src/fits/fitsbody.c:195-466): integer data are scaled by BSCALE/BZERO, floating-point
 * NaN/Inf pixels are replaced by -BIG so that they never pass a threshold.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fitsio.h"

#define FBLOCK 2880

static int card_value(const char *card, const char *key, double *val)
{
    size_t kl = strlen(key);
    if (strncmp(card, key, kl) || (card[kl] != ' ' && card[kl] != '=')) return 0;
    const char *eq = memchr(card, '=', 80);
    if (!eq) return 0;
    char buf[81];
    memcpy(buf, eq + 1, (size_t)(card + 80 - (eq + 1)));
    buf[card + 80 - (eq + 1)] = 0;
    *val = atof(buf);
    return 1;
}

float *sxf_read_image(const char *path, int *width, int *height, double *gain, double *satur,
                      const char **err)
{
    FILE *f = fopen(path, "rb");
    if (!f) { if (err) *err = "cannot open file"; return NULL; }
    char card[80];
    double bitpix = 0, naxis = 0, n1 = 0, n2 = 0, bscale = 1.0, bzero = 0.0, v;
    long ncards = 0;
    int end = 0;
    while (!end && fread(card, 1, 80, f) == 80) {
        ncards++;
        if (!strncmp(card, "END     ", 8) || !strncmp(card, "END ", 4)) end = 1;
        else if (card_value(card, "BITPIX", &v)) bitpix = v;
        else if (card_value(card, "NAXIS1", &v)) n1 = v;
        else if (card_value(card, "NAXIS2", &v)) n2 = v;
        else if (card_value(card, "NAXIS", &v)) naxis = v;
        else if (card_value(card, "BSCALE", &v)) bscale = v;
        else if (card_value(card, "BZERO", &v)) bzero = v;
        else if (gain && card_value(card, "GAIN", &v)) *gain = v;
        else if (satur && card_value(card, "SATURATE", &v)) *satur = v;
    }
    if (!end || naxis < 2 || n1 < 1 || n2 < 1) { fclose(f); if (err) *err = "bad FITS header"; return NULL; }
    long hdrbytes = ((ncards * 80 + FBLOCK - 1) / FBLOCK) * FBLOCK;
    fseek(f, hdrbytes, SEEK_SET);
    int W = (int)n1, H = (int)n2, bp = (int)bitpix, nb = abs(bp) / 8;
    size_t n = (size_t)W * H;
    unsigned char *raw = (unsigned char *)malloc(n * nb);
    float *img = (float *)malloc(n * sizeof(float));
    if (!raw || !img || fread(raw, nb, n, f) != n) {
        free(raw); free(img); fclose(f);
        if (err) *err = "short FITS data";
        return NULL;
    }
    fclose(f);
    for (size_t i = 0; i < n; i++) {
        unsigned char *b = raw + i * nb;
        double x = 0.0;
        switch (bp) {
        case 8: x = b[0] * bscale + bzero; break;
        case 16: x = (int16_t)((b[0] << 8) | b[1]) * bscale + bzero; break;
        case 32: x = (int32_t)(((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3]) * bscale + bzero; break;
        case -32: {
            uint32_t u = ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
            float fv;
            memcpy(&fv, &u, 4);
            if ((u & 0x7f800000u) == 0x7f800000u) { img[i] = -1e30f; continue; }
            img[i] = (bscale == 1.0 && bzero == 0.0) ? fv : (float)(fv * bscale + bzero);
            continue;
        }
        case -64: {
            uint64_t u = 0;
            for (int k = 0; k < 8; k++) u = (u << 8) | b[k];
            double dv;
            memcpy(&dv, &u, 8);
            if (!isfinite(dv)) { img[i] = -1e30f; continue; }
            x = dv * bscale + bzero;
            break;
        }
        default: free(raw); free(img); if (err) *err = "unsupported BITPIX"; return NULL;
        }
        img[i] = (float)x;
    }
    free(raw);
    *width = W;
    *height = H;
    return img;
}

static void put_card(FILE *f, const char *text)
{
    char card[80];
    memset(card, ' ', 80);
    memcpy(card, text, strlen(text) < 80 ? strlen(text) : 80);
    fwrite(card, 1, 80, f);
}

int sxf_write_image(const char *path, const void *data, int width, int height, int bitpix)
{
    if (bitpix != -32 && bitpix != 32) return -1;
    FILE *f = fopen(path, "wb");
    if (!f) return -2;
    char line[81];
    int ncard = 0;
    put_card(f, "SIMPLE  =                    T"); ncard++;
    snprintf(line, sizeof(line), "BITPIX  = %20d", bitpix); put_card(f, line); ncard++;
    put_card(f, "NAXIS   =                    2"); ncard++;
    snprintf(line, sizeof(line), "NAXIS1  = %20d", width); put_card(f, line); ncard++;
    snprintf(line, sizeof(line), "NAXIS2  = %20d", height); put_card(f, line); ncard++;
    put_card(f, "END"); ncard++;
    for (; ncard % 36; ncard++) put_card(f, "");
    size_t n = (size_t)width * height;
    const uint32_t *src = (const uint32_t *)data;
    unsigned char buf[4 * 1024];
    for (size_t i = 0; i < n;) {
        size_t k = 0;
        for (; k < 1024 && i < n; k++, i++) {
            uint32_t u = src[i];
            buf[4 * k] = u >> 24; buf[4 * k + 1] = u >> 16; buf[4 * k + 2] = u >> 8; buf[4 * k + 3] = u;
        }
        fwrite(buf, 4, k, f);
    }
    size_t bytes = n * 4, pad = (FBLOCK - bytes % FBLOCK) % FBLOCK;
    memset(buf, 0, sizeof(buf));
    while (pad) { size_t c = pad < sizeof(buf) ? pad : sizeof(buf); fwrite(buf, 1, c, f); pad -= c; }
    fclose(f);
    return 0;
}
