/*
 * fitsio.c -- minimal FITS image I/O (primary HDU, 2880-byte blocks, big-endian data).
 * Reading follows the conversions of the reference's FITS library
 * (src/fits/fitsbody.c:195-345, src/fits/fitshead.c:148-160): integer data are scaled by
 * BSCALE/BZERO in float, BLANK and NaN/Inf pixels are replaced by -BIG so that they never
 * pass a threshold.
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

/* The conversions of read_body (fitsbody.c:195-345): BSCALE and BZERO are rounded to
   float (PIXTYPE) and integer pixels are scaled in float arithmetic (BITPIX 8 bytes are
   signed unless BITSGN = 0, fitshead.c:159-160); pixels equal to BLANK, and IEEE NaN or
   Inf pixels, become -BIG; BITPIX -64 pixels are scaled in double, then rounded. */
float *sxf_read_image(const char *path, int *width, int *height, double *gain, double *satur,
                      const char **err)
{
    FILE *f = fopen(path, "rb");
    if (!f) { if (err) *err = "cannot open file"; return NULL; }
    char card[80];
    double bitpix = 0, naxis = 0, n1 = 0, n2 = 0, bscale = 1.0, bzero = 0.0, blank = 0.0,
           bitsgn = 1.0, v;
    int blankflag = 0;
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
        else if (card_value(card, "BLANK", &v)) { blank = v; blankflag = 1; }
        else if (card_value(card, "BITSGN", &v)) bitsgn = v;
        else if (gain && card_value(card, "GAIN", &v)) *gain = v;
        else if (satur && card_value(card, "SATURATE", &v)) *satur = v;
    }
    if (!end || naxis < 2 || n1 < 1 || n2 < 1) { fclose(f); if (err) *err = "bad FITS header"; return NULL; }
    long hdrbytes = ((ncards * 80 + FBLOCK - 1) / FBLOCK) * FBLOCK;
    fseek(f, hdrbytes, SEEK_SET);
    int W = (int)n1, H = (int)n2, bp = (int)bitpix, nb = abs(bp) / 8;
    if (bp != 8 && bp != 16 && bp != 32 && bp != -32 && bp != -64) {
        fclose(f);
        if (err) *err = "unsupported BITPIX";
        return NULL;
    }
    size_t n = (size_t)W * H;
    unsigned char *raw = (unsigned char *)malloc(n * nb);
    float *img = (float *)malloc(n * sizeof(float));
    if (!raw || !img || fread(raw, nb, n, f) != n) {
        free(raw); free(img); fclose(f);
        if (err) *err = "short FITS data";
        return NULL;
    }
    fclose(f);
    const float bs = (float)bscale, bz = (float)bzero, mbig = (float)-1e30;
    const long lblank = (long)blank;
    for (size_t i = 0; i < n; i++) {
        const unsigned char *b = raw + i * nb;
        switch (bp) {
        case 8: {
            const int cv = bitsgn != 0.0 ? (int)(signed char)b[0] : (int)b[0];
            const int bl = bitsgn != 0.0 ? (int)(signed char)lblank : (int)(unsigned char)lblank;
            img[i] = blankflag && cv == bl ? mbig : (float)cv * bs + bz;
            break;
        }
        case 16: {
            const uint16_t u = (uint16_t)((b[0] << 8) | b[1]);
            const int sv = bitsgn != 0.0 ? (int)(int16_t)u : (int)u;
            const int bl = bitsgn != 0.0 ? (int)(int16_t)lblank : (int)(uint16_t)lblank;
            img[i] = blankflag && sv == bl ? mbig : (float)sv * bs + bz;
            break;
        }
        case 32: {
            const uint32_t u = ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) |
                               ((uint32_t)b[2] << 8) | b[3];
            if (bitsgn != 0.0) {
                const int32_t iv = (int32_t)u;
                img[i] = blankflag && iv == (int32_t)lblank ? mbig : (float)iv * bs + bz;
            } else {
                img[i] = blankflag && u == (uint32_t)lblank ? mbig : (float)u * bs + bz;
            }
            break;
        }
        case -32: {
            const uint32_t u = ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) |
                               ((uint32_t)b[2] << 8) | b[3];
            float fv;
            memcpy(&fv, &u, 4);
            img[i] = (u & 0x7f800000u) == 0x7f800000u ? mbig : fv * bs + bz;
            break;
        }
        default: {   /* -64 */
            uint64_t u = 0;
            for (int k = 0; k < 8; k++) u = (u << 8) | b[k];
            double dv;
            memcpy(&dv, &u, 8);
            img[i] = (u & 0x7ff0000000000000ull) == 0x7ff0000000000000ull
                         ? mbig
                         : (float)(dv * (double)bs + (double)bz);
            break;
        }
        }
    }
    free(raw);
    *width = W;
    *height = H;
    return img;
}

/* the string value of a card (quotes and trailing blanks removed) */
static int card_string(const char *card, const char *key, char *out, int n)
{
    size_t kl = strlen(key);
    if (strncmp(card, key, kl) || (card[kl] != ' ' && card[kl] != '=')) return 0;
    const char *q = memchr(card, '\'', 80);
    if (!q) return 0;
    int k = 0;
    for (const char *c = q + 1; c < card + 80 && *c != '\'' && k < n - 1; c++) out[k++] = *c;
    while (k > 0 && out[k - 1] == ' ') k--;
    out[k] = 0;
    return 1;
}

int sxf_read_wcs(const char *path, char ctype[2][9], double crval[2], double crpix[2],
                 double cd[4], double *longpole, double *latpole, double *equinox)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    char card[80], s[80];
    int have_ct[2] = {0, 0}, have_cd = 0, have_pc = 0, have_eq = 0;
    double cdelt[2] = {1.0, 1.0}, cdm[4] = {1.0, 0.0, 0.0, 1.0}, pc[4] = {1.0, 0.0, 0.0, 1.0};
    double crota = 0.0, epoch = 2000.0, v;
    crval[0] = crval[1] = 0.0;
    crpix[0] = crpix[1] = 1.0;
    *longpole = *latpole = 999.0;
    *equinox = 2000.0;
    while (fread(card, 1, 80, f) == 80) {
        if (!strncmp(card, "END ", 4)) break;
        for (int l = 0; l < 2; l++) {
            char key[16];
            snprintf(key, sizeof(key), "CTYPE%d", l + 1);
            if (card_string(card, key, s, sizeof(s))) {
                const size_t m = strlen(s) < 8 ? strlen(s) : 8;
                memset(ctype[l], 0, 9);
                memcpy(ctype[l], s, m);
                have_ct[l] = 1;
            }
            snprintf(key, sizeof(key), "CRVAL%d", l + 1);
            if (card_value(card, key, &v)) crval[l] = v;
            snprintf(key, sizeof(key), "CRPIX%d", l + 1);
            if (card_value(card, key, &v)) crpix[l] = v;
            snprintf(key, sizeof(key), "CDELT%d", l + 1);
            if (card_value(card, key, &v)) cdelt[l] = v;
            for (int j = 0; j < 2; j++) {
                snprintf(key, sizeof(key), "CD%d_%d", l + 1, j + 1);
                if (card_value(card, key, &v)) { cdm[2 * l + j] = v; have_cd = 1; }
                snprintf(key, sizeof(key), "PC%03d%03d", l + 1, j + 1);
                if (card_value(card, key, &v)) { pc[2 * l + j] = v; have_pc = 1; }
            }
        }
        if (card_value(card, "CROTA2", &v)) crota = v;
        if (card_value(card, "LONPOLE", &v) || card_value(card, "LONGPOLE", &v)) *longpole = v;
        if (card_value(card, "LATPOLE", &v)) *latpole = v;
        if (card_value(card, "EPOCH", &v)) epoch = v;
        if (card_value(card, "EQUINOX", &v)) { *equinox = v; have_eq = 1; }
    }
    fclose(f);
    if (!have_eq) *equinox = epoch;
    if (have_cd) {
        for (int k = 0; k < 4; k++) cd[k] = cdm[k];
    } else if (have_pc) {
        for (int k = 0; k < 4; k++) cd[k] = pc[k] * cdelt[k / 2];
    } else {
        const double r = crota * 3.14159265358979323846 / 180.0;
        cd[0] = cos(r) * cdelt[0];
        cd[2] = sin(r) * cdelt[0];
        cd[1] = -sin(r) * cdelt[1];
        cd[3] = cos(r) * cdelt[1];
    }
    return have_ct[0] && have_ct[1];
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
