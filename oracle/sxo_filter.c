/*
 * sxo_filter.c -- oracle restatement of the reference convolution (src/filter.c).
 * TEST INFRASTRUCTURE ONLY (never linked into the product).
 *
 * convolve (filter.c:39-94): correlation of the frame with the mask, no flip; for each
 * output pixel the products mask*pixel are added to a zero-initialised float in mask
 * row-major order, taps that fall outside the frame being skipped; no FMA contraction.
 * getconv (filter.c:101-186): "CONV [NORM|NONORM]" header, rows of numbers, '#'
 * comments; NORM divides each element by the sum of absolute values computed in double
 * (by the quadratic norm when that sum is zero).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sxo.h"

void sxo_convolve(const float *in, int W, int H, const float *mask, int mw, int mh, float *out)
{
    const int mw2 = mw / 2, mh2 = mh / 2;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float acc = 0.0f;
            for (int i = 0; i < mh; i++) {
                int yy = y + i - mh2;
                if (yy < 0 || yy >= H) continue;
                const float *row = in + (size_t)yy * W;
                for (int j = 0; j < mw; j++) {
                    int xx = x + j - mw2;
                    if (xx < 0 || xx >= W) continue;
                    float prod = mask[i * mw + j] * row[xx];
                    acc = acc + prod;
                }
            }
            out[(size_t)y * W + x] = acc;
        }
}

int sxo_getconv(const char *filename, float *conv, int maxn, int *mw, int *mh)
{
    char str[512], *sstr;
    double sum, var, pix;
    int i, j, n, normflag;
    FILE *file = fopen(filename, "r");
    if (!file) return -1;
    if (!fgets(str, sizeof(str), file) || strncmp(str, "CONV", 4)) {
        fclose(file);
        return -2;
    }
    normflag = strstr(str, "NORM") ? (strstr(str, "NONORM") ? 0 : 1) : 1;
    for (i = 0, n = 0; fgets(str, sizeof(str), file);) {
        j = 0;
        sstr = strtok(str, " \t\n\r");
        if (sstr && sstr[0] != '#')
            do {
                j++;
                if (i >= maxn) { fclose(file); return -3; }
                conv[i++] = (float)atof(sstr);
            } while ((sstr = strtok(NULL, " \t\n\r")));
        if (j > n) n = j;
    }
    fclose(file);
    if (n < 1 || i % n) return -4;
    *mw = n;
    *mh = i / n;
    var = 0.0;
    for (j = 0, sum = 0.0; j < i; j++) {
        sum += fabs(pix = conv[j]);
        var += pix * pix;
    }
    var = sqrt(var);
    if (normflag) {
        if (sum == 0.0) sum = var;
        for (j = 0; j < i; j++) conv[j] = (float)(conv[j] / sum);
    }
    return 0;
}
