/*
 * sxo_wcs.c -- oracle restatement of the world coordinates of catalogue positions.
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker of the GPU astrometry.
 *
 * The reference computes ALPHA_J2000/DELTA_J2000 in computeastrom (src/astrom.c:86-140):
 * raw_to_wcs (src/fitswcs.c:1251-1268, the linear transform, projection and spherical
 * rotation of its WCSLIB 2.x, src/wcs/) of the 1-based position (X_IMAGE, Y_IMAGE), then
 * precess (src/fitswcs.c:1718-1755) to J2000 when the equinox is not 2000.  The GPU path of
 * the reference writes zeros instead (src/cuda/cudaanalyse.cu:3033-3047, :3127-3128).
 *
 * Restated here from the published definitions, not from the WCSLIB code:
 *   - intermediate world coordinates x_i = sum_j CD_ij (p_j - CRPIX_j), degrees;
 *   - the zenithal projections of Calabretta & Greisen (2002, A&A 395, 1077), sect. 5.1:
 *     native longitude phi = arg(-y, x) and latitude theta from R = sqrt(x^2 + y^2):
 *     TAN theta = atan(180 / (pi R)), SIN theta = acos(pi R / 180), ARC theta = 90 - R,
 *     STG theta = 90 - 2 atan(pi R / 360), ZEA theta = 90 - 2 asin(pi R / 360);
 *   - the native-to-celestial rotation of their eq. (2) with the native pole at the
 *     reference point (theta_0 = 90) and phi_p = LONPOLE, by default 180 (the reference's
 *     WCSLIB 2.x takes 180 for every zenithal projection, even at delta_0 = 90, where the
 *     2002 paper takes 0);
 *   - the IAU 1976 precession (Lieske et al. 1977, A&A 58, 1, eq. 6): the angles zeta, z,
 *     theta as polynomials in the start epoch and the interval (Julian centuries), then
 *     the rotation of the equatorial frame by them.
 */
#include <math.h>
#include <string.h>

#include "sxo.h"

#define WCS_D2R (3.14159265358979323846 / 180.0)

/* projection code of a CTYPE value ("RA---TAN" -> TAN), -1 if not a supported one */
static int proj_code(const char *ctype)
{
    static const char *codes[] = {"TAN", "SIN", "ARC", "STG", "ZEA"};
    if (strlen(ctype) < 8 || ctype[4] != '-') return -1;
    for (int i = 0; i < 5; i++)
        if (!strncmp(ctype + 5, codes[i], 3)) return i;
    return -1;
}

int sxo_wcs_setup(sxo_wcs *w, const char *ctype1, const char *ctype2, const double *crval,
                  const double *crpix, const double *cd, double longpole, double equinox)
{
    memset(w, 0, sizeof(*w));
    const int ra1 = !strncmp(ctype1, "RA--", 4), dec1 = !strncmp(ctype1, "DEC-", 4);
    const int ra2 = !strncmp(ctype2, "RA--", 4), dec2 = !strncmp(ctype2, "DEC-", 4);
    if (!((ra1 && dec2) || (dec1 && ra2))) return -1;
    w->lng = ra1 ? 0 : 1;
    w->lat = 1 - w->lng;
    w->proj = proj_code(ctype1);
    if (w->proj < 0 || w->proj != proj_code(ctype2)) return -1;
    for (int i = 0; i < 2; i++) {
        w->crval[i] = crval[i];
        w->crpix[i] = crpix[i];
    }
    for (int i = 0; i < 4; i++) w->cd[i] = cd[i];
    /* LONPOLE 999: the default of the reference's WCSLIB for a native pole at the
       reference point (src/wcs/cel.c:403-413), 180 whatever delta_0 */
    w->phip = longpole != 999.0 ? longpole : 180.0;
    w->equinox = equinox;
    return 0;
}

int sxo_wcs_world(const sxo_wcs *w, double px, double py, double *world, double *radec)
{
    const double dx = px - w->crpix[0], dy = py - w->crpix[1];
    double img[2];
    img[0] = w->cd[0] * dx + w->cd[1] * dy;
    img[1] = w->cd[2] * dx + w->cd[3] * dy;
    const double X = img[w->lng], Y = img[w->lat];
    const double R = sqrt(X * X + Y * Y);
    const double phi = R == 0.0 ? 0.0 : atan2(X, -Y) / WCS_D2R;
    double theta;
    switch (w->proj) {
    case 0: theta = atan2(180.0 / M_PI, R) / WCS_D2R; break;
    case 1:
        if (R * M_PI / 180.0 > 1.0) return -1;
        theta = acos(R * M_PI / 180.0) / WCS_D2R;
        break;
    case 2: theta = 90.0 - R; break;
    case 3: theta = 90.0 - 2.0 * atan(R * M_PI / 360.0) / WCS_D2R; break;
    default:
        if (R * M_PI / 360.0 > 1.0) return -1;
        theta = 90.0 - 2.0 * asin(R * M_PI / 360.0) / WCS_D2R;
        break;
    }
    const double ap = w->crval[w->lng] * WCS_D2R, dp = w->crval[w->lat] * WCS_D2R;
    const double t = theta * WCS_D2R, dphi = (phi - w->phip) * WCS_D2R;
    const double st = sin(t), ct = cos(t), sdp = sin(dp), cdp = cos(dp);
    double sd = st * sdp + ct * cdp * cos(dphi);
    if (sd > 1.0) sd = 1.0;
    if (sd < -1.0) sd = -1.0;
    const double delta = asin(sd) / WCS_D2R;
    double alpha = (ap + atan2(-ct * sin(dphi), st * cdp - ct * sdp * cos(dphi))) / WCS_D2R;
    alpha = fmod(alpha, 360.0);
    if (alpha < 0.0) alpha += 360.0;
    world[w->lng] = alpha;
    world[w->lat] = delta;
    if (fabs(w->equinox - 2000.0) > 0.003)
        sxo_precess(w->equinox, alpha, delta, 2000.0, &radec[0], &radec[1]);
    else {
        radec[0] = alpha;
        radec[1] = delta;
    }
    return 0;
}

void sxo_precess(double yin, double ain, double din, double yout, double *aout, double *dout)
{
    /* Lieske et al. (1977) eq. (6): the angles in arcseconds, T the start epoch from J2000
       and t the interval, both in Julian centuries */
    const double T = (yin - 2000.0) / 100.0, t = (yout - yin) / 100.0;
    const double as2r = WCS_D2R / 3600.0;
    const double zeta = ((2306.2181 + 1.39656 * T - 0.000139 * T * T) * t +
                         (0.30188 - 0.000344 * T) * t * t + 0.017998 * t * t * t) * as2r;
    const double z = ((2306.2181 + 1.39656 * T - 0.000139 * T * T) * t +
                      (1.09468 + 0.000066 * T) * t * t + 0.018203 * t * t * t) * as2r;
    const double th = ((2004.3109 - 0.85330 * T - 0.000217 * T * T) * t +
                       (-0.42665 - 0.000217 * T) * t * t - 0.041833 * t * t * t) * as2r;
    const double a = ain * WCS_D2R + zeta, d = din * WCS_D2R;
    const double A = cos(d) * sin(a);
    const double B = cos(th) * cos(d) * cos(a) - sin(th) * sin(d);
    const double Cz = sin(th) * cos(d) * cos(a) + cos(th) * sin(d);
    double ao = (atan2(A, B) + z) / WCS_D2R;
    ao = fmod(ao, 360.0);
    if (ao < 0.0) ao += 360.0;
    *aout = ao;
    *dout = asin(Cz > 1.0 ? 1.0 : Cz < -1.0 ? -1.0 : Cz) / WCS_D2R;
}
