/*
 * wcs.hip -- world coordinates of the catalogue positions (ALPHA_J2000, DELTA_J2000,
 * X_WORLD, Y_WORLD).
 *
 * The reference's GPU path writes zeros (src/cuda/cudaanalyse.cu:3033-3047, :3127-3128);
 * its CPU semantics are computeastrom (src/astrom.c:86-140): raw_to_wcs of the 1-based
 * position (X_IMAGE, Y_IMAGE) through the linear transform, the projection and the
 * spherical rotation of its WCSLIB (src/fitswcs.c:1251, src/wcs/), then precess to J2000
 * (src/fitswcs.c:1718-1755) when the equinox is not 2000.  Covered here: celestial pairs
 * (RA---xxx, DEC--xxx, either order) with the zenithal projections TAN, SIN, ARC, STG and
 * ZEA, whose native pole is the reference point (phi_p = LONPOLE, 180 by default as the
 * reference's WCSLIB sets it); precession by the IAU 1976 angles (Lieske et al. 1977).
 * One lane per object; a position outside the projection's domain gets 1e31 (the
 * reference's WCS_NOCOORD, src/define.h:158).
 */
#include <cstring>

#include "sxg_internal.h"
#include "sxg_objects.h"

#define WCS_NOCOORD 1e31

/* the transformation a kernel applies: projection 0 TAN, 1 SIN, 2 ARC, 3 STG, 4 ZEA */
struct WcsDev {
    int lng = 0, lat = 1, proj = -1;
    double crval[2] = {0.0, 0.0}, crpix[2] = {0.0, 0.0}, cd[4] = {0.0, 0.0, 0.0, 0.0};
    double phip = 180.0, equinox = 2000.0;
};

/* Byte offsets of the wcsstruct fields read here, as the reference's build lays them out
   on x86-64 (src/types.h:547-592, NAXIS 2); checked by a test compiled against the
   reference's headers through sxg_reference_layout. */
enum WcsLayout : long {
    WL_SIZE = 544, WL_NAXIS = 0, WL_CTYPE = 12, WL_CRVAL = 96, WL_CRPIX = 128, WL_CD = 176,
    WL_LONGPOLE = 224, WL_LATPOLE = 232, WL_LAT = 328, WL_LNG = 332, WL_EQUINOX = 408
};

/* the WCS layout entries of sxg_reference_layout (after the prefs/obj2 ones): sizeof,
   naxis, ctype, crval, crpix, cd, longpole, latpole, lat, lng, equinox */
int sxg_wcs_layout(long *out, int n)
{
    static const long v[] = {WL_SIZE, WL_NAXIS, WL_CTYPE, WL_CRVAL, WL_CRPIX, WL_CD,
                             WL_LONGPOLE, WL_LATPOLE, WL_LAT, WL_LNG, WL_EQUINOX};
    const int nv = (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; out && i < n && i < nv; i++) out[i] = v[i];
    return nv;
}

/* the projection of a CTYPE value, -1 if it is not one of the covered ones */
static int proj_of(const char *ct)
{
    static const char *codes[] = {"TAN", "SIN", "ARC", "STG", "ZEA"};
    if (strnlen(ct, 9) < 8 || ct[4] != '-') return -1;
    for (int i = 0; i < 5; i++)
        if (!strncmp(ct + 5, codes[i], 3)) return i;
    return -1;
}

/* a WcsDev from header values; false (error recorded) when it is not covered */
static bool wcs_setup(const char ct1[9], const char ct2[9], const double *crval,
                      const double *crpix, const double *cd, double longpole, double equinox,
                      WcsDev &w)
{
    char a[9] = {0}, b[9] = {0};
    memcpy(a, ct1, 8);
    memcpy(b, ct2, 8);
    const bool ra1 = !strncmp(a, "RA--", 4), dec1 = !strncmp(a, "DEC-", 4);
    const bool ra2 = !strncmp(b, "RA--", 4), dec2 = !strncmp(b, "DEC-", 4);
    if (!((ra1 && dec2) || (dec1 && ra2))) {
        sxg_set_error(std::string("WCS: CTYPE ") + a + "/" + b +
                      " is not a right ascension / declination pair");
        return false;
    }
    w.lng = ra1 ? 0 : 1;
    w.lat = 1 - w.lng;
    w.proj = proj_of(a);
    if (w.proj < 0 || w.proj != proj_of(b)) {
        sxg_set_error(std::string("WCS: projection of ") + a + " is not one of TAN, SIN, ARC, "
                      "STG, ZEA");
        return false;
    }
    for (int i = 0; i < 2; i++) {
        w.crval[i] = crval[i];
        w.crpix[i] = crpix[i];
    }
    for (int i = 0; i < 4; i++) w.cd[i] = cd[i];
    w.phip = longpole != 999.0 ? longpole : 180.0;
    w.equinox = equinox;
    return true;
}

extern "C" int sxg_set_wcs(const sxg_wcs *w)
{
    SxgState &S = sxg_state();
    sxg_clear_error();
    S.wcs_set = false;
    if (!w) return 0;
    WcsDev d;
    if (!wcs_setup(w->ctype[0], w->ctype[1], w->crval, w->crpix, w->cd, w->longpole, w->equinox,
                   d))
        return -1;
    static_assert(sizeof(WcsDev) <= sizeof(S.wcs_dev), "WCS parameter block");
    memcpy(S.wcs_dev, &d, sizeof(d));
    S.wcs_set = true;
    return 0;
}

/* the WCS of the reference's wcsstruct; false with no error when it holds no celestial
   pair (lng == lat: the reference computes nothing), false with an error when its pair is
   not covered */
static bool wcs_from_reference(const void *wcs, WcsDev &d, bool &have)
{
    const char *p = static_cast<const char *>(wcs);
    int naxis = 0, lat = 0, lng = 0;
    memcpy(&naxis, p + WL_NAXIS, sizeof(int));
    memcpy(&lat, p + WL_LAT, sizeof(int));
    memcpy(&lng, p + WL_LNG, sizeof(int));
    have = false;
    if (naxis != 2 || lat == lng || lat < 0 || lng < 0 || lat > 1 || lng > 1) return true;
    double crval[2], crpix[2], cd[4], longpole, equinox;
    memcpy(crval, p + WL_CRVAL, sizeof(crval));
    memcpy(crpix, p + WL_CRPIX, sizeof(crpix));
    memcpy(cd, p + WL_CD, sizeof(cd));
    memcpy(&longpole, p + WL_LONGPOLE, sizeof(double));
    memcpy(&equinox, p + WL_EQUINOX, sizeof(double));
    if (!wcs_setup(p + WL_CTYPE, p + WL_CTYPE + 9, crval, crpix, cd, longpole, equinox, d))
        return false;
    have = true;
    return true;
}

__device__ inline double wrap360(double a)
{
    a = fmod(a, 360.0);
    return a < 0.0 ? a + 360.0 : a;
}

/* IAU 1976 precession (Lieske et al. 1977, eq. 6), epochs in years */
__device__ inline void precess_dev(double yin, double ain, double din, double yout, double &aout,
                                   double &dout)
{
    const double r = M_PI / 180.0;
    const double T = (yin - 2000.0) / 100.0, t = (yout - yin) / 100.0, as2r = r / 3600.0;
    const double zeta = ((2306.2181 + 1.39656 * T - 0.000139 * T * T) * t +
                         (0.30188 - 0.000344 * T) * t * t + 0.017998 * t * t * t) * as2r;
    const double z = ((2306.2181 + 1.39656 * T - 0.000139 * T * T) * t +
                      (1.09468 + 0.000066 * T) * t * t + 0.018203 * t * t * t) * as2r;
    const double th = ((2004.3109 - 0.85330 * T - 0.000217 * T * T) * t +
                       (-0.42665 - 0.000217 * T) * t * t - 0.041833 * t * t * t) * as2r;
    const double a = ain * r + zeta, d = din * r;
    const double A = cos(d) * sin(a);
    const double B = cos(th) * cos(d) * cos(a) - sin(th) * sin(d);
    const double C = sin(th) * cos(d) * cos(a) + cos(th) * sin(d);
    aout = wrap360((atan2(A, B) + z) / r);
    dout = asin(fmin(1.0, fmax(-1.0, C))) / r;
}

__global__ void __launch_bounds__(256) k_astrom(ObjCols o, int n, WcsDev w, double *xw,
                                                double *yw, double *a2000, double *d2000)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double r = M_PI / 180.0;
    /* X_IMAGE, Y_IMAGE: 1-based (cudaanalyse.cu:2715-2716) */
    const double dx = o.mx[i] + 1.0 - w.crpix[0], dy = o.my[i] + 1.0 - w.crpix[1];
    const double img0 = w.cd[0] * dx + w.cd[1] * dy, img1 = w.cd[2] * dx + w.cd[3] * dy;
    const double X = w.lng == 0 ? img0 : img1, Y = w.lng == 0 ? img1 : img0;
    const double R = sqrt(X * X + Y * Y);
    const double phi = R == 0.0 ? 0.0 : atan2(X, -Y) / r;
    double theta = 0.0;
    bool ok = true;
    switch (w.proj) {
    case 0: theta = atan2(180.0 / M_PI, R) / r; break;
    case 1:
        ok = R * M_PI / 180.0 <= 1.0;
        theta = ok ? acos(R * M_PI / 180.0) / r : 0.0;
        break;
    case 2: theta = 90.0 - R; break;
    case 3: theta = 90.0 - 2.0 * atan(R * M_PI / 360.0) / r; break;
    default:
        ok = R * M_PI / 360.0 <= 1.0;
        theta = ok ? 90.0 - 2.0 * asin(R * M_PI / 360.0) / r : 0.0;
        break;
    }
    if (!ok) {
        xw[i] = yw[i] = a2000[i] = d2000[i] = WCS_NOCOORD;
        return;
    }
    const double ap = w.crval[w.lng] * r, dp = w.crval[w.lat] * r;
    const double t = theta * r, dphi = (phi - w.phip) * r;
    const double st = sin(t), ct = cos(t), sdp = sin(dp), cdp = cos(dp), cph = cos(dphi);
    const double delta = asin(fmin(1.0, fmax(-1.0, st * sdp + ct * cdp * cph))) / r;
    const double alpha = wrap360((ap + atan2(-ct * sin(dphi), st * cdp - ct * sdp * cph)) / r);
    xw[i] = w.lng == 0 ? alpha : delta;
    yw[i] = w.lng == 0 ? delta : alpha;
    double a = alpha, d = delta;
    if (fabs(w.equinox - 2000.0) > 0.003) precess_dev(w.equinox, alpha, delta, 2000.0, a, d);
    a2000[i] = a;
    d2000[i] = d;
}

/* four double arrays of n entries set to v (one lane per entry) */
__global__ void __launch_bounds__(256) k_fill_f64(double *a, double *b, double *c, double *d,
                                                  long n, double v)
{
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    if (k >= 4 * n) return;
    double *p = k < n ? a : k < 2 * n ? b : k < 3 * n ? c : d;
    p[k % n] = v;
}

/* The world coordinates of the current frame's objects into "wcs.*" (zeros without a
   WCS: sxg_set_wcs, else the reference's wcsstruct wcs when it holds a celestial pair).
   Queued on the context's stream; false (error recorded) on an uncovered WCS. */
bool sxg_astrometry(const void *wcs)
{
    SxgState &S = sxg_state();
    const int n = S.nobj;
    S.wcs_ready = false;
    double *xw = sxg_take<double>("wcs.xw", (size_t)(n > 0 ? n : 1));
    double *yw = sxg_take<double>("wcs.yw", (size_t)(n > 0 ? n : 1));
    double *a = sxg_take<double>("wcs.alpha", (size_t)(n > 0 ? n : 1));
    double *d = sxg_take<double>("wcs.delta", (size_t)(n > 0 ? n : 1));
    if (!xw || !yw || !a || !d) return false;
    WcsDev w;
    bool have = S.wcs_set;
    if (have)
        memcpy(&w, S.wcs_dev, sizeof(w));
    else if (wcs && !wcs_from_reference(wcs, w, have))
        return false;
    if (n > 0) {
        if (!have) {
            hipLaunchKernelGGL(k_fill_f64, dim3(sxg_div_up(4L * n, 256)), dim3(256), 0, S.stream,
                               xw, yw, a, d, (long)n, 0.0);
        } else {
            ObjCols o;
            if (!sxg_obj_view(o)) return false;
            hipLaunchKernelGGL(k_astrom, dim3(sxg_div_up(n, 256)), dim3(256), 0, S.stream, o, n,
                               w, xw, yw, a, d);
        }
        SXG_HIP_RET(hipGetLastError(), false);
    }
    S.wcs_ready = true;
    return true;
}

extern "C" int sxg_get_world(double *xw, double *yw, double *alpha2000, double *delta2000,
                             size_t count)
{
    SxgState &S = sxg_state();
    sxg_clear_error();
    if (!S.wcs_ready) {
        sxg_set_error("sxg_get_world: call after endobject_gpu");
        return -1;
    }
    if (count < (size_t)S.nobj) {
        sxg_set_error("sxg_get_world: output buffers too small");
        return -1;
    }
    if (S.nobj == 0) return 0;
    const char *names[4] = {"wcs.xw", "wcs.yw", "wcs.alpha", "wcs.delta"};
    double *dst[4] = {xw, yw, alpha2000, delta2000};
    for (int k = 0; k < 4; k++)
        if (dst[k])
            SXG_HIP_RET(hipMemcpyAsync(dst[k], sxg_take<double>(names[k], (size_t)S.nobj),
                                       sizeof(double) * S.nobj, hipMemcpyDeviceToHost, S.stream),
                        -1);
    SXG_HIP_RET(hipStreamSynchronize(S.stream), -1);
    return 0;
}
