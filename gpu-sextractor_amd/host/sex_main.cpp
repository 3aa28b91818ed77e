/*
 * sex_main.cpp -- the `sex` front-end of the MI355X detection path.
 *
 *   sex image.fits [-c config.sex] [-KEY value ...]      run
 *   sex -d                                               print the default configuration
 *
 * It reads the keys of the reference's configuration (src/preflist.h) that the detection
 * path uses, the .param list of catalogue columns and the .conv filter, reads the FITS
 * frame, calls the C-ABI in the order of the reference's src/scan.c:78-146, and writes
 * the catalogue (ASCII_HEAD with the header lines of catout.c:583-594, ASCII, or
 * FITS_LDAC as catout.c:720-748 lays it out: an empty primary HDU, LDAC_IMHEAD, and
 * LDAC_OBJECTS with big-endian columns; column names, formats and units as in the
 * reference's key table src/globals.c) and the check-images SEGMENTATION, -BACKGROUND,
 * BACKGROUND_RMS and FILTERED.  Catalogue NUMBER and the segmentation values are
 * 1-based.
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "fitsio.h"
#include "sexgpu.h"

namespace {

struct Config {
    std::map<std::string, std::string> kv;
    Config()
    {
        /* defaults of the reference's default.sex / preflist.h for the keys used here */
        kv = {{"CATALOG_NAME", "test.cat"}, {"CATALOG_TYPE", "ASCII_HEAD"},
              {"PARAMETERS_NAME", "default.param"}, {"DETECT_TYPE", "CCD"},
              {"DETECT_MINAREA", "5"}, {"THRESH_TYPE", "RELATIVE"}, {"DETECT_THRESH", "1.5"},
              {"ANALYSIS_THRESH", "1.5"}, {"FILTER", "Y"}, {"FILTER_NAME", "default.conv"},
              {"DEBLEND_NTHRESH", "32"}, {"DEBLEND_MINCONT", "0.005"}, {"CLEAN", "Y"},
              {"CLEAN_PARAM", "1.0"}, {"MEMORY_OBJSTACK", "3000"}, {"SATUR_LEVEL", "50000.0"},
              {"MAG_ZEROPOINT", "0.0"}, {"GAIN", "0.0"}, {"BACK_TYPE", "AUTO"},
              {"BACK_VALUE", "0.0"}, {"BACK_SIZE", "64"}, {"BACK_FILTERSIZE", "3"},
              {"BACK_FILTTHRESH", "0.0"}, {"WEIGHT_TYPE", "NONE"}, {"WEIGHT_IMAGE", "weight.fits"},
              {"WEIGHT_GAIN", "Y"},
              {"CHECKIMAGE_TYPE", "NONE"}, {"CHECKIMAGE_NAME", "check.fits"},
              {"VERBOSE_TYPE", "NORMAL"}, {"PHOT_APERTURES", "5"},
              {"PHOT_AUTOPARAMS", "2.5, 3.5"}, {"PHOT_AUTOAPERS", "0.0,0.0"},
              {"MASK_TYPE", "CORRECT"}, {"SEEING_FWHM", "1.2"}, {"PIXEL_SCALE", "1.0"},
              {"STARNNW_NAME", "default.nnw"}, {"CATALOG_ROWS", "PARAM"}};
    }
    std::string get(const std::string &k) const
    {
        auto it = kv.find(k);
        return it == kv.end() ? std::string() : it->second;
    }
    double num(const std::string &k) const { return atof(get(k).c_str()); }
    bool flag(const std::string &k) const
    {
        std::string v = get(k);
        return !v.empty() && (v[0] == 'Y' || v[0] == 'y' || v[0] == 'T' || v[0] == 't');
    }
    /* first element of a comma-separated list */
    std::string first(const std::string &k) const { return nth(k, 0); }
    /* element i of a comma-separated list (the last one when the list is shorter) */
    std::string nth(const std::string &k, int i) const
    {
        std::string v = get(k);
        for (; i > 0; i--) {
            size_t c = v.find(',');
            if (c == std::string::npos) break;
            v = v.substr(c + 1);
        }
        size_t c = v.find(',');
        return c == std::string::npos ? v : v.substr(0, c);
    }
};

std::string trim(const std::string &s)
{
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

bool read_config(const char *path, Config &cfg)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
        size_t h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        line = trim(line);
        if (line.empty()) continue;
        std::istringstream ss(line);
        std::string key, rest;
        ss >> key;
        std::getline(ss, rest);
        cfg.kv[key] = trim(rest);
    }
    return true;
}

struct Column {
    const char *name, *desc, *fmt, *unit;
};

/* the FITS binary-table type of a column (J int32, I int16, E float, D double) */
char column_tform(const std::string &n)
{
    static const char *dbl[] = {"X2_IMAGE", "Y2_IMAGE", "XY_IMAGE", "ERRX2_IMAGE",
                                "ERRY2_IMAGE", "ERRXY_IMAGE", "ALPHA_J2000", "DELTA_J2000",
                                "X_WORLD", "Y_WORLD"};
    static const char *ints[] = {"NUMBER", "ISOAREA_IMAGE", "XMIN_IMAGE", "YMIN_IMAGE",
                                 "XMAX_IMAGE", "YMAX_IMAGE", "XPEAK_IMAGE", "YPEAK_IMAGE"};
    for (const char *d : dbl)
        if (n == d) return 'D';
    for (const char *d : ints)
        if (n == d) return 'J';
    if (n == "FLAGS") return 'I';
    if (n.compare(0, 3, "ISO") == 0 && n.size() == 4) return 'J';
    return 'E';
}

/* TDISP of a printf format: %10.3f -> F10.3, %12.7g -> G12.7, %15.10e -> E15.10, %10d -> I10 */
std::string tdisp(const char *fmt)
{
    std::string f(fmt + 1);
    const char c = f.back();
    f.pop_back();
    const char t = c == 'f' ? 'F' : c == 'g' ? 'G' : c == 'e' ? 'E' : 'I';
    return std::string(1, t) + f;
}

/* one catalogue value: integer (J, I) or floating (E, D) */
struct Cell {
    char t;
    long long i;
    double d;
};

/* an 80-character FITS header card */
std::string card(const std::string &key, const std::string &value, const std::string &comment)
{
    char buf[81];
    if (value.empty())
        snprintf(buf, sizeof buf, "%-80s", key.c_str());
    else if (comment.empty())
        snprintf(buf, sizeof buf, "%-8.8s= %-70.70s", key.c_str(), value.c_str());
    else
        snprintf(buf, sizeof buf, "%-8.8s= %s / %s", key.c_str(), value.c_str(), comment.c_str());
    std::string c(buf);
    c.resize(80, ' ');
    return c;
}
std::string fits_num(long long v)
{
    char b[32];
    snprintf(b, sizeof b, "%20lld", v);
    return b;
}
std::string fits_real(double v)
{
    char b[32];
    snprintf(b, sizeof b, "%20.10G", v);
    return b;
}
std::string fits_str(const std::string &v)
{
    std::string q = "'" + v;
    while (q.size() < 9) q += ' ';
    return q + "'";
}
std::string fits_bool(bool v) { return std::string(19, ' ') + (v ? "T" : "F"); }

/* header and data of an HDU, each padded to whole 2880-byte records */
void write_hdu(FILE *f, const std::vector<std::string> &cards, const std::vector<char> &data)
{
    std::string h;
    for (const std::string &c : cards) h += c;
    h += card("END", "", "");
    h.resize((h.size() + 2879) / 2880 * 2880, ' ');
    fwrite(h.data(), 1, h.size(), f);
    if (!data.empty()) {
        std::vector<char> d(data);
        d.resize((d.size() + 2879) / 2880 * 2880, 0);
        fwrite(d.data(), 1, d.size(), f);
    }
}

void put_be(std::vector<char> &out, const void *p, int n)
{
    const char *c = static_cast<const char *>(p);
    for (int k = n - 1; k >= 0; k--) out.push_back(c[k]);
}

/* names, descriptions, formats and units as in the reference's key table (globals.c) */
const Column kColumns[] = {
    {"NUMBER", "Running object number", "%10d", ""},
    {"ALPHA_J2000", "Right ascension of barycenter (J2000)", "%11.7f", "deg"},
    {"DELTA_J2000", "Declination of barycenter (J2000)", "%+11.7f", "deg"},
    {"X_WORLD", "Barycenter position along world x axis", "%15e", "deg"},
    {"Y_WORLD", "Barycenter position along world y axis", "%15e", "deg"},
    {"X_IMAGE", "Object position along x", "%10.3f", "pixel"},
    {"Y_IMAGE", "Object position along y", "%10.3f", "pixel"},
    {"FLUX_ISO", "Isophotal flux", "%12.7g", "count"},
    {"FLUXERR_ISO", "RMS error for isophotal flux", "%12.7g", "count"},
    {"MAG_ISO", "Isophotal magnitude", "%8.4f", "mag"},
    {"MAGERR_ISO", "RMS error for isophotal magnitude", "%8.4f", "mag"},
    {"ISOAREA_IMAGE", "Isophotal area above Analysis threshold", "%9d", "pixel**2"},
    {"FLUX_MAX", "Peak flux above background", "%12.7g", "count"},
    {"BACKGROUND", "Background at centroid position", "%12.7g", "count"},
    {"THRESHOLD", "Detection threshold above background", "%12.7g", "count"},
    {"X2_IMAGE", "Variance along x", "%15.10e", "pixel**2"},
    {"Y2_IMAGE", "Variance along y", "%15.10e", "pixel**2"},
    {"XY_IMAGE", "Covariance between x and y", "%15.10e", "pixel**2"},
    {"A_IMAGE", "Profile RMS along major axis", "%9.3f", "pixel"},
    {"B_IMAGE", "Profile RMS along minor axis", "%9.3f", "pixel"},
    {"THETA_IMAGE", "Position angle (CCW/x)", "%5.1f", "deg"},
    {"CXX_IMAGE", "Cxx object ellipse parameter", "%12.7e", "pixel**(-2)"},
    {"CYY_IMAGE", "Cyy object ellipse parameter", "%12.7e", "pixel**(-2)"},
    {"CXY_IMAGE", "Cxy object ellipse parameter", "%12.7e", "pixel**(-2)"},
    {"ERRX2_IMAGE", "Variance of position along x", "%15.10e", "pixel**2"},
    {"ERRY2_IMAGE", "Variance of position along y", "%15.10e", "pixel**2"},
    {"ERRXY_IMAGE", "Covariance of position between x and y", "%15.10e", "pixel**2"},
    {"FLAGS", "Extraction flags", "%3d", ""},
    {"FWHM_IMAGE", "FWHM assuming a gaussian core", "%8.2f", "pixel"},
    {"ELONGATION", "A_IMAGE/B_IMAGE", "%8.3f", ""},
    {"ELLIPTICITY", "1 - B_IMAGE/A_IMAGE", "%8.3f", ""},
    {"XMIN_IMAGE", "Minimum x-coordinate among detected pixels", "%10d", "pixel"},
    {"YMIN_IMAGE", "Minimum y-coordinate among detected pixels", "%10d", "pixel"},
    {"XMAX_IMAGE", "Maximum x-coordinate among detected pixels", "%10d", "pixel"},
    {"YMAX_IMAGE", "Maximum y-coordinate among detected pixels", "%10d", "pixel"},
    {"XPEAK_IMAGE", "x-coordinate of the brightest pixel", "%10d", "pixel"},
    {"YPEAK_IMAGE", "y-coordinate of the brightest pixel", "%10d", "pixel"},
    {"ISO0", "Isophotal area at level 0", "%8d", "pixel**2"},
    {"ISO1", "Isophotal area at level 1", "%8d", "pixel**2"},
    {"ISO2", "Isophotal area at level 2", "%8d", "pixel**2"},
    {"ISO3", "Isophotal area at level 3", "%8d", "pixel**2"},
    {"ISO4", "Isophotal area at level 4", "%8d", "pixel**2"},
    {"ISO5", "Isophotal area at level 5", "%8d", "pixel**2"},
    {"ISO6", "Isophotal area at level 6", "%8d", "pixel**2"},
    {"ISO7", "Isophotal area at level 7", "%8d", "pixel**2"},
    {"FLUX_ISOCOR", "Corrected isophotal flux", "%12.7g", "count"},
    {"FLUXERR_ISOCOR", "RMS error for corrected isophotal flux", "%12.7g", "count"},
    {"FLUX_APER", "Flux vector within fixed circular aperture(s)", "%12.7g", "count"},
    {"FLUXERR_APER", "RMS error vector for aperture flux(es)", "%12.7g", "count"},
    {"MAG_APER", "Fixed aperture magnitude vector", "%8.4f", "mag"},
    {"MAGERR_APER", "RMS error vector for fixed aperture mag.", "%8.4f", "mag"},
    {"FLUX_AUTO", "Flux within a Kron-like elliptical aperture", "%12.7g", "count"},
    {"FLUXERR_AUTO", "RMS error for AUTO flux", "%12.7g", "count"},
    {"MAG_AUTO", "Kron-like elliptical aperture magnitude", "%8.4f", "mag"},
    {"MAGERR_AUTO", "RMS error for AUTO magnitude", "%8.4f", "mag"},
    {"KRON_RADIUS", "Kron apertures in units of A or B", "%5.2f", ""},
    {"CLASS_STAR", "S/G classifier output", "%5.2f", ""},
};

const Column *find_column(const std::string &name)
{
    for (const Column &c : kColumns)
        if (name == c.name) return &c;
    return nullptr;
}

bool read_conv(const std::string &path, std::vector<float> &mask, int &mw)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    if (!std::getline(f, line) || line.compare(0, 4, "CONV")) return false;
    bool norm = line.find("NORM") == std::string::npos || line.find("NONORM") == std::string::npos;
    int width = 0;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        std::string tok;
        int j = 0;
        while (ss >> tok) {
            if (j == 0 && tok[0] == '#') break;
            mask.push_back((float)atof(tok.c_str()));
            j++;
        }
        if (j > width) width = j;
    }
    if (width < 1 || mask.size() % width) return false;
    if (norm) {   /* filter.c:164-181 */
        double sum = 0.0, var = 0.0;
        for (float v : mask) {
            sum += std::fabs((double)v);
            var += (double)v * v;
        }
        if (sum == 0.0) sum = std::sqrt(var);
        for (float &v : mask) v = (float)(v / sum);
    }
    mw = width;
    return true;
}

/* a NNW weight table as the reference's getnnw reads it (src/neurro.c:96-180): after the
   "NNW" line and skipping comments and blank lines, the layer count and sizes, the input
   biases, the input scales, per layer and neuron its input weights then its bias, the
   output biases and the output scales */
bool read_nnw(const std::string &path, sxg_brain &br)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    if (!std::getline(f, line) || line.compare(0, 3, "NNW")) return false;
    std::vector<std::vector<double>> rows;
    while (std::getline(f, line)) {
        std::string t = trim(line);
        if (t.empty() || t[0] == '#') continue;
        std::istringstream ss(t);
        std::vector<double> r;
        double v;
        while (ss >> v) r.push_back(v);
        rows.push_back(r);
    }
    memset(&br, 0, sizeof(br));
    if (rows.empty() || rows[0].empty()) return false;
    br.layersnb = (int)rows[0][0];
    if (br.layersnb < 2 || br.layersnb > 3 || (int)rows[0].size() < 1 + br.layersnb) return false;
    for (int l = 0; l < br.layersnb; l++) {
        br.nn[l] = (int)rows[0][1 + l];
        if (br.nn[l] < 1 || br.nn[l] > 10) return false;
    }
    const int nin = br.nn[0], nout = br.nn[br.layersnb - 1];
    size_t r = 1;
    auto row = [&](size_t need) -> const std::vector<double> * {
        if (r >= rows.size() || rows[r].size() < need) return nullptr;
        return &rows[r++];
    };
    const std::vector<double> *ib = row(nin), *is = row(nin);
    if (!ib || !is) return false;
    for (int k = 0; k < nin; k++) {
        br.inbias[k] = (*ib)[k];
        br.inscale[k] = (*is)[k];
    }
    for (int l = 0; l < br.layersnb - 1; l++)
        for (int j = 0; j < br.nn[l + 1]; j++) {
            const std::vector<double> *w = row(br.nn[l] + 1);
            if (!w) return false;
            for (int k = 0; k < br.nn[l]; k++) br.w[l][k][j] = (*w)[k];
            br.b[l][j] = (*w)[br.nn[l]];
        }
    const std::vector<double> *ob = row(nout), *os = row(nout);
    if (!ob || !os) return false;
    for (int k = 0; k < nout; k++) {
        br.outbias[k] = (*ob)[k];
        br.outscale[k] = (*os)[k];
    }
    return true;
}

void dump_defaults()
{
    Config c;
    for (auto &kv : c.kv) printf("%-18s %s\n", kv.first.c_str(), kv.second.c_str());
}

bool fail(const char *what)
{
    const char *e = sxg_last_error();
    if (e) {
        fprintf(stderr, "sex: %s: %s\n", what, e);
        return true;
    }
    return false;
}

}  // namespace

/* One catalogue row as the reference writes it (printobj, scan.c:331-406): the fields of
   its outobjstruct in a fixed format, whatever the .param list holds. */
static void print_reference_row(FILE *s, const sxg_outobj &o)
{
    fprintf(s,
            "%10d %11.7f %+11.7f %10.3f %10.3f %12.7g %12.7g %8.4f %8.4f %12.7g %12.7g %12.7g "
            "%12.7g %8.4f %8.4f %12.7g %12.7g %8.4f %8.4f %5.1f %15.10e %15.10e %15.10e %3d "
            "%8.3f %8.3f %5.2f %12.7g %5.2f %12.7g %8.2f\n",
            o.number, o.alpha2000, o.delta2000, o.sposx, o.sposy, o.flux_iso, o.fluxerr_iso,
            o.mag_iso, o.magerr_iso, o.flux_isocor, o.fluxerr_isocor, o.flux_aper[0],
            o.fluxerr_aper[0], o.mag_aper[0], o.magerr_aper[0], o.flux_auto, o.fluxerr_auto,
            o.mag_auto, o.magerr_auto, o.theta, o.poserr_mx2, o.poserr_my2, o.poserr_mxy,
            (int)o.flag, o.elong, o.ellip, o.sprob, o.peak, o.kronfactor, o.bkg, o.fwhm);
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: sex image.fits [-c config.sex] [-KEY value ...] | sex -d\n");
        return 1;
    }
    if (!strcmp(argv[1], "-d") || !strcmp(argv[1], "-dd")) {
        dump_defaults();
        return 0;
    }
    Config cfg;
    std::string image;
    std::vector<std::pair<std::string, std::string>> overrides;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "-c") && i + 1 < argc) {
            if (!read_config(argv[++i], cfg)) {
                fprintf(stderr, "sex: cannot read configuration %s\n", argv[i]);
                return 1;
            }
        } else if (argv[i][0] == '-' && i + 1 < argc) {
            overrides.push_back({argv[i] + 1, argv[i + 1]});
            i++;
        } else
            image = argv[i];
    }
    for (auto &o : overrides) cfg.kv[o.first] = o.second;   /* main.c:104-107 order */
    if (image.empty()) {
        fprintf(stderr, "sex: no image given\n");
        return 1;
    }
    /* WEIGHT_TYPE (weight.c:35-126): the detection image's weighting, if any */
    const std::string wtname = cfg.first("WEIGHT_TYPE");
    int wtype = -1;
    if (wtname == "MAP_WEIGHT") wtype = SXG_WEIGHT_MAP_WEIGHT;
    else if (wtname == "MAP_RMS") wtype = SXG_WEIGHT_MAP_RMS;
    else if (wtname == "MAP_VAR") wtype = SXG_WEIGHT_MAP_VAR;
    else if (wtname == "BACKGROUND") wtype = SXG_WEIGHT_BACKGROUND;
    else if (wtname != "NONE") {
        fprintf(stderr, "sex: WEIGHT_TYPE %s is not supported by this build\n", wtname.c_str());
        return 1;
    }
    if (cfg.get("DETECT_TYPE") != "CCD") {
        fprintf(stderr, "sex: only DETECT_TYPE CCD is supported\n");
        return 1;
    }
    /* catalogue columns */
    std::vector<const Column *> cols;
    {
        std::ifstream pf(cfg.get("PARAMETERS_NAME"));
        if (!pf) {
            fprintf(stderr, "sex: cannot read %s\n", cfg.get("PARAMETERS_NAME").c_str());
            return 1;
        }
        std::string line;
        while (std::getline(pf, line)) {
            line = trim(line);
            if (line.empty() || line[0] == '#') continue;
            std::string name = line.substr(0, line.find_first_of(" \t("));
            const Column *c = find_column(name);
            if (c)
                cols.push_back(c);
            else
                fprintf(stderr, "sex: warning: catalogue parameter %s is not produced by the "
                                "detection path; skipped\n", name.c_str());
        }
    }
    /* frame */
    int W = 0, H = 0;
    double gain = cfg.num("GAIN"), satur = cfg.num("SATUR_LEVEL");
    const char *err = nullptr;
    float *img = sxf_read_image(image.c_str(), &W, &H, nullptr, nullptr, &err);
    if (!img) {
        fprintf(stderr, "sex: cannot read %s: %s\n", image.c_str(), err ? err : "?");
        return 1;
    }
    const int bsz = atoi(cfg.first("BACK_SIZE").c_str());
    const int bw = bsz < W ? bsz : W, bh = bsz < H ? bsz : H;
    if (bw != bh) {
        fprintf(stderr, "sex: only square background meshes are supported\n");
        return 1;
    }
    const int nbx = (W - 1) / bw + 1, nby = (H - 1) / bh + 1;
    const int nfs = atoi(cfg.first("BACK_FILTERSIZE").c_str());
    const int back_type = cfg.first("BACK_TYPE") == "MANUAL" ? 1 : 0;

    /* the weight map, as read by newweight (weight.c:35-83) */
    if (wtype >= 0) {
        const bool abs_t = cfg.first("THRESH_TYPE") == "ABSOLUTE";
        float *wmap = nullptr;
        int ww = 0, wh = 0;
        if (wtype != SXG_WEIGHT_BACKGROUND) {
            const char *werr = nullptr;
            wmap = sxf_read_image(cfg.first("WEIGHT_IMAGE").c_str(), &ww, &wh, nullptr, nullptr,
                                  &werr);
            if (!wmap) {
                fprintf(stderr, "sex: cannot read %s: %s\n", cfg.first("WEIGHT_IMAGE").c_str(),
                        werr ? werr : "?");
                return 1;
            }
            if (ww != W || wh != H) {
                fprintf(stderr, "sex: measured frame and weight map have different sizes\n");
                return 1;
            }
        }
        const int rc = sxg_set_weight_map(wmap, ww, wh, wtype, abs_t ? 1 : 0,
                                          atof(cfg.first("DETECT_THRESH").c_str()),
                                          atof(cfg.first("ANALYSIS_THRESH").c_str()),
                                          cfg.flag("WEIGHT_GAIN") ? 1 : 0);
        free(wmap);
        if (rc != 0) {
            fail("sxg_set_weight_map");
            return 1;
        }
    }
    /* makeit.c:189, back.c:104 */
    init_device(W, H, img);
    if (fail("init_device")) return 1;
    sxg_set_back_value((float)atof(cfg.first("BACK_VALUE").c_str()));
    sxg_set_mag_zeropoint(cfg.num("MAG_ZEROPOINT"));
    float backmean = 0, backsig = 0;
    back_estimate(bw, (float)cfg.num("BACK_FILTTHRESH"), &backmean, &backsig,
                  nbx > 1 ? nfs : 1, nby > 1 ? nfs : 1, back_type);
    if (fail("back_estimate")) return 1;
    /* back.c:155-192 */
    const bool absolute = cfg.first("THRESH_TYPE") == "ABSOLUTE";
    const double dt0 = atof(cfg.first("DETECT_THRESH").c_str());
    const double at0 = atof(cfg.first("ANALYSIS_THRESH").c_str());
    const float dthresh = absolute ? (float)dt0 : (float)(dt0 * backsig);
    const float thresh = absolute ? (float)at0 : (float)(at0 * backsig);
    if (!(dthresh > 0.0f) || !(thresh > 0.0f)) {
        fprintf(stderr, "sex: cannot deal with zero or negative thresholds\n");
        return 1;
    }
    /* scan.c:78-146 */
    init_detection();
    if (cfg.flag("FILTER")) {
        std::vector<float> mask;
        int mw = 0;
        if (!read_conv(cfg.get("FILTER_NAME"), mask, mw)) {
            fprintf(stderr, "sex: cannot read filter %s\n", cfg.get("FILTER_NAME").c_str());
            return 1;
        }
        if ((int)mask.size() != mw * mw) {
            fprintf(stderr, "sex: the filter mask must be square\n");
            return 1;
        }
        cudaFilter(mask.data(), (unsigned)mw);
        if (fail("cudaFilter")) return 1;
    }
    const int minarea = atoi(cfg.get("DETECT_MINAREA").c_str());
    const int nthresh = atoi(cfg.get("DEBLEND_NTHRESH").c_str());
    size_t nraw = 0, npix = 0;
    run_detection(dthresh, thresh, minarea, &nraw, &npix);
    if (fail("run_detection")) return 1;
    init_deblend(dthresh, nthresh, nraw, npix);
    const int nobj = parcel_out(dthresh, nthresh, minarea < 3 ? minarea : 3, nraw, npix,
                                cfg.num("DEBLEND_MINCONT"));
    clear_deblend(nthresh);
    if (fail("parcel_out") || nobj < 0) return 1;
    double ftrue;
    float ftruef;
    {
        uint64_t u = 1;
        uint32_t v = 1;
        memcpy(&ftrue, &u, 8);
        memcpy(&ftruef, &v, 4);
    }
    analyse_gpu(ftrue, 1, 1, ftruef, backsig, gain, 0.0, thresh, dthresh, satur, wtype >= 0 ? 1 : 0,
                0, (wtype >= 0 && cfg.flag("WEIGHT_GAIN")) ? 1 : 0, minarea, cfg.flag("CLEAN"), 0,
                nbx, nby, bw, bh, W, nobj);
    if (fail("analyse_gpu")) return 1;
    std::vector<uint32_t> master(nobj > 0 ? nobj : 1);
    run_clean(reinterpret_cast<int *>(master.data()), cfg.num("CLEAN_PARAM"),
              atoi(cfg.get("MEMORY_OBJSTACK").c_str()), nobj);
    if (fail("run_clean")) return 1;
    /* photometry settings (the reference's prefs) and the columns that request them */
    auto wants = [&](std::initializer_list<const char *> names) {
        for (const Column *c : cols)
            for (const char *n : names)
                if (!strcmp(c->name, n)) return 1;
        return 0;
    };
    sxg_photom_params pp;
    pp.apert = atof(cfg.first("PHOT_APERTURES").c_str());
    pp.autoparam[0] = atof(cfg.nth("PHOT_AUTOPARAMS", 0).c_str());
    pp.autoparam[1] = atof(cfg.nth("PHOT_AUTOPARAMS", 1).c_str());
    pp.autoaper[0] = atof(cfg.nth("PHOT_AUTOAPERS", 0).c_str());
    pp.autoaper[1] = atof(cfg.nth("PHOT_AUTOAPERS", 1).c_str());
    pp.seeing_fwhm = cfg.num("SEEING_FWHM");
    pp.mag_zeropoint = cfg.num("MAG_ZEROPOINT");
    pp.mask_correct = cfg.first("MASK_TYPE") == "CORRECT";
    pp.photo = cfg.first("DETECT_TYPE") == "PHOTO";
    /* CATALOG_ROWS REFERENCE: the rows of the reference's printobj, which hold every
       photometry group */
    const std::string rowfmt = cfg.first("CATALOG_ROWS");
    if (rowfmt != "PARAM" && rowfmt != "REFERENCE") {
        fprintf(stderr, "sex: CATALOG_ROWS must be PARAM or REFERENCE\n");
        return 1;
    }
    const bool refrows = rowfmt == "REFERENCE";
    pp.do_aper = refrows || wants({"FLUX_APER", "FLUXERR_APER", "MAG_APER", "MAGERR_APER"});
    pp.do_auto = refrows ||
                 wants({"FLUX_AUTO", "FLUXERR_AUTO", "MAG_AUTO", "MAGERR_AUTO", "KRON_RADIUS"});
    pp.do_class = refrows || wants({"CLASS_STAR"});
    if (sxg_set_photometry(&pp) != 0) {
        fail("sxg_set_photometry");
        return 1;
    }
    sxg_brain brain;
    if (pp.do_class && !read_nnw(cfg.get("STARNNW_NAME"), brain)) {
        fprintf(stderr, "sex: cannot read the NNW table %s\n", cfg.get("STARNNW_NAME").c_str());
        return 1;
    }
    const double pixscale = cfg.num("PIXEL_SCALE") > 0.0 ? cfg.num("PIXEL_SCALE") : 1.0;
    /* the frame's celestial WCS, for ALPHA_J2000/DELTA_J2000 and X_WORLD/Y_WORLD */
    {
        sxg_wcs wcs;
        memset(&wcs, 0, sizeof(wcs));
        const int hw = sxf_read_wcs(image.c_str(), wcs.ctype, wcs.crval, wcs.crpix, wcs.cd,
                                    &wcs.longpole, &wcs.latpole, &wcs.equinox);
        if (hw == 1 && sxg_set_wcs(&wcs) != 0) {
            fprintf(stderr, "sex: warning: %s; world coordinates are 0\n", sxg_last_error());
            sxg_set_wcs(nullptr);
        } else if (hw != 1) {
            sxg_set_wcs(nullptr);
        }
    }
    std::vector<sxg_outobj> rows(nobj > 0 ? nobj : 1);
    endobject_gpu(nullptr, pp.do_class ? &brain : nullptr, nullptr, rows.data(), nullptr, 0, H,
                  W, H, backsig, thresh, satur, pixscale, 0.0, gain, nobj);
    if (fail("endobject_gpu")) return 1;
    std::vector<double> xw(nobj), yw(nobj), ra(nobj), dec(nobj);
    if (nobj > 0 && sxg_get_world(xw.data(), yw.data(), ra.data(), dec.data(), nobj) != 0) {
        fail("sxg_get_world");
        return 1;
    }
    /* extra columns not in the reference's fixed row */
    std::vector<int> npixc(nobj), xmin(nobj), xmax(nobj), ymin(nobj), ymax(nobj), px(nobj),
        py(nobj), iso(8 * (size_t)nobj);
    std::vector<double> mx2(nobj), my2(nobj), mxy(nobj);
    std::vector<float> a(nobj), b(nobj), cxx(nobj), cyy(nobj), cxy(nobj), thr(nobj);
    if (nobj > 0) {
        sxg_get_column(SXG_COL_NPIX, npixc.data(), 4 * nobj);
        sxg_get_column(SXG_COL_XMIN, xmin.data(), 4 * nobj);
        sxg_get_column(SXG_COL_XMAX, xmax.data(), 4 * nobj);
        sxg_get_column(SXG_COL_YMIN, ymin.data(), 4 * nobj);
        sxg_get_column(SXG_COL_YMAX, ymax.data(), 4 * nobj);
        sxg_get_column(SXG_COL_PEAKX, px.data(), 4 * nobj);
        sxg_get_column(SXG_COL_PEAKY, py.data(), 4 * nobj);
        sxg_get_column(SXG_COL_ISO, iso.data(), 32 * (size_t)nobj);
        sxg_get_column(SXG_COL_MX2, mx2.data(), 8 * nobj);
        sxg_get_column(SXG_COL_MY2, my2.data(), 8 * nobj);
        sxg_get_column(SXG_COL_MXY, mxy.data(), 8 * nobj);
        sxg_get_column(SXG_COL_A, a.data(), 4 * nobj);
        sxg_get_column(SXG_COL_B, b.data(), 4 * nobj);
        sxg_get_column(SXG_COL_CXX, cxx.data(), 4 * nobj);
        sxg_get_column(SXG_COL_CYY, cyy.data(), 4 * nobj);
        sxg_get_column(SXG_COL_CXY, cxy.data(), 4 * nobj);
        sxg_get_column(SXG_COL_THRESH, thr.data(), 4 * nobj);
        if (fail("sxg_get_column")) return 1;
    }
    /* check-images */
    const std::string ctype = cfg.first("CHECKIMAGE_TYPE"), cname = cfg.first("CHECKIMAGE_NAME");
    std::vector<int32_t> seg((size_t)W * H);
    int nsurv = sxg_get_segmap(seg.data());
    if (fail("sxg_get_segmap")) return 1;
    if (ctype == "SEGMENTATION" && sxf_write_image(cname.c_str(), seg.data(), W, H, 32)) {
        fprintf(stderr, "sex: cannot write %s\n", cname.c_str());
        return 1;
    }
    if (ctype == "BACKGROUND_RMS") {
        std::vector<float> rms((size_t)W * H);
        if (sxg_get_back_rms(rms.data()) != 0) {
            fail("sxg_get_back_rms");
            return 1;
        }
        sxf_write_image(cname.c_str(), rms.data(), W, H, -32);
    }
    if (ctype == "FILTERED") {
        std::vector<float> cd((size_t)W * H);
        sxg_get_filtered(cd.data());
        sxf_write_image(cname.c_str(), cd.data(), W, H, -32);
    }
    clear_detection();
    clear_device(img);   /* the background-subtracted frame */
    if (ctype == "-BACKGROUND") sxf_write_image(cname.c_str(), img, W, H, -32);
    free(img);
    /* catalogue */
    const std::string cattype = cfg.first("CATALOG_TYPE");
    if (cattype != "ASCII_HEAD" && cattype != "ASCII" && cattype != "FITS_LDAC") {
        fprintf(stderr, "sex: CATALOG_TYPE %s is not supported by this build\n", cattype.c_str());
        return 1;
    }
    FILE *cat = cfg.get("CATALOG_NAME") == "STDOUT" ? stdout : fopen(cfg.get("CATALOG_NAME").c_str(), "wb");
    if (!cat) {
        fprintf(stderr, "sex: cannot write %s\n", cfg.get("CATALOG_NAME").c_str());
        return 1;
    }
    const std::string ctype_cat = cfg.first("CATALOG_TYPE");
    /* the value of column n for object i (catalogue NUMBER `number`) */
    auto cell = [&](const std::string &n, int i, int number) -> Cell {
        const sxg_outobj &r = rows[i];
        auto val = [&](auto v) -> Cell {
            Cell c;
            c.t = column_tform(n);
            c.i = (long long)v;
            c.d = (double)v;
            return c;
        };
        if (n == "NUMBER") return val(number);
        else if (n == "ALPHA_J2000") return val(ra[i]);
        else if (n == "DELTA_J2000") return val(dec[i]);
        else if (n == "X_WORLD") return val(xw[i]);
        else if (n == "Y_WORLD") return val(yw[i]);
        else if (n == "X_IMAGE") return val(r.sposx);
        else if (n == "Y_IMAGE") return val(r.sposy);
        else if (n == "FLUX_ISO") return val(r.flux_iso);
        else if (n == "FLUXERR_ISO") return val(r.fluxerr_iso);
        else if (n == "MAG_ISO") return val(r.mag_iso);
        else if (n == "MAGERR_ISO") return val(r.magerr_iso);
        else if (n == "ISOAREA_IMAGE") return val(npixc[i]);
        else if (n == "FLUX_MAX") return val(r.peak);
        else if (n == "BACKGROUND") return val(r.bkg);
        else if (n == "THRESHOLD") return val(thr[i]);
        else if (n == "X2_IMAGE") return val(mx2[i]);
        else if (n == "Y2_IMAGE") return val(my2[i]);
        else if (n == "XY_IMAGE") return val(mxy[i]);
        else if (n == "A_IMAGE") return val(a[i]);
        else if (n == "B_IMAGE") return val(b[i]);
        else if (n == "THETA_IMAGE") return val(r.theta);
        else if (n == "CXX_IMAGE") return val(cxx[i]);
        else if (n == "CYY_IMAGE") return val(cyy[i]);
        else if (n == "CXY_IMAGE") return val(cxy[i]);
        else if (n == "ERRX2_IMAGE") return val(r.poserr_mx2);
        else if (n == "ERRY2_IMAGE") return val(r.poserr_my2);
        else if (n == "ERRXY_IMAGE") return val(r.poserr_mxy);
        else if (n == "FLAGS") return val((int)r.flag);
        else if (n == "FWHM_IMAGE") return val(r.fwhm);
        else if (n == "ELONGATION") return val(r.elong);
        else if (n == "ELLIPTICITY") return val(r.ellip);
        else if (n == "FLUX_ISOCOR") return val(r.flux_isocor);
        else if (n == "FLUXERR_ISOCOR") return val(r.fluxerr_isocor);
        else if (n == "FLUX_APER") return val(r.flux_aper[0]);
        else if (n == "FLUXERR_APER") return val(r.fluxerr_aper[0]);
        else if (n == "MAG_APER") return val(r.mag_aper[0]);
        else if (n == "MAGERR_APER") return val(r.magerr_aper[0]);
        else if (n == "FLUX_AUTO") return val(r.flux_auto);
        else if (n == "FLUXERR_AUTO") return val(r.fluxerr_auto);
        else if (n == "MAG_AUTO") return val(r.mag_auto);
        else if (n == "MAGERR_AUTO") return val(r.magerr_auto);
        else if (n == "KRON_RADIUS") return val(r.kronfactor);
        else if (n == "CLASS_STAR") return val(r.sprob);
        else if (n == "XMIN_IMAGE") return val(xmin[i] + 1);
        else if (n == "YMIN_IMAGE") return val(ymin[i] + 1);
        else if (n == "XMAX_IMAGE") return val(xmax[i] + 1);
        else if (n == "YMAX_IMAGE") return val(ymax[i] + 1);
        else if (n == "XPEAK_IMAGE") return val(px[i]);
        else if (n == "YPEAK_IMAGE") return val(py[i]);
        else if (n.compare(0, 3, "ISO") == 0 && n.size() == 4)
            return val(iso[8 * (size_t)i + (n[3] - '0')]);
        return val(0);
    };
    if (ctype_cat == "ASCII_HEAD")   /* catout.c:583-594 */
        for (size_t k = 0; k < cols.size(); k++) {
            if (*cols[k]->unit)
                fprintf(cat, "# %3d %-22.22s %-58.58s [%s]\n", (int)k + 1, cols[k]->name,
                        cols[k]->desc, cols[k]->unit);
            else
                fprintf(cat, "# %3d %-22.22s %-58.58s\n", (int)k + 1, cols[k]->name,
                        cols[k]->desc);
        }
    int number = 0;
    std::vector<char> table;
    int rowbytes = 0;
    for (const Column *c : cols) {
        const char t = column_tform(c->name);
        rowbytes += t == 'D' ? 8 : t == 'I' ? 2 : 4;
    }
    const bool ldac = ctype_cat == "FITS_LDAC";
    if (refrows && ldac) {
        fprintf(stderr, "sex: CATALOG_ROWS REFERENCE writes ASCII rows only\n");
        return 1;
    }
    for (int i = 0; i < nobj; i++) {
        if (master[i] != 0xffffffffu) continue;   /* scan.c:148-157 */
        if (refrows) {
            /* scan.c:150-156: the survivors numbered from 0, one printobj row each */
            rows[i].number = number++;
            print_reference_row(cat, rows[i]);
            continue;
        }
        number++;
        for (size_t k = 0; k < cols.size(); k++) {
            const Cell v = cell(cols[k]->name, i, number);
            if (ldac) {
                if (v.t == 'J') { const int32_t x = (int32_t)v.i; put_be(table, &x, 4); }
                else if (v.t == 'I') { const int16_t x = (int16_t)v.i; put_be(table, &x, 2); }
                else if (v.t == 'E') { const float x = (float)v.d; put_be(table, &x, 4); }
                else { const double x = v.d; put_be(table, &x, 8); }
                continue;
            }
            if (k) fputc(' ', cat);
            if (v.t == 'J' || v.t == 'I')
                fprintf(cat, cols[k]->fmt, (int)v.i);
            else
                fprintf(cat, cols[k]->fmt, v.d);
        }
        if (!ldac) fputc('\n', cat);
    }
    if (ldac) {
        /* FITS_LDAC (catout.c:720-748): an empty primary HDU, LDAC_IMHEAD holding the
           frame's header cards as one string cell, LDAC_OBJECTS the catalogue */
        std::vector<std::string> im = {
            card("SIMPLE", fits_bool(true), ""), card("BITPIX", fits_num(-32), ""),
            card("NAXIS", fits_num(2), ""), card("NAXIS1", fits_num(W), ""),
            card("NAXIS2", fits_num(H), ""),
            card("SEXBKGND", fits_real(backmean), "Median background level (ADU)"),
            card("SEXBKDEV", fits_real(backsig), "Median background RMS (ADU)"),
            card("SEXTHLD", fits_real(dthresh), "Extraction threshold (ADU)"),
            card("SEXATHLD", fits_real(thresh), "Analysis threshold (ADU)"),
            card("SEXNDET", fits_num(nobj), "Number of raw detections"),
            card("SEXGAIN", fits_real(gain), "Gain used (e-/ADU)"), card("END", "", "")};
        std::vector<char> imdata;
        for (const std::string &c : im) imdata.insert(imdata.end(), c.begin(), c.end());
        const int ncard = (int)im.size();
        write_hdu(cat, {card("SIMPLE", fits_bool(true), ""), card("BITPIX", fits_num(8), ""),
                        card("NAXIS", fits_num(0), ""), card("EXTEND", fits_bool(true), "")},
                  {});
        write_hdu(cat, {card("XTENSION", fits_str("BINTABLE"), ""), card("BITPIX", fits_num(8), ""),
                        card("NAXIS", fits_num(2), ""), card("NAXIS1", fits_num(80 * ncard), ""),
                        card("NAXIS2", fits_num(1), ""), card("PCOUNT", fits_num(0), ""),
                        card("GCOUNT", fits_num(1), ""), card("TFIELDS", fits_num(1), ""),
                        card("TTYPE1", fits_str("Field Header Card"), ""),
                        card("TFORM1", fits_str(std::to_string(80 * ncard) + "A"), ""),
                        card("TDIM1", fits_str("(80, " + std::to_string(ncard) + ")"), ""),
                        card("EXTNAME", fits_str("LDAC_IMHEAD"), "")},
                  imdata);
        std::vector<std::string> th = {
            card("XTENSION", fits_str("BINTABLE"), ""), card("BITPIX", fits_num(8), ""),
            card("NAXIS", fits_num(2), ""), card("NAXIS1", fits_num(rowbytes), ""),
            card("NAXIS2", fits_num(number), ""), card("PCOUNT", fits_num(0), ""),
            card("GCOUNT", fits_num(1), ""), card("TFIELDS", fits_num((long long)cols.size()), "")};
        for (size_t k = 0; k < cols.size(); k++) {
            const std::string j = std::to_string(k + 1);
            th.push_back(card("TTYPE" + j, fits_str(cols[k]->name), cols[k]->desc));
            th.push_back(card("TFORM" + j, fits_str(std::string("1") + column_tform(cols[k]->name)), ""));
            if (*cols[k]->unit) th.push_back(card("TUNIT" + j, fits_str(cols[k]->unit), ""));
            th.push_back(card("TDISP" + j, fits_str(tdisp(cols[k]->fmt)), ""));
        }
        th.push_back(card("EXTNAME", fits_str("LDAC_OBJECTS"), ""));
        write_hdu(cat, th, table);
    }
    if (cat != stdout) fclose(cat);
    if (cfg.get("VERBOSE_TYPE") != "QUIET")
        fprintf(stderr, "Objects: detected %d  sextracted %d\n", nobj, nsurv);
    return 0;
}
