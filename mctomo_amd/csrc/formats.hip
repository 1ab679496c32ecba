// The data file formats of the likelihood's inputs and outputs, host code:
// read_sources / write_sources and read_times_3 (src/likelihood_settings.f90
// :198-257,332-420) and write_atimes (src/read_write.f90:484-530), with the
// reference's Fortran list-directed input rules: a READ starts on a new
// record, takes values separated by blanks, commas or line ends across as many
// records as it needs, discards the rest of its last record, and stops early
// at a slash (the remaining items keep their values).  Repeat counts r*v are
// accepted; a null value (r*) leaves its item unchanged.
#include "mct_internal.h"
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct ld_reader {
    FILE *f = nullptr;
    std::vector<std::string> tok;   // values of the current record; "/" ends a READ
    size_t pos = 0;
    std::string line;
    ~ld_reader() { if (f) fclose(f); }

    bool next_record()
    {
        tok.clear();
        pos = 0;
        line.clear();
        int ch;
        bool any = false;
        while ((ch = fgetc(f)) != EOF) {
            any = true;
            if (ch == '\n') break;
            line.push_back((char)ch);
        }
        if (!any) return false;
        std::string cur;
        bool quoted = false;
        char q = 0;
        auto flush = [&]() { if (!cur.empty()) { tok.push_back(cur); cur.clear(); } };
        for (char c : line) {
            if (quoted) {
                if (c == q) quoted = false;
                cur.push_back(c);
                continue;
            }
            if (c == '\'' || c == '"') { quoted = true; q = c; cur.push_back(c); continue; }
            if (c == ' ' || c == '\t' || c == '\r' || c == ',') { flush(); continue; }
            if (c == '/') { flush(); tok.push_back("/"); break; }
            cur.push_back(c);
        }
        flush();
        return true;
    }

    // One READ of n items: returns the tokens (an empty string = null value,
    // leave the item unchanged), fewer than n after a slash; false at end of file.
    bool read(int n, std::vector<std::string> &out)
    {
        out.clear();
        if (!next_record()) return false;
        while ((int)out.size() < n) {
            if (pos == tok.size()) {
                if (!next_record()) return false;
                continue;
            }
            const std::string &t = tok[pos++];
            if (t == "/") return true;
            size_t star = t.find('*');
            if (star != std::string::npos && star > 0) {
                long r = strtol(t.substr(0, star).c_str(), nullptr, 10);
                std::string v = t.substr(star + 1);
                for (long i = 0; i < r && (int)out.size() < n; i++) out.push_back(v);
            } else {
                out.push_back(t);
            }
        }
        return true;
    }
};

bool ld_real(const std::string &t, double &v)
{
    if (t.empty()) return true;                 // null value: unchanged
    std::string s = t;
    for (char &c : s)
        if (c == 'd' || c == 'D' || c == 'q' || c == 'Q') c = 'e';
    char *end = nullptr;
    errno = 0;
    double x = strtod(s.c_str(), &end);
    if (end == s.c_str() || *end != '\0') return false;
    v = x;
    return true;
}

bool ld_int(const std::string &t, int &v)
{
    if (t.empty()) return true;
    char *end = nullptr;
    long x = strtol(t.c_str(), &end, 10);
    if (end == t.c_str() || *end != '\0') return false;
    v = (int)x;
    return true;
}

// Shortest text that reads back to the same double (list-directed output has
// no fixed form; any form the reader parses to the same value is equivalent).
void put_real(FILE *f, double v)
{
    char b[40];
    for (int p = 15; p <= 17; p++) {
        snprintf(b, sizeof b, "%.*E", p - 1, v);
        if (strtod(b, nullptr) == v) break;
    }
    fprintf(f, " %s", b);
}

} // namespace

extern "C" {

int mct_read_sources(const char *path, const double *bnd, int nbnd, long cap, int *n, int *ndim, double *src)
{
    if (!path || !n || !ndim || (bnd && nbnd < 0)) return MCT_ERR_ARG;
    ld_reader R;
    R.f = fopen(path, "r");
    if (!R.f) return MCT_ERR_IO;
    std::vector<std::string> t;
    int ns = 0, nd = 0;
    if (!R.read(2, t) || t.size() < 2 || !ld_int(t[0], ns) || !ld_int(t[1], nd) || ns < 0 || nd < 1)
        return MCT_ERR_IO;
    *n = ns;
    *ndim = nd;
    if (!src) return MCT_OK;                    // sizes only
    // src holds cap doubles; out_bnd compares each coordinate with its (min,
    // max) pair, so bnd must hold one pair per coordinate of the file
    if ((long)ns * nd > cap || (bnd && nbnd != 2 * nd)) return MCT_ERR_ARG;
    for (int i = 0; i < ns; i++) {
        if (!R.read(nd, t)) return MCT_ERR_IO;
        for (int k = 0; k < (int)t.size(); k++)
            if (!ld_real(t[k], src[(long)i * nd + k])) return MCT_ERR_IO;
        if (bnd) {   // out_bnd (src/settings.f90:143-154): bnd = (min, max) per coordinate
            for (int k = 0; k < nd; k++)
                if (src[(long)i * nd + k] < bnd[2 * k] || src[(long)i * nd + k] > bnd[2 * k + 1]) return MCT_ERR_OUTSIDE;
        }
    }
    return MCT_OK;
}

int mct_read_times(const char *path, int nsrc, int nrev, int cap_np, int *np, double *freqs, double *ttime,
                   int32_t *raystat, int *nrays)
{
    if (!path || !np || nsrc < 1 || nrev < 1) return MCT_ERR_ARG;
    ld_reader R;
    R.f = fopen(path, "r");
    if (!R.f) return MCT_ERR_IO;
    std::vector<std::string> t;
    int nf = 0;
    if (!R.read(1, t) || t.size() < 1 || !ld_int(t[0], nf) || nf < 0) return MCT_ERR_IO;
    *np = nf;
    if (!freqs || !ttime || !raystat) return MCT_OK;   // sizes only
    if (nf > cap_np) return MCT_ERR_ARG;
    if (!R.read(nf, t)) return MCT_ERR_IO;
    for (int k = 0; k < (int)t.size(); k++)
        if (!ld_real(t[k], freqs[k])) return MCT_ERR_IO;
    const long npair = (long)nsrc * nrev;
    // ttime(npair,3,np) and raystat(npair,2,np) start at zero (:370-377)
    memset(ttime, 0, sizeof(double) * 3 * npair * nf);
    memset(raystat, 0, sizeof(int32_t) * 2 * npair * nf);
    const double EPS = (double)0.000000001f;    // m_utils EPS, a default-real literal (src/utils.f90:57)
    int rays = 0;
    long npairs = 0;
    double tv = 0, nv = 0;                      // kept across READs, as the reference's locals
    int validity = 0;
    for (int i = 1; i <= nsrc; i++)
        for (int j = 1; j <= nrev; j++) {
            npairs++;
            const long q = npairs - 1;
            if (!R.read(1, t) || (t.size() == 1 && !ld_int(t[0], validity))) return MCT_ERR_IO;
            const double dssval = 1;            // no ds file (read_data passes none)
            if (validity != 0) {
                for (int k = 1; k <= nf; k++) {
                    if (!R.read(2, t)) return MCT_ERR_IO;
                    if ((t.size() > 0 && !ld_real(t[0], tv)) || (t.size() > 1 && !ld_real(t[1], nv))) return MCT_ERR_IO;
                    ttime[((long)(k - 1) * 3 + 0) * npair + q] = tv;
                    ttime[((long)(k - 1) * 3 + 1) * npair + q] = nv;
                    ttime[((long)(k - 1) * 3 + 2) * npair + q] = dssval;
                    if (fabs(tv + 1) > EPS) {
                        raystat[((long)(k - 1) * 2 + 0) * npair + q] = 1;
                        raystat[((long)(k - 1) * 2 + 1) * npair + q] = (int32_t)npairs;
                        rays++;
                    }
                }
            } else {
                for (int k = 1; k <= nf; k++) {
                    raystat[((long)(k - 1) * 2 + 0) * npair + q] = 0;
                    raystat[((long)(k - 1) * 2 + 1) * npair + q] = 0;
                    ttime[((long)(k - 1) * 3 + 0) * npair + q] = -1;
                    ttime[((long)(k - 1) * 3 + 1) * npair + q] = 1;
                }
            }
        }
    if (nrays) *nrays = rays;
    return MCT_OK;
}

int mct_write_sources(const char *path, int n, int ndim, const double *src)
{
    if (!path || n < 0 || ndim < 1 || (n > 0 && !src)) return MCT_ERR_ARG;
    FILE *f = fopen(path, "w");
    if (!f) return MCT_ERR_IO;
    fprintf(f, " %d %d\n", n, ndim);
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < ndim; k++) put_real(f, src[(long)i * ndim + k]);
        fputc('\n', f);
    }
    return fclose(f) == 0 ? MCT_OK : MCT_ERR_IO;
}

int mct_write_times(const char *path, int np, const double *freqs, int nsrc, int nrev, const int32_t *raystat,
                    const double *phase_time, const double *sigma, const double *locations)
{
    if (!path || np < 1 || nsrc < 1 || nrev < 1 || !freqs || !raystat || !phase_time || !sigma) return MCT_ERR_ARG;
    FILE *f = fopen(path, "w");
    if (!f) return MCT_ERR_IO;
    const long npair = (long)nsrc * nrev;
    fprintf(f, " %d\n", np);
    for (int k = 0; k < np; k++) put_real(f, freqs[k]);
    fputc('\n', f);
    long npairs = 0;
    for (int i = 1; i <= nsrc; i++)
        for (int j = 1; j <= nrev; j++) {
            npairs++;
            const long q = npairs - 1;
            bool any = false;
            for (int k = 0; k < np; k++) any = any || raystat[((long)k * 2) * npair + q] == 1;
            fprintf(f, " %d\n", any ? 1 : 0);
            if (!any) continue;
            for (int k = 0; k < np; k++) {
                double t = phase_time[((long)k * nsrc + (i - 1)) * nrev + (j - 1)];
                if (raystat[((long)k * 2) * npair + q] == 1 && locations) t = t + locations[(long)(i - 1) * 4 + 3];
                put_real(f, t);
                put_real(f, sigma[(long)k * npair + q]);
                fputc('\n', f);
            }
        }
    return fclose(f) == 0 ? MCT_OK : MCT_ERR_IO;
}

} // extern "C"
