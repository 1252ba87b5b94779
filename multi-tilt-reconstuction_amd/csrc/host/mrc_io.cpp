// MRC and TXBR file boundary (host code).
#include "mtr_io.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static_assert(sizeof(mtr_mrc_header_t) == 1024, "MRC header must be 1024 bytes");

namespace {

struct File {
    FILE *f = nullptr;
    File(const char *p, const char *m) : f(std::fopen(p, m)) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

int64_t file_size(FILE *f)
{
    if (std::fseek(f, 0, SEEK_END) != 0) return -1;
    const long long n = ftello(f);
    std::fseek(f, 0, SEEK_SET);
    return n;
}

size_t mode_bytes(int mode)
{
    switch (mode) {
        case MTR_MRC_BYTE: return 1;
        case MTR_MRC_SHORT: return 2;
        case MTR_MRC_FLOAT: return 4;
        case MTR_MRC_USHORT: return 2;
        default: return 0;
    }
}

void fill_header(mtr_mrc_header_t &h, int nx, int ny, int nz, int xs, int ys, int zs, const float *data)
{
    std::memset(&h, 0, sizeof h);
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    h.mode = MTR_MRC_FLOAT;
    h.nxstart = xs;
    h.nystart = ys;
    h.nzstart = zs;
    h.mx = nx;
    h.my = ny;
    h.mz = nz;
    h.xlen = (float)nx;
    h.ylen = (float)ny;
    h.zlen = (float)nz;
    h.alpha = h.beta = h.gamma = 90.f;
    h.mapc = 1;
    h.mapr = 2;
    h.maps = 3;
    h.creatid = 1000;
    std::memcpy(h.cmap, "MAP ", 4);
    h.stamp[0] = 0x44;  // little-endian machine stamp
    h.stamp[1] = 0x41;
    const int64_t n = (int64_t)nx * ny * nz;
    if (n > 0) {
        long double mn = data[0], mx = data[0], sum = 0, sq = 0;
        for (int64_t i = 0; i < n; i++) {
            const long double v = data[i];
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
            sum += v;
            sq += v * v;
        }
        const long double mean = sum / n;
        h.amin = (float)mn;
        h.amax = (float)mx;
        h.amean = (float)mean;
        const long double var = sq / n - mean * mean;
        h.rms = (float)std::sqrt((double)(var > 0 ? var : 0));
    }
}

int write_file(const char *path, const mtr_mrc_header_t &h, const float *data, int64_t n)
{
    File f(path, "wb");
    if (!f.f) return MTR_ERR_ARG;
    if (std::fwrite(&h, sizeof h, 1, f.f) != 1) return MTR_ERR_ARG;
    if (n > 0 && std::fwrite(data, sizeof(float), (size_t)n, f.f) != (size_t)n) return MTR_ERR_ARG;
    return MTR_OK;
}

std::vector<std::string> split_ws(const std::string &s)
{
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
        size_t j = i;
        while (j < s.size() && !std::isspace((unsigned char)s[j])) j++;
        if (j > i) out.push_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

}  // namespace

extern "C" int mtr_mrc_read_header(const char *path, mtr_mrc_header_t *hdr)
{
    if (!path || !hdr) return MTR_ERR_ARG;
    File f(path, "rb");
    if (!f.f) return MTR_ERR_ARG;
    if (std::fread(hdr, sizeof *hdr, 1, f.f) != 1) return MTR_ERR_SHAPE;
    if (hdr->nx <= 0 || hdr->ny <= 0 || hdr->nz <= 0 || hdr->next < 0) return MTR_ERR_SHAPE;
    if (!mode_bytes(hdr->mode)) return MTR_ERR_SHAPE;
    return MTR_OK;
}

extern "C" int mtr_mrc_read_data(const char *path, float *out, int64_t count)
{
    if (!path || !out) return MTR_ERR_ARG;
    mtr_mrc_header_t h;
    int rc = mtr_mrc_read_header(path, &h);
    if (rc != MTR_OK) return rc;
    const int64_t n = (int64_t)h.nx * h.ny * h.nz;
    if (count < n) return MTR_ERR_SHAPE;
    const size_t eb = mode_bytes(h.mode);
    File f(path, "rb");
    if (!f.f) return MTR_ERR_ARG;
    const int64_t off = 1024 + (int64_t)h.next;
    if (file_size(f.f) < off + n * (int64_t)eb) return MTR_ERR_SHAPE;
    if (std::fseek(f.f, (long)off, SEEK_SET) != 0) return MTR_ERR_ARG;
    std::vector<unsigned char> buf((size_t)(n * eb));
    if (n > 0 && std::fread(buf.data(), eb, (size_t)n, f.f) != (size_t)n) return MTR_ERR_SHAPE;
    for (int64_t i = 0; i < n; i++) {
        const unsigned char *b = buf.data() + i * eb;
        switch (h.mode) {
            case MTR_MRC_BYTE: out[i] = (float)b[0]; break;
            case MTR_MRC_SHORT: {
                int16_t v;
                std::memcpy(&v, b, 2);
                out[i] = (float)v;
                break;
            }
            case MTR_MRC_USHORT: {
                uint16_t v;
                std::memcpy(&v, b, 2);
                out[i] = (float)v;
                break;
            }
            default: std::memcpy(&out[i], b, 4); break;
        }
    }
    return MTR_OK;
}

extern "C" int mtr_mrc_write_volume(const char *path, const mtr_volume_size *vol, const float *data)
{
    if (!path || !vol || !data || vol->X <= 0 || vol->Y <= 0 || vol->Z <= 0) return MTR_ERR_ARG;
    mtr_mrc_header_t h;
    fill_header(h, vol->X, vol->Y, vol->Z, vol->Xstart, vol->Ystart, vol->Zstart, data);
    return write_file(path, h, data, (int64_t)vol->X * vol->Y * vol->Z);
}

extern "C" int mtr_mrc_write_stack(const char *path, int nx, int ny, int nz, const float *data)
{
    if (!path || !data || nx <= 0 || ny <= 0 || nz <= 0) return MTR_ERR_ARG;
    mtr_mrc_header_t h;
    fill_header(h, nx, ny, nz, 0, 0, 0, data);
    return write_file(path, h, data, (int64_t)nx * ny * nz);
}

extern "C" int mtr_txbr_read(const char *path, mtr_volume_size *vol, double *xc, double *yc, int max_views,
                             int *views)
{
    if (!path || !vol || !xc || !yc || !views || max_views <= 0) return MTR_ERR_ARG;
    File f(path, "r");
    if (!f.f) return MTR_ERR_ARG;
    char line[4096];
    int nx = 0, ny = 0;
    const int cap = 4 * max_views;
    while (std::fgets(line, sizeof line, f.f)) {
        const std::string s(line);
        if (s.size() >= 3 && (s[0] == 'x' || s[0] == 'y' || s[0] == 'z') && s[1] == '-' && s[2] == '>') {
            // range line  a->START:STEP:END  (END exclusive)
            int a = 0, st = 0, en = 0;
            if (std::sscanf(s.c_str() + 3, "%d:%d:%d", &a, &st, &en) != 3) return MTR_ERR_SHAPE;
            int *lo = s[0] == 'x' ? &vol->Xstart : (s[0] == 'y' ? &vol->Ystart : &vol->Zstart);
            int *hi = s[0] == 'x' ? &vol->Xend : (s[0] == 'y' ? &vol->Yend : &vol->Zend);
            *lo = a < *lo ? a : *lo;
            *hi = en > *hi ? en : *hi;
        } else if (s.size() >= 2 && (s[0] == 'x' || s[0] == 'y') && s[1] == '-') {
            // coefficient line: first token is the tag, the rest are numbers
            auto tok = split_ws(s);
            for (size_t t = 1; t < tok.size(); t++) {
                char *end = nullptr;
                const double v = std::strtod(tok[t].c_str(), &end);
                if (end == tok[t].c_str()) return MTR_ERR_SHAPE;
                if (s[0] == 'x') {
                    if (nx >= cap) return MTR_ERR_RANGE;
                    xc[nx++] = v;
                } else {
                    if (ny >= cap) return MTR_ERR_RANGE;
                    yc[ny++] = v;
                }
            }
        }
    }
    if (nx != ny || nx % 4 != 0) return MTR_ERR_SHAPE;
    vol->X = vol->Xend - vol->Xstart;
    vol->Y = vol->Yend - vol->Ystart;
    vol->Z = vol->Zend - vol->Zstart;
    *views = nx / 4;
    return MTR_OK;
}

extern "C" int mtr_txbr_write(const char *path, const mtr_volume_size *vol, const double *xc, const double *yc,
                              int views)
{
    if (!path || !vol || !xc || !yc || views <= 0) return MTR_ERR_ARG;
    File f(path, "w");
    if (!f.f) return MTR_ERR_ARG;
    std::fprintf(f.f, "x->%d:1:%d\ny->%d:1:%d\nz->%d:1:%d\n", vol->Xstart, vol->Xend, vol->Ystart, vol->Yend,
                 vol->Zstart, vol->Zend);
    for (int a = 0; a < views; a++) {
        std::fprintf(f.f, "l-%d\n", a + 1);
        std::fprintf(f.f, "x-%d %.17g %.17g %.17g %.17g\n", a + 1, xc[4 * a], xc[4 * a + 1], xc[4 * a + 2],
                     xc[4 * a + 3]);
        std::fprintf(f.f, "y-%d %.17g %.17g %.17g %.17g\n", a + 1, yc[4 * a], yc[4 * a + 1], yc[4 * a + 2],
                     yc[4 * a + 3]);
    }
    return MTR_OK;
}
