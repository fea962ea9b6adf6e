// Host file formats of the BriefMatch tools: .flo / .bin flow files, the
// Middlebury colour wheel, PNG and PGM/PPM input, PNG output, EPE.
#include "../../include/bm_io.h"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>

namespace bmio {

namespace {

const float kFloTag = 202021.25f;  // "PIEH" read as a little-endian float

std::vector<uint8_t> read_file(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw IOError("cannot open '" + path + "'");
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

FILE *open_w(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw IOError("Invalid output path. Cannot write to file '" + path + "'.");
    return f;
}

bool ends_with(const std::string &s, const char *ext)
{
    const size_t n = std::strlen(ext);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)s[s.size() - n + i]) != ext[i]) return false;
    return true;
}

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

void put_be32(std::vector<uint8_t> &v, uint32_t x)
{
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

int paeth(int a, int b, int c)
{
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

void read_png(const std::string &path, const std::vector<uint8_t> &d, int &W, int &H, std::vector<uint8_t> &rgb)
{
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    if (d.size() < 8 || std::memcmp(d.data(), sig, 8) != 0) throw IOError("'" + path + "' is not a PNG file");
    size_t pos = 8;
    int depth = 0, ctype = -1;
    std::vector<uint8_t> idat, plte;
    W = H = 0;
    while (pos + 12 <= d.size()) {
        const uint32_t len = be32(&d[pos]);
        if (pos + 12 + (size_t)len > d.size()) throw IOError("truncated PNG '" + path + "'");
        const char *type = (const char *)&d[pos + 4];
        const uint8_t *data = &d[pos + 8];
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len < 13) throw IOError("bad PNG header in '" + path + "'");
            W = (int)be32(data);
            H = (int)be32(data + 4);
            depth = data[8];
            ctype = data[9];
            if (data[10] != 0 || data[11] != 0) throw IOError("unsupported PNG compression/filter in '" + path + "'");
            if (data[12] != 0) throw IOError("interlaced PNG not supported: '" + path + "'");
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(data, data + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + len;
    }
    int ch;
    switch (ctype) {
    case 0: ch = 1; break;
    case 2: ch = 3; break;
    case 3: ch = 1; break;
    case 4: ch = 2; break;
    case 6: ch = 4; break;
    default: throw IOError("unsupported PNG colour type in '" + path + "'");
    }
    if (!(depth == 8 || (depth == 16 && ctype != 3))) throw IOError("unsupported PNG bit depth in '" + path + "'");
    if (W <= 0 || H <= 0 || W > 65535 || H > 65535) throw IOError("bad PNG size in '" + path + "'");
    const size_t bpp = (size_t)ch * depth / 8, row = (size_t)W * bpp;
    // deflate expands at most 1032:1, so a header that promises more pixel data
    // than the compressed bytes can hold is corrupt; checked before allocating,
    // so a damaged header cannot request gigabytes
    if ((row + 1) * H > (idat.size() + 1) * 1032 + 64) throw IOError("corrupt PNG data in '" + path + "'");
    std::vector<uint8_t> raw((row + 1) * H);
    uLongf out = (uLongf)raw.size();
    if (uncompress(raw.data(), &out, idat.data(), (uLong)idat.size()) != Z_OK || out != raw.size())
        throw IOError("corrupt PNG data in '" + path + "'");
    std::vector<uint8_t> img(row * H);
    for (int y = 0; y < H; ++y) {
        const uint8_t f = raw[(row + 1) * y];
        const uint8_t *src = &raw[(row + 1) * y + 1];
        uint8_t *cur = &img[row * y];
        const uint8_t *up = y ? &img[row * (y - 1)] : nullptr;
        for (size_t i = 0; i < row; ++i) {
            const int a = i >= bpp ? cur[i - bpp] : 0, b = up ? up[i] : 0, c = (up && i >= bpp) ? up[i - bpp] : 0;
            int v = src[i];
            switch (f) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) / 2; break;
            case 4: v += paeth(a, b, c); break;
            default: throw IOError("bad PNG filter in '" + path + "'");
            }
            cur[i] = (uint8_t)v;
        }
    }
    rgb.resize((size_t)W * H * 3);
    const size_t step = depth / 8;  // take the high byte of 16-bit samples
    for (size_t p = 0; p < (size_t)W * H; ++p) {
        const uint8_t *s = &img[p * bpp];
        uint8_t r, g, b;
        if (ctype == 3) {
            const size_t k = (size_t)s[0] * 3;
            if (k + 2 >= plte.size()) throw IOError("bad PNG palette index in '" + path + "'");
            r = plte[k];
            g = plte[k + 1];
            b = plte[k + 2];
        } else if (ch <= 2) {
            r = g = b = s[0];
        } else {
            r = s[0];
            g = s[step];
            b = s[2 * step];
        }
        rgb[3 * p] = r;
        rgb[3 * p + 1] = g;
        rgb[3 * p + 2] = b;
    }
}

void read_pnm(const std::string &path, const std::vector<uint8_t> &d, int &W, int &H, std::vector<uint8_t> &rgb)
{
    size_t pos = 2;
    auto next_int = [&]() {
        while (pos < d.size()) {
            if (d[pos] == '#')
                while (pos < d.size() && d[pos] != '\n') ++pos;
            else if (std::isspace(d[pos]))
                ++pos;
            else
                break;
        }
        long v = 0;
        bool any = false;
        while (pos < d.size() && std::isdigit(d[pos])) {
            v = v * 10 + (d[pos++] - '0');
            any = true;
            if (v > 100000000L) throw IOError("bad PNM header in '" + path + "'");  // no such image fits a file
        }
        if (!any) throw IOError("bad PNM header in '" + path + "'");
        return v;
    };
    const bool color = d[1] == '6';
    W = (int)next_int();
    H = (int)next_int();
    const long maxv = next_int();
    if (maxv <= 0 || maxv > 255 || W <= 0 || H <= 0) throw IOError("unsupported PNM in '" + path + "'");
    ++pos;
    const size_t n = (size_t)W * H * (color ? 3 : 1);
    if (pos + n > d.size()) throw IOError("truncated PNM '" + path + "'");
    rgb.resize((size_t)W * H * 3);
    for (size_t p = 0; p < (size_t)W * H; ++p)
        for (int c = 0; c < 3; ++c) rgb[3 * p + c] = d[pos + (color ? 3 * p + c : p)];
}

}  // namespace

bool unknown_flow(float u, float v)
{
    return std::fabs(u) > 1e9f || std::fabs(v) > 1e9f || std::isnan(u) || std::isnan(v);
}

void write_flo(const std::string &path, int W, int H, const float *U, const float *V)
{
    if (!ends_with(path, ".flo")) throw IOError("filename '" + path + "' should have extension '.flo'");
    FILE *f = open_w(path);
    std::fwrite("PIEH", 1, 4, f);
    const int32_t w = W, h = H;
    std::fwrite(&w, 4, 1, f);
    std::fwrite(&h, 4, 1, f);
    std::vector<float> row(2 * (size_t)W);
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            row[2 * x] = U[(size_t)y * W + x];
            row[2 * x + 1] = V[(size_t)y * W + x];
        }
        if (std::fwrite(row.data(), sizeof(float), row.size(), f) != row.size()) {
            std::fclose(f);
            throw IOError("problem writing '" + path + "'");
        }
    }
    std::fclose(f);
}

void read_flo(const std::string &path, int &W, int &H, std::vector<float> &U, std::vector<float> &V)
{
    if (!ends_with(path, ".flo")) throw IOError("extension .flo expected: '" + path + "'");
    const std::vector<uint8_t> d = read_file(path);
    if (d.size() < 12) throw IOError("problem reading '" + path + "'");
    float tag;
    int32_t w, h;
    std::memcpy(&tag, d.data(), 4);
    std::memcpy(&w, d.data() + 4, 4);
    std::memcpy(&h, d.data() + 8, 4);
    if (tag != kFloTag) throw IOError("wrong tag in '" + path + "' (possibly due to big-endian machine?)");
    if (w < 1 || w > 99999) throw IOError("illegal width in '" + path + "'");
    if (h < 1 || h > 99999) throw IOError("illegal height in '" + path + "'");
    const size_t n = (size_t)w * h;
    if (d.size() < 12 + 8 * n) throw IOError("'" + path + "' is too short");
    if (d.size() > 12 + 8 * n) throw IOError("'" + path + "' is too long");
    W = w;
    H = h;
    U.resize(n);
    V.resize(n);
    const float *p = reinterpret_cast<const float *>(d.data() + 12);
    for (size_t i = 0; i < n; ++i) {
        std::memcpy(&U[i], p + 2 * i, 4);
        std::memcpy(&V[i], p + 2 * i + 1, 4);
    }
}

void write_raw(const std::string &path, int W, int H, const float *U, const float *V)
{
    FILE *f = open_w(path);
    const size_t n = (size_t)W * H;
    const bool ok = std::fwrite(U, sizeof(float), n, f) == n && std::fwrite(V, sizeof(float), n, f) == n;
    std::fclose(f);
    if (!ok) throw IOError("problem writing '" + path + "'");
}

// Middlebury colour wheel (Baker et al., "A Database and Evaluation Methodology
// for Optical Flow"): 55 hues with transition lengths RY 15, YG 6, GC 4, CB 11,
// BM 13, MR 6; inside the unit circle saturation grows with the radius, outside
// the colour is darkened to 75%.
void flow_to_rgb(int W, int H, const float *U, const float *V, float maxmot, uint8_t *rgb)
{
    static const int seg[6] = {15, 6, 4, 11, 13, 6};
    std::vector<float> wheel;  // ncols x 3
    for (int s = 0; s < 6; ++s)
        for (int i = 0; i < seg[s]; ++i) {
            const float t = (float)(255 * i / seg[s]);
            float c[3];
            switch (s) {
            case 0: c[0] = 255; c[1] = t; c[2] = 0; break;
            case 1: c[0] = 255 - t; c[1] = 255; c[2] = 0; break;
            case 2: c[0] = 0; c[1] = 255; c[2] = t; break;
            case 3: c[0] = 0; c[1] = 255 - t; c[2] = 255; break;
            case 4: c[0] = t; c[1] = 0; c[2] = 255; break;
            default: c[0] = 255; c[1] = 0; c[2] = 255 - t; break;
            }
            wheel.insert(wheel.end(), c, c + 3);
        }
    const int ncols = (int)wheel.size() / 3;
    const size_t n = (size_t)W * H;
    float maxrad = -1.0f;
    for (size_t i = 0; i < n; ++i)
        if (!unknown_flow(U[i], V[i])) maxrad = std::max(maxrad, std::sqrt(U[i] * U[i] + V[i] * V[i]));
    if (maxmot > 0) maxrad = maxmot;
    if (maxrad == 0) maxrad = 1;
    for (size_t i = 0; i < n; ++i) {
        uint8_t *px = rgb + 3 * i;
        if (unknown_flow(U[i], V[i])) {
            px[0] = px[1] = px[2] = 0;
            continue;
        }
        const float fx = U[i] / maxrad, fy = V[i] / maxrad;
        const float rad = std::sqrt(fx * fx + fy * fy);
        const float a = std::atan2(-fy, -fx) / (float)M_PI;
        const float fk = (a + 1.0f) / 2.0f * (ncols - 1);
        const int k0 = (int)fk, k1 = (k0 + 1) % ncols;
        const float f = fk - k0;
        for (int b = 0; b < 3; ++b) {
            float col = (1 - f) * wheel[3 * k0 + b] / 255.0f + f * wheel[3 * k1 + b] / 255.0f;
            col = rad <= 1 ? 1 - rad * (1 - col) : col * 0.75f;
            px[b] = (uint8_t)(int)(255.0f * col);
        }
    }
}

void read_image_rgb(const std::string &path, int &W, int &H, std::vector<uint8_t> &rgb)
{
    const std::vector<uint8_t> d = read_file(path);
    if (d.size() >= 8 && d[0] == 0x89 && d[1] == 'P') return read_png(path, d, W, H, rgb);
    if (d.size() >= 3 && d[0] == 'P' && (d[1] == '5' || d[1] == '6')) return read_pnm(path, d, W, H, rgb);
    throw IOError("unsupported image format: '" + path + "' (PNG, PGM or PPM)");
}

void write_png_rgb(const std::string &path, int W, int H, const uint8_t *rgb)
{
    std::vector<uint8_t> raw(((size_t)W * 3 + 1) * H);
    for (int y = 0; y < H; ++y) {
        raw[((size_t)W * 3 + 1) * y] = 0;
        std::memcpy(&raw[((size_t)W * 3 + 1) * y + 1], rgb + (size_t)y * W * 3, (size_t)W * 3);
    }
    uLongf zn = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 6) != Z_OK) throw IOError("PNG compression failed");
    z.resize(zn);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    auto chunk = [&](const char *type, const std::vector<uint8_t> &data) {
        put_be32(out, (uint32_t)data.size());
        const size_t start = out.size();
        out.insert(out.end(), type, type + 4);
        out.insert(out.end(), data.begin(), data.end());
        put_be32(out, (uint32_t)crc32(0L, &out[start], (uInt)(out.size() - start)));
    };
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)W);
    put_be32(ihdr, (uint32_t)H);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    chunk("IHDR", ihdr);
    chunk("IDAT", z);
    chunk("IEND", {});
    FILE *f = open_w(path);
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    std::fclose(f);
    if (!ok) throw IOError("problem writing '" + path + "'");
}

void epe(const float *U, const float *V, const float *gtU, const float *gtV, const uint8_t *occ, size_t n,
         double &epe_all, double &epe_noc)
{
    double sa = 0.0, sn = 0.0;
    size_t na = 0, nn = 0;
    for (size_t i = 0; i < n; ++i) {
        if (unknown_flow(gtU[i], gtV[i])) continue;
        const double du = (double)gtU[i] - U[i], dv = (double)gtV[i] - V[i];
        const double e = std::sqrt(du * du + dv * dv);
        sa += e;
        ++na;
        if (!occ || !occ[i]) {
            sn += e;
            ++nn;
        }
    }
    epe_all = na ? sa / (double)na : 0.0;
    epe_noc = nn ? sn / (double)nn : 0.0;
}

}  // namespace bmio

namespace {
thread_local std::string g_io_error;

template <class F>
int guarded(F f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_io_error = e.what();
        return 4;  // BM_E_IO
    }
}
}  // namespace

extern "C" {

const char *bm_io_last_error(void) { return g_io_error.c_str(); }

int bm_write_flo(const char *path, int W, int H, const float *U, const float *V)
{
    return guarded([&] { bmio::write_flo(path, W, H, U, V); });
}

int bm_read_flo(const char *path, int *W, int *H, float *U, float *V)
{
    return guarded([&] {
        std::vector<float> u, v;
        bmio::read_flo(path, *W, *H, u, v);
        if (U && V) {
            std::memcpy(U, u.data(), u.size() * sizeof(float));
            std::memcpy(V, v.data(), v.size() * sizeof(float));
        }
    });
}

int bm_write_raw(const char *path, int W, int H, const float *U, const float *V)
{
    return guarded([&] { bmio::write_raw(path, W, H, U, V); });
}

int bm_flow_to_rgb(int W, int H, const float *U, const float *V, float maxmot, uint8_t *rgb)
{
    return guarded([&] { bmio::flow_to_rgb(W, H, U, V, maxmot, rgb); });
}

int bm_read_image_rgb(const char *path, int *W, int *H, uint8_t *rgb)
{
    return guarded([&] {
        std::vector<uint8_t> v;
        bmio::read_image_rgb(path, *W, *H, v);
        if (rgb) std::memcpy(rgb, v.data(), v.size());
    });
}

int bm_write_png_rgb(const char *path, int W, int H, const uint8_t *rgb)
{
    return guarded([&] { bmio::write_png_rgb(path, W, H, rgb); });
}

}  // extern "C"
