// Robustness of the image and flow readers (briefmatch_amd/csrc/bm_io.cpp)
// under AddressSanitizer and UBSan: every truncation and seeded byte mutations
// of valid PNG, PPM, PGM and .flo files, and hand-made headers that promise
// more data than the file holds.  Every input must either parse or throw
// bmio::IOError; a crash, a sanitizer report or any other exception fails.
//
//   g++ -O1 -g -fsanitize=address,undefined -fno-sanitize-recover=all -std=c++17 \
//       -Iinclude tests/io_sanitize.cpp briefmatch_amd/csrc/bm_io.cpp -lz -o io_sanitize
//   ./io_sanitize <scratch directory>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "bm_io.h"

namespace {

uint32_t g_state = 2463534242u;
uint32_t next_u32()
{
    g_state ^= g_state << 13;
    g_state ^= g_state >> 17;
    g_state ^= g_state << 5;
    return g_state;
}

std::vector<uint8_t> slurp(const std::string &p)
{
    std::ifstream f(p, std::ios::binary);
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

void spill(const std::string &p, const std::vector<uint8_t> &d)
{
    std::ofstream f(p, std::ios::binary);
    f.write((const char *)d.data(), (std::streamsize)d.size());
}

long g_parsed = 0, g_rejected = 0;

// 0: parsed or rejected with IOError; 1: anything else
int try_image(const std::string &p)
{
    try {
        int W = 0, H = 0;
        std::vector<uint8_t> rgb;
        bmio::read_image_rgb(p, W, H, rgb);
        if (W <= 0 || H <= 0 || rgb.size() != (size_t)W * H * 3) {
            std::fprintf(stderr, "io_sanitize: inconsistent image from %s\n", p.c_str());
            return 1;
        }
        ++g_parsed;
    } catch (const bmio::IOError &) {
        ++g_rejected;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "io_sanitize: unexpected exception on %s: %s\n", p.c_str(), e.what());
        return 1;
    }
    return 0;
}

int try_flo(const std::string &p)
{
    try {
        int W = 0, H = 0;
        std::vector<float> U, V;
        bmio::read_flo(p, W, H, U, V);
        if (U.size() != (size_t)W * H || V.size() != U.size()) return 1;
        ++g_parsed;
    } catch (const bmio::IOError &) {
        ++g_rejected;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "io_sanitize: unexpected exception on %s: %s\n", p.c_str(), e.what());
        return 1;
    }
    return 0;
}

void put32be(std::vector<uint8_t> &d, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) d.push_back((uint8_t)(v >> s));
}

// a PNG whose IHDR claims W x H with the given colour type / depth and whose
// single IDAT chunk holds `idat` (CRCs are not checked by the reader)
std::vector<uint8_t> png_with(uint32_t W, uint32_t H, uint8_t depth, uint8_t ctype, const std::vector<uint8_t> &idat)
{
    std::vector<uint8_t> d = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    put32be(d, 13);
    for (char c : std::string("IHDR")) d.push_back((uint8_t)c);
    put32be(d, W);
    put32be(d, H);
    d.insert(d.end(), {depth, ctype, 0, 0, 0});
    put32be(d, 0);
    put32be(d, (uint32_t)idat.size());
    for (char c : std::string("IDAT")) d.push_back((uint8_t)c);
    d.insert(d.end(), idat.begin(), idat.end());
    put32be(d, 0);
    put32be(d, 0);
    for (char c : std::string("IEND")) d.push_back((uint8_t)c);
    put32be(d, 0);
    return d;
}

}  // namespace

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    int bad = 0;
    // valid sources
    const int W = 7, H = 5;
    std::vector<uint8_t> rgb((size_t)W * H * 3);
    for (auto &v : rgb) v = (uint8_t)next_u32();
    std::vector<float> U((size_t)W * H), V(U.size());
    for (size_t i = 0; i < U.size(); ++i) {
        U[i] = (float)(int)(next_u32() % 200) - 100.0f;
        V[i] = (float)(int)(next_u32() % 200) - 100.0f;
    }
    const std::string png = dir + "/v.png", flo = dir + "/v.flo";
    bmio::write_png_rgb(png, W, H, rgb.data());
    bmio::write_flo(flo, W, H, U.data(), V.data());
    std::vector<uint8_t> ppm = {'P', '6', '\n', '7', ' ', '5', '\n', '2', '5', '5', '\n'};
    ppm.insert(ppm.end(), rgb.begin(), rgb.end());
    std::vector<uint8_t> pgm = {'P', '5', ' ', '#', 'c', '\n', '7', ' ', '5', ' ', '2', '5', '5', '\n'};
    for (int i = 0; i < W * H; ++i) pgm.push_back(rgb[3 * i]);
    struct Src {
        std::vector<uint8_t> bytes;
        std::string ext;
        bool flow;
    } srcs[] = {{slurp(png), ".png", false}, {ppm, ".ppm", false}, {pgm, ".pgm", false}, {slurp(flo), ".flo", true}};
    for (const Src &s : srcs) {
        const std::string p = dir + "/t" + s.ext;
        // every truncation
        for (size_t n = 0; n <= s.bytes.size(); ++n) {
            spill(p, std::vector<uint8_t>(s.bytes.begin(), s.bytes.begin() + (long)n));
            bad |= s.flow ? try_flo(p) : try_image(p);
        }
        // seeded mutations: 1..4 random bytes set to random or boundary values
        for (int m = 0; m < 3000; ++m) {
            std::vector<uint8_t> d = s.bytes;
            const int k = 1 + (int)(next_u32() % 4);
            for (int j = 0; j < k; ++j) {
                const size_t at = next_u32() % d.size();
                const uint32_t r = next_u32();
                d[at] = (r & 3) == 0 ? 0xFF : (r & 3) == 1 ? 0x00 : (uint8_t)(r >> 8);
            }
            spill(p, d);
            bad |= s.flow ? try_flo(p) : try_image(p);
        }
    }
    // headers that promise far more data than the file holds
    const std::string p = dir + "/h.png";
    const std::vector<uint8_t> tiny = {0x78, 0x9c, 0x63, 0x00, 0x00, 0x00, 0x01, 0x00, 0x01};
    for (uint8_t ct : {0, 2, 3, 4, 6})
        for (uint8_t dp : {8, 16}) {
            spill(p, png_with(65535, 65535, dp, ct, tiny));
            bad |= try_image(p);
            spill(p, png_with(65535, 1, dp, ct, tiny));
            bad |= try_image(p);
            spill(p, png_with(0x80000000u, 3, dp, ct, tiny));
            bad |= try_image(p);
        }
    for (const char *hdr : {"P6\n99999999999999999999999 5\n255\n", "P6\n5 4294967301\n255\n",
                            "P5\n-3 4\n255\n", "P6\n4 4\n65535\n", "P5\n100000 100000\n255\n", "P6", "P5 7"}) {
        const std::string q = dir + "/h.ppm";
        std::vector<uint8_t> d(hdr, hdr + std::string(hdr).size());
        d.resize(d.size() + 48, 7);
        spill(q, d);
        bad |= try_image(q);
    }
    std::printf("io_sanitize: %ld parsed, %ld rejected, %s\n", g_parsed, g_rejected, bad ? "FAILED" : "ok");
    return bad;
}
