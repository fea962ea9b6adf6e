// AddressSanitizer / UBSan driver of the CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// Runs every exported oracle function on small frames of odd sizes (a few
// pixels up to 70 x 45) and descriptor lengths 2..512, including median
// windows wider than the frame, trilateral row ranges and EPE with unknown and
// occluded ground truth, so that the sanitizers see every indexing path.
// Built by `make -C oracle san`; exits non-zero on any sanitizer report
// (-fno-sanitize-recover) or failed check.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "oracle_bm.h"

namespace {

uint32_t g_state = 12345u;
uint32_t next_u32()
{
    g_state ^= g_state << 13;
    g_state ^= g_state >> 17;
    g_state ^= g_state << 5;
    return g_state;
}

int fail(const char *what, int W, int H)
{
    std::fprintf(stderr, "oracle_san: %s failed at %dx%d\n", what, W, H);
    return 1;
}

int run_case(int W, int H, float up, int N, int lanczos)
{
    orc_params p;
    if (orc_derive_params((unsigned)W, (unsigned)H, up, up, 3.0f, 0.025f, 1.0f, 0.85f, &p) != 0)
        return fail("derive_params", W, H);
    p.N = (unsigned)N;
    std::vector<uint8_t> rgb1((size_t)W * H * 3), rgb2(rgb1.size());
    for (size_t i = 0; i < rgb1.size(); ++i) {
        rgb1[i] = (uint8_t)next_u32();
        rgb2[i] = (uint8_t)((i % 7 == 0) ? next_u32() : rgb1[(i + 3) % rgb1.size()]);
    }
    const size_t M = (size_t)p.sx * p.sy, Win = (size_t)W * H;
    std::vector<uint32_t> off(M);
    std::vector<uint16_t> cost(M);
    std::vector<float> U(Win), V(Win);
    if (orc_flow_pair_up(&p, 7u, 3u, rgb1.data(), rgb2.data(), lanczos, off.data(), cost.data(), U.data(),
                         V.data()) != 0)
        return fail("flow_pair_up", W, H);
    for (size_t i = 0; i < Win; ++i)
        if (!std::isfinite(U[i]) || !std::isfinite(V[i])) return fail("finite flow", W, H);

    // stages on their own: temporal init, levels of both directions, random search
    const int NS = (N + 31) / 32;
    std::vector<uint8_t> g1(Win), g2(Win);
    orc_gray(rgb1.data(), W, H, g1.data());
    orc_gray(rgb2.data(), W, H, g2.data());
    std::vector<float> L1(M), L2(M);
    orc_upsample(g1.data(), W, H, (int)p.sx, (int)p.sy, L1.data());
    orc_upsample_lanczos(g2.data(), W, H, (int)p.sx, (int)p.sy, L2.data());
    std::vector<int16_t> pat((size_t)4 * N);
    orc_make_pattern(11u, (unsigned)N, p.patchArea, pat.data());
    std::vector<uint32_t> F1(M * NS), F2(M * NS);
    orc_brief(L1.data(), (int)p.sx, (int)p.sy, pat.data(), N, F1.data());
    orc_brief(L2.data(), (int)p.sx, (int)p.sy, pat.data(), N, F2.data());
    std::vector<uint32_t> prev(M), o2(M);
    std::vector<uint16_t> c2(M);
    for (size_t i = 0; i < M; ++i) {
        const int dx = (int)(next_u32() % 61) - 30, dy = (int)(next_u32() % 61) - 30;
        prev[i] = (uint32_t)(uint16_t)(int16_t)dx | ((uint32_t)(uint16_t)(int16_t)dy << 16);
    }
    orc_init(F1.data(), F2.data(), (int)p.sx, (int)p.sy, NS, 0, prev.data(), off.data(), cost.data());
    for (int lvl : {256, 17, 2, 1})
        for (int pd : {-1, 1}) {
            orc_propagate_level(F1.data(), F2.data(), (int)p.sx, (int)p.sy, NS, lvl, pd, pd > 0, off.data(),
                                cost.data(), o2.data(), c2.data());
            off.swap(o2);
            cost.swap(c2);
        }
    orc_random_search(F1.data(), F2.data(), (int)p.sx, (int)p.sy, NS, orc_rs_sigma_q16(p.sx), 2u, 5u, 9u,
                      off.data(), cost.data());

    // refinement pieces with windows wider than the frame and partial row ranges
    std::vector<float> kernel(64);
    unsigned P = 0;
    orc_gauss_kernel(p.sigma, kernel.data(), &P);
    orc_downsample(off.data(), (int)p.sx, (int)p.sy, W, H, kernel.data(), (int)P, U.data(), V.data());
    std::vector<float> Um(Win), Vm(Win), Uo(Win), Vo(Win), I(Win);
    for (size_t i = 0; i < Win; ++i) I[i] = (float)g1[i];
    const int fs = 2 * std::max(W, H) + 1;
    orc_median1d(U.data(), Um.data(), W, H, 1, fs);
    orc_median1d(Um.data(), Vm.data(), W, H, 0, 3);
    orc_trilateral(U.data(), V.data(), Um.data(), Vm.data(), I.data(), W, H, 0.16f, 0.01f, 30.0f, H / 3,
                   H, Uo.data(), Vo.data());
    std::vector<uint8_t> occ(Win);
    for (size_t i = 0; i < Win; ++i) occ[i] = (uint8_t)(i % 3 == 0);
    std::vector<float> gU(U), gV(V);
    gU[0] = 2e9f;                      // unknown ground truth
    gV[Win - 1] = std::nanf("");
    double ea = 0, en = 0;
    orc_epe(U.data(), V.data(), gU.data(), gV.data(), occ.data(), (int64_t)Win, &ea, &en);
    if (!std::isfinite(ea) || !std::isfinite(en)) return fail("epe", W, H);
    return 0;
}

}  // namespace

int main()
{
    struct Case {
        int W, H;
        float up;
        int N, lanczos;
    };
    const Case cases[] = {{5, 3, 1.0f, 2, 0},    {16, 9, 2.0f, 32, 1},   {37, 23, 2.9f, 72, 0},
                          {64, 48, 1.5f, 128, 1}, {70, 45, 4.9f, 512, 0}, {13, 40, 3.1f, 100, 1}};
    int bad = 0;
    for (const Case &c : cases) {
        bad |= run_case(c.W, c.H, c.up, c.N, c.lanczos);
        std::printf("oracle_san: %dx%d x%.1f N=%d lanczos=%d ok\n", c.W, c.H, c.up, c.N, c.lanczos);
    }
    return bad;
}
