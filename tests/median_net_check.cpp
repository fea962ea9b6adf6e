// Checks the median-selection networks of bm_median_net.h against a sort: all
// 0/1 inputs for windows up to 19 (0-1 principle: a comparator network that
// selects correctly on every 0/1 input selects correctly on every input), and
// random floats (with ties) for every window up to 47.  Exit status 0 = pass.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>
#include "bm_median_net.h"

template <int FS>
float med(const float *x)
{
    float v[64];
    for (int i = 0; i < FS; ++i) v[i] = x[i];
    constexpr auto &net = bm::MedianNet<FS>::net;
    for (int i = 0; i < net.count; ++i) {
        const bm::MedOp o = net.op[i];
        if (o.kind == 0) { v[o.a] = v[o.b]; continue; }
        const float lo = std::min(v[o.a], v[o.b]), hi = std::max(v[o.a], v[o.b]);
        if (o.kind & 1) v[o.a] = lo;
        if (o.kind & 2) v[o.b] = hi;
    }
    return v[FS / 2];
}

template <int FS>
int check()
{
    std::mt19937 rng(FS);
    float x[64], y[64];
    long bad = 0;
    if (FS <= 19) {
        for (uint32_t m = 0; m < (1u << FS); ++m) {
            int ones = 0;
            for (int i = 0; i < FS; ++i) { x[i] = (float)((m >> i) & 1u); ones += (m >> i) & 1u; }
            const float want = ones > FS / 2 ? 1.0f : 0.0f;
            if (med<FS>(x) != want) ++bad;
        }
    }
    std::uniform_real_distribution<float> u(-50.f, 50.f);
    std::uniform_int_distribution<int> d(-3, 3);
    for (int t = 0; t < 50000; ++t) {
        for (int i = 0; i < FS; ++i) x[i] = (t & 1) ? u(rng) : (float)d(rng);
        std::copy(x, x + FS, y);
        std::nth_element(y, y + FS / 2, y + FS);
        if (med<FS>(x) != y[FS / 2]) ++bad;
    }
    std::printf("FS %2d ops %3d (partial sort %3d) mismatches %ld\n", FS, bm::MedianNet<FS>::net.minmax,
                bm::MedianNet<FS>::partial_sort_ops, bad);
    return bad != 0;
}

int main()
{
    int r = 0;
    r |= check<1>(); r |= check<3>(); r |= check<5>(); r |= check<7>(); r |= check<9>(); r |= check<11>();
    r |= check<13>(); r |= check<15>(); r |= check<17>(); r |= check<19>(); r |= check<21>(); r |= check<23>();
    r |= check<25>(); r |= check<27>(); r |= check<29>(); r |= check<31>(); r |= check<33>(); r |= check<35>();
    r |= check<37>(); r |= check<39>(); r |= check<41>(); r |= check<43>(); r |= check<45>(); r |= check<47>();
    return r;
}
