// Median-selection networks built at compile time.
//
// Batcher's odd-even merge sorting network on the next power of two >= FS
// wires, with wires FS.. as +infinity: a comparator between two +inf wires is
// dropped, one between a value and a +inf wire only moves the value (a "move",
// free after unrolling), and a backward pass from the median wire FS / 2 keeps
// only comparators whose outputs are used, each as min, max or both.  The
// result selects the exact middle element (any comparator network that sorts
// does, and pruning removes only work whose result is never read).
#ifndef BM_MEDIAN_NET_H
#define BM_MEDIAN_NET_H

namespace bm {

struct MedOp {
    int a, b;
    int kind;  // 0: move b -> a; 1: a = min; 2: b = max; 3: both
};

template <int FS>
struct MedianNet {
    static constexpr int N = FS <= 2 ? 2 : FS <= 4 ? 4 : FS <= 8 ? 8 : FS <= 16 ? 16 : FS <= 32 ? 32 : 64;
    static constexpr int kMaxComps = 600;  // Batcher on 64 wires has 543 comparators

    struct Net {
        MedOp op[kMaxComps];
        int count;
        int minmax;  // min/max instructions of the pruned network
    };

    static constexpr Net build()
    {
        int ca[kMaxComps] = {}, cb[kMaxComps] = {};
        int nc = 0;
        for (int p = 1; p < N; p *= 2)
            for (int k = p; k >= 1; k /= 2)
                for (int j = k % p; j + k < N; j += 2 * k)
                    for (int i = 0; i < k && i < N - j - k; ++i)
                        if ((i + j) / (2 * p) == (i + j + k) / (2 * p)) {
                            ca[nc] = i + j;
                            cb[nc] = i + j + k;
                            ++nc;
                        }
        // forward: +inf wires
        bool inf[64] = {};
        for (int w = FS; w < N; ++w) inf[w] = true;
        MedOp live[kMaxComps] = {};
        int nl = 0;
        for (int c = 0; c < nc; ++c) {
            const int a = ca[c], b = cb[c];
            if (inf[b]) continue;  // min(a, inf) = a, max = inf: nothing moves
            if (inf[a]) {          // min = b lands on a, b becomes +inf
                inf[a] = false;
                inf[b] = true;
                live[nl++] = MedOp{a, b, 0};
                continue;
            }
            live[nl++] = MedOp{a, b, 3};
        }
        // backward from the median wire
        bool need[64] = {};
        need[FS / 2] = true;
        Net net = {};
        int n = 0;
        for (int c = nl - 1; c >= 0; --c) {
            MedOp o = live[c];
            if (o.kind == 0) {
                if (need[o.a]) {
                    need[o.a] = false;
                    need[o.b] = true;
                    net.op[n++] = o;
                }
                continue;
            }
            const bool na = need[o.a], nb = need[o.b];
            if (!na && !nb) continue;
            o.kind = (na ? 1 : 0) | (nb ? 2 : 0);
            need[o.a] = need[o.b] = true;
            net.op[n++] = o;
            net.minmax += (na ? 1 : 0) + (nb ? 1 : 0);
        }
        // reverse into execution order
        for (int i = 0; i < n / 2; ++i) {
            const MedOp t = net.op[i];
            net.op[i] = net.op[n - 1 - i];
            net.op[n - 1 - i] = t;
        }
        net.count = n;
        return net;
    }

    static constexpr Net net = build();
    // min/max instructions of the partial selection sort it replaces
    static constexpr int partial_sort_ops = (FS / 2 + 1) * (2 * FS - 2 - FS / 2);
};

}  // namespace bm

#endif
