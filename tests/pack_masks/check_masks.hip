// Host check of the packed engine's lane-mask steps (include/rbf_wavepack.h):
// WavePack::segments_tied and WavePack::pivot_lanes, which work on whole
// 64-bit ballots at once (carries and borrows confined to each segment),
// against a segment-by-segment brute force: the lowest set bit of each
// segment's hit lanes (or of its active lanes when it has none, that segment
// then reported in nocand), and "some segment has two or more hit lanes".
// Random and structured masks, 16- and 32-lane segments. Prints "ok <count>"
// or the first mismatch; exit status 0 only when every case agrees.
#include <cstdint>
#include <cstdio>

#include "rbf_wavepack.h"

using namespace rbf_operators::detail;

namespace {

uint64_t g_state = 0x9E3779B97F4A7C15ull;
uint64_t next() {  // xorshift64
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return g_state;
}

int popcount64(uint64_t x) {
    int c = 0;
    for (; x; x &= x - 1) ++c;
    return c;
}

template <int SL>
uint64_t lowest_per_segment(uint64_t x) {
    const uint64_t segm = SL == 64 ? ~0ull : ((1ull << SL) - 1);
    uint64_t r = 0;
    for (int g = 0; g < 64 / SL; ++g) {
        const uint64_t s = (x >> (g * SL)) & segm;
        if (s) r |= (s & (~s + 1)) << (g * SL);
    }
    return r;
}

// one random mask per segment, drawn from several shapes
template <int SL>
uint64_t draw(uint64_t within) {
    const uint64_t segm = (1ull << SL) - 1;
    uint64_t m = 0;
    for (int g = 0; g < 64 / SL; ++g) {
        const uint64_t w = (within >> (g * SL)) & segm;
        uint64_t s = 0;
        switch (next() % 6) {
            case 0: s = 0; break;                                // empty
            case 1: s = w & (1ull << (next() % SL)); break;      // one lane (or none)
            case 2: s = w; break;                                // all
            case 3: s = w & next(); break;                       // random
            case 4: s = w & next() & next(); break;              // sparse
            default: s = w & (3ull << (next() % (SL - 1))); break;  // two neighbours
        }
        m |= s << (g * SL);
    }
    return m;
}

template <class E, int SL>
long check(long iters, const char* name) {
    const uint64_t segm = (1ull << SL) - 1;
    long bad = 0;
    for (long it = 0; it < iters; ++it) {
        // active rows: nonempty in every segment (the elimination removes one
        // row per segment per step, and stops before any segment runs out)
        uint64_t act = draw<SL>(E::ROWS);
        for (int g = 0; g < 64 / SL; ++g)
            if (!((act >> (g * SL)) & segm)) act |= (1ull << (next() % E::NT)) << (g * SL);
        const uint64_t hit = draw<SL>(act);
        bool tied = false;
        uint64_t nc_ref = 0;
        for (int g = 0; g < 64 / SL; ++g) {
            const uint64_t h = (hit >> (g * SL)) & segm;
            if (popcount64(h) >= 2) tied = true;
            if (!h) nc_ref |= segm << (g * SL);
        }
        const uint64_t pm_ref = lowest_per_segment<SL>(hit) | lowest_per_segment<SL>(act & nc_ref);
        unsigned long long nc = ~0ull;
        const uint64_t pm = E::pivot_lanes(hit, act, nc);
        if (E::segments_tied(hit) != tied || pm != pm_ref || nc != nc_ref) {
            if (bad++ < 5)
                std::printf("%s mismatch: act %016llx hit %016llx: tied %d/%d pm %016llx/%016llx "
                            "nocand %016llx/%016llx\n",
                            name, (unsigned long long)act, (unsigned long long)hit,
                            (int)E::segments_tied(hit), (int)tied, (unsigned long long)pm,
                            (unsigned long long)pm_ref, (unsigned long long)nc,
                            (unsigned long long)nc_ref);
        }
    }
    return bad;
}

}  // namespace

int main() {
    const long n = 2000000;
    long bad = 0;
    bad += check<WavePack<2, 9, 2, 3, 9, true, 16>, 16>(n, "16-lane, nt 15");
    bad += check<WavePack<2, 13, 1, 3, 3, false, 16>, 16>(n, "16-lane, nt 16");
    bad += check<WavePack<2, 13, 2, 3, 3, false, 32>, 32>(n, "32-lane, nt 19");
    bad += check<WavePack<2, 21, 2, 3, 3, false, 32>, 32>(n, "32-lane, nt 27");
    if (bad) {
        std::printf("FAILED %ld\n", bad);
        return 1;
    }
    std::printf("ok %ld\n", 4 * n);
    return 0;
}
