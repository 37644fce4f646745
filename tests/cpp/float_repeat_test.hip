// Host check of knp_amd/csrc/float_repeat.h: repeat_add and replay_uniform against the additions done one by
// one, bit for bit, over random and constructed cases (binade crossings, ties to even, sign changes, zeros,
// subnormals, overflow, NaN). Prints "ok <cases>" or the first mismatch.
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <random>

#include "../../knp_amd/csrc/float_repeat.h"

using knp_mi355x::repeat_add;
using knp_mi355x::replay_uniform;

static uint32_t bits(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

static float plain(float x, float c, uint64_t n) {
    for (; n; --n) x += c;
    return x;
}

static int check(float x, float c, uint64_t n, uint64_t &cases) {
    ++cases;
    const float want = plain(x, c, n), got = repeat_add(x, c, n);
    if (bits(want) != bits(got) && !(std::isnan(want) && std::isnan(got))) {
        std::printf("mismatch: x=%a c=%a n=%" PRIu64 " want=%a got=%a\n", x, c, n, want, got);
        return 1;
    }
    return 0;
}

int main(int argc, char **argv) {
    const uint64_t rounds = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 200000;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    uint64_t cases = 0;
    auto sgn = [&]() { return (rng() & 1) ? -1.0f : 1.0f; };
    auto n_of = [&]() -> uint64_t {
        switch (rng() % 4) {
            case 0: return rng() % 8;
            case 1: return rng() % 100;
            case 2: return rng() % 2000;
            default: return rng() % 70000;
        }
    };
    for (uint64_t r = 0; r < rounds; ++r) {
        // random magnitudes: x over a wide range, c relative to it
        const int ex = (int)(rng() % 80) - 40, rel = (int)(rng() % 40) - 30;
        const float x = sgn() * std::ldexp((float)(1.0 + U(rng)), ex);
        float c = sgn() * std::ldexp((float)(1.0 + U(rng)), ex + rel);
        if (rng() % 8 == 0) {  // a tie: c an odd multiple of half the ulp of x's binade
            int e;
            (void)std::frexp(std::fabs(x), &e);
            c = sgn() * std::ldexp((float)(2 * (rng() % 1000) + 1), e - 25);
        }
        if (check(x, c, n_of(), cases)) return 1;
    }
    const float special_x[] = {0.0f, -0.0f, 1e-45f, -1e-45f, 1e-40f, 1.1754944e-38f, 1.0f, -1.0f, 3e38f, -3e38f,
                               INFINITY, -INFINITY, NAN, 0.5f, 0.75f, 1024.0f, 16777216.0f, 16777215.0f};
    const float special_c[] = {1e-45f, -1e-45f, 1e-40f, -3e-39f, 1e-30f, 0x1p-100f, -0x1p-101f, 0.0625f, -0.0625f,
                               0.1f, -0.1f, 1.0f, -1.0f, 0.5f, 1e38f, -1e38f, 3e-8f, -3e-8f, 0x1p100f, 0x1p101f,
                               NAN, 0.0f, -0.0f, 0.3f, -0.7f};
    for (float x : special_x)
        for (float c : special_c)
            for (uint64_t n : {0ull, 1ull, 2ull, 3ull, 4ull, 5ull, 7ull, 64ull, 100ull, 785ull, 5000ull, 100000ull})
                if (check(x, c, n, cases)) return 1;
    // runs that cross zero and binades in both directions
    for (int k = 0; k < 20000; ++k) {
        const float x = sgn() * (float)U(rng) * 4.0f, c = sgn() * std::ldexp((float)(1.0 + U(rng)), -(int)(rng() % 24));
        if (check(x, c, rng() % 20000, cases)) return 1;
    }
    // replay_uniform against the lane-by-lane order
    for (int k = 0; k < 200000; ++k) {
        uint64_t active = rng();
        if (rng() % 4 == 0) active = ~0ull;
        if (rng() % 4 == 0) active &= (rng() % 64 == 0) ? 1ull : (~0ull >> (rng() % 64));
        const uint64_t contrib = active & rng() & rng();
        const float x = sgn() * (float)U(rng) * std::ldexp(1.0f, (int)(rng() % 10) - 5);
        const float du = sgn() * std::ldexp((float)(1.0 + U(rng)), -(int)(rng() % 12));
        const float delta = sgn() * std::ldexp((float)(1.0 + U(rng)), -(int)(rng() % 12));
        float want = x;
        for (int l = 0; l < 64; ++l) {
            if (!((active >> l) & 1)) continue;
            want += du;
            if ((contrib >> l) & 1) want += delta;
        }
        const float got = replay_uniform(x, active, contrib, du, delta);
        ++cases;
        if (bits(want) != bits(got)) {
            std::printf("replay mismatch: x=%a active=%" PRIx64 " contrib=%" PRIx64 " du=%a delta=%a want=%a got=%a\n",
                        x, active, contrib, du, delta, want, got);
            return 1;
        }
    }
    std::printf("ok %" PRIu64 "\n", cases);
    return 0;
}
