// float_repeat.h - n successive float additions of one value, in a number of operations that grows with the
// binades the running sum crosses instead of with n. Host and device.
//
// The synaptic-resource rule sums a neuron's free resource synapse by synapse in float (blifat_stdp.h:105-131:
// free += d_u for every synapse of a spiking neuron; blifat_stdp.h:133-156: free -= change for every synapse
// that takes dopamine), so its float result depends on the order and number of the additions and a sum of the
// terms computed any other way differs in its last bits. When the added value is the same for a run of
// synapses (one d_u for the whole projection, one change for a dopamine step), the run has a closed form:
//
//   Inside one binade [2^E, 2^(E+1)) every float is an integer multiple A*u of u = 2^(E-23), and
//   RN(A*u + c) = (A + D)*u, with D = round(c/u) to nearest, ties to the even result. The rounding
//   direction depends only on the fractional part of c/u, and for a tie on the parity of A + floor(c/u):
//   after one addition rounded in the binade A is even when a tie occurs, so from the second addition on,
//   every addition inside the binade adds the same D. Three plain additions whose results lie strictly inside
//   one binade (at least u from both ends) establish D (the last two results' difference, exact by Sterbenz);
//   the following m additions then give A + m*D for as long as every result stays at least u inside the
//   binade (the exact sum is within u/2 of its rounded result, so it lies inside the binade too and was
//   rounded with granularity u). Outside those conditions (signs, zero, subnormals, a binade boundary) the
//   additions are done one by one.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace knp_mi355x {

// x after `n` successive float additions x += c, each rounded to nearest even, bit for bit.
__host__ __device__ inline float repeat_add(float x, float c, uint64_t n) {
    const float ac = fabsf(c);
    if (!(ac >= 0x1p-100f && ac <= 0x1p100f)) {  // zero, tiny, huge or NaN increments: one by one
        for (; n; --n) x += c;
        return x;
    }
    while (n >= 3) {
        const float x1 = x + c, x2 = x1 + c, x3 = x2 + c;
        n -= 3;
        x = x3;
        if (!(fabsf(x3) <= 0x1.fffffep127f)) return x3;  // an infinity stays (finite c added); NaN stays NaN
        if (x3 == x2) return x3;       // x2 + c rounds back to x2 (non-zero): every later addition does too
        if ((copysignf(1.0f, x1) != copysignf(1.0f, x3)) || (copysignf(1.0f, x2) != copysignf(1.0f, x3))) continue;
        const float a1 = fabsf(x1), a2 = fabsf(x2), a3 = fabsf(x3);
        if (!(a1 >= 0x1p-120f && a2 >= 0x1p-120f && a3 >= 0x1p-120f)) continue;  // normal, away from subnormals
        int e;
        (void)frexpf(a3, &e);  // a3 in [2^(e-1), 2^e); u = 2^(e-24)
        const float u = ldexpf(1.0f, e - 24);
        const float lo = ldexpf(1.0f, e - 1) + u;  // the binade's second float
        const float hi = ldexpf(1.0f, e) - u;      // and its last
        if (!(a1 >= lo && a1 <= hi && a2 >= lo && a2 <= hi && a3 >= lo && a3 <= hi)) continue;
        const int32_t A = (int32_t)ldexpf(a3, 24 - e);           // significand of x3, in [2^23 + 1, 2^24 - 1]
        const int32_t D = (int32_t)ldexpf(a3 - a2, 24 - e);      // step in units of u (|x| grows if positive)
        // further additions while A + m*D stays in [2^23 + 1, 2^24 - 1]
        const uint32_t room = D > 0 ? (uint32_t)((1 << 24) - 1 - A) : (uint32_t)(A - (1 << 23) - 1);
        const uint32_t dd = (uint32_t)(D > 0 ? D : -D);
        uint64_t m = room / dd;
        if (m > n) m = n;
        if (m == 0) continue;
        const int32_t Am = A + (int32_t)m * D;
        x = copysignf(ldexpf((float)Am, e - 24), x3);
        n -= m;
    }
    for (; n; --n) x += c;
    return x;
}

// The free-resource additions of one 64-synapse chunk whose synapses all carry the same non-zero d_u: for each
// active lane in ascending order, x += du, and for each contributing lane (a subset) x += delta after it.
__host__ __device__ inline float replay_uniform(float x, uint64_t active, uint64_t contrib, float du, float delta) {
    uint64_t rest = active;
    while (contrib) {
        const int l = __builtin_ctzll(contrib);
        contrib &= contrib - 1;
        const uint64_t upto = l == 63 ? ~0ull : ((2ull << l) - 1);  // lanes 0..l
        x = repeat_add(x, du, (uint64_t)__builtin_popcountll(rest & upto));
        rest &= ~upto;
        x += delta;
    }
    return repeat_add(x, du, (uint64_t)__builtin_popcountll(rest));
}

}  // namespace knp_mi355x
