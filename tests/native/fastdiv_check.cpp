// Host build of the engine's fast divide (csrc/fastdiv.hpp) checked against
// IEEE f32 division.  Compiled by the test with -ffp-contract=off.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../total-viewsheds_amd/csrc/fastdiv.hpp"

namespace {
uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
float bits(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
bool same(float x, float y) { return std::memcmp(&x, &y, 4) == 0; }
}  // namespace

extern "C" {

// Random a with |a| in [2^-40, 2^40] (and zeros), d in [2^-60, 2^60].
uint64_t fastdiv_random(uint64_t n, uint64_t seed) {
    uint64_t bad = 0, s = seed;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t r = splitmix(s);
        const uint32_t ea = 127 - 40 + static_cast<uint32_t>((r >> 32) % 81);
        const uint32_t ed = 127 - 60 + static_cast<uint32_t>((r >> 40) % 121);
        const uint32_t ma = static_cast<uint32_t>(r) & 0x7fffffu;
        const uint32_t md = static_cast<uint32_t>(r >> 9) & 0x7fffffu;
        float a = bits((ea << 23) | ma);
        if ((r >> 62) & 1u) a = -a;
        if ((r & 0xffff) == 7) a = 0.0f;
        float d = bits((ed << 23) | md);
        if (d < 0x1p-60f || d > 0x1p60f) continue;
        const tvs::Reciprocal y = tvs::split_reciprocal(d);
        if (!same(tvs::div_by_reciprocal(a, d, y.hi, y.lo), a / d)) ++bad;
    }
    return bad;
}

// The walk's own domain: a = fl32(h - fl32(hp + obs)) for integer heights,
// d from a given list (e.g. every band distance of a workload).
uint64_t fastdiv_heights(int32_t h_lo, int32_t h_hi, int32_t hp_lo, int32_t hp_hi,
                         int32_t hp_step, float obs, const float *d, uint64_t nd) {
    uint64_t bad = 0;
    for (uint64_t j = 0; j < nd; ++j) {
        const float dj = d[j];
        if (!tvs::regular_distance(dj)) continue;
        const tvs::Reciprocal y = tvs::split_reciprocal(dj);
        for (int32_t hp = hp_lo; hp <= hp_hi; hp += hp_step) {
            const float h0 = static_cast<float>(hp) + obs;
            for (int32_t h = h_lo; h <= h_hi; ++h) {
                const float a = static_cast<float>(h) - h0;
                if (!same(tvs::div_by_reciprocal(a, dj, y.hi, y.lo), a / dj)) ++bad;
            }
        }
    }
    return bad;
}

// Hard quotients: a/d within 2^-12 ulp of a rounding midpoint.  m is the
// midpoint above a random q; p = d*m is exact in f64 (24 + 25 bits); keep
// the cases where p lies within 2^-12 f32-ulp of its f32 rounding a, so
// a/d is that close to m and only a correct residual step rounds it right.
uint64_t fastdiv_hard(uint64_t want, uint64_t seed, uint64_t *found) {
    uint64_t bad = 0, s = seed, got = 0;
    for (uint64_t tries = 0; got < want && tries < want * 200000ull; ++tries) {
        const uint64_t r = splitmix(s);
        const uint32_t ed = 127 - 30 + static_cast<uint32_t>((r >> 32) % 61);
        const float d = bits((ed << 23) | (static_cast<uint32_t>(r) & 0x7fffffu));
        const uint32_t eq = 127 - 30 + static_cast<uint32_t>((r >> 40) % 61);
        const float q = bits((eq << 23) | (static_cast<uint32_t>(r >> 9) & 0x7fffffu));
        const double ulp = std::ldexp(1.0, static_cast<int>(eq) - 127 - 23);
        const double m = static_cast<double>(q) + 0.5 * ulp;
        const double p = static_cast<double>(d) * m;
        const float a = static_cast<float>(p);
        if (!std::isfinite(a) || a == 0.0f) continue;
        int ea;
        std::frexp(a, &ea);
        const double ulp_a = std::ldexp(1.0, ea - 24);
        if (std::fabs(static_cast<double>(a) - p) > ulp_a * 0x1p-12) continue;
        if (std::fabs(a) < 0x1p-40f || std::fabs(a) > 0x1p40f || !tvs::regular_distance(d))
            continue;
        ++got;
        const tvs::Reciprocal y = tvs::split_reciprocal(d);
        if (!same(tvs::div_by_reciprocal(a, d, y.hi, y.lo), a / d)) ++bad;
        if (!same(tvs::div_by_reciprocal(-a, d, y.hi, y.lo), -a / d)) ++bad;
    }
    *found = got;
    return bad;
}

// Quotients within a few 2^-24 ulp of a midpoint, built exactly: for an odd
// 24-bit significand D and a small odd r, the odd N with D*N = r (mod 2^k)
// makes A = (D*N - r) / 2^k an integer, and A/D = N/2^k - r/(2^k D): the
// midpoint N/2^k of two floats, missed by |r| / D half-ulps.
uint64_t fastdiv_midpoints(uint64_t want, uint64_t seed, uint64_t *found) {
    uint64_t bad = 0, s = seed, got = 0;
    for (uint64_t tries = 0; got < want && tries < want * 64ull; ++tries) {
        const uint64_t rnd = splitmix(s);
        const uint64_t D = (rnd & 0x7fffffull) | 0x800001ull;          // odd, 24 bits
        const int k = ((rnd >> 23) & 1u) ? 25 : 24;
        int64_t r = static_cast<int64_t>((rnd >> 24) % 64) * 2 + 1;    // odd
        if ((rnd >> 31) & 1u) r = -r;
        uint64_t inv = D;                                              // D^-1 mod 2^64
        for (int it = 0; it < 6; ++it) inv *= 2 - D * inv;
        const uint64_t mask = (1ull << k) - 1;
        uint64_t N = (inv * static_cast<uint64_t>(r)) & mask;
        if (k == 24) N += 1ull << 24;
        if (N < (1ull << 24) || N >= (1ull << 25)) continue;
        const uint64_t P = D * N;
        if (((P - static_cast<uint64_t>(r)) & mask) != 0) continue;
        const uint64_t A = (P - static_cast<uint64_t>(r)) >> k;
        if (A < (1ull << 23) || A >= (1ull << 24)) continue;
        const int ea = -23 + static_cast<int>((rnd >> 40) % 41) - 20;
        const int ed = -23 + static_cast<int>((rnd >> 48) % 41) - 20;
        const float a = std::ldexp(static_cast<float>(A), ea);
        const float d = std::ldexp(static_cast<float>(D), ed);
        ++got;
        const tvs::Reciprocal y = tvs::split_reciprocal(d);
        if (!same(tvs::div_by_reciprocal(a, d, y.hi, y.lo), a / d)) ++bad;
        if (!same(tvs::div_by_reciprocal(-a, d, y.hi, y.lo), -a / d)) ++bad;
    }
    *found = got;
    return bad;
}

// One pair through all three divides: out[0] = IEEE, out[1] = split
// reciprocal, out[2] = short divide.
void fastdiv_pair(uint32_t a_bits, uint32_t d_bits, float *out) {
    const float a = bits(a_bits), d = bits(d_bits);
    const tvs::Reciprocal y = tvs::split_reciprocal(d);
    out[0] = a / d;
    out[1] = tvs::div_by_reciprocal(a, d, y.hi, y.lo);
    out[2] = tvs::div_short(a, y.hi, y.lo);
}

// Hard divisors: significands near 1 and near 2 (all ones), powers of two.
uint64_t fastdiv_edges(uint64_t n, uint64_t seed) {
    uint64_t bad = 0, s = seed;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t r = splitmix(s);
        const uint32_t ed = 127 - 20 + static_cast<uint32_t>((r >> 32) % 41);
        const uint32_t tail = static_cast<uint32_t>(r >> 50) & 0xfffu;
        const uint32_t md = ((r >> 49) & 1u) ? (0x7fffffu - tail) : tail;
        const float d = bits((ed << 23) | md);
        const uint32_t ea = 127 - 20 + static_cast<uint32_t>((r >> 8) % 41);
        const float a = bits((ea << 23) | (static_cast<uint32_t>(r) & 0x7fffffu));
        const tvs::Reciprocal y = tvs::split_reciprocal(d);
        if (!same(tvs::div_by_reciprocal(a, d, y.hi, y.lo), a / d)) ++bad;
        if (!same(tvs::div_by_reciprocal(-a, d, y.hi, y.lo), -a / d)) ++bad;
    }
    return bad;
}
}
