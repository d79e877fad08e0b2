// fastdiv.hpp -- correctly rounded f32 division by a per-step constant.
//
// The walk divides a per-lane height difference a by a distance d that is
// the same for every lane of a step (kernel.rs:175).  With the step's
// reciprocal split in two terms,
//     y_hi = RN(1/d),   y_lo = RN((1 - d*y_hi) * y_hi)     (1 - d*y_hi exact),
// y_hi + y_lo = (1/d)(1 + e) with |e| <= 2^-47, so
//     q0 = RN(a*y_hi + RN(a*y_lo))
// lies within 0.5 ulp + 2^-46.4 |a/d| of a/d: within one ulp.  Then the
// residual r = a - d*q0 is exact in one FMA, and Markstein's theorem
// (y_hi within half an ulp of 1/d, q0 within one ulp of a/d) gives
//     RN(q0 + r*y_hi) = RN(a/d).
// Valid while no intermediate leaves the normal range: d in [2^-60, 2^60]
// and |a| in [2^-40, 2^40] or a = 0 (the host checks both before choosing
// this path).  Four FMA-pipe operations instead of the ten of the general
// IEEE sequence (scale, reciprocal, Newton steps, fixup).
#pragma once

#if defined(__HIPCC__)
#define TVS_HD __host__ __device__ __forceinline__
#else
#define TVS_HD inline
#endif

#include <cmath>

namespace tvs {

struct Reciprocal {
    float hi;
    float lo;
};

TVS_HD Reciprocal split_reciprocal(float d) {
    const float hi = 1.0f / d;
    const float e = std::fma(-hi, d, 1.0f);
    return Reciprocal{hi, e * hi};
}

TVS_HD float div_by_reciprocal(float a, float d, float y_hi, float y_lo) {
    const float q0 = std::fma(a, y_hi, a * y_lo);
    const float r = std::fma(-q0, d, a);
    return std::fma(r, y_hi, q0);
}

// Two operations: q0 of div_by_reciprocal without its residual step.  q0 is
// within 0.5 ulp + 2^-46.4 |a/d| of a/d, so it is the correctly rounded
// quotient unless a/d lies within ~2^-22 ulp of a rounding midpoint -- which
// does happen (e.g. a = -1463.65039, d = 126.479172) -- so the walk uses it
// for a sector only after the device has compared it with div_by_reciprocal
// for every height difference the DEM can produce at every regular step of
// that sector.
TVS_HD float div_short(float a, float y_hi, float y_lo) {
    return std::fma(a, y_hi, a * y_lo);
}

// Distances the fast divide is proven for.
TVS_HD bool regular_distance(float d) { return d >= 0x1p-60f && d <= 0x1p60f; }

}  // namespace tvs
