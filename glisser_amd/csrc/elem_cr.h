// Correctly rounded atan2, acos and asin in double precision (round to
// nearest) for the orbital elements of the track output.
//
// The reference's Cartesian-to-element conversion (src/convert.cpp:310-463,
// called by the track writer, src/data.cpp:770-812) uses glibc's atan2, acos,
// asin, sin and cos, which return the correctly rounded value in all but rare
// cases.  The conversion amplifies a last-bit difference: the argument of
// pericentre is acos of a quotient formed with the sine and cosine of the
// node longitude (itself an atan2), ill-conditioned where that quotient nears
// +-1, so one ulp of difference in those functions can change a float32 of
// the track record.  With these routines and sincos_cr the device's elements
// are the host's wherever glibc rounds correctly.
//
// Method: atan2(y, x) from t = min(|x|, |y|) / max(|x|, |y|) in [0, 1] as a
// double-double quotient (atan t, pi/2 - atan t, and pi minus either for
// x < 0); table-based reduction (J.-M. Muller, "Elementary Functions:
// Algorithms and Implementation", the table-based methods): c = k/32 nearest
// to t, atan t = atan c + atan u with u = (t - c) / (1 + t c), |u| <= 1/64,
// and atan u by its Taylor series through u^17 (truncation below 2^-108
// relative), everything in double-double (the error-free transformations of
// sincos_cr.h), so the unrounded result is within about 2^-100 relative of
// the exact value and its rounding to double is the correctly rounded result
// unless the exact value lies that close to a rounding boundary.
// acos x = atan2(sqrt((1 - x)(1 + x)), x), asin x = atan2(x, sqrt((1 - x)(1 +
// x))), with 1 - x and 1 + x exact pairs.  Zeros, infinities, NaN, |x| >= 1
// and operands whose quotient is below 2^-900 take the math library: there
// the result is a special value (+-0, +-pi, +-pi/2, +-pi/4, +-3pi/4 rounded,
// NaN), or +-pi rounded, or the quotient min/max itself, from which atan
// differs by a factor 1 - t^2/3 that is 1 to far below an ulp.
// The constants are atan(k/32), (-1)^n / (2n + 1), pi/2 and pi as pairs
// (rounded value, rounded remainder) from 300-bit arithmetic
// (gen_elem_cr_table.py).  FMA contraction is off: every fused operation is
// an explicit fma(), so the host twin rounds as the device does.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

#include "sincos_cr.h"

namespace glx {
namespace cr {

constexpr Pair kAtanTable[33] = {
    {0x0.0p+0, 0x0.0p+0},                             // atan(0/32)
    {0x1.ffd55bba97625p-6, -0x1.5ec431444912cp-60},   // atan(1/32)
    {0x1.ff55bb72cfdeap-5, -0x1.c934d86d23f1dp-60},   // atan(2/32)
    {0x1.7ee182602f10fp-4, -0x1.cfb654c0c3d98p-58},   // atan(3/32)
    {0x1.fd5ba9aac2f6ep-4, -0x1.cd37686760c17p-59},   // atan(4/32)
    {0x1.3d6eee8c6626cp-3, 0x1.61a3b0ce9281bp-57},    // atan(5/32)
    {0x1.7b97b4bce5b02p-3, 0x1.347b0b4f881cap-58},    // atan(6/32)
    {0x1.b90d7529260a2p-3, 0x1.17b10d2e0e5abp-61},    // atan(7/32)
    {0x1.f5b75f92c80ddp-3, 0x1.8ab6e3cf7afbdp-57},    // atan(8/32)
    {0x1.18bf5a30bf178p-2, 0x1.30ca4748b1bf9p-57},    // atan(9/32)
    {0x1.362773707ebccp-2, -0x1.963a544b672d8p-57},   // atan(10/32)
    {0x1.530ad9951cd4ap-2, -0x1.2566480884082p-57},   // atan(11/32)
    {0x1.6f61941e4def1p-2, -0x1.c63aae6f6e918p-56},   // atan(12/32)
    {0x1.8b24d394a1b25p-2, 0x1.b6d0ba3748fa8p-56},    // atan(13/32)
    {0x1.a64eec3cc23fdp-2, -0x1.24dec1b50b7ffp-56},   // atan(14/32)
    {0x1.c0db4c94ec9f0p-2, -0x1.cc1ce70934c34p-56},   // atan(15/32)
    {0x1.dac670561bb4fp-2, 0x1.a2b7f222f65e2p-56},    // atan(16/32)
    {0x1.f40dd0b541418p-2, -0x1.a3992dc382a23p-57},   // atan(17/32)
    {0x1.0657e94db30d0p-1, -0x1.d5b495f6349e6p-56},   // atan(18/32)
    {0x1.1255d9bfbd2a9p-1, -0x1.2bdaee1c0ee35p-58},   // atan(19/32)
    {0x1.1e00babdefeb4p-1, -0x1.928df287a668fp-58},   // atan(20/32)
    {0x1.2958e59308e31p-1, -0x1.09e73b0c6c087p-56},   // atan(21/32)
    {0x1.345f01cce37bbp-1, 0x1.1021137c71102p-55},    // atan(22/32)
    {0x1.3f13fb89e96f4p-1, 0x1.ecf8b492644f0p-56},    // atan(23/32)
    {0x1.4978fa3269ee1p-1, 0x1.2419a87f2a458p-56},    // atan(24/32)
    {0x1.538f57b89061fp-1, -0x1.1bb74abda520cp-55},   // atan(25/32)
    {0x1.5d58987169b18p-1, 0x1.0028e4bc5e7cap-57},    // atan(26/32)
    {0x1.66d663923e087p-1, -0x1.6ea6febe8bbbap-56},   // atan(27/32)
    {0x1.700a7c5784634p-1, -0x1.8c34d25aadef6p-56},   // atan(28/32)
    {0x1.78f6bbd5d315ep-1, 0x1.406a089803740p-55},    // atan(29/32)
    {0x1.819d0b7158a4dp-1, -0x1.bf76229d3b917p-56},   // atan(30/32)
    {0x1.89ff5ff57f1f8p-1, -0x1.55b9a5e177a1bp-55},   // atan(31/32)
    {0x1.921fb54442d18p-1, 0x1.1a62633145c07p-55},    // atan(32/32)
};
constexpr Pair kAtanSeries[9] = {
    {0x1p+0, 0x0p+0},
    {-0x1.5555555555555p-2, -0x1.5555555555555p-56},  // -1/3
    {0x1.999999999999ap-3, -0x1.999999999999ap-57},   // 1/5
    {-0x1.2492492492492p-3, -0x1.2492492492492p-57},  // -1/7
    {0x1.c71c71c71c71cp-4, 0x1.c71c71c71c71cp-58},    // 1/9
    {-0x1.745d1745d1746p-4, 0x1.745d1745d1746p-59},   // -1/11
    {0x1.3b13b13b13b14p-4, -0x1.3b13b13b13b14p-58},   // 1/13
    {-0x1.1111111111111p-4, -0x1.1111111111111p-60},  // -1/15
    {0x1.e1e1e1e1e1e1ep-5, 0x1.e1e1e1e1e1e1ep-61},    // 1/17
};
constexpr Pair kHalfPi = {0x1.921fb54442d18p+0, 0x1.1a62633145c07p-54};
constexpr Pair kPiPair = {0x1.921fb54442d18p+1, 0x1.1a62633145c07p-53};

// a + b without large cancellation (|a + b| >= |a|/2 or so), relative error ~2^-105.
__host__ __device__ inline Pair add_pp(Pair a, Pair b) {
#pragma clang fp contract(off)
    const Pair s = two_sum(a.hi, b.hi);
    return fast_two_sum(s.hi, s.lo + (a.lo + b.lo));
}

// a / b for b > 0, relative error ~2^-104: q1 = a.hi / b.hi, then the
// remainder a - q1 b (q1 b.hi split exactly by an fma; a.hi - q1 b.hi is
// exact by Sterbenz's lemma) divided once more.
__host__ __device__ inline Pair div_pp(Pair a, Pair b) {
#pragma clang fp contract(off)
    const double q1 = a.hi / b.hi;
    const double p = q1 * b.hi;
    const double pe = fma(q1, b.hi, -p);
    const double r = (((a.hi - p) - pe) + a.lo) - q1 * b.lo;
    return fast_two_sum(q1, r / b.hi);
}

// sqrt of w > 0, relative error ~2^-104: h = sqrt(w.hi), then one
// correction from the exact residual w.hi - h^2.
__host__ __device__ inline Pair sqrt_pp(Pair w) {
#pragma clang fp contract(off)
    const double h = sqrt(w.hi);
    const double e = fma(-h, h, w.hi);
    return fast_two_sum(h, (e + w.lo) / (2.0 * h));
}

// atan t for t in [0, 1] (t.hi <= 1 + 2^-52).
__host__ __device__ inline Pair atan_unit(Pair t) {
#pragma clang fp contract(off)
    // nearest k to 32 t (32 t.hi is exact, so |32 t.hi - k| <= 1/2 with no
    // intermediate rounding): |t - k/32| <= 1/64 and, for k >= 1, t >= k/64
    const int k = (int)rint(t.hi * 32.0);
    Pair u = t;
    if (k > 0) {
        const double c = k * 0x1p-5;
        // t - c: t.hi - c is exact (c/2 <= t.hi <= 2c), then t.lo added.
        const Pair num = two_sum(t.hi - c, t.lo);
        // 1 + t c: the product split exactly, then summed with 1.
        const double p = t.hi * c;
        const double pe = fma(t.hi, c, -p);
        const Pair d = two_sum(1.0, p);
        const Pair den = fast_two_sum(d.hi, d.lo + (pe + t.lo * c));
        u = div_pp(num, den);
    }
    const Pair z = mul_pp(u, u);
    Pair acc = kAtanSeries[8];
#pragma unroll
    for (int j = 7; j >= 0; j--) acc = horner_pp(kAtanSeries[j], z, acc);
    const Pair s = mul_pp(u, acc);
    return k > 0 ? add_pp(kAtanTable[k], s) : s;
}

// The angle in [0, pi] of the point (+-ax, ay) for ax, ay > 0 given as
// pairs; x_neg selects the left half-plane.
__host__ __device__ inline Pair angle_pp(Pair ay, Pair ax, bool x_neg) {
    Pair r;
    if (ay.hi <= ax.hi) {
        r = atan_unit(div_pp(ay, ax));
    } else {
        const Pair a = atan_unit(div_pp(ax, ay));
        r = add_pp(kHalfPi, {-a.hi, -a.lo});
    }
    if (x_neg) r = add_pp(kPiPair, {-r.hi, -r.lo});
    return r;
}

__host__ __device__ inline double atan2_cr(double y, double x) {
    const double ax = fabs(x), ay = fabs(y);
    // comparisons, not fmax/fmin, so that a NaN operand reaches the library
    const double big = ax > ay ? ax : ay, small = ax > ay ? ay : ax;
    if (!(small > 0.0) || !(big <= 0x1.fffffffffffffp+1023) || small < big * 0x1p-900) return atan2(y, x);
    // common power-of-two scale: the larger operand in [1, 2), the smaller
    // (at least 2^-900 of it) still normal, both exact
    int e;
    (void)frexp(big, &e);
    const Pair r = angle_pp({ldexp(ay, 1 - e), 0.0}, {ldexp(ax, 1 - e), 0.0}, x < 0.0);
    return y < 0.0 ? -r.hi : r.hi;
}

__host__ __device__ inline double acos_cr(double x) {
    const double ax = fabs(x);
    if (!(ax < 1.0)) return acos(x);  // +-1 (0, pi), |x| > 1 and NaN (NaN)
    const Pair w = mul_pp(two_sum(1.0, -ax), two_sum(1.0, ax));
    return angle_pp(sqrt_pp(w), {ax, 0.0}, x < 0.0).hi;
}

__host__ __device__ inline double asin_cr(double x) {
    const double ax = fabs(x);
    if (!(ax < 1.0) || x == 0.0) return asin(x);  // +-0, +-1 (+-pi/2), |x| > 1 and NaN
    const Pair w = mul_pp(two_sum(1.0, -ax), two_sum(1.0, ax));
    const double r = angle_pp({ax, 0.0}, sqrt_pp(w), false).hi;
    return x < 0.0 ? -r : r;
}

}  // namespace cr
}  // namespace glx
