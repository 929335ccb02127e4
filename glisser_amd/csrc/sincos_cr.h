// Correctly rounded sin and cos in double precision (round to nearest) for
// the reference-arithmetic (fp64) form of the Kepler drift.
//
// The reference CPU path calls glibc's sin/cos (src/wh.cpp:46-65, the guess
// at src/wh.cpp:928); on the arguments of the Kepler solve (|x| <~ 0.1)
// glibc returns the correctly rounded value except in a few calls per
// million.  A device routine that is correctly rounded therefore reproduces
// the CPU path's sin/cos in all but those calls, where the device math
// library differs far more often.
//
// Method: Ziv's strategy (A. Ziv, "Fast evaluation of elementary
// mathematical functions with correctly rounded last bit", ACM TOMS 17(3),
// 1991) with error-free transformations (Knuth's TwoSum, Dekker's
// Fast2Sum, the FMA two-product):
//   * quick phase, |x| <= 2^-2, no argument reduction: sin x = x + t and
//     cos x = 1 + u with the leading Taylor term of t (x^3 / 6) and of u
//     (x^2 / 2) exact in double-double and the rest in double; the sum is
//     held as an unevaluated pair (hi, lo) whose error bound D is proven
//     below; if hi + (lo - D) and hi + (lo + D) round to the same double,
//     that double is the correctly rounded result (rounding is monotonic);
//   * accurate phase (every other |x| < 2^20, and any quick-phase
//     ambiguity): x - k pi/2 in double-double (pi/2 split in three
//     doubles, k = nearest integer of 2x/pi), then the Taylor series of
//     sin and cos on |y| <= pi/4 with the first terms in double-double,
//     relative error below 2^-100;
//   * |x| >= 2^20, infinities and NaN: the device math library (never
//     reached by the Kepler solve, whose arguments are below 2 pi + 1).
// The coefficients are (-1)^k / n! rounded to double (pairs: the rounded
// value and the rounded remainder), computed for this file with 400-bit
// arithmetic.  FMA contraction is off: every fused operation is an explicit
// fma(), so the host twin (the same code compiled for the CPU) rounds
// exactly as the device does.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

namespace glx {
namespace cr {

struct Pair {
    double hi, lo;
};

// a + b = s + e exactly (Knuth's TwoSum).
__host__ __device__ inline Pair two_sum(double a, double b) {
#pragma clang fp contract(off)
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

// a + b = s + e exactly when |a| >= |b| or a == 0 (Dekker's Fast2Sum).
__host__ __device__ inline Pair fast_two_sum(double a, double b) {
#pragma clang fp contract(off)
    const double s = a + b;
    const double e = b - (s - a);
    return {s, e};
}

// (a.hi + a.lo) * (b.hi + b.lo), relative error ~2^-104 (the lo*lo term
// and the rounding of the cross terms).
__host__ __device__ inline Pair mul_pp(Pair a, Pair b) {
#pragma clang fp contract(off)
    const double p = a.hi * b.hi;
    const double e = fma(a.hi, b.hi, -p);
    const double c = fma(a.hi, b.lo, a.lo * b.hi);
    return fast_two_sum(p, e + c);
}

// c + z * acc for |c| >= |z * acc| (one double-double Horner step).
__host__ __device__ inline Pair horner_pp(Pair c, Pair z, Pair acc) {
#pragma clang fp contract(off)
    const Pair p = mul_pp(z, acc);
    const Pair s = fast_two_sum(c.hi, p.hi);
    return fast_two_sum(s.hi, s.lo + (c.lo + p.lo));
}

// Taylor coefficients: sin (odd powers 3..27), cos (even powers 2..28);
// pairs for the terms whose double rounding would matter in the accurate
// phase.
constexpr double kS3h = -0x1.5555555555555p-3, kS3l = -0x1.5555555555555p-57;
constexpr double kS5h = 0x1.1111111111111p-7, kS5l = 0x1.1111111111111p-63;
constexpr double kS7h = -0x1.a01a01a01a01ap-13, kS7l = -0x1.a01a01a01a01ap-73;
constexpr double kS9h = 0x1.71de3a556c734p-19, kS9l = -0x1.c154f8ddc6c00p-73;
constexpr double kS11h = -0x1.ae64567f544e4p-26, kS11l = 0x1.c062e06d1f209p-80;
constexpr double kS13h = 0x1.6124613a86d09p-33, kS13l = 0x1.f28e0cc748ebep-87;
constexpr double kS15h = -0x1.ae7f3e733b81fp-41, kS15l = -0x1.1d8656b0ee8cbp-97;
constexpr double kS17 = 0x1.952c77030ad4ap-49;
constexpr double kS19 = -0x1.2f49b46814157p-57;
constexpr double kS21 = 0x1.71b8ef6dcf572p-66;
constexpr double kS23 = -0x1.761b41316381ap-75;
constexpr double kS25 = 0x1.3f3ccdd165fa9p-84;
constexpr double kS27 = -0x1.d1ab1c2dccea3p-94;

constexpr double kC4h = 0x1.5555555555555p-5, kC4l = 0x1.5555555555555p-59;
constexpr double kC6h = -0x1.6c16c16c16c17p-10, kC6l = 0x1.f49f49f49f49fp-65;
constexpr double kC8h = 0x1.a01a01a01a01ap-16, kC8l = 0x1.a01a01a01a01ap-76;
constexpr double kC10h = -0x1.27e4fb7789f5cp-22, kC10l = -0x1.cbbc05b4fa99ap-76;
constexpr double kC12h = 0x1.1eed8eff8d898p-29, kC12l = -0x1.2aec959e14c06p-83;
constexpr double kC14h = -0x1.93974a8c07c9dp-37, kC14l = -0x1.05d6f8a2efd1fp-92;
constexpr double kC16h = 0x1.ae7f3e733b81fp-45, kC16l = 0x1.1d8656b0ee8cbp-101;
constexpr double kC18 = -0x1.6827863b97d97p-53;
constexpr double kC20 = 0x1.e542ba4020225p-62;
constexpr double kC22 = -0x1.0ce396db7f853p-70;
constexpr double kC24 = 0x1.f2cf01972f578p-80;
constexpr double kC26 = -0x1.88e85fc6a4e5ap-89;
constexpr double kC28 = 0x1.0a18a2635085dp-98;

// pi/2 = kP1 + kP2 + kP3 + O(2^-160); 2/pi rounded.
constexpr double kP1 = 0x1.921fb54442d18p+0;
constexpr double kP2 = 0x1.1a62633145c07p-54;
constexpr double kP3 = -0x1.f1976b7ed8fbcp-110;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;

// Quick-phase bounds and error bounds (see sincos_quick): three tiers of
// polynomial length, |x| <= kQuickShort, <= kQuickTight (both with the tight
// error bounds) and <= kQuickMax (one more term, looser bounds).
constexpr double kQuickMax = 0x1p-2;
constexpr double kQuickTight = 0x1p-4;
constexpr double kQuickShort = 0x1p-5;
constexpr double kQuickSinErr = 0x1p-70;  // relative to |x|
constexpr double kQuickCosErr = 0x1p-69;  // absolute (cos x ~ 1)
constexpr double kWideSinErr = 0x1p-64;   // the same for 2^-4 < |x| <= 2^-2
constexpr double kWideCosErr = 0x1p-62;
// Rounding-test factors of the quick phase for those bounds (sincos_quick).
constexpr double kZivSin = 1.0 + 0x1p-15;
constexpr double kZivCos = 1.0 + 0x1p-14;
constexpr double kZivWideSin = 1.0 + 0x1p-9;
constexpr double kZivWideCos = 1.0 + 0x1p-7;
// 2^54 eps < 1 - 1 / z, with eps the error bound relative to the result
// (|sin x| >= 0.9896 |x|, cos x >= 0.9689 for |x| <= 2^-2)
static_assert(0x1p54 * kQuickSinErr / 0.9896 < 1.0 - 1.0 / kZivSin, "sin rounding test");
static_assert(0x1p54 * kQuickCosErr / 0.9689 < 1.0 - 1.0 / kZivCos, "cos rounding test");
static_assert(0x1p54 * kWideSinErr / 0.9896 < 1.0 - 1.0 / kZivWideSin, "wide sin rounding test");
static_assert(0x1p54 * kWideCosErr / 0.9689 < 1.0 - 1.0 / kZivWideCos, "wide cos rounding test");
// Accurate-phase error bounds: relative to the result (series and pair
// arithmetic) plus the absolute error of the argument reduction.
constexpr double kAccErr = 0x1p-98;
constexpr double kRedErr = 0x1p-102;

// Ziv's rounding test in its one-operation form: when the exact value y lies
// within eps |hi| of hi + lo, RN(hi + lo z) == hi for z = 1 + 2^-k with
// 2^54 eps < 1 - 1 / z proves RN(y) == hi.  (hi + lo z then lies in hi's
// rounding interval, whose half-widths are at least 2^-54 |hi|, so |lo| is
// at most that half-width / z, and |y - hi| <= half-width (1 / z + 2^54
// eps) < half-width.)  Needs no normalisation of the pair.
__host__ __device__ inline bool rounds_to_hi(double hi, double lo, double z) {
#pragma clang fp contract(off)
    return fma(lo, z, hi) == hi;
}

// RN(hi + lo) when the exact value lies within err of hi + lo and both ends
// of that interval round alike; otherwise false.  hi + lo is normalised
// (|lo| <= ulp(hi)), err > 0.
__host__ __device__ inline bool round_decided(double hi, double lo, double err, double& out) {
#pragma clang fp contract(off)
    const double up = hi + (lo + err);
    const double dn = hi + (lo - err);
    out = up;
    return up == dn;
}

// Quick phase for |x| <= 2^-2 (x not tiny).  Error analysis for |x| <= 2^-4, sin:
// t = x^3 s3 + x^5 P(x^2), the first term as a pair from the exact x^2
// (x2 + x2l), the exact-enough x^3 (x3 + x3l) and s3 as a pair (errors
// ~2^-104 |x^3|); P(z) = s5 + z (s7 + z (s9 + z (s11 + z s13))) in double,
// relative error < 6 * 2^-53 with x^5 = x3 * x2 (< 3 * 2^-53), so
// |err| < 9 * 2^-53 * |x|^5 / 119 < 2^-72.6 |x| at |x| = 2^-4; the dropped
// x^15 / 15! term is < 2^-95 |x|; the final pair additions are exact (Fast2Sum)
// or round at 2^-53 of quantities below ulp(x).  Total < 2^-72 |x|, and
// kQuickSinErr = 2^-70 |x| is used.  cos: u = -x^2/2 (exact pair) + x^4
// C(x^2), C = c4 + z (c6 + z (c8 + z (c10 + z c12))) in double: |err| <
// 6 * 2^-53 * x^4 / 24 < 2^-71 at |x| = 2^-4, dropped x^14 / 14! < 2^-100;
// kQuickCosErr = 2^-69.  For |x| <= 2^-5 the polynomials stop at s9 and c8:
// the dropped sin terms are below x^11 / 11! < 2^-75 |x| and the dropped cos
// terms below x^10 / 10! < 2^-71.8, the rounding errors shrink with x, so
// the same two bounds hold.  For 2^-4 < |x| <= 2^-2 the polynomials run to
// s15 and c14: the same sums give |err| < 9 * 2^-53 * |x|^5 / 119 <
// 2^-64.7 |x| for sin (dropped x^17 / 17! < 2^-80 |x|) and 7 * 2^-53 * x^4 /
// 24 < 2^-62.8 for cos (dropped x^16 / 16! < 2^-76), so kWideSinErr =
// 2^-64 |x| and kWideCosErr = 2^-62.  (A decided result is the correctly
// rounded value whichever polynomial produced it, so the choice never
// changes the returned bits.)  Rounding test (rounds_to_hi): relative to
// the result, |sin x| >= 0.9896 |x| and cos x >= 0.9689 on |x| <= 2^-2, so
// eps <= 2^-69.98 (sin) and 2^-68.95 (cos), or 2^-63.98 and 2^-61.95 in the
// wide tier; kZiv* = 1 + 2^-15, 1 + 2^-14, 1 + 2^-9 and 1 + 2^-7 satisfy
// 2^54 eps < 1 - 1 / z, and leave about 2^-15, 2^-14, 2^-9 and 2^-7 of the
// calls undecided.  Returns false when either result is undecided.
// The short tier's leading polynomial coefficients and rounding-test
// factors, passed by value so a kernel can hold them in vector registers: a
// fused multiply-add with two constant operands (or with a literal, which a
// three-operand FP64 instruction cannot encode on this target) becomes a
// register copy plus the two-operand form, whose accumulator is the
// constant.
struct QuickLead {
    double s9, c8, zs, zc;
};
constexpr QuickLead kQuickLead = {kS9h, kC8h, kZivSin, kZivCos};

// fma(a, b, c), on the device as one three-operand v_fma_f64 with every
// operand in a vector register: where the addend is a constant the compiler
// otherwise copies it into the destination for the two-operand form.
__host__ __device__ inline double fma3(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r;
    __asm__("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
#else
    return fma(a, b, c);
#endif
}

// One tier of the quick phase (0: |x| <= 2^-5, 1: |x| <= 2^-4, 2: |x| <=
// 2^-2), evaluated straight through: each tier is its own code path, so its
// polynomial coefficients and rounding-test factors stay constants (tier 0:
// those of L, the values of kQuickLead).
template <int kTier>
__host__ __device__ inline bool sincos_quick_tier(double x, double& s, double& c, const QuickLead& L) {
#pragma clang fp contract(off)
    const double x2 = x * x;
    const double x2l = fma(x, x, -x2);
    // sin
    const double x3 = x * x2;
    const double x3l = fma(x, x2, -x3) + x * x2l;
    const double t1 = kS3h * x3;
    const double t1l = fma(kS3h, x3, -t1) + fma(kS3h, x3l, kS3l * x3);
    double ps, pc;
    if constexpr (kTier == 0) {
        ps = fma3(fma3(x2, L.s9, kS7h), x2, kS5h);
        pc = fma3(fma3(x2, L.c8, kC6h), x2, kC4h);
    } else if constexpr (kTier == 1) {
        ps = fma(x2, kS13h, kS11h);
        ps = fma(ps, x2, kS9h);
        ps = fma(ps, x2, kS7h);
        ps = fma(ps, x2, kS5h);
        pc = fma(x2, kC12h, kC10h);
        pc = fma(pc, x2, kC8h);
        pc = fma(pc, x2, kC6h);
        pc = fma(pc, x2, kC4h);
    } else {
        ps = fma(x2, kS15h, kS13h);
        ps = fma(ps, x2, kS11h);
        ps = fma(ps, x2, kS9h);
        ps = fma(ps, x2, kS7h);
        ps = fma(ps, x2, kS5h);
        pc = fma(x2, kC14h, kC12h);
        pc = fma(pc, x2, kC10h);
        pc = fma(pc, x2, kC8h);
        pc = fma(pc, x2, kC6h);
        pc = fma(pc, x2, kC4h);
    }
    const double zs = kTier == 2 ? kZivWideSin : (kTier == 1 ? kZivSin : L.zs);
    const double zc = kTier == 2 ? kZivWideCos : (kTier == 1 ? kZivCos : L.zc);
    const double t2 = (x3 * x2) * ps;
    const Pair th = fast_two_sum(t1, t2);
    const Pair sh = fast_two_sum(x, th.hi);
    const double slo = sh.lo + (th.lo + t1l);
    // cos
    const double h = -0.5 * x2, hl = -0.5 * x2l;
    const double u2 = (x2 * x2) * pc;
    const Pair uh = fast_two_sum(h, u2);
    const Pair ch = fast_two_sum(1.0, uh.hi);
    const double clo = ch.lo + (uh.lo + hl);
    s = sh.hi;
    c = ch.hi;
    // both tests evaluated, no branch between them
    return (static_cast<unsigned>(rounds_to_hi(sh.hi, slo, zs)) & static_cast<unsigned>(rounds_to_hi(ch.hi, clo, zc))) != 0u;
}

// The quick phase for 2^-27 <= |x| <= 2^-2 (the tiers above).
__host__ __device__ inline bool sincos_quick(double x, double& s, double& c, const QuickLead& L = kQuickLead) {
    const double ax = fabs(x);
    if (ax <= kQuickShort) return sincos_quick_tier<0>(x, s, c, L);
    if (ax <= kQuickTight) return sincos_quick_tier<1>(x, s, c, L);
    return sincos_quick_tier<2>(x, s, c, L);
}

// Accurate phase for |x| < 2^20: reduction to y = x - k pi/2 as a pair,
// then sin y = y (1 + z S(z)) and cos y = 1 + z C(z), z = y^2, with the
// coefficients up to x^15 / x^16 as pairs.  Returns false if the rounding
// is still undecided at kAccErr (probability ~2^-45 per call); s and c then
// hold the leading doubles of the pairs.
__host__ __device__ inline bool sincos_accurate(double x, double& s, double& c) {
#pragma clang fp contract(off)
    const double k = rint(x * kTwoOverPi);
    // y = x - k (P1 + P2 + P3): the products k P1 and k P2 are split exactly
    // (p + pl), the three largest terms are added with exact sums, and the
    // remaining terms (each below 2^-52 for |y| < 1) are summed in double:
    // absolute error below 2^-103 (plus k 2^-160 for the truncation of
    // pi/2), which the rounding test below adds to its bound.
    const double p1 = k * kP1;
    const double p1l = fma(k, kP1, -p1);
    const Pair a = two_sum(x, -p1);
    const double p2 = k * kP2;
    const double p2l = fma(k, kP2, -p2);
    const Pair b = two_sum(a.hi, -p2);
    const Pair d = two_sum(b.hi, -p1l);
    const double rest = ((d.lo + b.lo) + a.lo) - (p2l + k * kP3);
    const Pair y = fast_two_sum(d.hi, rest);
    // z = y^2 as a pair
    const double zh = y.hi * y.hi;
    const double zl = fma(y.hi, y.hi, -zh) + 2.0 * y.hi * y.lo;
    const Pair z = fast_two_sum(zh, zl);
    // sin: S(z) = s3 + z s5 + ... ; tail from s17 in double
    double ts = fma(z.hi, kS27, kS25);
    ts = fma(ts, z.hi, kS23);
    ts = fma(ts, z.hi, kS21);
    ts = fma(ts, z.hi, kS19);
    ts = fma(ts, z.hi, kS17);
    Pair as = {ts, 0.0};
    as = horner_pp({kS15h, kS15l}, z, as);
    as = horner_pp({kS13h, kS13l}, z, as);
    as = horner_pp({kS11h, kS11l}, z, as);
    as = horner_pp({kS9h, kS9l}, z, as);
    as = horner_pp({kS7h, kS7l}, z, as);
    as = horner_pp({kS5h, kS5l}, z, as);
    as = horner_pp({kS3h, kS3l}, z, as);
    const Pair ps = mul_pp(z, as);       // z S(z), |.| < 0.11
    const Pair ys = mul_pp(y, ps);       // y z S(z)
    const Pair s0 = fast_two_sum(y.hi, ys.hi);
    const Pair sn = fast_two_sum(s0.hi, s0.lo + (y.lo + ys.lo));
    // cos: C(z) = -1/2 + z c4 + ... ; tail from c18 in double
    double tc = fma(z.hi, kC28, kC26);
    tc = fma(tc, z.hi, kC24);
    tc = fma(tc, z.hi, kC22);
    tc = fma(tc, z.hi, kC20);
    tc = fma(tc, z.hi, kC18);
    Pair ac = {tc, 0.0};
    ac = horner_pp({kC16h, kC16l}, z, ac);
    ac = horner_pp({kC14h, kC14l}, z, ac);
    ac = horner_pp({kC12h, kC12l}, z, ac);
    ac = horner_pp({kC10h, kC10l}, z, ac);
    ac = horner_pp({kC8h, kC8l}, z, ac);
    ac = horner_pp({kC6h, kC6l}, z, ac);
    ac = horner_pp({kC4h, kC4l}, z, ac);
    ac = horner_pp({-0.5, 0.0}, z, ac);
    const Pair pc = mul_pp(z, ac);       // z C(z), |.| < 0.3
    const Pair c0 = fast_two_sum(1.0, pc.hi);
    const Pair cs = fast_two_sum(c0.hi, c0.lo + pc.lo);
    // quadrant: (sin, cos) of x from those of y
    const int q = (int)((int64_t)k & 3);
    Pair rs = sn, rc = cs;
    if (q & 1) {
        rs = cs;
        rc = {-sn.hi, -sn.lo};
    }
    if (q & 2) {
        rs = {-rs.hi, -rs.lo};
        rc = {-rc.hi, -rc.lo};
    }
    const bool ok_s = round_decided(rs.hi, rs.lo, fma(kAccErr, fabs(rs.hi), kRedErr), s);
    const bool ok_c = round_decided(rc.hi, rc.lo, fma(kAccErr, fabs(rc.hi), kRedErr), c);
    if (!ok_s) s = rs.hi;
    if (!ok_c) c = rc.hi;
    return ok_s && ok_c;
}

// Everything but the quick phase, as a call: {sin x, cos x}.  Kept out of
// line so the rarely taken accurate phase (a few hundred operations and its
// coefficients) does not occupy the caller's registers at every call site.
__host__ __device__ __attribute__((noinline)) inline Pair sincos_slow(double x) {
    double s, c;
    if (fabs(x) < 0x1p-27) {
        // sin x rounds to x and cos x to 1
        s = x;
        c = 1.0;
    } else if (fabs(x) < 0x1p20) {
        sincos_accurate(x, s, c);
    } else {
        s = sin(x);
        c = cos(x);
    }
    return {s, c};
}

// sin x and cos x, correctly rounded for |x| < 2^20 (except the undecided
// accurate-phase cases above, never met in the tests).  Below 2^-27 sin x
// rounds to x and cos x to 1 (in sincos_slow: kept off the common path, so
// its results are not set up at every call).
// L: kQuickLead, or a copy of it held in registers.
__host__ __device__ inline void sincos_cr(double x, double* s, double* c, const QuickLead& L = kQuickLead) {
    const double ax = fabs(x);
    if ((ax >= 0x1p-27) & (ax <= kQuickMax) && sincos_quick(x, *s, *c, L)) return;
    const Pair r = sincos_slow(x);
    *s = r.hi;
    *c = r.lo;
}

}  // namespace cr
}  // namespace glx
