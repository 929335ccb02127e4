// FAST and MIXED forms of the particle step.
//
// Same map as the reference (kick-drift-kick, eccentric-anomaly-difference
// Kepler solve, direct planet sum + indirect term, identical flag logic),
// restructured for the gfx950 FP64 VALU:
//   * one reciprocal square root of |r|^2 and of 1/a replace the drift's
//     divisions and square roots; 1/a = -2E/mu needs no division at all;
//   * reciprocals / reciprocal square roots start from v_rcp_f64 /
//     v_rsq_f64 (relative error < 6e-8) and take two FMA Newton steps where
//     the result feeds the state, none where it only scales a Newton
//     correction;
//   * sin/cos of the mean-anomaly step come from a five-term Taylor
//     polynomial (|dM| <= 0.1); the first guess and each Newton correction
//     rotate (sin E, cos E - 1) by the small angle with a short series
//     instead of re-evaluating sin/cos; cos E - 1 and E - dM are carried
//     explicitly, so f, g and the Kepler residual avoid cancellation;
//   * the Kepler solve starts from one Newton step off E = dM (instead of
//     the reference's series guess) and takes exactly two more corrections,
//     accepted when the last is below 1e-14 (the convergence test of the
//     reference's early-exit mode, src/wh.cu:60-71);
//   * every shortcut is checked per lane; the few lanes outside them
//     (large steps, slow convergence, unbound orbits) redo the step with the
//     reference-arithmetic drift, so results and flags stay defined for any
//     input;
//   * FMA contraction is allowed.
// MIXED additionally evaluates the r^-3 of every planet but the heaviest
// with the FP32 reciprocal square root (|dr|^2 and the encounter test stay
// FP64) and accumulates the acceleration in FP64.
#pragma once

#include "wh_device.h"

namespace glx {

// Reciprocal of x (x normal, positive or negative), two Newton steps from
// the hardware estimate: faithful to about one ulp.
__device__ __forceinline__ double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// Reciprocal with one Newton step (relative error below 3e-15).
__device__ __forceinline__ double rcp_nr1(double x) {
    const double y = __builtin_amdgcn_rcp(x);
    return fma(y, fma(-x, y, 1.0), y);
}

// 1/sqrt(x) for x > 0, two Newton steps from the hardware estimate.
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    double e = fma(-h, y * y, 0.5);
    y = fma(y, e, y);
    e = fma(-h, y * y, 0.5);
    return fma(y, e, y);
}

// Taylor coefficients of sin (x^3..x^11) and of cos - 1 (x^2..x^10):
// (-1)^k / n!.  Kept in constant memory so they reach the FMAs as scalar
// operands (literal FP64 constants would be copied into VGPRs first).
__constant__ double kSinTaylor[5] = {-1.0 / 6.0, 1.0 / 120.0, -1.0 / 5040.0, 1.0 / 362880.0, -1.0 / 39916800.0};
__constant__ double kCosTaylor[5] = {-1.0 / 2.0, 1.0 / 24.0, -1.0 / 720.0, 1.0 / 40320.0, -1.0 / 3628800.0};

// sin x and cos x - 1 from the first N Taylor terms (Horner in x^2).  With
// N = 5 the truncation is below 1e-20 for |x| <= 0.1; with N = 4 below
// 1e-24 for |x| <= 0.02.
template <int N>
__device__ __forceinline__ void sincos_taylor(double x, double& s, double& cm1) {
    const double x2 = x * x;
    double ps = kSinTaylor[N - 1];
    double pc = kCosTaylor[N - 1];
#pragma unroll
    for (int k = N - 2; k >= 0; k--) {
        ps = fma(ps, x2, kSinTaylor[k]);
        pc = fma(pc, x2, kCosTaylor[k]);
    }
    s = fma(ps * x2, x, x);
    cm1 = pc * x2;
}

// Advance (s, cm1) = (sin E, cos E - 1) to E + d (|d| <= 0.02; larger
// corrections are routed to the fallback by the caller).
__device__ __forceinline__ void rotate_taylor(double d, double& s, double& cm1) {
    double sd, cd;
    sincos_taylor<4>(d, sd, cd);
    // sin(E+d) = s (1 + cd) + (1 + cm1) sd ; cos(E+d) - 1 = cm1 + cd + cm1 cd - s sd
    const double s_new = s + sd + fma(s, cd, cm1 * sd);
    const double c_new = cm1 + cd + fma(cm1, cd, -s * sd);
    s = s_new;
    cm1 = c_new;
}

// Kepler drift, FAST form.  inv_mu, sqrt_mu, inv_sqrt_mu: host constants.
//
// Straight-line path for the common case, then one wave vote: lanes for
// which a shortcut does not hold (|dM| > 0.1, a start above 0.02, a first
// correction above 1e-5, a second correction not below 1e-14, an unbound
// orbit) redo the step with the reference-arithmetic drift from their saved
// state, which also sets their 0x04 / 0x08 flags.
__device__ __forceinline__ void drift_fast(Particle& p, double dt, double mu, double inv_mu, double sqrt_mu,
                                           double inv_sqrt_mu) {
    const double rx0 = p.rx, ry0 = p.ry, rz0 = p.rz;
    const double vx0 = p.vx, vy0 = p.vy, vz0 = p.vz;
    const double r2 = fma(rz0, rz0, fma(ry0, ry0, rx0 * rx0));
    const double v2 = fma(vz0, vz0, fma(vy0, vy0, vx0 * vx0));
    const double vdotr = fma(vz0, rz0, fma(vy0, ry0, vx0 * rx0));
    const double idist = rsqrt_nr(r2);
    const double dist = r2 * idist;
    const double energy = fma(v2, 0.5, -mu * idist);
    const double ia = -2.0 * energy * inv_mu;      // 1/a
    const double a = rcp_nr(ia);
    const double rs_ia = rsqrt_nr(ia);              // sqrt(a)
    const double n_ = sqrt_mu * ia * (ia * rs_ia);  // sqrt(mu / a^3)
    const double in_ = inv_sqrt_mu * a * rs_ia;     // 1/n
    const double ecosEo = fma(-dist, ia, 1.0);
    const double esinEo = vdotr * in_ * ia * ia;
    // |dM| <= 0.1 on this path, so no full turns are removed and the time
    // left to advance, dM / n, is dt itself
    const double dM = dt * n_;
    double s, cm1;
    sincos_taylor<5>(dM, s, cm1);
    const double omec = 1.0 - ecosEo;
    // Kepler residual F(E) = E - ecosEo sin E + esinEo (1 - cos E) - dM and
    // F'(E) = 1 - ecosEo cos E + esinEo sin E, written with dd = E - dM and
    // (s, cm1) = (sin E, cos E - 1).  The start is one Newton step from
    // E = dM (error ~ e^3 dM^2 / 2) instead of the reference's series guess
    // (error ~ e^2 dM), so two more corrections reach the 1e-14 tolerance.
    double f = fma(-esinEo, cm1, -ecosEo * s);
    double fp = fma(esinEo, s, fma(-ecosEo, cm1, omec));
    double dd = -f * rcp_nr(fp);
    bool bad = !(energy < 0.0) || !(fabs(dM) <= 0.1) || !(fabs(dd) <= 0.02);
    rotate_taylor(dd, s, cm1);
    // first correction: the reciprocal with one Newton step (relative error
    // ~2e-15) is exact enough for a correction of at most 1e-5
    f = fma(-esinEo, cm1, fma(-ecosEo, s, dd));
    fp = fma(esinEo, s, fma(-ecosEo, cm1, omec));
    const double d1 = -f * rcp_nr1(fp);
    dd += d1;
    bad |= !(fabs(d1) <= 1e-5);
    {
        // sin d1 = d1 - d1^3/6, cos d1 - 1 = -d1^2/2 + d1^4/24 (exact to rounding for |d1| <= 1e-5)
        const double d2_ = d1 * d1;
        const double sd = fma(d1 * d2_, -1.0 / 6.0, d1);
        const double cd = d2_ * fma(d2_, 1.0 / 24.0, -0.5);
        const double s_new = s + sd + fma(s, cd, cm1 * sd);
        const double c_new = cm1 + cd + fma(cm1, cd, -s * sd);
        s = s_new;
        cm1 = c_new;
    }
    // second correction: tiny, the raw reciprocal estimate suffices
    f = fma(-esinEo, cm1, fma(-ecosEo, s, dd));
    fp = fma(esinEo, s, fma(-ecosEo, cm1, omec));
    const double d2 = -f * __builtin_amdgcn_rcp(fp);
    dd += d2;
    bad |= !(fabs(d2) < kTolKep);
    // |d2| < 1e-14: the first-order rotation is exact to rounding
    const double s2 = fma(cm1, d2, s + d2);
    cm1 = fma(-s, d2, cm1);
    s = s2;
    const double dE = dM + dd;
    fp = fma(esinEo, s, fma(-ecosEo, cm1, omec));
    const double ifp = rcp_nr(fp);
    const double fg = fma(a * cm1, idist, 1.0);
    const double g = fma(s - dE, in_, dt);
    const double fdot = -n_ * s * a * idist * ifp;
    const double gdot = fma(cm1, ifp, 1.0);
    p.rx = fma(rx0, fg, vx0 * g);
    p.ry = fma(ry0, fg, vy0 * g);
    p.rz = fma(rz0, fg, vz0 * g);
    p.vx = fma(rx0, fdot, vx0 * gdot);
    p.vy = fma(ry0, fdot, vy0 * gdot);
    p.vz = fma(rz0, fdot, vz0 * gdot);
    p.flags = 0;
    if (__any(bad)) {
        if (bad) {
            p.rx = rx0;
            p.ry = ry0;
            p.rz = rz0;
            p.vx = vx0;
            p.vy = vy0;
            p.vz = vz0;
            drift_fp64(p, dt, mu);
        }
    }
}

// Planet sum, FAST form: r^-3 from one refined FP64 reciprocal square root.
template <int NB>
__device__ __forceinline__ void accel_fast(Particle& p, const double* __restrict__ row, const BlockArgs& A,
                                           uint32_t nb) {
    double ax = row[0], ay = row[1], az = row[2];
    const uint32_t nbodies = NB > 0 ? (uint32_t)NB : nb;
#pragma unroll
    for (uint32_t j = 1; j < (NB > 0 ? (uint32_t)NB : (uint32_t)GLX_MAX_BODIES); j++) {
        if (NB == 0 && j >= nbodies) break;
        const double dx = p.rx - row[3 * j];
        const double dy = p.ry - row[3 * j + 1];
        const double dz = p.rz - row[3 * j + 2];
        const double rad = fma(dz, dz, fma(dy, dy, dx * dx));
        if (rad < A.pc.rh2[j] && p.flags == 0) p.flags = ((A.pc.id[j] << 8) | kFlagEncounter) & 0xFFFFu;
        const double y = rsqrt_nr(rad);
        const double fac = A.pc.m[j] * (y * y * y);
        ax = fma(-dx, fac, ax);
        ay = fma(-dy, fac, ay);
        az = fma(-dz, fac, az);
    }
    p.ax = ax;
    p.ay = ay;
    p.az = az;
    const double rad = fma(p.rz, p.rz, fma(p.ry, p.ry, p.rx * p.rx));
    if (rad < A.pc.rh2[0]) p.flags = kFlagEncounter;
    if (rad > A.outer2) p.flags = kFlagOut;
}

// Planet sum, MIXED form: FP32 r^-3 for every planet except the dominant
// one (the heaviest, chosen on the host), whose term is evaluated with the
// refined FP64 reciprocal square root; all terms accumulate in FP64 and
// |dr|^2 and the encounter test stay FP64.
template <int NB>
__device__ __forceinline__ void accel_mixed(Particle& p, const double* __restrict__ row, const BlockArgs& A,
                                            uint32_t nb) {
    double ax = row[0], ay = row[1], az = row[2];
    const uint32_t nbodies = NB > 0 ? (uint32_t)NB : nb;
#pragma unroll
    for (uint32_t j = 1; j < (NB > 0 ? (uint32_t)NB : (uint32_t)GLX_MAX_BODIES); j++) {
        if (NB == 0 && j >= nbodies) break;
        const double dx = p.rx - row[3 * j];
        const double dy = p.ry - row[3 * j + 1];
        const double dz = p.rz - row[3 * j + 2];
        const double rad = fma(dz, dz, fma(dy, dy, dx * dx));
        if (rad < A.pc.rh2[j] && p.flags == 0) p.flags = ((A.pc.id[j] << 8) | kFlagEncounter) & 0xFFFFu;
        double y3;
        if (j == A.dominant) {  // wave-uniform branch
            const double y = rsqrt_nr(rad);
            y3 = y * y * y;
        } else {
            const float yf = __builtin_amdgcn_rsqf((float)rad);
            y3 = (double)(yf * yf * yf);
        }
        const double fac = A.pc.m[j] * y3;
        ax = fma(-dx, fac, ax);
        ay = fma(-dy, fac, ay);
        az = fma(-dz, fac, az);
    }
    p.ax = ax;
    p.ay = ay;
    p.az = az;
    const double rad = fma(p.rz, p.rz, fma(p.ry, p.ry, p.rx * p.rx));
    if (rad < A.pc.rh2[0]) p.flags = kFlagEncounter;
    if (rad > A.outer2) p.flags = kFlagOut;
}

__device__ __forceinline__ void kick_fma(Particle& p, double half_dt) {
    p.vx = fma(p.ax, half_dt, p.vx);
    p.vy = fma(p.ay, half_dt, p.vy);
    p.vz = fma(p.az, half_dt, p.vz);
}

}  // namespace glx
