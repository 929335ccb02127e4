// FAST and MIXED forms of the particle step.
//
// Same map as the reference (kick-drift-kick, eccentric-anomaly-difference
// Kepler solve, direct planet sum + indirect term, identical flag logic),
// restructured for the gfx950 FP64 VALU:
//   * one reciprocal square root of |r|^2 and of 1/a replace the drift's
//     divisions and square roots; 1/a = -2E/mu needs no division at all;
//   * reciprocals / reciprocal square roots start from v_rcp_f64 /
//     v_rsq_f64 (relative error < 6e-8) and take one FMA Newton step where
//     the result feeds the state (relative error < 6e-15, a few ulp of the
//     values it scales, far below the step's truncation error), none where
//     it only scales a Newton correction;
//   * sin/cos of the mean-anomaly step and of the start E come from a
//     five-term Taylor polynomial (|dM|, |E| <= 0.2); each later Newton
//     correction rotates (sin E, cos E - 1) by the small angle with a short
//     series instead of re-evaluating sin/cos; cos E - 1 and E - dM are
//     carried explicitly, so f, g and the Kepler residual avoid
//     cancellation;
//   * the Kepler solve starts from one Newton step off E = dM (instead of
//     the reference's series guess) and takes one more correction (two in
//     waves where some lane needs it), accepted when the error it can leave
//     is below the 1e-14 tolerance of the reference's convergence test
//     (src/wh.cu:60-71);
//   * every shortcut is checked per lane; the few lanes outside them
//     (large steps, slow convergence, unbound orbits) redo the step with the
//     reference-arithmetic drift, so results and flags stay defined for any
//     input;
//   * FMA contraction is allowed.
// MIXED additionally starts each planet's r^-3 from the FP32 reciprocal
// square root (corrected in FP64); everything else, the accumulation
// included, stays FP64.
#pragma once

#include "wh_device.h"

namespace glx {

// Reciprocal of x (x normal) with one Newton step from v_rcp_f64: the
// estimate's relative error (< 6e-8) is squared, leaving < 3e-15 (measured
// on gfx950: 2.2e-15).
__device__ __forceinline__ double rcp_nr1(double x) {
    const double y = __builtin_amdgcn_rcp(x);
    return fma(y, fma(-x, y, 1.0), y);
}

// 1/sqrt(x) for x > 0 with one Newton step from v_rsq_f64: relative error
// about 1.5 times the estimate's squared, < 6e-15 (measured: 4.3e-15).
__device__ __forceinline__ double rsqrt_nr1(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    return fma(y, fma(-0.5 * x, y * y, 0.5), y);
}

// Taylor coefficients of sin (x^3..x^11) and of cos - 1 (x^2..x^10):
// (-1)^k / n!.  Kept in constant memory so they reach the FMAs as scalar
// operands (literal FP64 constants would be copied into VGPRs first).
__constant__ double kSinTaylor[5] = {-1.0 / 6.0, 1.0 / 120.0, -1.0 / 5040.0, 1.0 / 362880.0, -1.0 / 39916800.0};
__constant__ double kCosTaylor[5] = {-1.0 / 2.0, 1.0 / 24.0, -1.0 / 720.0, 1.0 / 40320.0, -1.0 / 3628800.0};

// sin x and cos x - 1 from the first N Taylor terms (Horner in x^2).  With
// N = 5 the truncation is below 9e-18 for |x| <= 0.2.
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

// Validity bound of the FAST drift's shortcut: the five-term Taylor sin/cos
// of dM and of the start E (truncation below 9e-18 absolute at 0.2).
constexpr double kMaxDM = 0.2;

// One Newton correction of the Kepler solve in the (dd, sin E, cos E - 1)
// form; returns whether the error it leaves is provably below 1e-14
// (d^2 < 2e-14 F').  Corrections above 1e-3 are left unapplied and
// reported as not converged (the two-term rotation is exact to rounding
// only below that).
__device__ __forceinline__ bool newton_step(double esinEo, double ecosEo, double omec, double& dd, double& s,
                                            double& cm1) {
    const double f = fma(-esinEo, cm1, fma(-ecosEo, s, dd));
    const double fp = fma(esinEo, s, fma(-ecosEo, cm1, omec));
    double d = -f * rcp_nr1(fp);
    const bool small = fabs(d) <= 1e-3;
    d = small ? d : 0.0;
    dd += d;
    // sin d = d - d^3/6, cos d - 1 = -d^2/2 + d^4/24 (truncation below 1e-17 for |d| <= 1e-3)
    const double d2 = d * d;
    const double sd = fma(d * d2, -1.0 / 6.0, d);
    const double cd = d2 * fma(d2, 1.0 / 24.0, -0.5);
    const double s_new = s + sd + fma(s, cd, cm1 * sd);
    const double c_new = cm1 + cd + fma(cm1, cd, -s * sd);
    s = s_new;
    cm1 = c_new;
    return small && d2 < (2.0 * kTolKep) * fp;
}

// Kepler drift, FAST form.  inv_mu, sqrt_mu, inv_sqrt_mu: host constants;
// p.r2 must hold |r|^2 (fma(z, z, fma(y, y, x * x))) of the current position.
//
// Straight-line path for the common case, then one wave vote: lanes for
// which a shortcut does not hold (|dM| or the start |E| > 0.2, a solve not
// provably within 1e-14 after two corrections, an unbound orbit) redo
// the step with the reference-arithmetic drift from their saved state,
// which also sets their 0x04 / 0x08 flags.
__device__ __forceinline__ void drift_fast(Particle& p, double dt, double mu, double inv_mu, double sqrt_mu,
                                           double inv_sqrt_mu) {
    const double rx0 = p.rx, ry0 = p.ry, rz0 = p.rz;
    const double vx0 = p.vx, vy0 = p.vy, vz0 = p.vz;
    const double r2 = p.r2;  // |r|^2, left by the previous step's planet sum
    const double v2 = fma(vz0, vz0, fma(vy0, vy0, vx0 * vx0));
    const double vdotr = fma(vz0, rz0, fma(vy0, ry0, vx0 * rx0));
    const double idist = rsqrt_nr1(r2);
    const double dist = r2 * idist;
    // 1/a = -2E/mu = 2/|r| - v^2/mu; bound orbits have 1/a > 0
    const double ia = fma(-v2, inv_mu, 2.0 * idist);
    const double rs_ia = rsqrt_nr1(ia);              // sqrt(a)
    const double a = rs_ia * rs_ia;
    const double n_ = sqrt_mu * ia * (ia * rs_ia);  // sqrt(mu / a^3)
    const double in_ = inv_sqrt_mu * a * rs_ia;     // 1/n
    const double ecosEo = fma(-dist, ia, 1.0);
    const double esinEo = vdotr * in_ * ia * ia;
    // |dM| <= 0.2 on this path, so no full turns are removed and the time
    // left to advance, dM / n, is dt itself
    const double dM = dt * n_;
    double s, cm1;
    sincos_taylor<5>(dM, s, cm1);
    const double omec = 1.0 - ecosEo;
    // Kepler residual F(E) = E - ecosEo sin E + esinEo (1 - cos E) - dM and
    // F'(E) = 1 - ecosEo cos E + esinEo sin E, written with dd = E - dM and
    // (s, cm1) = (sin E, cos E - 1).  The start is one Newton step from
    // E = dM (error ~ e^3 dM^2 / 2, against ~ e^2 dM for the reference's
    // series guess); the raw reciprocal suffices there because the next
    // Newton step corrects any error of the start quadratically.
    double f = fma(-esinEo, cm1, -ecosEo * s);
    double fp = fma(esinEo, s, fma(-ecosEo, cm1, omec));
    double dd = -f * __builtin_amdgcn_rcp(fp);
    // sin / cos - 1 of the start E = dM + dd from the same polynomial
    const double E1 = dM + dd;
    bool bad = !(ia > 0.0) || !(fabs(dM) <= kMaxDM) || !(fabs(E1) <= kMaxDM);
    sincos_taylor<5>(E1, s, cm1);
    // Newton corrections.  One leaves an error of at most |F''| d^2 / (2 F')
    // <= d^2 / (2 F') (|F''| <= e < 1), so the solve is accepted as soon as
    // that bound is below the reference's 1e-14 tolerance.  One correction
    // suffices when e dM is small (the config of most swarms); a second runs
    // only in waves where some lane needs it.  The reciprocal with one
    // Newton step (relative error ~2e-15) is exact enough for corrections
    // that small.
    bool conv = newton_step(esinEo, ecosEo, omec, dd, s, cm1);
    if (__any(!conv)) conv = newton_step(esinEo, ecosEo, omec, dd, s, cm1);
    bad |= !conv;
    const double dE = dM + dd;
    fp = fma(esinEo, s, fma(-ecosEo, cm1, omec));
    const double ifp = rcp_nr1(fp);
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

// Planet sum and flag tests, FAST and MIXED forms.
//
// r^-3 of each planet: the hardware reciprocal square root y of |dr|^2
// (FAST: v_rsq_f64; MIXED: v_rsq_f32 of the rounded |dr|^2, trading the
// FP64 transcendental for two conversions) and the first-order correction
// of its cube, y^3 (1 + 1.5 e) with e = 1 - |dr|^2 y^2, in FP64: relative
// error 1.875 e^2, about 2e-14 (FAST) and 1e-13 (MIXED) at the estimates'
// ~1e-7, far below the planet term's share of the step error.  The
// terms, their accumulation, |dr|^2 and every test are FP64.
//
// The encounter, Sun and outer-boundary tests keep the reference's
// precedence (src/wh.cu:151-170: the first planet inside its radius wins if
// the particle is still unflagged, then the Sun and the outer test
// overwrite), but the main path only ORs "some radius crossed" into one
// wave vote; the exact assignment runs in the rare waves where a lane
// crossed one.  |r|^2 of the new position is left in p.r2 for the next
// drift (the kick in between does not move the particle).
template <int NB, bool MIXED, int NMAX = (NB > 0 ? NB : GLX_MAX_BODIES)>
__device__ __forceinline__ void accel_vote(Particle& p, const double* __restrict__ row, const PlanetConsts& pc,
                                           double outer2, uint32_t nb) {
    double ax = row[0], ay = row[1], az = row[2];
    const uint32_t nbodies = NB > 0 ? (uint32_t)NB : nb;
    uint64_t close = 0;  // wave mask: lanes inside some planet's radius
#pragma unroll
    for (uint32_t j = 1; j < (uint32_t)NMAX; j++) {
        if (NB == 0 && j >= nbodies) continue;  // wave-uniform
        const double dx = p.rx - row[3 * j];
        const double dy = p.ry - row[3 * j + 1];
        const double dz = p.rz - row[3 * j + 2];
        const double rad = fma(dz, dz, fma(dy, dy, dx * dx));
        close |= __ballot(rad < pc.rh2[j]);
        // y ~ rad^-1/2 with relative error d; e = 1 - rad y^2 ~ -2d and
        // rad^-3/2 = y^3 (1 - e)^-3/2 = y^3 (1 + 1.5 e + 1.875 e^2 + ...)
        double y;
        if constexpr (MIXED) {
            y = (double)__builtin_amdgcn_rsqf((float)rad);
        } else {
            y = __builtin_amdgcn_rsq(rad);
        }
        const double y2 = y * y;
        const double e = fma(-rad, y2, 1.0);
        const double my3 = pc.m[j] * (y2 * y);
        const double fac = fma(my3, 1.5 * e, my3);
        ax = fma(-dx, fac, ax);
        ay = fma(-dy, fac, ay);
        az = fma(-dz, fac, az);
    }
    p.ax = ax;
    p.ay = ay;
    p.az = az;
    const double r2 = fma(p.rz, p.rz, fma(p.ry, p.ry, p.rx * p.rx));
    p.r2 = r2;
    if ((close | __ballot(r2 < pc.rh2[0] || r2 > outer2)) != 0) {
        if (((close >> __lane_id()) & 1) && p.flags == 0) {
            for (uint32_t j = 1; j < nbodies; j++) {
                const double dx = p.rx - row[3 * j];
                const double dy = p.ry - row[3 * j + 1];
                const double dz = p.rz - row[3 * j + 2];
                if (fma(dz, dz, fma(dy, dy, dx * dx)) < pc.rh2[j]) {
                    p.flags = ((pc.id[j] << 8) | kFlagEncounter) & 0xFFFFu;
                    break;
                }
            }
        }
        if (r2 < pc.rh2[0]) p.flags = kFlagEncounter;
        if (r2 > outer2) p.flags = kFlagOut;
    }
}

__device__ __forceinline__ void kick_fma(Particle& p, double half_dt) {
    p.vx = fma(p.ax, half_dt, p.vx);
    p.vy = fma(p.ay, half_dt, p.vy);
    p.vz = fma(p.az, half_dt, p.vz);
}

}  // namespace glx
