// FAST and MIXED forms of the particle step.
//
// Same map as the reference (kick-drift-kick, eccentric-anomaly-difference
// Kepler solve, direct planet sum + indirect term, identical flag logic),
// restructured for the gfx950 FP64 VALU:
//   * one reciprocal square root of |r|^2 and of 1/a replace the drift's
//     divisions and square roots; 1/a = -2E/mu needs no division at all;
//   * reciprocals / reciprocal square roots start from v_rcp_f64 /
//     v_rsq_f64 and take two FMA Newton steps;
//   * sin/cos of the mean-anomaly step use a five-term Taylor polynomial
//     when every lane of the wave has |x| <= 0.1 (the device library
//     otherwise); the first guess and every Newton correction rotate
//     (sin E, cos E - 1) by the small angle with a short series instead of
//     re-evaluating sin/cos; cos E - 1 and E - dM are carried explicitly, so
//     f, g and the Kepler residual avoid cancellation;
//   * the Newton iteration stops once every lane of the wave has
//     |delta| < 1e-14 after applying it (the early-exit mode the reference
//     offers under CUDA_KEPEQ_CHECK_CONVERGENCE, src/wh.cu:60-71), at most
//     five iterations;
//   * FMA contraction is allowed.
// MIXED additionally evaluates each planet's r^-3 with the FP32 reciprocal
// square root (|dr|^2 and the encounter test stay FP64), accumulates the
// acceleration in FP64, and refines the lane's dominant planet term with one
// FP64 Newton step.
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

// sin x and cos x - 1 from the first N Taylor terms (Horner in x^2).
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

// Device-library sin/cos for large arguments, kept out of line so its
// range-reduction constants do not occupy registers in the hot loop.
__device__ __noinline__ void sincos_library(double x, double* s, double* cm1) {
    double c;
    sincos(x, s, &c);
    *cm1 = c - 1.0;
}

// sin x and cos x - 1: five Taylor terms when every lane of the wave has
// |x| <= 0.1 (truncation below 1e-20 absolute), the device library
// otherwise.
__device__ __forceinline__ void sincos_m1(double x, double& s, double& cm1) {
    if (__all(fabs(x) <= 0.1)) {
        sincos_taylor<5>(x, s, cm1);
    } else {
        sincos_library(x, &s, &cm1);
    }
}

// Advance (s, cm1) = (sin E, cos E - 1) to E + d.  sin d and cos d - 1 come
// from four Taylor terms each, exact to rounding for |d| <= 0.02; when a
// lane's correction is larger the wave re-evaluates sin/cos of e_new.
__device__ __forceinline__ void rotate_m1(double e_new, double d, double& s, double& cm1) {
    if (!__all(fabs(d) <= 0.02)) {
        sincos_m1(e_new, s, cm1);
        return;
    }
    double sd, cd;
    sincos_taylor<4>(d, sd, cd);
    // sin(E+d) = s (1 + cd) + (1 + cm1) sd ; cos(E+d) - 1 = cm1 + cd + cm1 cd - s sd
    const double s_new = s + sd + fma(s, cd, cm1 * sd);
    const double c_new = cm1 + cd + fma(cm1, cd, -s * sd);
    s = s_new;
    cm1 = c_new;
}

// Kepler drift, FAST form.  inv_mu, sqrt_mu, inv_sqrt_mu: host constants.
__device__ __forceinline__ void drift_fast(Particle& p, double dt, double mu, double inv_mu, double sqrt_mu,
                                           double inv_sqrt_mu) {
    const double r2 = fma(p.rz, p.rz, fma(p.ry, p.ry, p.rx * p.rx));
    const double v2 = fma(p.vz, p.vz, fma(p.vy, p.vy, p.vx * p.vx));
    const double vdotr = fma(p.vz, p.rz, fma(p.vy, p.ry, p.vx * p.rx));
    const double idist = rsqrt_nr(r2);
    const double dist = r2 * idist;
    const double energy = fma(v2, 0.5, -mu * idist);
    if (energy >= 0.0) {
        p.flags = kFlagUnbound;
        return;
    }
    const double ia = -2.0 * energy * inv_mu;       // 1/a
    const double a = rcp_nr(ia);
    const double rs_ia = rsqrt_nr(ia);               // sqrt(a)
    const double n_ = sqrt_mu * ia * (ia * rs_ia);   // sqrt(mu / a^3)
    const double in_ = inv_sqrt_mu * a * rs_ia;      // 1/n
    const double ecosEo = fma(-dist, ia, 1.0);
    const double esinEo = vdotr * in_ * ia * ia;
    const double dtn = dt * n_;
    // the reference's (int)(dt n / (pi/2)) full-turn subtraction; 0 for dt n < pi/2
    const double dM = fma(-(kPi * 2), (double)(int)(dtn * (2.0 / kPi)), dtn);
    const double adv_dt = dM * in_;
    double s, cm1;
    sincos_m1(dM, s, cm1);
    // first guess dE = dM + esinEo (cos dM - 1) + ecosEo sin dM; dd = dE - dM
    double dd = fma(esinEo, cm1, ecosEo * s);
    double dE = dM + dd;
    rotate_m1(dE, dd, s, cm1);
    const double omec = 1.0 - ecosEo;
    double delta = 0.0;
    for (int k = 0; k < kKepIts; k++) {
        // F(E) = E - ecosEo sin E + esinEo (1 - cos E) - dM,  F' = 1 - ecosEo cos E + esinEo sin E
        const double f = fma(-esinEo, cm1, fma(-ecosEo, s, dd));
        const double fp = fma(esinEo, s, fma(-ecosEo, cm1, omec));
        // after the first correction the step is tiny; the raw reciprocal
        // estimate is then accurate enough for it
        const double ifp = k == 0 ? rcp_nr(fp) : __builtin_amdgcn_rcp(fp);
        delta = -f * ifp;
        dd += delta;
        dE = dM + dd;
        rotate_m1(dE, delta, s, cm1);
        // written so that a NaN correction counts as not converged
        if (k > 0 && !__any(!(fabs(delta) < kTolKep))) break;
    }
    p.flags = fabs(delta) <= kTolKep ? 0u : kFlagKepler;
    const double fp = fma(esinEo, s, fma(-ecosEo, cm1, omec));
    const double ifp = rcp_nr(fp);
    const double f = fma(a * cm1, idist, 1.0);
    const double g = fma(s - dE, in_, adv_dt);
    const double fdot = -n_ * s * a * idist * ifp;
    const double gdot = fma(cm1, ifp, 1.0);
    const double rx0 = p.rx, ry0 = p.ry, rz0 = p.rz;
    p.rx = fma(rx0, f, p.vx * g);
    p.ry = fma(ry0, f, p.vy * g);
    p.rz = fma(rz0, f, p.vz * g);
    p.vx = fma(rx0, fdot, p.vx * gdot);
    p.vy = fma(ry0, fdot, p.vy * gdot);
    p.vz = fma(rz0, fdot, p.vz * gdot);
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
