// FAST and MIXED forms of the particle step.
//
// Same map as the reference (kick-drift-kick, eccentric-anomaly-difference
// Kepler solve, direct planet sum + indirect term, identical flag logic),
// restructured for the gfx950 FP64 VALU:
//   * one reciprocal square root of |r|^2 and of 1/a replace the drift's
//     divisions and square roots; 1/a = -2E/mu needs no division at all;
//   * reciprocals / reciprocal square roots start from v_rcp_f64 /
//     v_rsq_f64 and take two FMA Newton steps;
//   * sin/cos of the mean-anomaly step and of the first guess use a Taylor
//     polynomial (|x| <= 0.8; the device library beyond that);
//   * Newton corrections rotate (sin, cos) by the small correction with a
//     short series instead of re-evaluating sin/cos, and the iteration stops
//     once every lane of the wave has |delta| < 1e-14 (the early-exit mode
//     the reference offers under CUDA_KEPEQ_CHECK_CONVERGENCE,
//     src/wh.cu:60-71), at most five iterations;
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

// sin and cos by Taylor series in x^2 (terms through x^19 / x^18), accurate
// to below 1e-17 absolute for |x| <= 0.8; larger arguments (any lane) take
// the device library.
__device__ __forceinline__ void sincos_small(double x, double* s, double* c) {
    if (fabs(x) > 0.8) {
        sincos(x, s, c);
        return;
    }
    const double x2 = x * x;
    // sin: x * (1 - x2/3! + x2^2/5! - ... + x2^9/19!)
    double ps = -1.0 / 121645100408832000.0;  // -1/19!
    ps = fma(ps, x2, 1.0 / 355687428096000.0);  // 1/17!
    ps = fma(ps, x2, -1.0 / 1307674368000.0);   // -1/15!
    ps = fma(ps, x2, 1.0 / 6227020800.0);       // 1/13!
    ps = fma(ps, x2, -1.0 / 39916800.0);        // -1/11!
    ps = fma(ps, x2, 1.0 / 362880.0);           // 1/9!
    ps = fma(ps, x2, -1.0 / 5040.0);            // -1/7!
    ps = fma(ps, x2, 1.0 / 120.0);              // 1/5!
    ps = fma(ps, x2, -1.0 / 6.0);               // -1/3!
    *s = fma(ps * x2, x, x);
    // cos: 1 - x2/2! + x2^2/4! - ... + x2^9/18!
    double pc = 1.0 / 6402373705728000.0;       // 1/18!
    pc = fma(pc, x2, -1.0 / 20922789888000.0);  // -1/16!
    pc = fma(pc, x2, 1.0 / 87178291200.0);      // 1/14!
    pc = fma(pc, x2, -1.0 / 479001600.0);       // -1/12!
    pc = fma(pc, x2, 1.0 / 3628800.0);          // 1/10!
    pc = fma(pc, x2, -1.0 / 40320.0);           // -1/8!
    pc = fma(pc, x2, 1.0 / 720.0);              // 1/6!
    pc = fma(pc, x2, -1.0 / 24.0);              // -1/4!
    pc = fma(pc, x2, 0.5);                      // 1/2!  (applied with a minus)
    *c = fma(-pc, x2, 1.0);
}

// Rotate (s, c) = (sin E, cos E) to E + d for |d| < 1e-4:
// sin d = d - d^3/6 + d^5/120, cos d = 1 - d^2/2 + d^4/24 (next terms
// below 1e-26).  Larger corrections re-evaluate sin/cos of the new angle.
__device__ __forceinline__ void rotate_small(double e_new, double d, double& s, double& c) {
    if (fabs(d) >= 1e-4) {
        sincos_small(e_new, &s, &c);
        return;
    }
    const double d2 = d * d;
    const double sd = d * fma(d2, fma(d2, 1.0 / 120.0, -1.0 / 6.0), 1.0);
    const double cm1 = d2 * fma(d2, 1.0 / 24.0, -0.5);  // cos d - 1
    const double s0 = s, c0 = c;
    s = fma(s0, cm1, fma(c0, sd, s0));
    c = fma(c0, cm1, fma(-s0, sd, c0));
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
    const double ia = -2.0 * energy * inv_mu;    // 1/a
    const double a = rcp_nr(ia);
    const double rs_ia = rsqrt_nr(ia);            // sqrt(a)
    const double n_ = sqrt_mu * ia * (ia * rs_ia);  // sqrt(mu / a^3)
    const double in_ = inv_sqrt_mu * a * rs_ia;     // 1/n
    const double ecosEo = fma(-dist, ia, 1.0);
    const double esinEo = vdotr * in_ * ia * ia;
    const double dtn = dt * n_;
    // the reference's (int)(dt n / (pi/2)) full-turn subtraction; 0 for dt n < pi/2
    const double dM = fma(-(kPi * 2), (double)(int)(dtn * (2.0 / kPi)), dtn);
    const double adv_dt = dM * in_;
    double sM, cM;
    sincos_small(dM, &sM, &cM);
    double dE = dM - esinEo + esinEo * cM + ecosEo * sM;
    double s, c;
    sincos_small(dE, &s, &c);
    double delta = 0.0;
    for (int k = 0; k < kKepIts; k++) {
        const double f = fma(esinEo, 1.0 - c, fma(-ecosEo, s, dE)) - dM;
        const double fp = fma(esinEo, s, fma(-ecosEo, c, 1.0));
        delta = -f * rcp_nr(fp);
        dE += delta;
        rotate_small(dE, delta, s, c);
        if (!__any(fabs(delta) >= kTolKep)) break;
    }
    p.flags = fabs(delta) > kTolKep ? kFlagKepler : 0u;
    const double fp = fma(esinEo, s, fma(-ecosEo, c, 1.0));
    const double ifp = rcp_nr(fp);
    const double cm1 = c - 1.0;
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
