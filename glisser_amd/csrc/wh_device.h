// Device-side pieces of the test-particle Wisdom-Holman step for gfx950.
//
// One lane owns one particle for a whole time-block; its heliocentric state
// (r, v and the acceleration carried between steps) stays in VGPRs.  The
// per-step planet row (indirect term h0 followed by the heliocentric
// positions of bodies 1..NB-1) is read from LDS with wave-uniform addresses,
// so every read is a broadcast.
//
// Reference semantics (src/wh.cu:42-177): kick-drift-kick; the drift is the
// elliptic eccentric-anomaly-difference Kepler solve; the kick sums the
// planets and the indirect term; encounter (first planet, planet id in the
// high byte), Sun-cull and outer-limit flags are tested every step and the
// flagging step completes its second half kick.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "glisser_amd.h"
#include "sincos_cr.h"

namespace glx {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTolKep = 1e-14;  // src/wh.cu:13
constexpr int kKepIts = 5;         // src/wh.cu:12

// Steps of a block between two recomputations of the FAST / MIXED forms'
// carried radius from the position.  A launch also recomputes it at its
// first step, so launches that start on a multiple of this (glx_run_blocks
// splits blocks there) give the same bits as one launch per block.
constexpr uint32_t kRadiusPeriod = 64;
// Relative widening of the radii the carried |r| is voted against (the Sun
// cull radius and the outer limit), far beyond the carried value's error
// (< 1e-12 over kRadiusPeriod steps).
constexpr double kVoteMargin = 1e-9;

// Flag bits (src/data.h:169-180).
constexpr uint32_t kFlagEncounter = 0x01;
constexpr uint32_t kFlagOut = 0x02;
constexpr uint32_t kFlagKepler = 0x04;
constexpr uint32_t kFlagUnbound = 0x08;

struct PlanetConsts {
    double m[GLX_MAX_BODIES];    // GM of each body, m[0] = Sun
    double rh2[GLX_MAX_BODIES];  // squared encounter radius (rh[0]: Sun cull radius)
    uint32_t id[GLX_MAX_BODIES]; // planet id written into the flag high byte
};

// Arguments of one time-block launch.  Particle arrays are SoA; `log` is
// the interleaved planet block [n_steps][n_bodies][3]: h0 first, then the
// heliocentric positions of bodies 1..n_bodies-1.
struct BlockArgs {
    double* rx;
    double* ry;
    double* rz;
    double* vx;
    double* vy;
    double* vz;
    double* ax;
    double* ay;
    double* az;
    uint16_t* flags;
    uint32_t* dti;
    const double* log;
    uint32_t first;        // particles [first, n) of this launch
    uint32_t n;
    uint32_t step_offset;  // steps of the launch before this block (death index base)
    uint32_t block_step0;  // index within its block of the launch's first step
    uint32_t n_steps;
    uint32_t chunk;     // steps staged in LDS at a time
    uint32_t n_bodies;  // used by the generic (NB == 0) instantiation
    double dt;
    double half_dt;
    double mu;
    double outer2;
    double inv_mu;       // 1/mu, dt sqrt(mu), 1/sqrt(mu), 1/dt: FAST / MIXED drift
    double dt_sqrt_mu;
    double inv_sqrt_mu;
    double inv_dt;
    double sun_vote;     // FAST / MIXED: carried |r| below this or above
    double outer_vote;   // this triggers the exact Sun / outer tests
    PlanetConsts pc;
};

struct Particle {
    double rx, ry, rz, vx, vy, vz, ax, ay, az;
    double rl, ir;  // |r| and 1/|r| carried between steps (FAST / MIXED forms only)
    uint32_t flags;
};

// f and g functions of the Kepler drift in the reference's arithmetic
// (src/wh.cu:74-110 with kepeq at src/wh.cu:42-72): fixed five Newton
// iterations, IEEE divides and square roots, correctly rounded sin/cos
// (sincos_cr.h: glibc's sin/cos in the reference CPU build return the
// correctly rounded value on these arguments in all but a few calls per
// million), and no FMA contraction, so every basic operation rounds exactly
// as in the reference CPU build (the particle orbits amplify a one-ulp
// difference to ~1e-10 over 1e5 steps).  The new state is
// r = f r0 + g v0, v = fdot r0 + gdot v0.  Returns the flags of the drift:
// 0x04 if the last Newton correction exceeds 1e-14; for an unbound orbit
// (energy >= 0) 0x08 with the identity map, so r, v stay unchanged (the
// reference GPU code computes a NaN orbit there instead, and its CPU twin
// throws).  kCrTrig false: sin/cos from the device math library instead
// (the FAST / MIXED forms' fallback lanes, which need no bit-equality with
// the CPU path and keep their registers for the common path).
template <bool kCrTrig = true>
__device__ __forceinline__ uint32_t kepler_fg_fp64(double rx, double ry, double rz, double vx, double vy,
                                                   double vz, double dt, double mu, double& f, double& g,
                                                   double& fdot, double& gdot) {
#pragma clang fp contract(off)
    const double dist = sqrt(rx * rx + ry * ry + rz * rz);
    const double vdotr = vx * rx + vy * ry + vz * rz;
    const double energy = (vx * vx + vy * vy + vz * vz) * 0.5 - mu / dist;
    if (energy >= 0.0) {
        f = 1.0;
        g = 0.0;
        fdot = 0.0;
        gdot = 1.0;
        return kFlagUnbound;
    }
    const double a = -0.5 * mu / energy;
    const double n_ = sqrt(mu / (a * a * a));
    const double ecosEo = 1.0 - dist / a;
    const double esinEo = vdotr / (n_ * a * a);
    // dt*n - (pi*2)*(int)((dt*n/pi)*2): the reference's M_2PI macro
    // expansion.  For 0 <= dt*n < 1.5 the quotient lies below 0.955, the
    // integer part is 0 and the expression is dt*n itself (dt*n - 0.0), so
    // the divide only runs outside that range.
    const double dtn = dt * n_;
    double dM = dtn;
    if (!(dtn >= 0.0 && dtn < 1.5)) dM = dtn - kPi * 2 * (double)(int)(dtn / kPi * 2);
    const double adv_dt = dM / n_;
    double sM, cM;
    if constexpr (kCrTrig) cr::sincos_cr(dM, &sM, &cM);
    else sincos(dM, &sM, &cM);
    double dE = dM - esinEo + esinEo * cM + ecosEo * sM;
    double sindE, cosdE;
    if constexpr (kCrTrig) cr::sincos_cr(dE, &sindE, &cosdE);
    else sincos(dE, &sindE, &cosdE);
    double delta = 0.0;
#pragma unroll
    for (int k = 0; k < kKepIts; k++) {
        const double fk = dE - ecosEo * sindE + esinEo * (1. - cosdE) - dM;
        const double fpk = 1. - ecosEo * cosdE + esinEo * sindE;
        delta = -fk / fpk;
        dE += delta;
        if constexpr (kCrTrig) cr::sincos_cr(dE, &sindE, &cosdE);
        else sincos(dE, &sindE, &cosdE);
    }
    const double fp = 1.0 - ecosEo * cosdE + esinEo * sindE;
    f = 1.0 + a * (cosdE - 1.0) / dist;
    g = adv_dt + (sindE - dE) / n_;
    fdot = -n_ * sindE * a / (dist * fp);
    gdot = 1.0 + (cosdE - 1.0) / fp;
    return fabs(delta) > kTolKep ? kFlagKepler : 0u;
}

// Kepler drift in the reference's arithmetic: kepler_fg_fp64 applied
// without FMA contraction, as the reference does.
__device__ __forceinline__ void drift_fp64(Particle& p, double dt, double mu) {
#pragma clang fp contract(off)
    double f, g, fdot, gdot;
    p.flags = kepler_fg_fp64(p.rx, p.ry, p.rz, p.vx, p.vy, p.vz, dt, mu, f, g, fdot, gdot);
    if (p.flags == kFlagUnbound) return;
    const double rx0 = p.rx, ry0 = p.ry, rz0 = p.rz;
    p.rx = rx0 * f + p.vx * g;
    p.ry = ry0 * f + p.vy * g;
    p.rz = rz0 * f + p.vz * g;
    p.vx = rx0 * fdot + p.vx * gdot;
    p.vy = ry0 * fdot + p.vy * gdot;
    p.vz = rz0 * fdot + p.vz * gdot;
}

// Half kick v += a * (dt/2), rounded as the reference CPU build (no FMA).
__device__ __forceinline__ void kick(Particle& p, double half_dt) {
#pragma clang fp contract(off)
    p.vx = p.vx + p.ax * half_dt;
    p.vy = p.vy + p.ay * half_dt;
    p.vz = p.vz + p.az * half_dt;
}

// Acceleration from the step's planet row plus the encounter / cull /
// outer-limit tests (src/wh.cu:142-170), in FP64 with IEEE divide/sqrt.
// NB > 0: compile-time body count; NB == 0: nb bodies at run time, at most
// NMAX.
template <int NB, int NMAX = (NB > 0 ? NB : GLX_MAX_BODIES)>
__device__ __forceinline__ void accel_fp64(Particle& p, const double* __restrict__ row, const PlanetConsts& pc,
                                           double outer2, uint32_t nb) {
#pragma clang fp contract(off)
    p.ax = row[0];
    p.ay = row[1];
    p.az = row[2];
    const uint32_t nbodies = NB > 0 ? (uint32_t)NB : nb;
#pragma unroll
    for (uint32_t j = 1; j < (uint32_t)NMAX; j++) {
        if (NB == 0 && j >= nbodies) break;
        const double dx = p.rx - row[3 * j];
        const double dy = p.ry - row[3 * j + 1];
        const double dz = p.rz - row[3 * j + 2];
        const double rad = dx * dx + dy * dy + dz * dz;
        if (rad < pc.rh2[j] && p.flags == 0) p.flags = ((pc.id[j] << 8) | kFlagEncounter) & 0xFFFFu;
        const double inv3 = 1. / (rad * sqrt(rad));
        const double fac = pc.m[j] * inv3;
        p.ax -= dx * fac;
        p.ay -= dy * fac;
        p.az -= dz * fac;
    }
    const double rad = p.rx * p.rx + p.ry * p.ry + p.rz * p.rz;
    if (rad < pc.rh2[0]) p.flags = kFlagEncounter;
    if (rad > outer2) p.flags = kFlagOut;
}

}  // namespace glx
