// FAST and MIXED forms of the particle step.
//
// Same map as the reference (kick-drift-kick, eccentric-anomaly-difference
// Kepler solve, direct planet sum + indirect term, identical flag logic),
// restructured for the gfx950 FP64 VALU:
//   * one reciprocal square root of 1/a replaces the drift's divisions and
//     square roots; 1/a = -2E/mu needs no division at all; |r| and 1/|r|
//     come out of the previous step's Kepler solution (|r(E)| = a F'(E))
//     and are recomputed from the position every 64 steps of a block and at
//     the start of a launch;
//   * reciprocals / reciprocal square roots start from v_rcp_f64 /
//     v_rsq_f64 (relative error < 6e-8) and take one FMA Newton step where
//     the result feeds the state (relative error < 6e-15, a few ulp of the
//     values it scales, far below the step's truncation error), none where
//     it only scales a Newton correction;
//   * the Kepler solve starts from the series reversion of the Kepler
//     equation in E - E0 to third order (no sin/cos of the mean-anomaly
//     step), evaluates sin/cos of that start with a three-term Taylor
//     polynomial (|E| <= 0.02: long-period orbits) and takes one short
//     Newton correction (raw reciprocal, first-order rotation of
//     (sin E, cos E - 1) by the correction), accepted only when the error it
//     can leave is provably below one ulp; lanes outside that bound, or with
//     a larger start, restart from a five-term polynomial at the corrected E
//     (|E| <= 0.2) and take up to two full corrections, accepted when the
//     error is below the 1e-14 tolerance of the reference's convergence
//     test (src/wh.cu:60-71); cos E - 1 and E - dM are carried explicitly,
//     so f, g and the Kepler residual avoid cancellation;
//   * every shortcut is checked per lane, and every choice that changes a
//     lane's arithmetic is made per lane (a particle's result does not
//     depend on its wave-mates); the few lanes outside the shortcuts
//     (large steps, slow convergence, unbound orbits) take the f and g
//     functions of the reference-arithmetic drift, so results and flags
//     stay defined for any input;
//   * FMA contraction is allowed.
// MIXED additionally starts each planet's r^-3 from the FP32 reciprocal
// square root (corrected in FP64); everything else, the accumulation
// included, stays FP64.
#pragma once

#include "wh_device.h"

namespace glx {

// Wave votes on a lane predicate, taken straight from its compare mask
// (HIP's int-typed __any / __ballot move the predicate through a VGPR and
// compare it back: two VALU instructions per vote).  Inactive lanes vote
// false.
__device__ __forceinline__ uint64_t wave_mask(bool pred) { return __builtin_amdgcn_ballot_w64(pred); }
__device__ __forceinline__ bool wave_any(bool pred) { return __builtin_amdgcn_ballot_w64(pred) != 0; }
// Empty register constraint: x is materialised here (an optimisation
// barrier for its computation; no instruction is emitted).
__device__ __forceinline__ void pin(double& x) { __asm__ volatile("" : "+v"(x)); }

// This lane's bit of a wave mask.
__device__ __forceinline__ bool lane_in(uint64_t mask) { return (mask >> __lane_id()) & 1; }

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

// Taylor coefficients of sin (x^3..x^11, row 0) and of cos - 1 (x^2..x^10,
// row 1): (-1)^k / n!.  Kept in constant memory so they reach the FMAs as
// scalar operands (literal FP64 constants would be copied into VGPRs first).
__constant__ double kTaylor[2][5] = {
    {-1.0 / 6.0, 1.0 / 120.0, -1.0 / 5040.0, 1.0 / 362880.0, -1.0 / 39916800.0},
    {-1.0 / 2.0, 1.0 / 24.0, -1.0 / 720.0, 1.0 / 40320.0, -1.0 / 3628800.0}};

// sin x and cos x - 1 from the first N Taylor terms (Horner in x^2), with
// the coefficient table c (kTaylor).  With N = 5 the truncation is below
// 9e-18 for |x| <= 0.2.
template <int N>
__device__ __forceinline__ void sincos_taylor(double x, double& s, double& cm1,
                                              const double (*__restrict__ c)[5] = kTaylor) {
    const double x2 = x * x;
    double ps = c[0][N - 1];
    double pc = c[1][N - 1];
#pragma unroll
    for (int k = N - 2; k >= 0; k--) {
        ps = fma(ps, x2, c[0][k]);
        pc = fma(pc, x2, c[1][k]);
    }
    s = fma(ps * x2, x, x);
    cm1 = pc * x2;
}

// Highest coefficients of the three-term start polynomials (x^7 of sin,
// x^6 of cos - 1), passed by value so a kernel can hold them in vector
// registers: a fused multiply-add reads one scalar operand, so with both
// Horner operands in scalar registers one of them is copied to a vector
// register at every step.
struct TaylorLead {
    double s7, c6;
    cr::QuickLead q;  // fp64 form: the correctly rounded sin/cos's short-tier constants
};
__device__ __forceinline__ TaylorLead taylor_lead() { return {kTaylor[0][2], kTaylor[1][2], cr::kQuickLead}; }

// Holds a kernel's lead constants in vector registers: the fast forms'
// start-polynomial coefficients, or the fp64 form's sin/cos constants.
template <int FORM>
__device__ __forceinline__ void pin_lead(TaylorLead& lead) {
    if constexpr (FORM != GLX_FORM_FP64) {
        pin(lead.s7);
        pin(lead.c6);
    } else {
        pin(lead.q.s9);
        pin(lead.q.c8);
        pin(lead.q.zs);
        pin(lead.q.zc);
    }
}

// sincos_taylor<3> with the highest coefficients from `lead`, returning
// sin x - x (smx, whose rounding then stays relative to x^3 / 6) instead of
// sin x.
__device__ __forceinline__ void sincos_taylor3(double x, const TaylorLead& lead, double& smx, double& cm1) {
    const double x2 = x * x;
    const double ps = fma(fma(lead.s7, x2, kTaylor[0][1]), x2, kTaylor[0][0]);
    const double pc = fma(fma(lead.c6, x2, kTaylor[1][1]), x2, kTaylor[1][0]);
    smx = (ps * x2) * x;
    cm1 = pc * x2;
}

// Validity bound of the FAST drift's shortcut: |dM| (no full turns to
// remove) and |E| wherever the five-term Taylor sin/cos is evaluated
// (truncation below 9e-18 absolute at 0.2).
constexpr double kMaxDM = 0.2;
// Bound of the start |E| for the three-term Taylor sin/cos (truncation below
// 1e-18 absolute at 0.02: x^9 / 9! for sin, x^8 / 8! for cos - 1).
constexpr double kMaxStart = 0.02;

// One Newton correction of the Kepler solve in the (dd, sin E, cos E - 1)
// form; returns whether the error it leaves is provably below 1e-14
// (d^2 < 2e-14 F').  Corrections above 1e-3 are left unapplied and
// reported as not converged (the two-term rotation is exact to rounding
// only below that).  r receives the reciprocal of F' it used.
__device__ __forceinline__ bool newton_step(double esinEo, double ecosEo, double omec, double& dd, double& s,
                                            double& cm1, double& r) {
    const double f = fma(-esinEo, cm1, fma(-ecosEo, s, dd));
    const double fp = fma(esinEo, s, fma(-ecosEo, cm1, omec));
    r = rcp_nr1(fp);
    double d = -f * r;
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

// Acceptance bound of the first (short) Kepler correction: |d| < 7e-11 and
// F' > 0.49 (every orbit with e < 0.51 has F' = 1 - e cos E > 0.49), so
// d^2 < 4.9e-21 < 1e-20 F'.  Within it the error the correction leaves
// (d^2 / (2 F') < 5e-21), the raw reciprocal's share (|d| x 6e-8 < 1e-17)
// and the terms its rotation drops (s d^2 / 2 and d^3 / 6 from sin, d^2 / 2
// from cos - 1: below 1.2e-20) are all under one ulp of the quantities they
// touch (cos E - 1 only enters f, g-dot and F', all of order one).
constexpr double kShortD = 7e-11;
constexpr double kShortFp = 0.49;

// |r| and 1/|r| of the current position, from the position itself.
__device__ __forceinline__ void radius_from_position(Particle& p) {
    const double r2 = fma(p.rz, p.rz, fma(p.ry, p.ry, p.rx * p.rx));
    p.ir = rsqrt_nr1(r2);
    p.rl = r2 * p.ir;
}

// Kepler drift, FAST form.  inv_mu, sqrt_mu, inv_sqrt_mu: host constants;
// p.rl, p.ir must hold |r| and 1/|r| of the current position (carried from
// the previous drift or radius_from_position); the drift leaves those of the
// new position: |r(E)| = a F'(E) = |r0| + a (esinEo sin dE - ecosEo (cos dE
// - 1)), accumulated as that increment so an error of a (the refined
// reciprocal square root's, ~4e-15) only scales the increment, and
// 1/|r(E)| = (1/a) / F'(E).  A relative error d of the carried values moves
// the new state by O(d dE^2) only (f and g depend on a and e at that order).
//
// Straight-line path for the common case, then wave votes: lanes whose
// first correction is not provably converged recompute sin / cos of E and
// take up to two full corrections; lanes for which a shortcut does not
// hold (|dM| or |E| > 0.2 where the polynomial is evaluated, a solve not
// provably within 1e-14, an unbound orbit) take the f and g functions of
// the reference-arithmetic drift, which also sets their 0x04 / 0x08 flags.
// Returns the wave mask of those lanes (the only ones whose flags the drift
// can set).
__device__ __forceinline__ uint64_t drift_fast(Particle& p, double dt, double mu, double inv_mu,
                                               double dt_sqrt_mu, double inv_sqrt_mu, double inv_dt,
                                               const TaylorLead& lead) {
    const double rx0 = p.rx, ry0 = p.ry, rz0 = p.rz;
    const double vx0 = p.vx, vy0 = p.vy, vz0 = p.vz;
    const double v2 = fma(vz0, vz0, fma(vy0, vy0, vx0 * vx0));
    const double vdotr = fma(vz0, rz0, fma(vy0, ry0, vx0 * rx0));
    const double idist = p.ir;
    // 1/a = -2E/mu = 2/|r| - v^2/mu; bound orbits have 1/a > 0
    const double ia = fma(-v2, inv_mu, 2.0 * idist);
    const double rs_ia = rsqrt_nr1(ia);            // sqrt(a)
    const double a = rs_ia * rs_ia;
    const double sq_ia = ia * rs_ia;               // sqrt(1/a)
    const double in_ = inv_sqrt_mu * a * rs_ia;    // 1/n, n = sqrt(mu / a^3)
    const double omec = p.rl * ia;                 // 1 - e cos E0 = |r| / a
    const double ecosEo = 1.0 - omec;
    const double esinEo = vdotr * (inv_sqrt_mu * sq_ia);  // r.v / (n a^2)
    const double q = a * idist;                    // a / |r| = 1 / (1 - e cos E0)
    // |dM| <= 0.2 on this path, so no full turns are removed and the time
    // left to advance, dM / n, is dt itself
    const double dM = dt_sqrt_mu * ia * sq_ia;     // n dt
    // Kepler residual F(E) = E - ecosEo sin E + esinEo (1 - cos E) - dM and
    // F'(E) = 1 - ecosEo cos E + esinEo sin E, written with dd = E - dM and
    // (s, cm1) = (sin E, cos E - 1).  Start: the series of F in E,
    // (1 - c) E + (s0 / 2) E^2 + (c / 6) E^3 + O(E^4) = dM (c = ecosEo,
    // s0 = esinEo), reverted to third order: with t = dM / (1 - c),
    // A2 = s0 / (2 (1 - c)), A3 = c / (6 (1 - c)),
    // E = t - A2 t^2 + (2 A2^2 - A3) t^3 + O(t^4).  Its error is ~ e t^4 / 100
    // (below 3e-11 at t = 0.008, e = 0.1), so one correction usually ends
    // the solve; any error of the start is corrected by the Newton steps and
    // only decides how many run.
    const double h = esinEo * q;
    const double A2 = 0.5 * h;
    // 2 A2^2 - A3, with -A3 = -c q / 6 = (1 - q) / 6 (c q = q - 1, as
    // (1 - c) q = 1): one fused operation on the coefficient -1/6
    const double b3 = fma(A2, h, fma(q, kTaylor[0][0], -kTaylor[0][0]));
    const double t = dM * q;
    const double tu = t * fma(t, b3, -A2);
    const double E1 = fma(tu, t, t);
    // exact (Sterbenz) while E1 / dM lies in [1/2, 2], which holds for
    // e < 1/3; otherwise its rounding is one more part of the start's error
    double dd = E1 - dM;
    // sin / cos - 1 of the start: three Taylor terms suffice for |E1| <= 0.02
    // (the regime of the long-period swarms this path is built for); a lane
    // with a larger start takes the retry below, which evaluates five terms
    double smE, cm1;
    sincos_taylor3(E1, lead, smE, cm1);
    double s = E1 + smE;
    // First correction, short form: raw reciprocal and the rotation
    // sin(E + d) = s + d (1 + cm1), cos(E + d) - 1 = cm1 - s d, accepted only
    // within kShortD, kShortFp.  A Newton correction d leaves an error of at most
    // |F''| d^2 / (2 F') <= d^2 / (2 F') (|F''| <= e < 1).
    double f = fma(-esinEo, cm1, fma(-ecosEo, s, dd));
    double fp = fma(esinEo, s, fma(-ecosEo, cm1, omec));
    double r = __builtin_amdgcn_rcp(fp);
    const double d = -f * r;
    // sin E - E at the corrected E: (s + d + cm1 d) - (E1 + d)
    smE = fma(cm1, d, smE);
    const double s1 = fma(cm1, d, s + d);
    cm1 = fma(-s, d, cm1);
    s = s1;
    // A lane inside these bounds needs no test of |dM|: its corrected E is
    // the Kepler equation's root to within 1e-20 and |E| <= 0.02 + 7e-11,
    // so |dM| = |E - ecosEo sin E + esinEo (1 - cos E)| <= 0.0402 (e < 1
    // for 1/a > 0; an unbound orbit makes every quantity here NaN, which
    // fails the first test).
    const uint64_t retry =
        wave_mask(!(fabs(d) < kShortD)) | wave_mask(!(fp > kShortFp)) | wave_mask(!(fabs(E1) <= kMaxStart));
    // Lanes outside that bound (large steps or eccentric orbits) restart from
    // sin / cos of the corrected E and take full corrections (accurate
    // reciprocal, third-order rotation, the reference's 1e-14 tolerance as a
    // proven bound); these run only in waves where some lane needs them.
    // Lanes for the reference-arithmetic drift (bad): |dM| or the corrected
    // |E| above kMaxDM, or no proven convergence.
    uint64_t bad = 0;
    if (retry != 0) {
        bool conv = true, far = false;
        if (lane_in(retry)) {
            dd += d;
            const double E2 = dM + dd;
            far = !(fabs(E2) <= kMaxDM) || !(fabs(dM) <= kMaxDM);
            sincos_taylor<5>(E2, s, cm1);
            conv = newton_step(esinEo, ecosEo, omec, dd, s, cm1, r);
        }
        if (wave_any(!conv)) {
            if (!conv) conv = newton_step(esinEo, ecosEo, omec, dd, s, cm1, r);
        }
        if (lane_in(retry)) smE = s - (dM + dd);
        bad = wave_mask(far) | wave_mask(!conv);
    }
    // F'(E) - (1 - e cos E0) = esinEo sin dE - ecosEo (cos dE - 1)
    const double dfp = fma(esinEo, s, -ecosEo * cm1);
    fp = omec + dfp;
    // 1/F' at the solution: one Newton step from the reciprocal of the last
    // correction's F', which differs from it by |d| e relative at most
    const double ifp = fma(r, fma(-fp, r, 1.0), r);
    p.rl = fma(a, dfp, p.rl);
    p.ir = ia * ifp;
    double fg = fma(cm1, q, 1.0);
    double g = fma(smE, in_, dt);
    double fdot = -(t * inv_dt) * s * ifp;         // n q = t / dt
    double gdot = fma(cm1, ifp, 1.0);
    // Lanes outside the shortcuts take the f and g functions of the
    // reference-arithmetic drift (and its 0x04 / 0x08 flags) instead; the
    // state itself is then updated once, in place, for every lane.
    if (bad != 0) {
        if (lane_in(bad)) p.flags = kepler_fg_fp64<false>(rx0, ry0, rz0, vx0, vy0, vz0, dt, mu, fg, g, fdot, gdot);
    }
    p.vx = fma(rx0, fdot, vx0 * gdot);
    p.vy = fma(ry0, fdot, vy0 * gdot);
    p.vz = fma(rz0, fdot, vz0 * gdot);
    p.rx = fma(rx0, fg, vx0 * g);
    p.ry = fma(ry0, fg, vy0 * g);
    p.rz = fma(rz0, fg, vz0 * g);
    if (bad != 0) {
        if (lane_in(bad)) radius_from_position(p);
    }
    return bad;
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
// crossed one.  The Sun and outer tests vote on the carried |r| against
// radii widened by 1e-9 relative (sun_vote, outer_vote), far beyond its
// error, and the exact tests of |r|^2 run in the voting waves, so the flags
// are those of the exact tests.  hm, hrh2: the GM and squared radii read on
// the main path (the values of pc.m, pc.rh2).
// Returns the wave mask of lanes inside some radius (the only ones it can
// flag).
template <int NB, bool MIXED, int NMAX = (NB > 0 ? NB : GLX_MAX_BODIES)>
__device__ __forceinline__ uint64_t accel_vote(Particle& p, const double* __restrict__ row, const PlanetConsts& pc,
                                               const double* __restrict__ hm, const double* __restrict__ hrh2,
                                               double sun_vote, double outer_vote, double outer2, uint32_t nb) {
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
        close |= wave_mask(rad < hrh2[j]);
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
        // m (1 + 1.5 e): 1.5 m is loop-invariant and stays in a register
        const double fac = (y2 * y) * fma(e, 1.5 * hm[j], hm[j]);
        ax = fma(-dx, fac, ax);
        ay = fma(-dy, fac, ay);
        az = fma(-dz, fac, az);
    }
    p.ax = ax;
    p.ay = ay;
    p.az = az;
    const uint64_t event = close | wave_mask(p.rl < sun_vote) | wave_mask(p.rl > outer_vote);
    if (event != 0) {
        const double r2 = fma(p.rz, p.rz, fma(p.ry, p.ry, p.rx * p.rx));
        if (lane_in(close) && p.flags == 0) {
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
    return event;
}

__device__ __forceinline__ void kick_fma(Particle& p, double half_dt) {
    p.vx = fma(p.ax, half_dt, p.vx);
    p.vy = fma(p.ay, half_dt, p.vy);
    p.vz = fma(p.az, half_dt, p.vz);
}

// Bodies up to which a FAST / MIXED step copies its planet row from LDS into
// registers ahead of the drift (3 doubles per body); with more bodies the
// copy would cost more registers than the read latency it hides.
constexpr int kRowRegBodies = 5;

// One FAST / MIXED step of every active lane (none of them flagged).  C:
// the launch's constants (dt, half_dt, mu, inv_mu, sqrt_mu, inv_sqrt_mu,
// pc, outer2); hm, hrh2: the bodies' GM and squared radii as the planet sum
// reads them every step (pc's own arrays, or copies held in registers);
// rowp: the step's planet row in LDS; death: the death index a particle
// flagged in this step records.  kick_in: the kick before the drift,
// dt / 2 at the first step of a run of steps, else dt (the previous step's
// closing half kick and this step's opening one in one operation: the
// FAST / MIXED forms leave a step's closing half kick pending, see
// chunk_steps).  A lane flagged in this step gets its closing half kick at
// once, as the reference completes the flagging step.  Only the rare
// branches that can set a flag (the reference-arithmetic drift, a radius
// crossed) look at the flags.  Returns whether some lane of the wave was
// flagged.
template <int NB, int FORM, int NMAX = (NB > 0 ? NB : GLX_MAX_BODIES), class C>
__device__ __forceinline__ bool fast_step(Particle& p, uint32_t& dti, const C& A, const double* hm,
                                          const double* hrh2, const double* rowp, uint32_t nb, uint32_t death,
                                          double kick_in, const TaylorLead& lead) {
    uint64_t event;
    if constexpr (NB > 0 && NB <= kRowRegBodies) {
        // the step's planet row is read from LDS before the drift, so the
        // read latency hides behind it
        double rowv[3 * NB];
#pragma unroll
        for (int k = 0; k < 3 * NB; k++) rowv[k] = rowp[k];
        kick_fma(p, kick_in);
        event = drift_fast(p, A.dt, A.mu, A.inv_mu, A.dt_sqrt_mu, A.inv_sqrt_mu, A.inv_dt, lead);
        event |= accel_vote<NB, FORM == GLX_FORM_MIXED, NMAX>(p, rowv, A.pc, hm, hrh2, A.sun_vote, A.outer_vote,
                                                              A.outer2, nb);
    } else {
        kick_fma(p, kick_in);
        event = drift_fast(p, A.dt, A.mu, A.inv_mu, A.dt_sqrt_mu, A.inv_sqrt_mu, A.inv_dt, lead);
        event |= accel_vote<NB, FORM == GLX_FORM_MIXED, NMAX>(p, rowp, A.pc, hm, hrh2, A.sun_vote, A.outer_vote,
                                                              A.outer2, nb);
    }
    if (__builtin_expect(event != 0, 0)) {
        if (p.flags != 0) {
            dti = death;
            kick_fma(p, A.half_dt);
        }
        return wave_mask(p.flags != 0) != 0;
    }
    return false;
}

// One step of one particle, any form (skipped once it is flagged); the
// arguments are those of fast_step (kick_in is not used by the fp64 form,
// which kicks twice per step as the reference does).
template <int NB, int FORM, int NMAX = (NB > 0 ? NB : GLX_MAX_BODIES), class C>
__device__ __forceinline__ void particle_step(Particle& p, uint32_t& dti, const C& A, const double* hm,
                                              const double* hrh2, const double* rowp, uint32_t nb,
                                              uint32_t death, double kick_in, const TaylorLead& lead) {
    if (p.flags != 0) return;
    if constexpr (FORM == GLX_FORM_FP64) {
        kick(p, A.half_dt);
        drift_fp64(p, A.dt, A.mu, A.fp64_core != 0, lead.q);
        accel_fp64_step<NB, NMAX>(p, rowp, A.pc, A.outer2, nb, A.fp64_core != 0);
        kick(p, A.half_dt);
        if (p.flags != 0) dti = death;
    } else {
        fast_step<NB, FORM, NMAX>(p, dti, A, hm, hrh2, rowp, nb, death, kick_in, lead);
    }
}

// The closing half kick left pending by the FAST / MIXED steps, for every
// lane not flagged (flagged lanes completed their last step when flagged,
// and lanes flagged before the launch did not step).
__device__ __forceinline__ void close_kick(Particle& p, double half_dt, bool& pending) {
    if (pending && p.flags == 0) kick_fma(p, half_dt);
    pending = false;
}

// Steps [0, cn) of a staged chunk whose rows start at `rows` (row: doubles
// per step); death0 + s + 1 is step s's death index, g0 + s its index within
// its block.  In the FAST / MIXED forms a wave with no flagged lane steps
// without the per-lane skip, two steps per iteration: with neither the
// skip's merge nor a loop-carried copy, the state registers of consecutive
// steps alternate.  Once one of its lanes is flagged (and for an odd last
// step) the wave finishes the chunk step by step.
//
// FAST / MIXED: every step of the block that is a multiple of
// kRadiusPeriod starts a run of steps: the pending closing half kick of
// the step before is applied and the carried radius recomputed from the
// position (wave-uniform tests).  Within a run, a step's closing half kick
// and the next step's opening one are one kick by dt.  `pending` (wave
// uniform) says whether the last step's closing half kick is still due; the
// caller applies it (close_kick) at the end of the launch, so a launch ends,
// like every run, on completed steps, and launches split on multiples of
// kRadiusPeriod give the same bits as one launch.
template <int NB, int FORM, int NMAX = (NB > 0 ? NB : GLX_MAX_BODIES), class C>
__device__ __forceinline__ void chunk_steps(Particle& p, uint32_t& dti, const C& A, const double* hm,
                                            const double* hrh2, const double* rows, uint32_t row, uint32_t cn,
                                            uint32_t nb, uint32_t death0, uint32_t g0, bool& pending,
                                            const TaylorLead& lead) {
    constexpr uint32_t kMask = kRadiusPeriod - 1;
    static_assert((kRadiusPeriod & kMask) == 0, "power of two");
    uint32_t s = 0;
    if constexpr (FORM != GLX_FORM_FP64) {
        if (wave_mask(p.flags != 0) == 0) {
            while (s + 1 < cn) {
                // the pair's row offset held in a vector register: both
                // steps read their rows from it (the second through the
                // LDS instruction's offset field), one register move per
                // pair instead of one per step
                uint32_t off = s * row;
                __asm__ volatile("" : "+v"(off));
                const double* pair = rows + off;
                if (((g0 + s) & kMask) == 0) {
                    close_kick(p, A.half_dt, pending);
                    radius_from_position(p);
                }
                bool flagged = fast_step<NB, FORM, NMAX>(p, dti, A, hm, hrh2, pair, nb, death0 + s + 1,
                                                         pending ? A.dt : A.half_dt, lead);
                pending = true;
                s++;
                if (flagged) break;
                if (((g0 + s) & kMask) == 0) {
                    close_kick(p, A.half_dt, pending);
                    radius_from_position(p);
                }
                flagged = fast_step<NB, FORM, NMAX>(p, dti, A, hm, hrh2, pair + row, nb, death0 + s + 1,
                                                    pending ? A.dt : A.half_dt, lead);
                pending = true;
                s++;
                if (flagged) break;
            }
        }
    }
    for (; s < cn; s++) {
        if constexpr (FORM != GLX_FORM_FP64) {
            if (((g0 + s) & kMask) == 0) {
                close_kick(p, A.half_dt, pending);
                radius_from_position(p);
            }
        }
        particle_step<NB, FORM, NMAX>(p, dti, A, hm, hrh2, rows + s * row, nb, death0 + s + 1,
                                      pending ? A.dt : A.half_dt, lead);
        if constexpr (FORM != GLX_FORM_FP64) pending = true;
    }
}

}  // namespace glx
