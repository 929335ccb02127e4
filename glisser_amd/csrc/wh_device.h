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
    uint32_t fp64_core;  // fp64 form: fp64_core_constants(mu, dt)
    double inv_mu;       // 1/mu, dt sqrt(mu), 1/sqrt(mu), 1/dt: FAST / MIXED drift
    double dt_sqrt_mu;
    double inv_sqrt_mu;
    double inv_dt;
    double sun_vote;     // FAST / MIXED: carried |r| below this or above
    double outer_vote;   // this triggers the exact Sun / outer tests
    PlanetConsts pc;
};

// Queued multi-block launch (one persistent grid for several consecutive
// blocks).  A task is one 256-particle tile through one segment of steps.
// Tasks wait in a FIFO (`ring`) that holds only runnable tasks: it starts
// with every tile's first segment, and the workgroup that finishes a tile's
// segment k appends segment k + 1.  Entries are task + 1 (task = segment x
// n_tiles + tile); 0 marks a slot not written yet.
struct QueueSeg {
    uint32_t block;    // index into QueueArgs::blocks
    uint32_t r0;       // first step of the segment within its block (a multiple of kRadiusPeriod)
    uint32_t n_steps;
    uint32_t pad;
};
// What differs between the blocks of one queued launch.
struct QueueBlock {
    const double* log;     // the block's planet rows
    uint32_t step_offset;  // steps of the launch before this block
    uint32_t pad;
    double sun_vote;
    PlanetConsts pc;
};
struct QueueArgs {
    BlockArgs base;          // particle arrays and the constants the blocks share
    const QueueSeg* segs;
    const QueueBlock* blocks;
    uint32_t* ctl;           // [0] next slot to take, [1] next slot to fill, [2] nonzero once a wait gave up
    uint32_t* ring;          // n_segs x n_tiles slots
    uint32_t n_segs;
    uint32_t n_tiles;
    uint64_t wait_ticks;     // bound of a slot wait, in ticks of the 100 MHz real-time counter
};

struct Particle {
    double rx, ry, rz, vx, vy, vz, ax, ay, az;
    double rl, ir;  // |r| and 1/|r| carried between steps (FAST / MIXED forms only)
    uint32_t flags;
};

// Correctly rounded FP64 divide and square root (round to nearest) for
// operands in a range where the compiler's IEEE expansions rescale and fix
// up nothing.  For gfx950 hipcc expands a / b into v_div_scale_f64 of b and
// of a, r = v_rcp_f64(b), two Newton steps r = r + r (1 - b r),
// q = a r, q = q + (a - b q) r (v_div_fmas_f64) and v_div_fixup_f64, and
// sqrt(x) into a range test with v_ldexp_f64, y = v_rsq_f64(x), the
// Goldschmidt iteration below and a class test for zero and infinity.  The
// scaling steps change nothing while a, b, a / b and 1 / b are normal, |a| >=
// 2^-969 and the exponents of a and b differ by less than 768; the fix-up
// changes nothing for finite nonzero normal results; the square root's
// rescaling and class test change nothing for normal x >= 2^-767.  Inside
// those bounds the cores below are the same operations, so they return
// the same correctly rounded results with 3 (divide) and 8 (square root)
// instructions fewer.  The fp64 form uses them only where guards on the
// particle's state bound every operand well inside (kepler_fg_fp64,
// accel_fp64) and recomputes any other lane with the IEEE operators.

// a / b for b > 0 and a = 0 or a, b, a / b normal inside the bounds above
// (used for |a| in [2^-400, 2^300], b in [2^-300, 2^300]).  The expansion is
// run on -a and its result negated (round to nearest is symmetric), which
// also gives the IEEE sign of a zero quotient for either sign of a zero a.
__device__ __forceinline__ double core_div_pos(double a, double b) {
#pragma clang fp contract(off)
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    const double q = (-a) * r;
    const double rem = fma(-b, q, -a);
    return fma(-rem, r, -q);
}

// sqrt(x) for normal x inside the bounds above (used for x in [2^-700,
// 2^700]).
__device__ __forceinline__ double core_sqrt(double x) {
#pragma clang fp contract(off)
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    double d = fma(-g, g, x);
    h = fma(h, r, h);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}

// Whether a launch's constants admit the guarded cores: mu (GM of the
// central body) and dt in [2^-100, 2^100].  Every bound in the guards of
// kepler_fg_fp64 and accel_fp64 starts from these.
__host__ __device__ inline bool fp64_core_constants(double mu, double dt) {
    return mu >= 0x1p-100 && mu <= 0x1p100 && dt >= 0x1p-100 && dt <= 0x1p100;
}

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
//
// kCore: the divides and square roots whose operands the guards below bound
// run as core_div_pos / core_sqrt (same results there); ok is cleared for a
// lane outside the guards, whose results are then to be recomputed with
// kCore false.  The guards, given mu and dt in [2^-100, 2^100]
// (fp64_core_constants): |r|^2 in [2^-180, 2^200] and |v|^2 <= 2^200, so
// |r| in [2^-90, 2^100], mu / |r| in [2^-200, 2^190] and a nonzero energy
// (a difference of doubles of at least 2^-201) in [2^-253, 2^200]; then
// a = 0.5 mu / -energy in [|r| / 2, 2^352], guarded a <= 2^100, so a^3 in
// [2^-274, 2^300], mu / a^3 in [2^-400, 2^374], n in [2^-200, 2^187],
// |r| / a in [2^-191, 2], dM = dt n (no turns removed; the rare other
// branch clears ok) in [2^-300, 1.5) and dM / n near dt.  Each Newton
// denominator F' is guarded >= 2^-50, which keeps |dE| below 2^254 over
// five iterations (|F| <= |dE| + 5); a nonzero residual F = X - dM is then
// at least 2^-353 (Sterbenz, or larger than dM / 2).  sin dE - dE is 0 (sin
// rounds to dE below 2^-27) or at least 2^-80, cos dE - 1 is 0 or at least
// 2^-54 in magnitude, and the final F' is guarded >= 2^-50.  e sin E0 =
// r.v / (n a^2) (r.v may be arbitrarily small) and fdot's quotient (sin dE
// may be) keep the IEEE divide.
// ql: the correctly rounded sin/cos's short-tier constants (cr::QuickLead),
// possibly held in registers by the caller.
template <bool kCrTrig = true, bool kCore = false>
__device__ __forceinline__ uint32_t kepler_fg_fp64(double rx, double ry, double rz, double vx, double vy,
                                                   double vz, double dt, double mu, double& f, double& g,
                                                   double& fdot, double& gdot, bool* ok = nullptr,
                                                   const cr::QuickLead& ql = cr::kQuickLead) {
#pragma clang fp contract(off)
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double v2 = vx * vx + vy * vy + vz * vz;
    double dist;
    if constexpr (kCore) {
        *ok = *ok && r2 >= 0x1p-180 && r2 <= 0x1p200 && v2 <= 0x1p200;
        dist = core_sqrt(r2);
    } else {
        dist = sqrt(r2);
    }
    const double vdotr = vx * rx + vy * ry + vz * rz;
    const double energy = v2 * 0.5 - (kCore ? core_div_pos(mu, dist) : mu / dist);
    if (energy >= 0.0) {
        // kCore (called by drift_fp64 only, which leaves an unbound orbit's
        // state as it is): the identity f and g need not be set
        if constexpr (!kCore) {
            f = 1.0;
            g = 0.0;
            fdot = 0.0;
            gdot = 1.0;
        }
        return kFlagUnbound;
    }
    // -0.5 mu / energy: the same quotient as 0.5 mu / -energy
    const double a = kCore ? core_div_pos(0.5 * mu, -energy) : -0.5 * mu / energy;
    if constexpr (kCore) *ok = *ok && a <= 0x1p100;
    const double n_ = kCore ? core_sqrt(core_div_pos(mu, a * a * a)) : sqrt(mu / (a * a * a));
    const double ecosEo = 1.0 - (kCore ? core_div_pos(dist, a) : dist / a);
    const double esinEo = vdotr / (n_ * a * a);
    // dt*n - (pi*2)*(int)((dt*n/pi)*2): the reference's M_2PI macro
    // expansion.  For 0 <= dt*n < 1.5 the quotient lies below 0.955, the
    // integer part is 0 and the expression is dt*n itself (dt*n - 0.0), so
    // the divide only runs outside that range.
    const double dtn = dt * n_;
    double dM = dtn;
    if (!(dtn >= 0.0 && dtn < 1.5)) {
        dM = dtn - kPi * 2 * (double)(int)(dtn / kPi * 2);
        if constexpr (kCore) *ok = false;
    }
    const double adv_dt = kCore ? core_div_pos(dM, n_) : dM / n_;
    double sM, cM;
    if constexpr (kCrTrig) cr::sincos_cr(dM, &sM, &cM, ql);
    else sincos(dM, &sM, &cM);
    double dE = dM - esinEo + esinEo * cM + ecosEo * sM;
    double sindE, cosdE;
    if constexpr (kCrTrig) cr::sincos_cr(dE, &sindE, &cosdE, ql);
    else sincos(dE, &sindE, &cosdE);
    double delta = 0.0;
#pragma unroll
    for (int k = 0; k < kKepIts; k++) {
        const double fk = dE - ecosEo * sindE + esinEo * (1. - cosdE) - dM;
        const double fpk = 1. - ecosEo * cosdE + esinEo * sindE;
        if constexpr (kCore) {
            *ok = *ok && fpk >= 0x1p-50;
            delta = core_div_pos(-fk, fpk);
        } else {
            delta = -fk / fpk;
        }
        const double prev = dE;
        dE += delta;
        // sin and cos depend on the argument's bits only: in the last two
        // iterations the iterate has mostly stopped moving, and a lane whose
        // dE kept its bits keeps its sin and cos (a wave skips the
        // evaluation when all of its lanes did)
        if (k < kKepIts - 2 || __double_as_longlong(dE) != __double_as_longlong(prev)) {
            if constexpr (kCrTrig) cr::sincos_cr(dE, &sindE, &cosdE, ql);
            else sincos(dE, &sindE, &cosdE);
        }
    }
    const double fp = 1.0 - ecosEo * cosdE + esinEo * sindE;
    if constexpr (kCore) {
        *ok = *ok && fp >= 0x1p-50;
        f = 1.0 + core_div_pos(a * (cosdE - 1.0), dist);
        g = adv_dt + core_div_pos(sindE - dE, n_);
        gdot = 1.0 + core_div_pos(cosdE - 1.0, fp);
    } else {
        f = 1.0 + a * (cosdE - 1.0) / dist;
        g = adv_dt + (sindE - dE) / n_;
        gdot = 1.0 + (cosdE - 1.0) / fp;
    }
    fdot = -n_ * sindE * a / (dist * fp);
    return fabs(delta) > kTolKep ? kFlagKepler : 0u;
}

// Kepler drift in the reference's arithmetic: kepler_fg_fp64 applied
// without FMA contraction, as the reference does.  core (wave-uniform:
// fp64_core_constants of the launch): the guarded divide and square-root
// cores, with the lanes outside their guards recomputed with the IEEE
// operators (the same bits either way).
__device__ __forceinline__ void drift_fp64(Particle& p, double dt, double mu, bool core,
                                           const cr::QuickLead& ql = cr::kQuickLead) {
#pragma clang fp contract(off)
    double f, g, fdot, gdot;
    bool ok = core;
    uint32_t fl = 0;
    if (core)
        fl = kepler_fg_fp64<true, true>(p.rx, p.ry, p.rz, p.vx, p.vy, p.vz, dt, mu, f, g, fdot, gdot, &ok, ql);
    if (!ok) fl = kepler_fg_fp64(p.rx, p.ry, p.rz, p.vx, p.vy, p.vz, dt, mu, f, g, fdot, gdot, nullptr, ql);
    p.flags = fl;
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
// NMAX.  kCore: each planet's 1 / (|dr|^2 sqrt(|dr|^2)) with core_sqrt and
// core_div_pos, guarded |dr|^2 in [2^-200, 2^200] (so |dr|^3 in [2^-300,
// 2^300]); ok is cleared for a lane outside the guard.
template <int NB, int NMAX = (NB > 0 ? NB : GLX_MAX_BODIES), bool kCore = false>
__device__ __forceinline__ void accel_fp64(Particle& p, const double* __restrict__ row, const PlanetConsts& pc,
                                           double outer2, uint32_t nb, bool* ok = nullptr) {
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
        double inv3;
        if constexpr (kCore) {
            *ok = *ok && rad >= 0x1p-200 && rad <= 0x1p200;
            inv3 = core_div_pos(1., rad * core_sqrt(rad));
        } else {
            inv3 = 1. / (rad * sqrt(rad));
        }
        const double fac = pc.m[j] * inv3;
        p.ax -= dx * fac;
        p.ay -= dy * fac;
        p.az -= dz * fac;
    }
    const double rad = p.rx * p.rx + p.ry * p.ry + p.rz * p.rz;
    if (rad < pc.rh2[0]) p.flags = kFlagEncounter;
    if (rad > outer2) p.flags = kFlagOut;
}

// accel_fp64 of one fp64 step: with the guarded cores when core (wave
// uniform), a lane outside their guards recomputed with the IEEE operators
// from its flags before the call (the same bits either way).
template <int NB, int NMAX = (NB > 0 ? NB : GLX_MAX_BODIES)>
__device__ __forceinline__ void accel_fp64_step(Particle& p, const double* __restrict__ row, const PlanetConsts& pc,
                                                double outer2, uint32_t nb, bool core) {
    const uint32_t flags0 = p.flags;
    bool ok = core;
    if (core) accel_fp64<NB, NMAX, true>(p, row, pc, outer2, nb, &ok);
    if (!ok) {
        p.flags = flags0;
        accel_fp64<NB, NMAX>(p, row, pc, outer2, nb);
    }
}

}  // namespace glx
