// CPU oracle for the GLISSER Wisdom-Holman particle path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain C++, the CPU
// algorithm of the reference (src/wh.cpp, src/convert.cpp, src/data.cpp) so
// that the HIP path can be checked against it.  Only tests/, the smoke entry
// point and the CPU-baseline leg of bench.py may load the library built from
// it.  Nothing in the product links it.
//
// Every function keeps the reference's floating-point evaluation order
// (left-to-right sums, the same association of products, the M_2PI macro
// expansion) because the oracle is pinned bit-for-bit against the reference's
// own CPU sources compiled in this container (oracle/_ref) and against the
// golden vectors under tests/golden.  It must therefore be compiled like the
// reference's CPU build: g++ -O3 with FP contraction off (no FMA).
//
// Pinning status per function (what checks it besides the reference build):
//   kepler_fixed / kepler_var  : closed-form two-body propagation, golden vectors
//   drift                      : composition, reversibility, E and L invariance
//   particle_accel             : brute-force direct sum on tiny inputs, golden
//   step_particles             : golden vectors, zero-mass limit == pure Kepler
//   planet step / h0           : golden vectors, energy/L conservation
//   stable_partition           : std::stable_partition property tests
//   to_elements/from_elements  : round trip, circular/planar special cases
//   hermite_log                : exact for cubics, node reproduction,
//                                fourth-order convergence on Kepler orbits
//   correctly rounded sin/cos  : (the checker's alternative particle libm)
//                                200-bit mpmath values, tests/test_sincos.py
//   planet history (plhist_*)  : writer bytes and interpolated blocks equal
//                                the reference build's and its golden blocks;
//                                intervals end on the records' own elements;
//                                Keplerian records give Kepler motion
//
// Error behaviour: where the reference throws std::runtime_error (unbound
// orbit, unconverged Kepler solve; src/wh.cpp:894-908,942-945) the oracle
// returns a non-zero status and the offending index instead.

#include <quadmath.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <fstream>
#include <limits>
#include <numeric>

extern "C" void orc_load_h0(uint32_t n_bodies, const double* m, const double* r_log, uint32_t n_steps,
                            double* h0_log);

namespace {

// The reference's pi macro and its unparenthesised "M_PI * 2" twin
// (src/types.h:7-8).  The 2-pi expression is spelled out at each use so the
// evaluation order matches the macro expansion exactly.
constexpr double kPi = 3.14159265358979323846;

struct V3 {
    double x, y, z;
};

inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
inline V3 add(const V3& a, const V3& b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 sub(const V3& a, const V3& b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 scale(const V3& a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 divs(const V3& a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
inline V3 neg(const V3& a) { return v3(-a.x, -a.y, -a.z); }
inline void add_to(V3& a, const V3& b) { a.x += b.x; a.y += b.y; a.z += b.z; }
inline void sub_from(V3& a, const V3& b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; }
// x*x + y*y + z*z evaluated left to right (src/types.h:98-101)
inline double lensq(const V3& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline V3 cross(const V3& a, const V3& b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

inline V3 load3(const double* p, size_t i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
inline void store3(double* p, size_t i, const V3& a) {
    p[3 * i] = a.x;
    p[3 * i + 1] = a.y;
    p[3 * i + 2] = a.z;
}

constexpr double kTolKep = 1e-14;     // src/wh.cpp:15
constexpr uint32_t kMaxKepVar = 10;   // src/wh.cpp:14
constexpr uint32_t kParticleKepIts = 5;  // src/wh.cpp:934

// The reference's Kepler solvers return sin/cos through pointers, so its
// x86 build calls glibc sin() and cos() separately there, while the drift's
// initial guess (local variables) is fused into one sincos() call.  glibc's
// sincos() and sin()/cos() round differently for some arguments, so the
// oracle reproduces both call forms explicitly: separate calls through
// pointers the compiler cannot fuse, and a direct sincos() for the guess.
double (*volatile const g_sin)(double) = ::sin;
double (*volatile const g_cos)(double) = ::cos;
inline void sin_cos_separate(double x, double& s, double& c) {
    s = g_sin(x);
    c = g_cos(x);
}
inline void sin_cos_fused(double x, double& s, double& c) { ::sincos(x, &s, &c); }

// The C library the particle drift (the fixed-iteration path) calls, per
// calling thread: 0 (default) glibc as above, the reference's; 1 correctly
// rounded sin and cos -- glibc's long-double sinl / cosl, or libquadmath's
// 113-bit sinq / cosq where those cannot decide, rounded to the nearest
// double: a correctly rounded evaluation independent of the device's -- with a
// count of the calls in which glibc would have returned another double.
// The planet drift always calls glibc.  A checker's option: it shows which
// differences between the device's correctly rounded sin/cos and the
// reference build come from glibc's rounding alone.
// The long-double values (glibc's sinl / cosl, within a few units of 2^-64
// relative) decide the rounding whenever both ends of a 2^-60 relative
// interval around them round to the same double (over 99 % of calls); the
// rest take the 113-bit evaluation.
thread_local int tl_particle_libm = 0;
thread_local uint64_t tl_libm_calls = 0, tl_libm_differ = 0;
inline bool round_decided_ld(long double v, double& out) {
    const long double b = fabsl(v) * 0x1p-60L;
    const double lo = (double)(v - b), hi = (double)(v + b);
    out = lo;
    return lo == hi;
}
inline double sin_cr(double x) {
    double d;
    if (round_decided_ld(sinl((long double)x), d)) return d;
    return (double)sinq((__float128)x);
}
inline double cos_cr(double x) {
    double d;
    if (round_decided_ld(cosl((long double)x), d)) return d;
    return (double)cosq((__float128)x);
}
inline void particle_sin_cos(double x, double& s, double& c, bool fused) {
    if (tl_particle_libm == 0) {
        if (fused) sin_cos_fused(x, s, c);
        else sin_cos_separate(x, s, c);
        return;
    }
    s = sin_cr(x);
    c = cos_cr(x);
    double gs, gc;
    if (fused) sin_cos_fused(x, gs, gc);
    else sin_cos_separate(x, gs, gc);
    tl_libm_calls += 2;
    tl_libm_differ += (gs != s) + (gc != c);
}

// Newton iteration on the eccentric-anomaly difference form of Kepler's
// equation with a fixed iteration count; src/wh.cpp:46-65.
// Returns true on "error" (last correction not below the tolerance).
bool kepler_fixed(double dM, double esinEo, double ecosEo, double& dE, double& sindE, double& cosdE,
                  uint32_t its) {
    double delta = 1;
    particle_sin_cos(dE, sindE, cosdE, false);
    for (uint32_t i = 0; i < its; i++) {
        double f = dE - ecosEo * sindE + esinEo * (1. - cosdE) - dM;
        double fp = 1. - ecosEo * cosdE + esinEo * sindE;
        delta = -f / fp;
        dE += delta;
        particle_sin_cos(dE, sindE, cosdE, false);
    }
    return std::fabs(delta) >= kTolKep;
}

// Same iteration with early exit; src/wh.cpp:19-43 (used by the planet drift).
bool kepler_var(double dM, double esinEo, double ecosEo, double& dE, double& sindE, double& cosdE) {
    sin_cos_separate(dE, sindE, cosdE);
    for (uint32_t i = 0; i < kMaxKepVar; i++) {
        double f = dE - ecosEo * sindE + esinEo * (1. - cosdE) - dM;
        double fp = 1. - ecosEo * cosdE + esinEo * sindE;
        double delta = -f / fp;
        if (std::fabs(delta) < kTolKep) return false;
        dE += delta;
        sin_cos_separate(dE, sindE, cosdE);
    }
    return true;
}

enum Status : int { kOk = 0, kUnbound = 1, kUnconverged = 2, kBadInput = 3 };

// Heliocentric Kepler drift of entries [start, start+n) by time t about the
// per-entry central mass mu[i]; masked entries keep r and v.  Restates
// WHIntegrator::drift<fixed> (src/wh.cpp:875-957), including its two passes
// (distances/energies for every entry first, masked or not).
int drift(double t, V3* r, V3* v, size_t start, size_t n, const double* mu, const uint8_t* mask, bool fixed,
          size_t* bad) {
    std::vector<double> dist(n), energy(n), vdotr(n);
    for (size_t k = 0; k < n; k++) {
        size_t i = start + k;
        dist[k] = std::sqrt(lensq(r[i]));
        energy[k] = lensq(v[i]);
        vdotr[k] = v[i].x * r[i].x + v[i].y * r[i].y + v[i].z * r[i].z;
    }
    for (size_t k = 0; k < n; k++) {
        size_t i = start + k;
        energy[k] *= 0.5;
        energy[k] -= mu[i] / dist[k];
    }
    for (size_t k = 0; k < n; k++) {
        size_t i = start + k;
        if (mask && mask[i]) continue;
        if (energy[k] >= 0) {
            if (bad) *bad = i;
            return kUnbound;
        }
        V3 r0 = r[i];
        V3 v0 = v[i];
        double a = -0.5 * mu[i] / energy[k];
        double n_ = std::sqrt(mu[i] / (a * a * a));
        double ecosEo = 1.0 - dist[k] / a;
        double esinEo = vdotr[k] / (n_ * a * a);
        // "t * n_ - M_2PI * (int) (t * n_ / M_2PI)" with M_2PI = M_PI * 2
        double dM = t * n_ - kPi * 2 * (int)(t * n_ / kPi * 2);
        double adv_dt = dM / n_;
        double sM, cM;
        if (fixed) particle_sin_cos(dM, sM, cM, true);
        else sin_cos_fused(dM, sM, cM);
        double dE = dM - esinEo + esinEo * cM + ecosEo * sM;
        double sindE, cosdE;
        bool err = fixed ? kepler_fixed(dM, esinEo, ecosEo, dE, sindE, cosdE, kParticleKepIts)
                         : kepler_var(dM, esinEo, ecosEo, dE, sindE, cosdE);
        if (err) {
            if (bad) *bad = i;
            return kUnconverged;
        }
        double fp = 1.0 - ecosEo * cosdE + esinEo * sindE;
        double f = 1.0 + a * (cosdE - 1.0) / dist[k];
        double g = adv_dt + (sindE - dE) / n_;
        double fdot = -n_ * sindE * a / (dist[k] * fp);
        double gdot = 1.0 + (cosdE - 1.0) / fp;
        r[i] = add(scale(r0, f), scale(v0, g));
        v[i] = add(scale(r0, fdot), scale(v0, gdot));
    }
    return kOk;
}

// Acceleration and encounter/cull flags of one particle from one planet-log
// row; WHIntegrator::helio_acc_particle (src/wh.cpp:611-657) without FMA.
// planet_r holds bodies 1..n_bodies-1 (the Sun is not logged).
// deathtime (indexed by particle id) receives (float)time when a flag is set.
void particle_accel(uint32_t n_bodies, const double* m, const double* rh, double outer, const V3* planet_r,
                    const V3& h0, double time, const V3& r, V3& a, uint16_t& flags, uint32_t id,
                    double* deathtime) {
    a = h0;
    for (uint32_t j = 1; j < n_bodies; j++) {
        V3 dr = sub(r, planet_r[j - 1]);
        double rji2 = lensq(dr);
        double irij3 = 1. / (rji2 * std::sqrt(rji2));
        double fac = m[j] * irij3;
        sub_from(a, scale(dr, fac));
        if (rji2 < rh[j] * rh[j]) {
            if (deathtime) deathtime[id] = static_cast<float>(time);
            flags = static_cast<uint16_t>((j << 8) | 0x0001);
        }
    }
    double rji2 = lensq(r);
    if (rji2 < rh[0] * rh[0]) {
        if (deathtime) deathtime[id] = static_cast<float>(time);
        flags = 0x0001;
    }
    if (rji2 > outer * outer) {
        if (deathtime) deathtime[id] = static_cast<float>(time);
        flags = 0x0002;
    }
}

// ---- planets (the producer of the particle kernel's input) ----------------

// Jacobi <-> heliocentric transforms; src/convert.cpp:9-95.
void helio_to_jacobi_r(size_t n, const V3* r, const double* m, const double* eta, V3* rj) {
    rj[0] = v3(0, 0, 0);
    rj[1] = r[1];
    V3 sum = v3(0, 0, 0);
    for (size_t i = 2; i < n; i++) {
        add_to(sum, scale(r[i - 1], m[i - 1]));
        V3 bary = divs(sum, eta[i - 1]);
        rj[i] = sub(r[i], bary);
    }
}

void helio_to_jacobi_v(size_t n, const V3* v, const double* m, const double* eta, V3* vj) {
    vj[0] = v3(0, 0, 0);
    vj[1] = v[1];
    V3 psum = v3(0, 0, 0);
    for (size_t i = 2; i < n; i++) {
        add_to(psum, scale(v[i - 1], m[i - 1]));
        V3 vsum = divs(psum, eta[i - 1]);
        vj[i] = sub(v[i], vsum);
    }
}

void jacobi_to_helio(size_t n, const double* m, const double* eta, const V3* rj, const V3* vj, V3* r, V3* v) {
    r[0] = v3(0, 0, 0);
    v[0] = v3(0, 0, 0);
    r[1] = rj[1];
    v[1] = vj[1];
    V3 rsum = v3(0, 0, 0), vsum = v3(0, 0, 0);
    for (size_t i = 2; i < n; i++) {
        add_to(rsum, scale(rj[i - 1], m[i - 1] / eta[i - 1]));
        add_to(vsum, scale(vj[i - 1], m[i - 1] / eta[i - 1]));
        r[i] = add(rj[i], rsum);
        v[i] = add(vj[i], vsum);
    }
}

// Osculating elements; src/convert.cpp:310-463 (all outputs).
void to_elements(double mu, V3 r, V3 v, int* esignout, double* aout, double* eout, double* iout,
                 double* capomout, double* omout, double* fout) {
    double a = 0, e = 0, i = 0, capom = 0, om = 0, f = 0;
    int esign;
    double x = r.x, y = r.y, z = r.z;
    double vx = v.x, vy = v.y, vz = v.z;
    const double pi = 2.0 * std::asin(1.0);
    const double prec = 1.0e-13;
    double hx = y * vz - z * vy;
    double hy = z * vx - x * vz;
    double hz = x * vy - y * vx;
    double hsq = hx * hx + hy * hy + hz * hz;
    if (hsq > prec) {
        i = std::acos(hz / std::sqrt(hsq));
        if (std::fabs(i) < prec) {
            capom = 0.0;
        } else if (std::fabs(pi - std::fabs(i)) < prec) {
            capom = 0.0;
        } else {
            capom = std::atan2(hx, -hy);
        }
        double vsq = vx * vx + vy * vy + vz * vz;
        double vdotr = x * vx + y * vy + z * vz;
        double rr = std::sqrt(x * x + y * y + z * z);
        double xhat = x / rr, yhat = y / rr, zhat = z / rr;
        double nx = std::cos(capom), ny = std::sin(capom);
        double fac = vsq * rr - mu;
        double Px = fac * xhat - vdotr * vx;
        double Py = fac * yhat - vdotr * vy;
        double Pz = fac * zhat - vdotr * vz;
        double modP = std::sqrt(Px * Px + Py * Py + Pz * Pz);
        e = modP / mu;
        if (std::fabs(e) < prec) {
            om = 0.0;
        } else {
            if ((i < prec) || (pi - i < prec)) {
                om = std::atan2(Py, Px);
            } else {
                double ecosw = (nx * Px + ny * Py) / mu;
                om = std::acos(ecosw / e);
                if (std::fabs(Pz) > prec) om *= std::fabs(Pz) / Pz;
            }
        }
        double energy = vsq / 2.0 - mu / rr;
        if (std::fabs(energy) < prec) {
            esign = 0;
            a = 0.5 * hsq / mu;
            if (std::fabs(vdotr) < prec) f = 0.0;
            else f = 2.0 * std::acos(std::sqrt(a / rr)) * vdotr / std::fabs(vdotr);
        } else if (energy > 0.0) {
            esign = 1;
            a = -0.5 * mu / energy;
            if (std::fabs(vdotr) < prec) {
                f = 0.0;
            } else {
                fac = a * (1.0 - e * e) / rr - 1.0;
                f = std::acos(fac / e) * vdotr / std::fabs(vdotr);
            }
        } else {
            esign = -1;
            a = -0.5 * mu / energy;
            if (std::fabs(e) > prec) {
                if (std::fabs(vdotr) < prec) {
                    f = (rr < a) ? 0.0 : pi;
                } else {
                    fac = a * (1.0 - e * e) / rr - 1.0;
                    f = std::acos(fac / e) * vdotr / std::fabs(vdotr);
                    if (std::isnan(f)) f = (rr < a) ? 0.0 : pi;
                }
            } else {
                fac = (x * nx + y * ny) / rr;
                f = std::acos(fac);
                if (std::fabs(z) > prec) {
                    f *= std::fabs(z) / z;
                } else if ((i < prec) || (pi - i < prec)) {
                    f = std::atan2(y, x) * std::cos(i);
                }
            }
        }
    } else {
        esign = 1;
        a = std::sqrt(x * x + y * y + z * z);
        e = HUGE_VAL;
        i = std::asin(z / std::sqrt(x * x + y * y + z * z));
        capom = std::atan2(y, x);
        om = HUGE_VAL;
        f = HUGE_VAL;
    }
    if (esignout) *esignout = esign;
    if (aout) *aout = a;
    if (eout) *eout = e;
    if (iout) *iout = i;
    if (capomout) *capomout = capom;
    if (omout) *omout = om;
    if (fout) *fout = f;
}

// Cartesian state from elements with true anomaly; src/convert.cpp:250-308.
void from_elements(double mu, double a, double e, double i, double capom, double om, double f, V3* rv, V3* vv) {
    const double prec = 1.0e-13;
    double u = om + f;
    double xhat = std::cos(u) * std::cos(capom) - std::cos(i) * std::sin(capom) * std::sin(u);
    double yhat = std::cos(u) * std::sin(capom) + std::cos(i) * std::cos(capom) * std::sin(u);
    double zhat = std::sin(i) * std::sin(u);
    double hx = std::sin(capom) * std::sin(i);
    double hy = -std::cos(capom) * std::sin(i);
    double hz = std::cos(i);
    double r, h;
    if (std::fabs(e - 1.0) > prec) {
        r = a * (1.0 - e * e) / (1.0 + e * std::cos(f));
        h = std::sqrt(mu * a * (1.0 - e * e));
    } else {
        h = std::sqrt(2.0 * mu * a);
        r = a / (std::cos(f / 2.0) * std::cos(f / 2.0));
    }
    rv->x = r * xhat;
    rv->y = r * yhat;
    rv->z = r * zhat;
    double thx = hy * zhat - hz * yhat;
    double thy = hz * xhat - hx * zhat;
    double thz = hx * yhat - hy * xhat;
    double thdot = h / (r * r);
    double rdot = e * mu * std::sin(f) / h;
    vv->x = r * thdot * thx + rdot * xhat;
    vv->y = r * thdot * thy + rdot * yhat;
    vv->z = r * thdot * thz + rdot * zhat;
}

// State of the planet side of WHIntegrator: heliocentric and Jacobi
// coordinates, eta, the planet accelerations carried between steps and the
// encounter radii (src/wh.h:19-147).
struct Planets {
    size_t n = 0;
    std::vector<double> m, eta, mu, rh;
    std::vector<uint32_t> id;
    std::vector<V3> r, v, rj, vj, acc;
    std::vector<double> inv_helio3, inv_jacobi3;
    std::vector<uint8_t> mask;
    double cull_radius = 0.5, rh_factor = 0, outer = 300;
};

// WHIntegrator::recalculate_rh; src/wh.cpp:472-495.
void recalculate_rh(Planets& p) {
    if (std::abs(p.rh_factor) > 1e-6) {
        p.rh[0] = p.cull_radius;
        for (size_t i = 1; i < p.n; i++) {
            double a, e;
            to_elements(p.m[0] + p.m[i], p.r[i], p.v[i], nullptr, &a, &e, nullptr, nullptr, nullptr, nullptr);
            p.rh[i] = p.rh_factor * a * (1 - e) * std::pow(p.m[i] / (3 * p.m[0]), 1. / 3);
        }
    } else {
        for (size_t i = 0; i < p.n; i++) p.rh[i] = p.cull_radius;
    }
}

// WHIntegrator::helio_acc_planets; src/wh.cpp:696-756.  Returns the h0
// (indirect-term) vector that the reference stores in the h0 log.
V3 planet_accels(Planets& p) {
    for (size_t i = 1; i < p.n; i++) {
        double r2 = lensq(p.r[i]);
        p.inv_helio3[i] = 1. / (std::sqrt(r2) * r2);
        r2 = lensq(p.rj[i]);
        p.inv_jacobi3[i] = 1. / (std::sqrt(r2) * r2);
    }
    V3 a_common = v3(0, 0, 0);
    for (size_t i = 2; i < p.n; i++) {
        double mfac = p.m[i] * p.inv_helio3[i];
        sub_from(a_common, scale(p.r[i], mfac));
    }
    for (size_t i = 1; i < p.n; i++) p.acc[i] = a_common;
    V3 h0 = sub(a_common, scale(scale(p.r[1], p.m[1]), p.inv_helio3[1]));
    for (size_t i = 2; i < p.n; i++) {
        add_to(p.acc[i], scale(sub(scale(p.rj[i], p.inv_jacobi3[i]), scale(p.r[i], p.inv_helio3[i])), p.m[0]));
    }
    V3 a_accum = v3(0, 0, 0);
    for (size_t i = 2; i < p.n; i++) {
        double mfac = p.m[i] * p.m[0] * p.inv_jacobi3[i] / p.eta[i - 1];
        add_to(a_accum, scale(p.rj[i], mfac));
        add_to(p.acc[i], a_accum);
    }
    for (size_t i = 1; i + 1 < p.n; i++) {
        for (size_t j = i + 1; j < p.n; j++) {
            V3 dr = sub(p.r[j], p.r[i]);
            double r2 = lensq(dr);
            double irij3 = 1. / (r2 * std::sqrt(r2));
            double mfac = p.m[i] * irij3;
            sub_from(p.acc[j], scale(dr, mfac));
            mfac = p.m[j] * irij3;
            add_to(p.acc[i], scale(dr, mfac));
        }
    }
    return h0;
}

// WHIntegrator::step_planets; src/wh.cpp:1027-1067.  Writes the log row of
// this step (positions/velocities of bodies 1..n-1 after the drift and the
// new h0).
int step_planets(Planets& p, double dt, V3* r_row, V3* v_row, V3* h0_out, size_t* bad) {
    for (size_t i = 1; i < p.n; i++) add_to(p.v[i], scale(p.acc[i], dt / 2));
    helio_to_jacobi_v(p.n, p.v.data(), p.m.data(), p.eta.data(), p.vj.data());
    for (size_t i = 1; i < p.n; i++) {
        p.mu[i] = p.m[0] * p.eta[i] / p.eta[i - 1];
        p.mask[i] = 0;
    }
    int st = drift(dt, p.rj.data(), p.vj.data(), 1, p.n - 1, p.mu.data(), p.mask.data(), false, bad);
    if (st != kOk) return st;
    jacobi_to_helio(p.n, p.m.data(), p.eta.data(), p.rj.data(), p.vj.data(), p.r.data(), p.v.data());
    V3 h0 = planet_accels(p);
    for (size_t i = 1; i < p.n; i++) {
        r_row[i - 1] = p.r[i];
        v_row[i - 1] = p.v[i];
    }
    *h0_out = h0;
    for (size_t i = 1; i < p.n; i++) add_to(p.v[i], scale(p.acc[i], dt / 2));
    return kOk;
}

thread_local char g_err[256];

// ---- planet history file: element-linear interpolation --------------------

// Eccentric anomaly from mean anomaly: ehie for e >= 0.8, eget<1|2> below
// (src/convert.cpp:156-228).  Local sin/cos pairs are left for the compiler
// to fuse, as in the reference build.
double ehie(double e, double m) {
    int iflag = 0;
    int nper = static_cast<int>(m / (2 * kPi));
    m = m - nper * 2 * kPi;
    if (m < 0) m += 2 * kPi;
    if (m > kPi) {
        m = 2 * kPi - m;
        iflag = 1;
    }
    double x = std::pow(6 * m, 1. / 3) - m;
    for (int it = 0; it < 3; it++) {
        double sa = std::sin(x + m);
        double ca = std::cos(x + m);
        double esa = e * sa;
        double eca = e * ca;
        double f = x - esa;
        double fp = 1 - eca;
        double dx = -f / fp;
        dx = -f / (fp + dx * esa / 2);
        dx = -f / (fp + dx * (esa + eca * dx / 3) / 2);
        x = x + dx;
    }
    return iflag ? 2 * kPi - m - x : m + x;
}

template <int N>
double eget(double e, double m) {
    double sm = std::sin(m);
    double cm = std::cos(m);
    double x = m + e * sm * (1 + e * (cm + e * (1 - 1.5 * sm * sm)));
    for (int it = 0; it < N; it++) {
        double sx = std::sin(x);
        double cx = std::cos(x);
        double es = e * sx;
        double ec = e * cx;
        double f = x - es - m;
        double fp = 1 - ec;
        double fpp = es;
        double fppp = ec;
        double dx = -f / fp;
        dx = -f / (fp + dx * fpp / 2);
        dx = -f / (fp + dx * fpp / 2 + dx * dx * fppp / 6);
        x = x + dx;
    }
    return x;
}

double ehybrid(double e, double m) {
    if (e < 0.18) return eget<1>(e, m);
    if (e < 0.8) return eget<2>(e, m);
    return ehie(e, m);
}

// from_elements_M (src/convert.cpp:230-248): true anomaly from the mean
// anomaly, sign taken from E reduced to (-pi, pi].
void from_elements_M(double mu, double a, double e, double i, double capom, double om, double M, V3* r, V3* v) {
    double E = ehybrid(e, M);
    double f = std::acos((std::cos(E) - e) / (1 - e * std::cos(E)));
    E = E - static_cast<int>(E / 2 / kPi) * 2 * kPi;
    if (E > kPi)
        E -= 2 * kPi;
    else if (E < -kPi)
        E += 2 * kPi;
    if (E < 0) f = -f;
    from_elements(mu, a, e, i, capom, om, f, r, v);
}

// get_mean_anomaly (src/convert.cpp:465-470).
double mean_anomaly(double e, double f) {
    double E = std::acos((e + std::cos(f)) / (1 + e * std::cos(f)));
    if (f < 0) E = -E;
    return E - e * std::sin(E);
}

template <typename T>
T rd(std::istream& in) {
    T t;
    in.read(reinterpret_cast<char*>(&t), sizeof(T));
    return t;
}

// The Interpolator (src/interp.cpp:13-293) together with the executor glue
// that drives it (src/executor.cu:93-130 at start-up, :235-292 per block,
// :700 time advance).  Slot layout follows the reference's planet phase
// space of interp_maxpl entries: slot 0 the Sun, slot k the planet whose id
// was k-th to appear in the file.
struct History {
    std::ifstream in;
    double user_dt = 0;
    size_t nmax = 0;
    std::vector<double> pm;   // pl.m()
    std::vector<uint32_t> pid;  // pl.id()
    std::vector<V3> pr, pv;   // pl.r(), pl.v()
    size_t pl_alive = 0;
    std::vector<V3> aei0, aei1, oom0, oom1;
    std::vector<double> m1;
    double t0 = 0, t1 = 0;
    size_t npl1 = 0;
    std::vector<V3> rdaei, rdoom, raei_i, room_i;
    std::vector<uint32_t> rids;
    std::vector<double> rm;
    size_t n_alive = 0;
    std::vector<std::pair<uint32_t, size_t>> idmap;  // insertion order = slot
    size_t n_ts = 0, cur_ts = 0;
    double eff_dt = 0, rel_t = 0;
    double t = 0;  // executor time at the start of the next block

    // idmap lookup; a new id takes the next slot (the map's size before insertion)
    bool slot_of(uint32_t id, size_t* ind, bool* fresh) {
        for (auto& kv : idmap)
            if (kv.first == id) {
                *ind = kv.second;
                *fresh = false;
                return true;
            }
        if (idmap.size() >= nmax) return false;
        *ind = idmap.size();
        idmap.emplace_back(id, *ind);
        *fresh = true;
        return true;
    }

    // Interpolator::Interpolator (src/interp.cpp:13-91)
    int open(const char* path, double t_start, double dt, size_t maxpl) {
        in.open(path, std::ios_base::binary);
        if (!in) {
            std::snprintf(g_err, sizeof g_err, "cannot open %s", path);
            return kBadInput;
        }
        user_dt = dt;
        nmax = maxpl;
        pm.assign(nmax, 0);
        pid.assign(nmax, 0);
        pr.assign(nmax, v3(0, 0, 0));
        pv.assign(nmax, v3(0, 0, 0));
        pm[0] = rd<double>(in);
        in.seekg(32 - 8, std::ios_base::cur);
        aei0 = aei1 = oom0 = oom1 = std::vector<V3>(nmax, v3(0, 0, 0));
        m1.assign(nmax, 0);
        rm.assign(nmax, 0);
        rdaei = rdoom = raei_i = room_i = std::vector<V3>(nmax, v3(0, 0, 0));
        rids.assign(nmax, 0);
        t0 = HUGE_VAL;
        t1 = rd<double>(in);
        npl1 = rd<uint32_t>(in) + 1;
        pl_alive = npl1;
        in.seekg(32 - 8 - 4, std::ios_base::cur);
        if (t1 > t_start) {
            std::snprintf(g_err, sizeof g_err, "time t0 not found in datafile: planet history starts after t0");
            return kBadInput;
        }
        idmap.clear();
        idmap.emplace_back(0u, 0);
        for (size_t i = 1; i < npl1; i++) {
            uint32_t id = rd<uint32_t>(in);
            size_t ind;
            bool fresh;
            if (!slot_of(id, &ind, &fresh) || i >= nmax) {
                std::snprintf(g_err, sizeof g_err, "planet history holds more planets than the table (%zu)", nmax);
                return kBadInput;
            }
            if (fresh) pid[ind] = id;
            m1[ind] = rd<float>(in);
            aei1[ind].x = rd<float>(in);
            aei1[ind].y = rd<float>(in);
            aei1[ind].z = rd<float>(in);
            oom1[ind].x = rd<float>(in);
            oom1[ind].y = rd<float>(in);
            oom1[ind].z = rd<float>(in);
            V3 r, v;
            from_elements_M(pm[0] + m1[ind], aei1[ind].x, aei1[ind].y, aei1[ind].z, oom1[ind].x, oom1[ind].y,
                            oom1[ind].z, &r, &v);
            pr[i] = r;
            pv[i] = v;
            pid[i] = id;
            pm[i] = m1[ind];
        }
        pr[0] = v3(0, 0, 0);
        pv[0] = v3(0, 0, 0);
        aei1[0] = v3(0, 0, 0);
        oom1[0] = v3(0, 0, 0);
        // Executor::init (src/executor.cu:99-129)
        t = t_start;
        while (t1 <= t) {
            int st = next();
            if (st != kOk) return st;
        }
        if (std::isinf(t0)) {
            std::snprintf(g_err, sizeof g_err, "lookup file starts past the given Initial-Time");
            return kBadInput;
        }
        cur_ts = 0;
        n_ts = std::max<size_t>(1, static_cast<size_t>(std::round((t1 - t) / dt)));
        eff_dt = (t1 - t) / static_cast<double>(n_ts);
        rel_t = t - t0;
        fill_one(rel_t);
        return kOk;
    }

    // Interpolator::next (src/interp.cpp:167-293)
    int next() {
        aei0 = aei1;
        oom0 = oom1;
        t0 = t1;
        rel_t = 0;
        t1 = rd<double>(in);
        if (!in) {
            std::snprintf(g_err, sizeof g_err, "EOS reached for lookup file.");
            return kBadInput;
        }
        npl1 = rd<uint32_t>(in) + 1;
        double dt = t1 - t0;
        n_ts = std::max<size_t>(1, static_cast<size_t>(std::round(dt / user_dt)));
        eff_dt = dt / static_cast<double>(n_ts);
        in.seekg(32 - 8 - 4, std::ios_base::cur);
        const double nan = std::numeric_limits<double>::quiet_NaN();
        for (size_t i = 1; i < nmax; i++) {
            aei1[i] = v3(nan, nan, nan);
            oom1[i] = v3(nan, nan, nan);
        }
        size_t k = 0;
        for (size_t i = 1; i < npl1; i++) {
            uint32_t id = rd<uint32_t>(in);
            size_t ind;
            bool fresh;
            if (!slot_of(id, &ind, &fresh)) {
                std::snprintf(g_err, sizeof g_err, "planet history holds more planets than the table (%zu)", nmax);
                return kBadInput;
            }
            if (fresh) pid[ind] = id;
            pm[ind] = rd<float>(in);
            aei1[ind].x = rd<float>(in);
            aei1[ind].y = rd<float>(in);
            aei1[ind].z = rd<float>(in);
            oom1[ind].x = rd<float>(in);
            oom1[ind].y = rd<float>(in);
            oom1[ind].z = rd<float>(in);
            if (std::isnan(aei0[ind].x)) continue;
            rids[k] = id;
            rm[k] = m1[ind];
            raei_i[k] = aei0[ind];
            room_i[k] = oom0[ind];
            rdaei[k] = divs(sub(aei1[ind], aei0[ind]), dt);
            V3 doom = sub(oom1[ind], oom0[ind]);
            if (doom.x > kPi)
                doom.x -= 2 * kPi;
            else if (doom.x < -kPi)
                doom.x += 2 * kPi;
            if (doom.y > kPi)
                doom.y -= 2 * kPi;
            else if (doom.y < -kPi)
                doom.y += 2 * kPi;
            doom = divs(doom, dt);
            double mmfreq = 2 * kPi * std::sqrt(1 + pm[ind] / pm[0]) *
                            (std::pow(aei0[ind].x, -1.5) + std::pow(aei1[ind].x, -1.5)) / 2;
            double cmfin = oom0[ind].z + mmfreq * dt;
            cmfin = std::fmod(cmfin, 2 * kPi);
            double corr = oom1[ind].z - cmfin;
            if (corr > kPi)
                corr -= 2 * kPi;
            else if (corr < -kPi)
                corr += 2 * kPi;
            mmfreq += corr / dt;
            doom.z = mmfreq;
            rdoom[k] = doom;
            k++;
        }
        n_alive = k;
        return kOk;
    }

    // one planet's heliocentric state at relative time rt
    void state_at(size_t j, double rt, V3* r, V3* v) const {
        V3 aei = add(raei_i[j], scale(rdaei[j], rt));
        V3 oom = add(room_i[j], scale(rdoom[j], rt));
        double gm = rm[j] + pm[0];
        from_elements_M(gm, aei.x, aei.y, aei.z, oom.x, oom.y, oom.z, r, v);
    }

    // Interpolator::fill_one (src/interp.cpp:93-111)
    void fill_one(double rt) {
        pl_alive = n_alive + 1;
        for (size_t j = 0; j < n_alive; j++) state_at(j, rt, &pr[j + 1], &pv[j + 1]);
    }

    // Executor::update_planets for an interpolated history
    // (src/executor.cu:235-292) with Interpolator::fill (src/interp.cpp:113-165)
    // and the time advance of src/executor.cu:700.
    int block(uint32_t tbsize, double* t_start, double* dt_out, uint32_t* n_steps, uint32_t* n_bodies, double* m,
              uint32_t* id, double* r_log, double* h0_log) {
        if (cur_ts == n_ts) {
            double diff = rel_t - (t1 - t0);
            if (!(std::abs(diff) < 1e-8)) {
                std::snprintf(g_err, sizeof g_err, "sanity check failed - adjusting t by too much: %g", diff);
                return kBadInput;
            }
            int st = next();
            if (st != kOk) return st;
            cur_ts = 0;
            t = t0;
        }
        const double cur_dt = eff_dt;
        const uint32_t n = std::min(tbsize, static_cast<uint32_t>(n_ts - cur_ts));
        double rt = rel_t + cur_dt;
        pl_alive = n_alive + 1;
        for (size_t j = 0; j < n_alive; j++) {
            pid[j + 1] = rids[j];
            pm[j + 1] = rm[j];
        }
        for (uint32_t s = 0; s < n; s++) {
            if (rt > t1 - t0 + 1e-8) {
                std::snprintf(g_err, sizeof g_err, "interpolation ran too long, timechunk exceeded lookup boundary");
                return kBadInput;
            }
            pr[0] = v3(0, 0, 0);
            pv[0] = v3(0, 0, 0);
            for (size_t j = 0; j < n_alive; j++) {
                state_at(j, rt, &pr[j + 1], &pv[j + 1]);
                store3(r_log, (size_t)s * n_alive + j, pr[j + 1]);
            }
            rt += cur_dt;
        }
        *t_start = t;
        *dt_out = cur_dt;
        *n_steps = n;
        *n_bodies = static_cast<uint32_t>(pl_alive);
        for (size_t k = 0; k < pl_alive; k++) {
            m[k] = pm[k];
            id[k] = pid[k];
        }
        if (h0_log) orc_load_h0(static_cast<uint32_t>(pl_alive), pm.data(), r_log, n, h0_log);
        rel_t += static_cast<double>(n) * cur_dt;
        cur_ts += n;
        t += cur_dt * static_cast<double>(n);
        return kOk;
    }
};

// save_swift_plhist (src/data.cpp:817-873): one record of heliocentric
// planets (slot 0 the Sun) as float32 elements with the mean anomaly.
int write_history(const char* path, int append, double time, uint32_t n, const double* m, const double* r,
                  const double* v, const uint32_t* id) {
    std::ofstream out(path, append ? (std::ios_base::binary | std::ios_base::app) : std::ios_base::binary);
    if (!out) {
        std::snprintf(g_err, sizeof g_err, "cannot open %s", path);
        return kBadInput;
    }
    auto wr = [&](const auto& x) { out.write(reinterpret_cast<const char*>(&x), sizeof(x)); };
    auto pad = [&](size_t k) {
        for (size_t i = 0; i < k; i++) out.put(0);
    };
    if (!append) {
        wr(static_cast<double>(m[0]));
        pad(32 - 8);
    }
    wr(static_cast<double>(time));
    wr(static_cast<uint32_t>(n > 0 ? n - 1 : 0));
    pad(32 - 8 - 4);
    for (uint32_t i = 1; i < n; i++) {
        double center_mass = m[0];
        V3 center_r = scale(load3(r, 0), m[0]);
        V3 center_v = scale(load3(v, 0), m[0]);
        center_r = divs(center_r, center_mass);
        center_v = divs(center_v, center_mass);
        double a, e, in, capom, om, f;
        to_elements(m[i] + center_mass, sub(load3(r, i), center_r), sub(load3(v, i), center_v), nullptr, &a, &e,
                    &in, &capom, &om, &f);
        wr(static_cast<uint32_t>(id[i]));
        wr(static_cast<float>(m[i]));
        wr(static_cast<float>(a));
        wr(static_cast<float>(e));
        wr(static_cast<float>(in));
        wr(static_cast<float>(capom));
        wr(static_cast<float>(om));
        wr(static_cast<float>(mean_anomaly(e, f)));
    }
    out.flush();
    return out ? kOk : kBadInput;
}

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err; }

// Library of the particle drift's sin/cos on the calling thread (0 glibc,
// 1 correctly rounded); resets the thread's call counters.
void orc_set_particle_libm(int mode) {
    tl_particle_libm = mode;
    tl_libm_calls = 0;
    tl_libm_differ = 0;
}
// The correctly rounded sin and cos of mode 1 over n arguments.
void orc_sincos_cr(uint64_t n, const double* x, double* s, double* c) {
    for (uint64_t i = 0; i < n; i++) {
        s[i] = sin_cr(x[i]);
        c[i] = cos_cr(x[i]);
    }
}
// Calls made in mode 1 on the calling thread, and those where glibc differs.
void orc_particle_libm_counts(uint64_t* calls, uint64_t* differ) {
    *calls = tl_libm_calls;
    *differ = tl_libm_differ;
}

// ---- particle path -------------------------------------------------------

// Accelerations and flags of particles [0, n) from one planet row and h0;
// WHIntegrator::helio_acc_particles (src/wh.cpp:659-673).
void orc_particle_accels(uint32_t n_bodies, const double* m, const double* rh, double outer,
                         const double* planet_r, const double* h0, double time, uint32_t n, const double* r,
                         double* a, uint16_t* flags, const uint32_t* id, double* deathtime) {
    const V3* pr = reinterpret_cast<const V3*>(planet_r);
    V3 h = load3(h0, 0);
    for (uint32_t i = 0; i < n; i++) {
        V3 ai;
        particle_accel(n_bodies, m, rh, outer, pr, h, time, load3(r, i), ai, flags[i], id[i], deathtime);
        store3(a, i, ai);
    }
}

// One time-block of the CPU particle path: WHIntegrator::
// integrate_particles_timeblock (src/wh.cpp:597-609) -> step_particles
// (src/wh.cpp:959-986) for each of n_steps log rows.  r_log is
// [n_steps][n_bodies-1][3], h0_log is [n_steps][3].  The particle arrays are
// AoS xyz.  deathtime_index and deathtime are indexed by particle id, as in
// the reference.  Returns 0, or 1 (unbound) / 2 (Kepler not converged) with
// *bad set to the particle slot, where the reference throws.
int orc_particle_block(uint32_t n_bodies, const double* m, const double* rh, double outer, const double* r_log,
                       const double* h0_log, uint32_t n_steps, double t, double dt, uint32_t n, double* r,
                       double* v, double* a, uint16_t* flags, const uint32_t* id, uint32_t* deathtime_index,
                       double* deathtime, uint64_t* bad) {
    V3* R = reinterpret_cast<V3*>(r);
    V3* Vv = reinterpret_cast<V3*>(v);
    V3* A = reinterpret_cast<V3*>(a);
    std::vector<double> mu(n, m[0]);
    std::vector<uint8_t> mask(n);
    const size_t row = n_bodies - 1;
    for (uint32_t s = 0; s < n_steps; s++) {
        for (uint32_t i = 0; i < n; i++) {
            mask[i] = flags[i] != 0;
            if (!mask[i]) add_to(Vv[i], scale(A[i], dt / 2));
        }
        size_t b = 0;
        int st = drift(dt, R, Vv, 0, n, mu.data(), mask.data(), true, &b);
        if (st != kOk) {
            if (bad) *bad = b;
            std::snprintf(g_err, sizeof g_err, "%s (particle slot %zu, step %u)",
                          st == kUnbound ? "unbound orbit" : "Unconverging kepler", b, s);
            return st;
        }
        const V3* pr = reinterpret_cast<const V3*>(r_log) + s * row;
        V3 h = load3(h0_log, s);
        for (uint32_t i = 0; i < n; i++) particle_accel(n_bodies, m, rh, outer, pr, h, t, R[i], A[i], flags[i], id[i], deathtime);
        for (uint32_t i = 0; i < n; i++) {
            if (!mask[i]) {
                add_to(Vv[i], scale(A[i], dt / 2));
                deathtime_index[id[i]] = s + 1;
            }
        }
        t += dt;
    }
    return kOk;
}

// Heliocentric Kepler drift of n bodies about mu, fixed (5 iterations) or
// variable Newton count; exposed for the closed-form / invariant tests.
int orc_drift(double t, uint32_t n, double* r, double* v, const double* mu, int fixed, uint64_t* bad) {
    size_t b = 0;
    int st = drift(t, reinterpret_cast<V3*>(r), reinterpret_cast<V3*>(v), 0, n, mu, nullptr, fixed != 0, &b);
    if (st != kOk && bad) *bad = b;
    return st;
}

// ---- planet producer -----------------------------------------------------

// WHIntegrator constructor, planet half (src/wh.cpp:498-531): eta, rh,
// Jacobi coordinates and the first planet accelerations.  Bodies in
// heliocentric coordinates, m already in GM units.  h0_0 receives the h0
// vector at the initial time (the reference stores it in h0_log.old[0]).
void* orc_planets_new(uint32_t n, const double* m, const double* r, const double* v, const uint32_t* id,
                      double cull_radius, double rh_factor, double outer, double* h0_0) {
    Planets* p = new Planets();
    p->n = n;
    p->m.assign(m, m + n);
    p->id.assign(id, id + n);
    p->r.resize(n);
    p->v.resize(n);
    for (uint32_t i = 0; i < n; i++) {
        p->r[i] = load3(r, i);
        p->v[i] = load3(v, i);
    }
    p->eta.resize(n);
    p->mu.resize(n);
    p->rh.resize(n);
    p->rj.resize(n);
    p->vj.resize(n);
    p->acc.resize(n);
    p->inv_helio3.resize(n);
    p->inv_jacobi3.resize(n);
    p->mask.resize(n);
    p->cull_radius = cull_radius;
    p->rh_factor = rh_factor;
    p->outer = outer;
    p->eta[0] = p->m[0];
    for (uint32_t i = 1; i < n; i++) p->eta[i] = p->eta[i - 1] + p->m[i];
    recalculate_rh(*p);
    helio_to_jacobi_r(n, p->r.data(), p->m.data(), p->eta.data(), p->rj.data());
    helio_to_jacobi_v(n, p->v.data(), p->m.data(), p->eta.data(), p->vj.data());
    V3 h0 = planet_accels(*p);
    if (h0_0) store3(h0_0, 0, h0);
    return p;
}

void orc_planets_free(void* h) { delete static_cast<Planets*>(h); }

// WHIntegrator::integrate_planets_timeblock (src/wh.cpp:584-595): n_steps
// planet steps writing r_log/v_log rows [n_steps][n-1][3] and h0_log
// [n_steps][3].
int orc_planets_block(void* h, uint32_t n_steps, double dt, double* r_log, double* v_log, double* h0_log) {
    Planets& p = *static_cast<Planets*>(h);
    const size_t row = p.n - 1;
    for (uint32_t s = 0; s < n_steps; s++) {
        size_t b = 0;
        V3 h0;
        int st = step_planets(p, dt, reinterpret_cast<V3*>(r_log) + s * row,
                              reinterpret_cast<V3*>(v_log) + s * row, &h0, &b);
        if (st != kOk) {
            std::snprintf(g_err, sizeof g_err, "%s (planet %zu, step %u)",
                          st == kUnbound ? "unbound orbit" : "Unconverging kepler", b, s);
            return st;
        }
        store3(h0_log, s, h0);
    }
    return kOk;
}

// Current heliocentric planet state and encounter radii.
void orc_planets_get(void* h, double* r, double* v, double* rh) {
    Planets& p = *static_cast<Planets*>(h);
    for (size_t i = 0; i < p.n; i++) {
        if (r) store3(r, i, p.r[i]);
        if (v) store3(v, i, p.v[i]);
        if (rh) rh[i] = p.rh[i];
    }
}

void orc_planets_recalculate_rh(void* h) { recalculate_rh(*static_cast<Planets*>(h)); }

// WHIntegrator::load_h0 (src/wh.cpp:993-1025): h0 rows from a position log.
void orc_load_h0(uint32_t n_bodies, const double* m, const double* r_log, uint32_t n_steps, double* h0_log) {
    const size_t row = n_bodies - 1;
    for (uint32_t s = 0; s < n_steps; s++) {
        V3 h0 = v3(0, 0, 0), a_common = v3(0, 0, 0);
        for (uint32_t i = 1; i < n_bodies; i++) {
            V3 r = load3(r_log, s * row + i - 1);
            double r2 = lensq(r);
            double inv_r3 = 1. / (std::sqrt(r2) * r2);
            if (i == 1) h0 = scale(scale(neg(r), m[1]), inv_r3);
            if (i > 1) {
                double mfac = m[i] * inv_r3;
                sub_from(a_common, scale(r, mfac));
            }
        }
        add_to(h0, a_common);
        store3(h0_log, s, h0);
    }
}

// Cubic Hermite interpolation of a sampled planet history into a planet log
// (no reference counterpart: the reference interpolates orbital elements
// linearly, src/interp.cpp:113-293; cubic Hermite in Cartesian coordinates
// is this project's history format).  Samples k = 0..n_samples-1 at strictly
// increasing times t[k] hold heliocentric r and v of bodies 1..n_bodies-1
// ([k][planet][3] each).  Row s of the output is the state at
// t_start + (s+1) dt -- the planet-log convention of src/wh.cpp:1060 -- and
// h0_log row s is the load_h0 indirect term (src/wh.cpp:993-1025) of that
// row.  On the interval [t_k, t_k+1] with h = t_k+1 - t_k and
// u = (t - t_k) / h:
//   r(t) = h00 r_k + h10 h v_k + h01 r_k+1 + h11 h v_k+1,
//   h00 = 2u^3 - 3u^2 + 1, h10 = u^3 - 2u^2 + u, h01 = 3u^2 - 2u^3,
//   h11 = u^3 - u^2.
// k is the last sample with t_k <= t, clamped to [0, n_samples-2]; the
// evaluation order written here is the one the device kernel uses, so the
// two agree bit for bit.  Pinned by: exactness for cubic trajectories, node
// reproduction, fourth-order convergence on Kepler orbits (tests).
int orc_hermite_log(uint32_t n_bodies, const double* m, uint64_t n_samples, const double* t_s, const double* r_s,
                    const double* v_s, double t_start, double dt, uint32_t n_steps, double* r_log, double* h0_log) {
    if (n_samples < 2 || n_bodies < 1) {
        std::snprintf(g_err, sizeof g_err, "hermite: need >= 2 samples and >= 1 body");
        return kBadInput;
    }
    const uint32_t np = n_bodies - 1;
    for (uint32_t s = 0; s < n_steps; s++) {
        const double t = t_start + (double)(s + 1) * dt;
        uint64_t lo = 0, hi = n_samples - 1;
        while (hi - lo > 1) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (t_s[mid] <= t)
                lo = mid;
            else
                hi = mid;
        }
        const double h = t_s[lo + 1] - t_s[lo];
        const double u = (t - t_s[lo]) / h;
        const double u2 = u * u, u3 = u2 * u;
        const double h00 = 2.0 * u3 - 3.0 * u2 + 1.0;
        const double h10 = u3 - 2.0 * u2 + u;
        const double h01 = 3.0 * u2 - 2.0 * u3;
        const double h11 = u3 - u2;
        const double w10 = h10 * h, w11 = h11 * h;
        for (uint32_t j = 0; j < np; j++) {
            const double* r0 = r_s + (lo * np + j) * 3;
            const double* v0 = v_s + (lo * np + j) * 3;
            const double* r1 = r0 + (size_t)np * 3;
            const double* v1 = v0 + (size_t)np * 3;
            double* out = r_log + ((size_t)s * np + j) * 3;
            for (int c = 0; c < 3; c++) out[c] = h00 * r0[c] + w10 * v0[c] + h01 * r1[c] + w11 * v1[c];
        }
        if (h0_log) orc_load_h0(n_bodies, m, r_log + (size_t)s * np * 3, 1, h0_log + (size_t)s * 3);
    }
    return kOk;
}

// calculate_planet_metrics (src/wh.cpp:422-470): barycentric energy and
// angular momentum of the bodies.
void orc_planet_metrics(uint32_t n, const double* m, const double* r, const double* v, double* energy, double* l) {
    V3 rsum = v3(0, 0, 0), vsum = v3(0, 0, 0);
    double msum = 0;
    for (uint32_t i = 0; i < n; i++) {
        add_to(rsum, scale(load3(r, i), m[i]));
        add_to(vsum, scale(load3(v, i), m[i]));
        msum += m[i];
    }
    V3 br = divs(rsum, msum), bv = divs(vsum, msum);
    std::vector<V3> R(n), Vv(n);
    for (uint32_t i = 0; i < n; i++) {
        R[i] = sub(load3(r, i), br);
        Vv[i] = sub(load3(v, i), bv);
    }
    if (energy) {
        double ke = 0.0, pe = 0.0;
        for (uint32_t i = 0; i < n; i++)
            ke += 0.5 * (Vv[i].x * Vv[i].x + Vv[i].y * Vv[i].y + Vv[i].z * Vv[i].z) * m[i];
        for (uint32_t i = 0; i + 1 < n; i++)
            for (uint32_t j = i + 1; j < n; j++) {
                double dx = R[i].x - R[j].x, dy = R[i].y - R[j].y, dz = R[i].z - R[j].z;
                pe -= m[i] * m[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
            }
        *energy = ke + pe;
    }
    if (l) {
        V3 L = v3(0, 0, 0);
        for (uint32_t i = 0; i < n; i++) add_to(L, scale(cross(R[i], Vv[i]), m[i]));
        store3(l, 0, L);
    }
}

// ---- compaction ----------------------------------------------------------

// Stable partition of slots [begin, begin+len) by a flag predicate, as
// HostParticlePhaseSpace::stable_partition_unflagged / _alive
// (src/data.cpp:17-39,105-123): mode 0 keeps (flags & 0xFF) == 0 first,
// mode 1 keeps (flags & 0xFE) == 0 first.  perm[k] receives the source slot
// (relative to begin) of destination k; returns the split point (absolute).
uint64_t orc_stable_partition(const uint16_t* flags, uint64_t begin, uint64_t len, int mode, uint64_t* perm) {
    std::vector<size_t> idx(len);
    std::iota(idx.begin(), idx.end(), 0);
    const uint16_t mask = mode == 0 ? 0x00FF : 0x00FE;
    auto it = std::stable_partition(idx.begin(), idx.end(),
                                    [&](size_t k) { return (flags[k + begin] & mask) == 0; });
    for (size_t k = 0; k < len; k++) perm[k] = idx[k];
    return static_cast<uint64_t>(it - idx.begin()) + begin;
}

// ---- conversions ---------------------------------------------------------

void orc_to_elements(double mu, const double* r, const double* v, double* out6) {
    double a, e, i, capom, om, f;
    to_elements(mu, load3(r, 0), load3(v, 0), nullptr, &a, &e, &i, &capom, &om, &f);
    out6[0] = a;
    out6[1] = e;
    out6[2] = i;
    out6[3] = capom;
    out6[4] = om;
    out6[5] = f;
}

void orc_from_elements(double mu, const double* el6, double* r, double* v) {
    V3 R, Vv;
    from_elements(mu, el6[0], el6[1], el6[2], el6[3], el6[4], el6[5], &R, &Vv);
    store3(r, 0, R);
    store3(v, 0, Vv);
}

// from_elements_M (mean anomaly in el6[5]).
void orc_from_elements_mean(double mu, const double* el6, double* r, double* v) {
    V3 R, Vv;
    from_elements_M(mu, el6[0], el6[1], el6[2], el6[3], el6[4], el6[5], &R, &Vv);
    store3(r, 0, R);
    store3(v, 0, Vv);
}

// ---- planet history file ----------------------------------------------------

// Start (append = 0) or extend a planet history file with one record of n
// heliocentric bodies (slot 0 the Sun): begin_swift_plhist +
// save_swift_plhist (src/data.cpp:817-873).
int orc_plhist_write(const char* path, int append, double time, uint32_t n, const double* m, const double* r,
                     const double* v, const uint32_t* id) {
    return write_history(path, append, time, n, m, r, v, id);
}

// Open a history for a run starting at t0 with Time-Step dt and a planet
// table of maxpl slots; NULL on error (message in orc_last_error).
void* orc_plhist_open(const char* path, double t0, double dt, uint32_t maxpl) {
    History* h = new History();
    if (h->open(path, t0, dt, maxpl) != kOk) {
        delete h;
        return nullptr;
    }
    return h;
}

void orc_plhist_free(void* h) { delete static_cast<History*>(h); }

// The planet table as the reference's phase space holds it (maxpl slots):
// *n_bodies alive entries, GM, ids, heliocentric r and v (any output may be NULL).
void orc_plhist_state(void* hp, uint32_t* n_bodies, double* m, uint32_t* id, double* r, double* v) {
    History& h = *static_cast<History*>(hp);
    if (n_bodies) *n_bodies = static_cast<uint32_t>(h.pl_alive);
    for (size_t k = 0; k < h.nmax; k++) {
        if (m) m[k] = h.pm[k];
        if (id) id[k] = h.pid[k];
        if (r) store3(r, k, h.pr[k]);
        if (v) store3(v, k, h.pv[k]);
    }
}

// Next planet block of at most tbsize steps: its start time, step, length,
// bodies (GM, ids), position log (n_steps x (n_bodies-1) x 3) and h0 log.
int orc_plhist_block(void* hp, uint32_t tbsize, double* t_start, double* dt, uint32_t* n_steps, uint32_t* n_bodies,
                     double* m, uint32_t* id, double* r_log, double* h0_log) {
    return static_cast<History*>(hp)->block(tbsize, t_start, dt, n_steps, n_bodies, m, id, r_log, h0_log);
}

}  // extern "C"
