// Planet producer: Jacobi Wisdom-Holman planet steps on the host.  The
// arithmetic follows the reference CPU path term by term so the planet log
// it produces is bit-identical to the reference's (built with the same
// no-contraction flags).
#include "planets.h"

#include <cmath>
#include <sstream>

#include "convert.h"

namespace glx {
namespace host {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTolKep = 1e-14;
constexpr int kMaxKep = 10;

// glibc's sincos() rounds differently from sin() + cos() for some
// arguments.  The reference CPU build calls sin() and cos() separately inside
// its Kepler solver and one fused sincos() for the drift's first guess; the
// producer does the same, through function pointers the compiler cannot
// fuse, so its planet log is bit-identical to the reference's.
double (*volatile const g_sin)(double) = ::sin;
double (*volatile const g_cos)(double) = ::cos;

// Kepler drift of one Jacobi body about central GM mu with the
// early-exit Newton solve (src/wh.cpp:19-43 inside drift<false>,
// src/wh.cpp:875-957).  Returns 0, 1 (unbound) or 2 (not converged).
int kepler_drift(double t, double mu, Vec3& r, Vec3& v) {
    const double dist = std::sqrt(r.lensq());
    double energy = v.lensq();
    const double vdotr = v.x * r.x + v.y * r.y + v.z * r.z;
    energy *= 0.5;
    energy -= mu / dist;
    if (energy >= 0) return 1;
    const Vec3 r0 = r, v0 = v;
    const double a = -0.5 * mu / energy;
    const double n_ = std::sqrt(mu / (a * a * a));
    const double ecosEo = 1.0 - dist / a;
    const double esinEo = vdotr / (n_ * a * a);
    const double dM = t * n_ - kPi * 2 * (int)(t * n_ / kPi * 2);
    const double adv_dt = dM / n_;
    double sM, cM;
    ::sincos(dM, &sM, &cM);
    double dE = dM - esinEo + esinEo * cM + ecosEo * sM;
    double sindE = g_sin(dE), cosdE = g_cos(dE);
    bool converged = false;
    for (int k = 0; k < kMaxKep; k++) {
        const double f = dE - ecosEo * sindE + esinEo * (1. - cosdE) - dM;
        const double fp = 1. - ecosEo * cosdE + esinEo * sindE;
        const double delta = -f / fp;
        if (std::fabs(delta) < kTolKep) {
            converged = true;
            break;
        }
        dE += delta;
        sindE = g_sin(dE);
        cosdE = g_cos(dE);
    }
    if (!converged) return 2;
    const double fp = 1.0 - ecosEo * cosdE + esinEo * sindE;
    const double f = 1.0 + a * (cosdE - 1.0) / dist;
    const double g = adv_dt + (sindE - dE) / n_;
    const double fdot = -n_ * sindE * a / (dist * fp);
    const double gdot = 1.0 + (cosdE - 1.0) / fp;
    r = r0 * f + v0 * g;
    v = r0 * fdot + v0 * gdot;
    return 0;
}

}  // namespace

PlanetIntegrator::PlanetIntegrator(std::vector<double> m, std::vector<Vec3> r, std::vector<Vec3> v,
                                   std::vector<uint32_t> id, double cull_radius, double rh_factor,
                                   double outer_radius)
    : m_(std::move(m)), id_(std::move(id)), r_(std::move(r)), v_(std::move(v)), cull_(cull_radius),
      rh_factor_(rh_factor), outer_(outer_radius) {
    const size_t n = m_.size();
    if (n == 0 || r_.size() != n || v_.size() != n) throw PlanetError("planet arrays must be non-empty and equal");
    if (id_.size() != n) {
        id_.resize(n);
        for (size_t i = 0; i < n; i++) id_[i] = static_cast<uint32_t>(i);
    }
    eta_.assign(n, 0.0);
    rh_.assign(n, 0.0);
    mu_.assign(n, 0.0);
    rj_.assign(n, Vec3());
    vj_.assign(n, Vec3());
    acc_.assign(n, Vec3());
    ih3_.assign(n, 0.0);
    ij3_.assign(n, 0.0);
    eta_[0] = m_[0];
    for (size_t i = 1; i < n; i++) eta_[i] = eta_[i - 1] + m_[i];
    recalculate_rh();
    helio_to_jacobi_r(r_, m_, eta_, rj_);
    helio_to_jacobi_v(v_, m_, eta_, vj_);
    h0_init_ = accelerations();
}

void encounter_radii(size_t n, const double* m, const Vec3* r, const Vec3* v, double cull_radius, double rh_factor,
                     double* rh) {
    if (std::abs(rh_factor) > 1e-6) {
        rh[0] = cull_radius;
        for (size_t i = 1; i < n; i++) {
            const Elements el = to_elements(m[0] + m[i], r[i], v[i]);
            rh[i] = rh_factor * el.a * (1 - el.e) * std::pow(m[i] / (3 * m[0]), 1. / 3);
        }
    } else {
        for (size_t i = 0; i < n; i++) rh[i] = cull_radius;
    }
}

void PlanetIntegrator::recalculate_rh() {
    encounter_radii(m_.size(), m_.data(), r_.data(), v_.data(), cull_, rh_factor_, rh_.data());
}

Vec3 PlanetIntegrator::accelerations() {
    const size_t n = m_.size();
    for (size_t i = 1; i < n; i++) {
        double r2 = r_[i].lensq();
        ih3_[i] = 1. / (std::sqrt(r2) * r2);
        r2 = rj_[i].lensq();
        ij3_[i] = 1. / (std::sqrt(r2) * r2);
    }
    // Acceleration common to every planet from the bodies beyond the first.
    Vec3 common;
    for (size_t i = 2; i < n; i++) common -= r_[i] * (m_[i] * ih3_[i]);
    for (size_t i = 1; i < n; i++) acc_[i] = common;
    Vec3 h0 = common;
    if (n > 1) h0 = common - r_[1] * m_[1] * ih3_[1];
    // Indirect terms (planet 1 receives none).
    for (size_t i = 2; i < n; i++) acc_[i] += (rj_[i] * ij3_[i] - r_[i] * ih3_[i]) * m_[0];
    Vec3 accum;
    for (size_t i = 2; i < n; i++) {
        accum += rj_[i] * (m_[i] * m_[0] * ij3_[i] / eta_[i - 1]);
        acc_[i] += accum;
    }
    // Direct planet-planet terms.
    for (size_t i = 1; i + 1 < n; i++) {
        for (size_t j = i + 1; j < n; j++) {
            const Vec3 dr = r_[j] - r_[i];
            const double r2 = dr.lensq();
            const double irij3 = 1. / (r2 * std::sqrt(r2));
            acc_[j] -= dr * (m_[i] * irij3);
            acc_[i] += dr * (m_[j] * irij3);
        }
    }
    return h0;
}

void PlanetIntegrator::drift_jacobi(double dt) {
    const size_t n = m_.size();
    for (size_t i = 1; i < n; i++) mu_[i] = m_[0] * eta_[i] / eta_[i - 1];
    for (size_t i = 1; i < n; i++) {
        const int st = kepler_drift(dt, mu_[i], rj_[i], vj_[i]);
        if (st != 0) {
            std::ostringstream ss;
            ss << (st == 1 ? "unbound orbit of planet " : "Unconverging kepler for planet ") << i;
            throw PlanetError(ss.str());
        }
    }
}

void PlanetIntegrator::step(double dt, Vec3* r_row, Vec3* v_row, Vec3* h0) {
    const size_t n = m_.size();
    for (size_t i = 1; i < n; i++) v_[i] += acc_[i] * (dt / 2);
    helio_to_jacobi_v(v_, m_, eta_, vj_);
    drift_jacobi(dt);
    jacobi_to_helio(m_, eta_, rj_, vj_, r_, v_);
    *h0 = accelerations();
    for (size_t i = 1; i < n; i++) {
        r_row[i - 1] = r_[i];
        v_row[i - 1] = v_[i];
    }
    for (size_t i = 1; i < n; i++) v_[i] += acc_[i] * (dt / 2);
}

void PlanetIntegrator::integrate_block(uint32_t n_steps, double dt, double* r_log, double* v_log, double* h0_log) {
    const size_t row = m_.size() - 1;
    std::vector<Vec3> rrow(row), vrow(row);
    for (uint32_t s = 0; s < n_steps; s++) {
        Vec3 h0;
        step(dt, rrow.data(), vrow.data(), &h0);
        for (size_t j = 0; j < row; j++) {
            if (r_log) rrow[j].store(r_log, s * row + j);
            if (v_log) vrow[j].store(v_log, s * row + j);
        }
        if (h0_log) h0.store(h0_log, s);
    }
}

void PlanetIntegrator::metrics(double* energy, Vec3* l) const {
    const size_t n = m_.size();
    Vec3 rs, vs;
    double ms = 0;
    for (size_t i = 0; i < n; i++) {
        rs += r_[i] * m_[i];
        vs += v_[i] * m_[i];
        ms += m_[i];
    }
    const Vec3 br = rs / ms, bv = vs / ms;
    std::vector<Vec3> R(n), V(n);
    for (size_t i = 0; i < n; i++) {
        R[i] = r_[i] - br;
        V[i] = v_[i] - bv;
    }
    if (energy) {
        double ke = 0.0, pe = 0.0;
        for (size_t i = 0; i < n; i++) ke += 0.5 * (V[i].x * V[i].x + V[i].y * V[i].y + V[i].z * V[i].z) * m_[i];
        for (size_t i = 0; i + 1 < n; i++)
            for (size_t j = i + 1; j < n; j++) {
                const double dx = R[i].x - R[j].x, dy = R[i].y - R[j].y, dz = R[i].z - R[j].z;
                pe -= m_[i] * m_[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
            }
        *energy = ke + pe;
    }
    if (l) {
        Vec3 L;
        for (size_t i = 0; i < n; i++) L += R[i].cross(V[i]) * m_[i];
        *l = L;
    }
}

void h0_from_log(uint32_t n_bodies, const double* m, const double* r_log, uint32_t n_steps, double* h0_log) {
    const size_t row = n_bodies - 1;
    for (uint32_t s = 0; s < n_steps; s++) {
        Vec3 h0, common;
        for (uint32_t i = 1; i < n_bodies; i++) {
            const Vec3 r = Vec3::load(r_log, s * row + i - 1);
            const double r2 = r.lensq();
            const double inv_r3 = 1. / (std::sqrt(r2) * r2);
            if (i == 1)
                h0 = -r * m[1] * inv_r3;
            else
                common -= r * (m[i] * inv_r3);
        }
        h0 += common;
        h0.store(h0_log, s);
    }
}

}  // namespace host
}  // namespace glx
