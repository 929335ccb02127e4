// Host producer of the particle kernel's input: the planets' Jacobi
// Wisdom-Holman integrator (the planet half of the reference's
// WHIntegrator, src/wh.h:19-147, src/wh.cpp:422-595,696-756,993-1067).
//
// Per time-block it steps the planets n_steps times and writes the block's
// planet log (heliocentric positions/velocities of bodies 1..n-1 at
// t0 + (s+1) dt) and the indirect-term log h0.  Few bodies (<= 16) and a
// strictly sequential chain of steps: this stays on the CPU and overlaps
// the GPU kernel of the previous block.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "vec3.h"

namespace glx {
namespace host {

struct PlanetError : std::runtime_error {
    explicit PlanetError(const std::string& s) : std::runtime_error(s) {}
};

class PlanetIntegrator {
public:
    // m in GM units (m[0] is the Sun), heliocentric r, v.  Computes eta,
    // encounter radii, Jacobi coordinates and the first accelerations; the
    // indirect term at the initial time is available from h0_initial().
    PlanetIntegrator(std::vector<double> m, std::vector<Vec3> r, std::vector<Vec3> v, std::vector<uint32_t> id,
                     double cull_radius, double rh_factor, double outer_radius);

    // n_steps planet steps of size dt.  r_log / v_log: n_steps x (n-1) x 3,
    // h0_log: n_steps x 3 (row s at t0 + (s+1) dt).  Throws PlanetError on
    // an unbound or unconverged planet orbit, as the reference does.
    void integrate_block(uint32_t n_steps, double dt, double* r_log, double* v_log, double* h0_log);

    // Encounter radii from the current osculating elements
    // (src/wh.cpp:472-495).
    void recalculate_rh();

    size_t n() const { return m_.size(); }
    const std::vector<double>& m() const { return m_; }
    const std::vector<uint32_t>& id() const { return id_; }
    const std::vector<double>& rh() const { return rh_; }
    const std::vector<Vec3>& r() const { return r_; }
    const std::vector<Vec3>& v() const { return v_; }
    Vec3 h0_initial() const { return h0_init_; }
    double outer_radius() const { return outer_; }

    // Barycentric energy and angular momentum (src/wh.cpp:422-470).
    void metrics(double* energy, Vec3* l) const;

private:
    Vec3 accelerations();  // returns h0; src/wh.cpp:696-756
    void drift_jacobi(double dt);
    void step(double dt, Vec3* r_row, Vec3* v_row, Vec3* h0);

    std::vector<double> m_, eta_, rh_, mu_;
    std::vector<uint32_t> id_;
    std::vector<Vec3> r_, v_, rj_, vj_, acc_;
    std::vector<double> ih3_, ij3_;
    double cull_, rh_factor_, outer_;
    Vec3 h0_init_;
};

// Encounter radii of bodies 0..n-1 (src/wh.cpp:472-495): with
// |rh_factor| > 1e-6 the Sun gets cull_radius and planet i
// rh_factor a (1 - e) (m_i / 3 m_0)^(1/3) from its heliocentric osculating
// elements; otherwise every body gets cull_radius.
void encounter_radii(size_t n, const double* m, const Vec3* r, const Vec3* v, double cull_radius, double rh_factor,
                     double* rh);

// Indirect-term rows from a position log (src/wh.cpp:993-1025).
void h0_from_log(uint32_t n_bodies, const double* m, const double* r_log, uint32_t n_steps, double* h0_log);

}  // namespace host
}  // namespace glx
