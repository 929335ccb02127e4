// Orbital-element and frame conversions used by the host side of the
// particle path: osculating elements for encounter radii and the precision
// report, Cartesian states from elements for the planet-history producer,
// heliocentric <-> Jacobi transforms for the planet integrator.
#pragma once

#include <cstddef>
#include <vector>

#include "vec3.h"

namespace glx {
namespace host {

struct Elements {
    double a = 0, e = 0, i = 0, capom = 0, om = 0, f = 0;
    int esign = 0;  // -1 elliptic, 0 parabolic, +1 hyperbolic
};

// Osculating elements about central GM mu (src/convert.cpp:310-463).
Elements to_elements(double mu, const Vec3& r, const Vec3& v);

// Cartesian state from elements with true anomaly f (src/convert.cpp:250-308).
void from_elements(double mu, double a, double e, double i, double capom, double om, double f, Vec3* r, Vec3* v);

// Same from mean anomaly M (src/convert.cpp:230-248): eccentric anomaly by
// the hybrid Newton/Danby solver of src/convert.cpp:156-228.
void from_elements_M(double mu, double a, double e, double i, double capom, double om, double M, Vec3* r, Vec3* v);

// Mean anomaly from true anomaly (src/convert.cpp:465-470).
double mean_anomaly(double e, double f);

// Heliocentric <-> Jacobi for a body list whose entry 0 is the Sun
// (src/convert.cpp:9-95).  eta[i] = m[0] + ... + m[i].
void helio_to_jacobi_r(const std::vector<Vec3>& r, const std::vector<double>& m, const std::vector<double>& eta,
                       std::vector<Vec3>& rj);
void helio_to_jacobi_v(const std::vector<Vec3>& v, const std::vector<double>& m, const std::vector<double>& eta,
                       std::vector<Vec3>& vj);
void jacobi_to_helio(const std::vector<double>& m, const std::vector<double>& eta, const std::vector<Vec3>& rj,
                     const std::vector<Vec3>& vj, std::vector<Vec3>& r, std::vector<Vec3>& v);

}  // namespace host
}  // namespace glx
