// Host-side 3-vector with the reference f64_3's element-wise semantics and
// evaluation order (src/types.h:78-213): products and sums are per
// component, lensq is x*x + y*y + z*z left to right.  Host code built on it
// is compiled without FP contraction so it rounds like the reference's CPU
// build.
#pragma once

#include <cstddef>

namespace glx {
namespace host {

struct Vec3 {
    double x = 0, y = 0, z = 0;

    Vec3() = default;
    Vec3(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}

    static Vec3 load(const double* p, size_t i) { return Vec3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
    void store(double* p, size_t i) const {
        p[3 * i] = x;
        p[3 * i + 1] = y;
        p[3 * i + 2] = z;
    }

    Vec3 operator+(const Vec3& b) const { return Vec3(x + b.x, y + b.y, z + b.z); }
    Vec3 operator-(const Vec3& b) const { return Vec3(x - b.x, y - b.y, z - b.z); }
    Vec3 operator-() const { return Vec3(-x, -y, -z); }
    Vec3 operator*(double s) const { return Vec3(x * s, y * s, z * s); }
    Vec3 operator/(double s) const { return Vec3(x / s, y / s, z / s); }
    Vec3& operator+=(const Vec3& b) {
        x += b.x;
        y += b.y;
        z += b.z;
        return *this;
    }
    Vec3& operator-=(const Vec3& b) {
        x -= b.x;
        y -= b.y;
        z -= b.z;
        return *this;
    }
    double lensq() const { return x * x + y * y + z * z; }
    Vec3 cross(const Vec3& b) const { return Vec3(y * b.z - z * b.y, z * b.x - x * b.z, x * b.y - y * b.x); }
};

}  // namespace host
}  // namespace glx
