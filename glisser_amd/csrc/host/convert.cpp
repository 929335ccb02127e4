// Orbital conversions; evaluation order follows the reference's
// src/convert.cpp so host results round identically.
#include "convert.h"

#include <cmath>

namespace glx {
namespace host {

namespace {

constexpr double kPi = 3.14159265358979323846;

// Eccentric anomaly for high eccentricity (SWIFT's orbel_ehie; reference
// src/convert.cpp:156-190): cube-root first guess, three quartic Newton
// corrections.
double ecc_anomaly_high(double e, double m) {
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
        double sa = std::sin(x + m), ca = std::cos(x + m);
        double esa = e * sa, eca = e * ca;
        double f = x - esa;
        double fp = 1 - eca;
        double dx = -f / fp;
        dx = -f / (fp + dx * esa / 2);
        dx = -f / (fp + dx * (esa + eca * dx / 3) / 2);
        x = x + dx;
    }
    return iflag ? 2 * kPi - m - x : m + x;
}

// Series start plus n quartic Newton corrections (src/convert.cpp:192-221).
double ecc_anomaly_series(double e, double m, int n) {
    double sm = std::sin(m), cm = std::cos(m);
    double x = m + e * sm * (1 + e * (cm + e * (1 - 1.5 * sm * sm)));
    for (int it = 0; it < n; it++) {
        double sx = std::sin(x), cx = std::cos(x);
        double es = e * sx, ec = e * cx;
        double f = x - es - m;
        double fp = 1 - ec;
        double fpp = es, fppp = ec;
        double dx = -f / fp;
        dx = -f / (fp + dx * fpp / 2);
        dx = -f / (fp + dx * fpp / 2 + dx * dx * fppp / 6);
        x = x + dx;
    }
    return x;
}

double ecc_anomaly(double e, double m) {
    if (e < 0.18) return ecc_anomaly_series(e, m, 1);
    if (e < 0.8) return ecc_anomaly_series(e, m, 2);
    return ecc_anomaly_high(e, m);
}

}  // namespace

Elements to_elements(double mu, const Vec3& r, const Vec3& v) {
    Elements el;
    const double x = r.x, y = r.y, z = r.z;
    const double vx = v.x, vy = v.y, vz = v.z;
    const double pi = 2.0 * std::asin(1.0);
    const double prec = 1.0e-13;
    const double hx = y * vz - z * vy;
    const double hy = z * vx - x * vz;
    const double hz = x * vy - y * vx;
    const double hsq = hx * hx + hy * hy + hz * hz;
    if (!(hsq > prec)) {  // radial orbit
        el.esign = 1;
        el.a = std::sqrt(x * x + y * y + z * z);
        el.e = HUGE_VAL;
        el.i = std::asin(z / std::sqrt(x * x + y * y + z * z));
        el.capom = std::atan2(y, x);
        el.om = HUGE_VAL;
        el.f = HUGE_VAL;
        return el;
    }
    double i = std::acos(hz / std::sqrt(hsq));
    double capom;
    if (std::fabs(i) < prec || std::fabs(pi - std::fabs(i)) < prec) {
        capom = 0.0;
    } else {
        capom = std::atan2(hx, -hy);
    }
    const double vsq = vx * vx + vy * vy + vz * vz;
    const double vdotr = x * vx + y * vy + z * vz;
    const double rr = std::sqrt(x * x + y * y + z * z);
    const double xhat = x / rr, yhat = y / rr, zhat = z / rr;
    const double nx = std::cos(capom), ny = std::sin(capom);
    double fac = vsq * rr - mu;
    const double Px = fac * xhat - vdotr * vx;
    const double Py = fac * yhat - vdotr * vy;
    const double Pz = fac * zhat - vdotr * vz;
    const double modP = std::sqrt(Px * Px + Py * Py + Pz * Pz);
    const double e = modP / mu;
    double om;
    if (std::fabs(e) < prec) {
        om = 0.0;
    } else if ((i < prec) || (pi - i < prec)) {
        om = std::atan2(Py, Px);
    } else {
        const double ecosw = (nx * Px + ny * Py) / mu;
        om = std::acos(ecosw / e);
        if (std::fabs(Pz) > prec) om *= std::fabs(Pz) / Pz;
    }
    const double energy = vsq / 2.0 - mu / rr;
    double a, f;
    int esign;
    if (std::fabs(energy) < prec) {
        esign = 0;
        a = 0.5 * hsq / mu;
        f = std::fabs(vdotr) < prec ? 0.0 : 2.0 * std::acos(std::sqrt(a / rr)) * vdotr / std::fabs(vdotr);
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
                f = rr < a ? 0.0 : pi;
            } else {
                fac = a * (1.0 - e * e) / rr - 1.0;
                f = std::acos(fac / e) * vdotr / std::fabs(vdotr);
                if (std::isnan(f)) f = rr < a ? 0.0 : pi;
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
    el.a = a;
    el.e = e;
    el.i = i;
    el.capom = capom;
    el.om = om;
    el.f = f;
    el.esign = esign;
    return el;
}

void from_elements(double mu, double a, double e, double i, double capom, double om, double f, Vec3* rv, Vec3* vv) {
    const double prec = 1.0e-13;
    const double u = om + f;
    const double xhat = std::cos(u) * std::cos(capom) - std::cos(i) * std::sin(capom) * std::sin(u);
    const double yhat = std::cos(u) * std::sin(capom) + std::cos(i) * std::cos(capom) * std::sin(u);
    const double zhat = std::sin(i) * std::sin(u);
    const double hx = std::sin(capom) * std::sin(i);
    const double hy = -std::cos(capom) * std::sin(i);
    const double hz = std::cos(i);
    double r, h;
    if (std::fabs(e - 1.0) > prec) {
        r = a * (1.0 - e * e) / (1.0 + e * std::cos(f));
        h = std::sqrt(mu * a * (1.0 - e * e));
    } else {
        h = std::sqrt(2.0 * mu * a);
        r = a / (std::cos(f / 2.0) * std::cos(f / 2.0));
    }
    *rv = Vec3(r * xhat, r * yhat, r * zhat);
    const double thx = hy * zhat - hz * yhat;
    const double thy = hz * xhat - hx * zhat;
    const double thz = hx * yhat - hy * xhat;
    const double thdot = h / (r * r);
    const double rdot = e * mu * std::sin(f) / h;
    *vv = Vec3(r * thdot * thx + rdot * xhat, r * thdot * thy + rdot * yhat, r * thdot * thz + rdot * zhat);
}

void from_elements_M(double mu, double a, double e, double i, double capom, double om, double M, Vec3* r, Vec3* v) {
    double E = ecc_anomaly(e, M);
    double f = std::acos((std::cos(E) - e) / (1 - e * std::cos(E)));
    E = E - static_cast<int>(E / 2 / kPi) * 2 * kPi;
    if (E > kPi)
        E -= 2 * kPi;
    else if (E < -kPi)
        E += 2 * kPi;
    if (E < 0) f = -f;
    from_elements(mu, a, e, i, capom, om, f, r, v);
}

double mean_anomaly(double e, double f) {
    double E = std::acos((e + std::cos(f)) / (1 + e * std::cos(f)));
    if (f < 0) E = -E;
    return E - e * std::sin(E);
}

void helio_to_jacobi_r(const std::vector<Vec3>& r, const std::vector<double>& m, const std::vector<double>& eta,
                       std::vector<Vec3>& rj) {
    const size_t n = r.size();
    rj[0] = Vec3();
    if (n < 2) return;
    rj[1] = r[1];
    Vec3 sum;
    for (size_t i = 2; i < n; i++) {
        sum += r[i - 1] * m[i - 1];
        rj[i] = r[i] - sum / eta[i - 1];
    }
}

void helio_to_jacobi_v(const std::vector<Vec3>& v, const std::vector<double>& m, const std::vector<double>& eta,
                       std::vector<Vec3>& vj) {
    const size_t n = v.size();
    vj[0] = Vec3();
    if (n < 2) return;
    vj[1] = v[1];
    Vec3 psum;
    for (size_t i = 2; i < n; i++) {
        psum += v[i - 1] * m[i - 1];
        vj[i] = v[i] - psum / eta[i - 1];
    }
}

void jacobi_to_helio(const std::vector<double>& m, const std::vector<double>& eta, const std::vector<Vec3>& rj,
                     const std::vector<Vec3>& vj, std::vector<Vec3>& r, std::vector<Vec3>& v) {
    const size_t n = r.size();
    r[0] = Vec3();
    v[0] = Vec3();
    if (n < 2) return;
    r[1] = rj[1];
    v[1] = vj[1];
    Vec3 rsum, vsum;
    for (size_t i = 2; i < n; i++) {
        const double w = m[i - 1] / eta[i - 1];
        rsum += rj[i - 1] * w;
        vsum += vj[i - 1] * w;
        r[i] = rj[i] + rsum;
        v[i] = vj[i] + vsum;
    }
}

}  // namespace host
}  // namespace glx
