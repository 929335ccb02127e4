#include "history.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <ostream>
#include <stdexcept>

#include "convert.h"
#include "planets.h"

namespace glx {
namespace host {

namespace {

constexpr double kPi = 3.14159265358979323846;

template <typename T>
T read_le(std::istream& in) {
    T x{};
    in.read(reinterpret_cast<char*>(&x), sizeof(T));
    return x;
}

template <typename T>
void write_le(std::ostream& out, T x) {
    out.write(reinterpret_cast<const char*>(&x), sizeof(T));
}

void zeros(std::ostream& out, size_t n) {
    for (size_t k = 0; k < n; k++) out.put('\0');
}

// angle difference along the shorter arc
double wrap_pi(double d) {
    if (d > kPi)
        d -= 2 * kPi;
    else if (d < -kPi)
        d += 2 * kPi;
    return d;
}

}  // namespace

void write_history_record(std::ostream& out, bool first, double time, uint32_t n, const double* m, const Vec3* r,
                          const Vec3* v, const uint32_t* id) {
    if (first) {
        write_le<double>(out, m[0]);
        zeros(out, 24);
    }
    write_le<double>(out, time);
    write_le<uint32_t>(out, n > 0 ? n - 1 : 0);
    zeros(out, 20);
    for (uint32_t i = 1; i < n; i++) {
        // heliocentric elements about the Sun (src/data.cpp:836-871)
        const double center_m = m[0];
        const Vec3 center_r = r[0] * m[0] / center_m;
        const Vec3 center_v = v[0] * m[0] / center_m;
        const Elements el = to_elements(m[i] + center_m, r[i] - center_r, v[i] - center_v);
        write_le<uint32_t>(out, id[i]);
        write_le<float>(out, static_cast<float>(m[i]));
        write_le<float>(out, static_cast<float>(el.a));
        write_le<float>(out, static_cast<float>(el.e));
        write_le<float>(out, static_cast<float>(el.i));
        write_le<float>(out, static_cast<float>(el.capom));
        write_le<float>(out, static_cast<float>(el.om));
        write_le<float>(out, static_cast<float>(mean_anomaly(el.e, el.f)));
    }
}

size_t PlanetHistory::slot(uint32_t id) {
    for (const auto& s : slots_)
        if (s.first == id) return s.second;
    const size_t k = slots_.size();
    if (k >= m_.size())
        throw std::runtime_error("planet history names more planets than Planet-History-Max-Planet-Count (" +
                                 std::to_string(m_.size()) + ")");
    slots_.emplace_back(id, k);
    id_[k] = id;
    return k;
}

void PlanetHistory::read_record_header() {
    t_end_ = read_le<double>(in_);
    if (!in_) {
        ended_ = true;
        throw HistoryEnd();
    }
    n_end_ = read_le<uint32_t>(in_) + 1;
    in_.seekg(20, std::ios_base::cur);
}

PlanetHistory::PlanetHistory(const std::string& path, double t0, double dt, uint32_t table_size)
    : in_(path, std::ios_base::binary), user_dt_(dt) {
    if (!in_) throw std::runtime_error("cannot open planet history " + path);
    if (table_size < 1) throw std::runtime_error("planet history: the table needs at least the Sun");
    if (!(dt > 0.0)) throw std::runtime_error("planet history: dt must be positive");
    const size_t n = table_size;
    m_.assign(n, 0.0);
    id_.assign(n, 0);
    r_.assign(n, Vec3());
    v_.assign(n, Vec3());
    aei_begin_.assign(n, Vec3());
    oom_begin_.assign(n, Vec3());
    aei_end_.assign(n, Vec3());
    oom_end_.assign(n, Vec3());
    gm_first_.assign(n, 0.0);
    m_[0] = read_le<double>(in_);
    in_.seekg(24, std::ios_base::cur);
    // first record (src/interp.cpp:13-91)
    t_begin_ = std::numeric_limits<double>::infinity();
    read_record_header();
    n_alive_ = n_end_;
    if (t_end_ > t0) throw std::runtime_error("time t0 not found in datafile: planet history starts after t0");
    slots_.assign(1, {0u, size_t(0)});
    if (n_end_ > n)
        throw std::runtime_error("planet history names more planets than Planet-History-Max-Planet-Count (" +
                                 std::to_string(n) + ")");
    for (uint32_t i = 1; i < n_end_; i++) {
        const uint32_t pid = read_le<uint32_t>(in_);
        const size_t k = slot(pid);
        gm_first_[k] = read_le<float>(in_);
        aei_end_[k].x = read_le<float>(in_);
        aei_end_[k].y = read_le<float>(in_);
        aei_end_[k].z = read_le<float>(in_);
        oom_end_[k].x = read_le<float>(in_);
        oom_end_[k].y = read_le<float>(in_);
        oom_end_[k].z = read_le<float>(in_);
        if (!in_) throw std::runtime_error("planet history: the record at t = " + std::to_string(t_end_) +
                                           " is truncated");
        from_elements_M(m_[0] + gm_first_[k], aei_end_[k].x, aei_end_[k].y, aei_end_[k].z, oom_end_[k].x,
                        oom_end_[k].y, oom_end_[k].z, &r_[i], &v_[i]);
        id_[i] = pid;
        m_[i] = gm_first_[k];
    }
    r_[0] = v_[0] = aei_end_[0] = oom_end_[0] = Vec3();
    // Executor::init (src/executor.cu:99-129): the interval holding t0, its
    // remaining steps, and the planets at t0
    t_ = t0;
    while (t_end_ <= t_) advance();
    if (std::isinf(t_begin_)) throw std::runtime_error("lookup file starts past the given Initial-Time");
    cur_ts_ = 0;
    n_ts_ = std::max<size_t>(1, static_cast<size_t>(std::round((t_end_ - t_) / dt)));
    eff_dt_ = (t_end_ - t_) / static_cast<double>(n_ts_);
    rel_t_ = t_ - t_begin_;
    n_alive_ = static_cast<uint32_t>(tracks_.size() + 1);
    for (size_t j = 0; j < tracks_.size(); j++) state_at(tracks_[j], rel_t_, &r_[j + 1], &v_[j + 1]);
}

// Interpolator::next (src/interp.cpp:167-293): the record that ends the new
// interval and the linear / mean-motion rates of every planet present at
// both of its ends.
void PlanetHistory::advance() {
    aei_begin_ = aei_end_;
    oom_begin_ = oom_end_;
    t_begin_ = t_end_;
    rel_t_ = 0;
    read_record_header();
    const double span = t_end_ - t_begin_;
    n_ts_ = std::max<size_t>(1, static_cast<size_t>(std::round(span / user_dt_)));
    eff_dt_ = span / static_cast<double>(n_ts_);
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (size_t k = 1; k < m_.size(); k++) aei_end_[k] = oom_end_[k] = Vec3(nan, nan, nan);
    tracks_.clear();
    for (uint32_t i = 1; i < n_end_; i++) {
        const uint32_t pid = read_le<uint32_t>(in_);
        const size_t k = slot(pid);
        m_[k] = read_le<float>(in_);
        aei_end_[k].x = read_le<float>(in_);
        aei_end_[k].y = read_le<float>(in_);
        aei_end_[k].z = read_le<float>(in_);
        oom_end_[k].x = read_le<float>(in_);
        oom_end_[k].y = read_le<float>(in_);
        oom_end_[k].z = read_le<float>(in_);
        if (!in_) {
            ended_ = true;
            throw std::runtime_error("planet history: the record at t = " + std::to_string(t_end_) +
                                     " is truncated");
        }
        if (std::isnan(aei_begin_[k].x)) continue;  // absent at the interval start
        Track tr;
        tr.id = pid;
        tr.gm = gm_first_[k];
        tr.aei = aei_begin_[k];
        tr.oom = oom_begin_[k];
        tr.daei = (aei_end_[k] - aei_begin_[k]) / span;
        Vec3 d = oom_end_[k] - oom_begin_[k];
        d.x = wrap_pi(d.x);
        d.y = wrap_pi(d.y);
        d = d / span;
        // two-body rate of the mean 1/a^1.5, corrected to land on the recorded anomaly
        double n_mean = 2 * kPi * std::sqrt(1 + m_[k] / m_[0]) *
                        (std::pow(aei_begin_[k].x, -1.5) + std::pow(aei_end_[k].x, -1.5)) / 2;
        const double landing = std::fmod(oom_begin_[k].z + n_mean * span, 2 * kPi);
        n_mean += wrap_pi(oom_end_[k].z - landing) / span;
        d.z = n_mean;
        tr.doom = d;
        tracks_.push_back(tr);
    }
}

void PlanetHistory::state_at(const Track& k, double rel, Vec3* r, Vec3* v) const {
    const Vec3 aei = k.aei + k.daei * rel;
    const Vec3 oom = k.oom + k.doom * rel;
    from_elements_M(k.gm + m_[0], aei.x, aei.y, aei.z, oom.x, oom.y, oom.z, r, v);
}

// Executor::update_planets with an interpolated history
// (src/executor.cu:235-292, Interpolator::fill src/interp.cpp:113-165) and
// the time advance of src/executor.cu:700.
void PlanetHistory::next_block(uint32_t tbsize, HistoryBlock* out) {
    if (tbsize < 1) throw std::runtime_error("planet history: tbsize must be positive");
    if (ended_) throw HistoryEnd();
    if (cur_ts_ == n_ts_) {
        const double diff = rel_t_ - (t_end_ - t_begin_);
        if (!(std::abs(diff) < 1e-8))
            throw std::runtime_error("sanity check failed - adjusting t by too much: " + std::to_string(diff));
        advance();
        cur_ts_ = 0;
        t_ = t_begin_;
    }
    const double dt = eff_dt_;
    const uint32_t n = std::min(tbsize, static_cast<uint32_t>(n_ts_ - cur_ts_));
    const size_t np = tracks_.size();
    n_alive_ = static_cast<uint32_t>(np + 1);
    for (size_t j = 0; j < np; j++) {
        id_[j + 1] = tracks_[j].id;
        m_[j + 1] = tracks_[j].gm;
    }
    out->r_log.resize(static_cast<size_t>(n) * np * 3);
    out->h0_log.resize(static_cast<size_t>(n) * 3);
    double rel = rel_t_ + dt;  // the first row is one step past the block start
    for (uint32_t s = 0; s < n; s++) {
        if (rel > t_end_ - t_begin_ + 1e-8)
            throw std::runtime_error("interpolation ran too long, timechunk exceeded lookup boundary");
        r_[0] = v_[0] = Vec3();
        for (size_t j = 0; j < np; j++) {
            state_at(tracks_[j], rel, &r_[j + 1], &v_[j + 1]);
            r_[j + 1].store(out->r_log.data(), s * np + j);
        }
        rel += dt;
    }
    h0_from_log(n_alive_, m_.data(), out->r_log.data(), n, out->h0_log.data());
    out->t_start = t_;
    out->dt = dt;
    out->n_steps = n;
    out->n_bodies = n_alive_;
    out->m.assign(m_.begin(), m_.begin() + n_alive_);
    out->id.assign(id_.begin(), id_.begin() + n_alive_);
    rel_t_ += static_cast<double>(n) * dt;
    cur_ts_ += n;
    t_ += dt * static_cast<double>(n);
}

}  // namespace host
}  // namespace glx
