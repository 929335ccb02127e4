// Planet history files: the reference's "planetary history" format
// (README.md:154-171; written by save_swift_plhist, src/data.cpp:817-873)
// and the element-linear interpolation that turns one into planet blocks
// (the Interpolator of src/interp.cpp, driven the way Executor::init and
// Executor::update_planets drive it, src/executor.cu:93-130,235-292,700).
//
// File layout (little-endian): f64 GM of the Sun, 24 bytes of padding; then
// records of f64 time, u32 planet count, 20 bytes of padding and per planet
// u32 id + 7 float32 (GM, a, e, i, Omega, omega, mean anomaly), heliocentric.
//
// Between two records every planet present in both moves with its a, e, i,
// Omega and omega linear in time (the angles along the shorter arc) and its
// mean anomaly at a constant rate: the two-body rate of the interval's
// mean semi-major axis, corrected so that the interval ends on the
// recorded anomaly.  An interval of length T holds n = max(1, round(T / dt))
// steps of T / n, so blocks end on record times; a block never crosses one.
//
// Built without FP contraction so the rows equal the reference CPU build's
// bit for bit.
#pragma once

#include <cstdint>
#include <fstream>
#include <iosfwd>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vec3.h"

namespace glx {
namespace host {

// Appends one record of n heliocentric bodies (entry 0 the Sun; the others
// as float32 elements about the Sun with GM m[0] + m[i]).  With first set
// the file header (the Sun's GM) is written before it.
void write_history_record(std::ostream& out, bool first, double time, uint32_t n, const double* m, const Vec3* r,
                          const Vec3* v, const uint32_t* id);

// Thrown when a block is asked for past the last record of the file.
struct HistoryEnd : std::runtime_error {
    HistoryEnd() : std::runtime_error("EOS reached for lookup file.") {}
};

// One planet block from a history: start time, step, number of steps,
// bodies (entry 0 the Sun), position log (n_steps x (n_bodies-1) x 3,
// row s at t_start + (s+1) dt) and indirect-term log (n_steps x 3).
struct HistoryBlock {
    double t_start = 0, dt = 0;
    uint32_t n_steps = 0, n_bodies = 0;
    std::vector<double> m;
    std::vector<uint32_t> id;
    std::vector<double> r_log, h0_log;
};

class PlanetHistory {
public:
    // Opens the file for a run starting at t0 with Time-Step dt;
    // table_size is Planet-History-Max-Planet-Count (slots for the Sun and
    // every planet id the file names).  Positions the reader on the interval
    // holding t0 and fills the planet table at t0.  Throws
    // std::runtime_error with the reference's messages: the file starts
    // after t0, ends before it, or names more planets than the table holds;
    // a record cut short is an error too (the reference would read past it).
    PlanetHistory(const std::string& path, double t0, double dt, uint32_t table_size);

    // The next block, at most tbsize steps; throws HistoryEnd at the end of
    // the file.
    void next_block(uint32_t tbsize, HistoryBlock* out);

    // Planet table (table_size slots; the first n_bodies() in use): GM, ids,
    // heliocentric state of the last interpolated step.
    uint32_t n_bodies() const { return n_alive_; }
    uint32_t table_size() const { return static_cast<uint32_t>(m_.size()); }
    const std::vector<double>& m() const { return m_; }
    const std::vector<uint32_t>& id() const { return id_; }
    const std::vector<Vec3>& r() const { return r_; }
    const std::vector<Vec3>& v() const { return v_; }
    // Start time of the next block.
    double time() const { return t_; }

private:
    // one planet present at both ends of the current interval
    struct Track {
        uint32_t id = 0;
        double gm = 0;  // planet GM used for the interval
        Vec3 aei, oom;     // a, e, i and Omega, omega, M at the interval start
        Vec3 daei, doom;   // their rates
    };

    size_t slot(uint32_t id);
    void read_record_header();
    void advance();  // Interpolator::next
    void state_at(const Track& k, double rel, Vec3* r, Vec3* v) const;

    std::ifstream in_;
    double user_dt_ = 0;
    // planet table
    std::vector<double> m_;
    std::vector<uint32_t> id_;
    std::vector<Vec3> r_, v_;
    uint32_t n_alive_ = 0;
    std::vector<std::pair<uint32_t, size_t>> slots_;  // id -> table slot, in order of appearance
    // records at both ends of the interval, by slot
    std::vector<Vec3> aei_begin_, oom_begin_, aei_end_, oom_end_;
    std::vector<double> gm_first_;  // each planet's GM from the first record
    double t_begin_ = 0, t_end_ = 0;
    uint32_t n_end_ = 0;  // bodies in the end record (Sun included)
    std::vector<Track> tracks_;
    size_t n_ts_ = 0, cur_ts_ = 0;
    double eff_dt_ = 0, rel_t_ = 0;
    double t_ = 0;
    bool ended_ = false;  // the last record has been passed
};

}  // namespace host
}  // namespace glx
