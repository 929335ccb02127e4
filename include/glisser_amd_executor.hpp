// glisser_amd_executor.hpp -- C++ host API of the time-block loop.
//
// glx::Executor is the C++ counterpart of the reference's Executor
// (src/executor.cu:598-877) driving the C ABI of glisser_amd.h: per
// time-block it launches the particle kernel on the planet block produced
// during the previous block, produces the next planet block on the host
// while the kernel runs (planet Jacobi WH, src/wh.cpp:584-595), then
// compacts the particles and records the newly flagged ones.  With
// resolve_encounters, particles flagged for a planet encounter are not
// killed: the stable three-way compaction keeps them alive, their state at
// the start of the block (the rollback copy) and at the flagging step go to
// the host close-encounter queue, and a resolver hands them back with
// return_particles() -- the hand-off of Executor::resync2 /
// handle_encounters (src/executor.cu:470-596,766-877); the resolver itself
// (SWIFT in the reference, src/swift.cpp) is the caller's.
//
// Errors throw std::runtime_error, as the reference host code does.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "glisser_amd.h"

namespace glx {

struct ExecutorConfig {
    double t0 = 0.0;
    double dt = 180.0;                // Time-Step
    uint32_t tbsize = 2048;           // Time-Block-Size
    double outer_radius = 300.0;      // Outer-Limit
    double cull_radius = 0.5;         // Cull-Radius
    double rh_factor = 0.0;           // Encounter-RH-Factor
    bool resolve_encounters = false;  // Resolve-Encounters
    int form = GLX_FORM_FP64;         // kernel arithmetic
    int device = 0;
    // Read-Planet-History: planets interpolated from this history file
    // (blocks then follow its records: each interval of length T holds
    // max(1, round(T / dt)) equal steps, at most tbsize per block) instead
    // of being integrated; empty = integrate the planets given.
    std::string planet_history;
    uint32_t history_table_size = 16;  // Planet-History-Max-Planet-Count
    // Resync-Interval: the active range shrinks, and the particles that left
    // it get their death times and log lines, every resync_every blocks (the
    // compaction itself runs every block, as the reference's presort does);
    // must be 1 with resolve_encounters, as the reference requires
    // (src/data.cpp:317-320, src/executor.cu:729-735).
    uint32_t resync_every = 1;
    // Death time as the reference's resync records it: the end time of the
    // block at whose resync the particle left the active range, cast to float
    // (src/executor.cu:816-822).  false: the time of the flagging step.
    bool reference_deathtime = false;
    // When not empty, the file of "id death t" / "id encounter t" lines the
    // reference writes to encounter.out at every resync for the particles
    // that left the active range ("death" when the 0x80 bit is set), t the
    // end time of that block with 17 significant digits
    // (src/executor.cu:205-208,828-838; targets/main.cpp:125-126).
    std::string encounter_log;
};

// One particle handed to the host close-encounter queue: id, flags (0x01,
// planet id in the high byte), step (steps of its block completed when
// flagged), t_block (time at the start of that block), r0/v0 (state at the
// start of the block, the rollback copy) and r/v (state at the flagging step).
using EncounterRecord = glx_encounter;

// Host copy of every particle slot (r, v: n x 3).
struct ParticleState {
    std::vector<double> r, v, deathtime;  // deathtime: see ExecutorConfig::reference_deathtime (NaN if not set)
    std::vector<uint16_t> flags;
    std::vector<uint32_t> id;
    uint64_t n_alive = 0;
};

class Executor {
public:
    // Planets: n_bodies GM values m (m[0] the Sun), heliocentric r, v
    // (n_bodies x 3), ids (NULL: 0..n-1); not used (may be NULL) when
    // cfg.planet_history names a history.  Particles: n heliocentric r, v
    // and ids.  Uploads everything, computes the initial accelerations and
    // flags, compacts, and produces the first planet block.
    Executor(const ExecutorConfig& cfg, uint32_t n_bodies, const double* m, const double* r, const double* v,
             const uint32_t* planet_ids, uint64_t n, const double* pr, const double* pv, const uint32_t* ids);
    ~Executor();
    Executor(const Executor&) = delete;
    Executor& operator=(const Executor&) = delete;

    // One time-block (Executor::loop, and resync2 every resync_every blocks).
    void step();
    void run(uint32_t n_blocks);

    double time() const;
    uint64_t n_alive() const;
    uint64_t n_total() const;
    uint64_t n_queued() const;  // records waiting in the encounter queue

    // Take every queued record (the queue is emptied).
    std::vector<EncounterRecord> take_encounters();
    // The same into out[0, n_queued()), which must hold them all (else
    // nothing is taken and an exception says so); returns the count.
    uint64_t take_encounters(EncounterRecord* out, uint64_t capacity);
    // Hand resolved particles back at the current time(): heliocentric r, v
    // (n x 3) of particles previously taken from the queue.  Their
    // accelerations are recomputed from the planets at time() and an
    // encounter flag is cleared (the next step re-detects a lasting one),
    // as handle_encounters does (src/executor.cu:558-578).
    void return_particles(uint64_t n, const uint32_t* ids, const double* r, const double* v);

    ParticleState download() const;
    // The planets at time(): bodies (Sun first; the set can change between
    // blocks of a planet history), their GM values and ids, and the
    // heliocentric positions of bodies 1..n_bodies()-1 ((n_bodies() - 1) x 3).
    uint32_t n_bodies() const;
    std::vector<double> planet_masses() const;
    std::vector<uint32_t> planet_ids() const;
    std::vector<double> planet_positions() const;

    // Device time of the last kernel and compaction, host time of the last
    // planet block (ms).
    void last_times(double* kernel_ms, double* partition_ms, double* planet_ms) const;
    // Device time of every block kernel since construction (ms); *n_blocks
    // (if given) the number of blocks that ran one.
    double kernel_ms_total(uint64_t* n_blocks = nullptr) const;

private:
    struct Impl;
    std::unique_ptr<Impl> p_;
};

}  // namespace glx
