// Host time-block loop over the C ABI (class glx::Executor) and its C entry
// points (glx_exec_*).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <fstream>
#include <iomanip>
#include <limits>
#include <new>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../errors.h"
#include "../staging.h"
#include "glisser_amd.h"
#include "glisser_amd_executor.hpp"
#include "history.h"
#include "planets.h"

namespace glx {

namespace {

void check(int status, const char* what) {
    if (status != GLX_OK) throw std::runtime_error(std::string(what) + ": " + glx_last_error());
}

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

}  // namespace

struct Executor::Impl {
    ExecutorConfig cfg;
    // planets: integrated on the host, or interpolated from a history file
    std::unique_ptr<host::PlanetIntegrator> planets;
    std::unique_ptr<host::PlanetHistory> history;
    glx_ctx* ctx = nullptr;
    uint32_t nb_max = 0;
    // n_alive: the active range the kernel runs on (shrunk at resyncs);
    // n_split: unflagged particles after the last compaction
    uint64_t n_total = 0, n_alive = 0, n_split = 0;
    uint64_t resync_counter = 0;
    std::ofstream log;
    double t = 0.0;
    int buf = 0;
    // one planet block: start time, step, length, bodies and their radii,
    // position / velocity / h0 logs
    struct Block {
        double t0 = 0, dt = 0;
        uint32_t n_steps = 0, nb = 0;
        std::vector<double> m, rh, log_r, log_v, log_h0;
        std::vector<uint32_t> id;
    };
    // next block, produced while the kernel runs the current one
    Block next;
    // the planets at time t (positions of bodies 1..nb-1, then h0), with the
    // bodies and radii they were stepped with: the input of glx_accel_range
    // for particles handed back at t
    uint32_t row_nb = 0;
    std::vector<double> row_r, row_h0, row_m, row_rh;
    std::vector<uint32_t> row_id;
    // the host close-encounter queue, one chunk per staging (no regrowth
    // of one large buffer while records arrive)
    std::vector<std::vector<EncounterRecord>> queue;
    uint64_t queued = 0;
    // time of the flagging step by particle id: a flat table when the ids
    // are dense (max id < 4 n), else a hash map
    std::vector<double> deathtime_flat;
    std::unordered_map<uint32_t, double> deathtime_map;

    void set_deathtime(uint32_t id, double t) {
        if (!deathtime_flat.empty())
            deathtime_flat[id] = t;
        else
            deathtime_map[id] = t;
    }
    void clear_deathtime(uint32_t id) {
        if (!deathtime_flat.empty())
            deathtime_flat[id] = std::numeric_limits<double>::quiet_NaN();
        else
            deathtime_map.erase(id);
    }
    double get_deathtime(uint32_t id) const {
        if (!deathtime_flat.empty()) return deathtime_flat[id];
        auto it = deathtime_map.find(id);
        return it == deathtime_map.end() ? std::numeric_limits<double>::quiet_NaN() : it->second;
    }
    double planet_ms = 0.0;
    std::string history_end;  // set once a planet history has no block left

    explicit Impl(const ExecutorConfig& c) : cfg(c) {
        if (cfg.resync_every < 1) throw std::runtime_error("Executor: resync_every must be at least 1");
        if (cfg.resync_every != 1 && cfg.resolve_encounters)
            throw std::runtime_error("Error: Resync-Interval must be 1 when Resolve-Encounters is enabled");
        if (!cfg.encounter_log.empty()) {
            log.open(cfg.encounter_log);
            if (!log) throw std::runtime_error("Executor: cannot open encounter log " + cfg.encounter_log);
            log << std::setprecision(17);
        }
    }

    ~Impl() {
        if (ctx) glx_destroy(ctx);
    }

    // The block that starts at t_start into `next`, with the radii of its
    // start (recalculate_rh before the planets move, as update_planets does,
    // src/executor.cu:226-299).
    void produce_block(double t_start) {
        const auto c0 = Clock::now();
        Block& b = next;
        if (history) {
            const host::PlanetHistory& h = *history;
            std::vector<double> rh_all(h.table_size());
            host::encounter_radii(h.table_size(), h.m().data(), h.r().data(), h.v().data(), cfg.cull_radius,
                                  cfg.rh_factor, rh_all.data());
            host::HistoryBlock hb;
            try {
                history->next_block(cfg.tbsize, &hb);
            } catch (const host::HistoryEnd& e) {
                // the run ends with the history; the block before still runs
                b.n_steps = 0;
                history_end = e.what();
                planet_ms = ms_since(c0);
                return;
            }
            b.t0 = hb.t_start;
            b.dt = hb.dt;
            b.n_steps = hb.n_steps;
            b.nb = hb.n_bodies;
            b.m = std::move(hb.m);
            b.id = std::move(hb.id);
            b.rh.assign(rh_all.begin(), rh_all.begin() + b.nb);
            b.log_r = std::move(hb.r_log);
            b.log_h0 = std::move(hb.h0_log);
        } else {
            planets->recalculate_rh();
            b.t0 = t_start;
            b.dt = cfg.dt;
            b.n_steps = cfg.tbsize;
            b.nb = static_cast<uint32_t>(planets->n());
            b.m = planets->m();
            b.id = planets->id();
            b.rh = planets->rh();
            const size_t np = b.nb - 1;
            b.log_r.resize((size_t)cfg.tbsize * np * 3);
            b.log_v.resize((size_t)cfg.tbsize * np * 3);
            b.log_h0.resize((size_t)cfg.tbsize * 3);
            planets->integrate_block(cfg.tbsize, cfg.dt, b.log_r.data(), b.log_v.data(), b.log_h0.data());
        }
        planet_ms = ms_since(c0);
    }

    void upload_block() {
        const Block& b = next;
        if (b.n_steps == 0) return;  // past the end of a planet history
        check(glx_upload_planet_block(ctx, buf, b.nb, b.m.data(), b.id.data(), b.rh.data(), b.log_r.data(),
                                      b.log_h0.data(), b.n_steps),
              "glx_upload_planet_block");
    }

    // the planets at the end of block b become the row at time t
    void take_row(const Block& b) {
        const size_t np = b.nb - 1;
        row_nb = b.nb;
        row_m = b.m;
        row_id = b.id;
        row_rh = b.rh;
        row_r.assign(b.log_r.begin() + (size_t)(b.n_steps - 1) * np * 3, b.log_r.begin() + (size_t)b.n_steps * np * 3);
        row_h0.assign(b.log_h0.begin() + (size_t)(b.n_steps - 1) * 3, b.log_h0.begin() + (size_t)b.n_steps * 3);
    }

    // Records of the particles a block flagged, packed on the device and on
    // their way to the context's pinned staging buffer, not yet moved to the
    // queue and the death-time table (the next block's kernel runs
    // meanwhile).
    uint64_t staged_n = 0, staged_enc = 0;
    // device time of every block kernel so far, and their count
    double kernel_ms_total = 0.0;
    uint64_t kernel_blocks = 0;
    double staged_t = 0.0, staged_dt = 0.0;
    bool staged_resync = false;  // the records are those a resync removed
    double staged_t_end = 0.0;   // end time of that block

    // Stages the records of particles [begin, end) after the compaction of
    // the block that started at t_block: the particles it flagged or, at a
    // resync, every particle that leaves the active range (those flagged in
    // it and in the blocks since the last resync).  In resolve mode the first
    // n_enc of them (the encounter class) go to the queue with their
    // rollback state.
    void stage_flagged(uint64_t begin, uint64_t end, uint64_t n_enc, double t_block, double dt_block, bool resync,
                       double t_end) {
        const uint64_t n = end - begin;
        if (n == 0) return;
        check(pack_encounters_begin(ctx, begin, n, t_block), "glx_pack_encounters");
        staged_n = n;
        staged_enc = n_enc;
        staged_t = t_block;
        staged_dt = dt_block;
        staged_resync = resync;
        staged_t_end = t_end;
    }

    // Moves staged records to the death-time table and the queue, read where
    // the device copy put them.  Death time: the flagging step's time, set
    // once (a particle flagged in an earlier block of the same resync already
    // has it), or with reference_deathtime the resync block's end time cast
    // to float; after a resync also the log lines (src/executor.cu:816-838).
    void take_staged() {
        if (staged_n == 0) return;
        const glx_encounter* staged = nullptr;
        check(pack_encounters_wait(ctx, &staged), "glx_pack_encounters");
        const float t_end_f = static_cast<float>(staged_t_end);
        for (uint64_t k = 0; k < staged_n; k++) {
            const glx_encounter& e = staged[k];
            if (!cfg.reference_deathtime) {
                if (std::isnan(get_deathtime(e.id))) set_deathtime(e.id, staged_t + (double)e.step * staged_dt);
            } else if (staged_resync) {
                set_deathtime(e.id, static_cast<double>(t_end_f));
            }
            if (staged_resync && log.is_open())
                log << e.id << ((e.flags & 0x80) ? " death " : " encounter ") << staged_t_end << '\n';
        }
        if (staged_resync && log.is_open()) log.flush();
        if (staged_enc) {
            queue.emplace_back(staged, staged + staged_enc);
            queued += staged_enc;
        }
        staged_n = staged_enc = 0;
        staged_resync = false;
    }

    // One time-block: its kernel; then, on the host while the kernel runs,
    // the records the previous block staged and the next planet block; then
    // the compaction, this block's records staged, and the next planet block
    // uploaded.
    void block() {
        if (next.n_steps == 0) {
            take_staged();
            throw std::runtime_error("planet history ended at t = " + std::to_string(t) + ": " + history_end);
        }
        const uint64_t prev = n_alive;
        const double t_block = next.t0, dt_block = next.dt;
        const uint32_t n_steps = next.n_steps;
        check(glx_set_counts(ctx, prev, n_total), "glx_set_counts");
        if (prev) {
            if (cfg.resolve_encounters) check(glx_snapshot_rollback(ctx), "glx_snapshot_rollback");
            check(glx_run_block(ctx, buf, dt_block, cfg.outer_radius, cfg.form), "glx_run_block");
        }
        take_staged();
        // the planets at the end of this block, then the next block
        take_row(next);
        produce_block(t_block + (double)n_steps * dt_block);
        if (prev) {
            // After the stable compaction the particles flagged in this block
            // sit at [new split, split before), followed by those flagged
            // in the blocks since the last resync; a resync removes all of
            // them from the active range.
            const uint64_t before = n_split;
            const uint64_t ne = compact(prev);
            float k_ms = 0.0f;  // the kernel ended before the compaction's counts came back
            check(glx_last_times(ctx, &k_ms, nullptr), "glx_last_times");
            kernel_ms_total += k_ms;
            kernel_blocks++;
            const bool resync = ++resync_counter % cfg.resync_every == 0;
            const double t_end = t_block + (double)n_steps * dt_block;
            stage_flagged(n_split, resync ? prev : before, cfg.resolve_encounters ? ne : 0, t_block, dt_block, resync,
                          t_end);
            if (resync) n_alive = n_split;
        }
        buf ^= 1;
        upload_block();
        t = next.n_steps ? next.t0 : t_block + (double)n_steps * dt_block;
    }

    // Stable compaction of [0, n) (kill first unless resolving): n_split
    // becomes the unflagged count; returns the encounter-class size.
    uint64_t compact(uint64_t n) {
        check(glx_set_counts(ctx, n, n_total), "glx_set_counts");
        if (!cfg.resolve_encounters) check(glx_kill_encounters(ctx), "glx_kill_encounters");
        uint64_t na = 0, ne = 0;
        check(glx_partition(ctx, cfg.resolve_encounters ? GLX_PART_THREE_WAY : GLX_PART_UNFLAGGED, &na, &ne),
              "glx_partition");
        n_split = na;
        return ne;
    }
};

Executor::Executor(const ExecutorConfig& cfg, uint32_t n_bodies, const double* m, const double* r, const double* v,
                   const uint32_t* planet_ids, uint64_t n, const double* pr, const double* pv, const uint32_t* ids) {
    const bool from_history = !cfg.planet_history.empty();
    if (!from_history) {
        if (n_bodies < 1 || n_bodies > GLX_MAX_BODIES) throw std::runtime_error("Executor: n_bodies outside [1, 16]");
        if (!m || !r || !v) throw std::runtime_error("Executor: null planet argument");
    } else if (cfg.history_table_size < 1 || cfg.history_table_size > GLX_MAX_BODIES) {
        throw std::runtime_error("Executor: history_table_size outside [1, 16]");
    }
    if (n && (!pr || !pv || !ids)) throw std::runtime_error("Executor: null particle argument");
    if (cfg.tbsize < 1 || !(cfg.dt > 0.0)) throw std::runtime_error("Executor: tbsize and dt must be positive");
    p_ = std::make_unique<Impl>(cfg);
    Impl& s = *p_;
    // planets at t0: bodies, positions, velocities, radii, h0
    std::vector<double> m0, rh0, row_r, h0(3);
    std::vector<uint32_t> id0;
    if (from_history) {
        // Executor::init with Read-Planet-History (src/executor.cu:93-133)
        s.history = std::make_unique<host::PlanetHistory>(cfg.planet_history, cfg.t0, cfg.dt, cfg.history_table_size);
        const host::PlanetHistory& h = *s.history;
        const uint32_t nb = h.n_bodies();
        std::vector<double> rh_all(h.table_size());
        host::encounter_radii(h.table_size(), h.m().data(), h.r().data(), h.v().data(), cfg.cull_radius, cfg.rh_factor,
                              rh_all.data());
        m0.assign(h.m().begin(), h.m().begin() + nb);
        id0.assign(h.id().begin(), h.id().begin() + nb);
        rh0.assign(rh_all.begin(), rh_all.begin() + nb);
        row_r.resize((size_t)(nb - 1) * 3);
        for (uint32_t i = 1; i < nb; i++) h.r()[i].store(row_r.data(), i - 1);
        if (nb > 1) host::h0_from_log(nb, m0.data(), row_r.data(), 1, h0.data());
        s.nb_max = cfg.history_table_size;
    } else {
        std::vector<double> mm(m, m + n_bodies);
        std::vector<host::Vec3> rr(n_bodies), vv(n_bodies);
        std::vector<uint32_t> pid(n_bodies);
        for (uint32_t i = 0; i < n_bodies; i++) {
            rr[i] = host::Vec3::load(r, i);
            vv[i] = host::Vec3::load(v, i);
            pid[i] = planet_ids ? planet_ids[i] : i;
        }
        s.planets = std::make_unique<host::PlanetIntegrator>(std::move(mm), std::move(rr), std::move(vv),
                                                             std::move(pid), cfg.cull_radius, cfg.rh_factor,
                                                             cfg.outer_radius);
        // the WHIntegrator constructor (src/wh.cpp:498-543)
        const host::PlanetIntegrator& pl = *s.planets;
        m0 = pl.m();
        id0 = pl.id();
        rh0 = pl.rh();
        row_r.resize((size_t)(n_bodies - 1) * 3);
        for (uint32_t i = 1; i < n_bodies; i++) pl.r()[i].store(row_r.data(), i - 1);
        const host::Vec3 hv = pl.h0_initial();
        h0 = {hv.x, hv.y, hv.z};
        s.nb_max = n_bodies;
    }
    s.row_nb = static_cast<uint32_t>(m0.size());
    s.row_m = m0;
    s.row_id = id0;
    s.row_rh = rh0;
    s.row_r = row_r;
    s.row_h0 = h0;
    s.n_total = n;
    s.t = cfg.t0;
    glx_params gp{};
    gp.n_particles = n;
    gp.n_bodies_max = s.nb_max;
    gp.tbsize_max = cfg.tbsize;
    gp.device = cfg.device;
    gp.keep_rollback = cfg.resolve_encounters ? 1 : 0;
    gp.n_planet_buffers = 2;
    check(glx_create(&gp, &s.ctx), "glx_create");
    uint32_t max_id = 0;
    for (uint64_t k = 0; k < n; k++) max_id = std::max(max_id, ids[k]);
    if (n && (uint64_t)max_id < 4 * n)
        s.deathtime_flat.assign((size_t)max_id + 1, std::numeric_limits<double>::quiet_NaN());
    if (n) {
        std::vector<uint16_t> flags(n, 0);
        std::vector<uint32_t> dti(n, 0);
        std::vector<double> a(3 * n, 0.0);
        check(glx_upload_particles(s.ctx, 0, n, pr, pv, a.data(), flags.data(), ids, dti.data()),
              "glx_upload_particles");
    }
    check(glx_set_counts(s.ctx, n, n), "glx_set_counts");
    if (n) {
        check(glx_init_accel(s.ctx, s.row_nb, s.row_m.data(), s.row_id.data(), s.row_rh.data(), s.row_r.data(),
                             s.row_h0.data(), cfg.outer_radius),
              "glx_init_accel");
        if (cfg.resolve_encounters) check(glx_snapshot_rollback(s.ctx), "glx_snapshot_rollback");
        if (cfg.reference_deathtime && !cfg.resolve_encounters) {
            // as Executor::init (src/executor.cu:180-195): particles flagged
            // for an encounter at t0 stay in the active range until the first
            // block's kill and resync; the others never get a death time
            uint64_t na = 0, ne = 0;
            check(glx_partition(s.ctx, GLX_PART_THREE_WAY, &na, &ne), "glx_partition");
            s.n_split = na;
            s.n_alive = na + ne;
        } else {
            const uint64_t ne = s.compact(n);
            s.n_alive = s.n_split;
            s.stage_flagged(s.n_alive, n, cfg.resolve_encounters ? ne : 0, cfg.t0, cfg.dt, false, cfg.t0);
            s.take_staged();
        }
    }
    s.produce_block(cfg.t0);
    s.upload_block();
    if (s.next.n_steps) s.t = s.next.t0;
}

Executor::~Executor() = default;

void Executor::step() { run(1); }

// Blocks in a row: each block's records are taken while the next block's
// kernel runs; all of them are in the queue when this returns.
void Executor::run(uint32_t n_blocks) {
    Impl& s = *p_;
    try {
        for (uint32_t k = 0; k < n_blocks; k++) s.block();
    } catch (...) {
        s.take_staged();
        throw;
    }
    s.take_staged();
}

double Executor::time() const { return p_->t; }
uint64_t Executor::n_alive() const { return p_->n_alive; }
uint64_t Executor::n_total() const { return p_->n_total; }
uint64_t Executor::n_queued() const { return p_->queued; }

std::vector<EncounterRecord> Executor::take_encounters() {
    Impl& s = *p_;
    std::vector<EncounterRecord> out;
    if (s.queue.size() == 1) {
        out.swap(s.queue[0]);
    } else {
        out.reserve(s.queued);
        for (const auto& q : s.queue) out.insert(out.end(), q.begin(), q.end());
    }
    s.queue.clear();
    s.queued = 0;
    return out;
}

uint64_t Executor::take_encounters(EncounterRecord* out, uint64_t capacity) {
    Impl& s = *p_;
    const uint64_t n = s.queued;
    if (n > capacity || (n && !out))
        throw std::invalid_argument("encounter queue holds " + std::to_string(n) + " records, capacity " +
                                    std::to_string(capacity));
    EncounterRecord* o = out;
    for (const auto& q : s.queue) {
        std::memcpy(o, q.data(), q.size() * sizeof(EncounterRecord));
        o += q.size();
    }
    s.queue.clear();
    s.queued = 0;
    return n;
}

void Executor::return_particles(uint64_t n, const uint32_t* ids, const double* r, const double* v) {
    Impl& s = *p_;
    if (n == 0) return;
    if (!ids || !r || !v) throw std::runtime_error("return_particles: null argument");
    const uint64_t begin = s.n_alive, rest = s.n_total - s.n_alive;
    std::vector<double> rr(rest * 3), vv(rest * 3);
    std::vector<uint16_t> flags(rest);
    std::vector<uint32_t> id(rest);
    check(glx_download_particles(s.ctx, begin, rest, rr.data(), vv.data(), nullptr, flags.data(), id.data(), nullptr),
          "glx_download_particles");
    std::unordered_map<uint32_t, uint64_t> slot;
    for (uint64_t k = 0; k < rest; k++) slot[id[k]] = k;
    for (uint64_t k = 0; k < n; k++) {
        auto it = slot.find(ids[k]);
        if (it == slot.end() || (flags[it->second] & 0xFE) != 0 || flags[it->second] == 0)
            throw std::runtime_error("return_particles: particle " + std::to_string(ids[k]) +
                                     " is not waiting in the encounter class");
        const uint64_t j = it->second;
        std::memcpy(&rr[3 * j], r + 3 * k, 3 * sizeof(double));
        std::memcpy(&vv[3 * j], v + 3 * k, 3 * sizeof(double));
        flags[j] = 0;
        s.clear_deathtime(ids[k]);
    }
    check(glx_upload_particles(s.ctx, begin, rest, rr.data(), vv.data(), nullptr, flags.data(), id.data(), nullptr),
          "glx_upload_particles");
    // the returned particles join the unflagged class right after the alive ones
    check(glx_set_counts(s.ctx, s.n_total, s.n_total), "glx_set_counts");
    uint64_t na = 0, ne = 0;
    check(glx_partition(s.ctx, GLX_PART_THREE_WAY, &na, &ne), "glx_partition");
    if (na != begin + n) throw std::runtime_error("return_particles: compaction count mismatch");
    check(glx_accel_range(s.ctx, begin, n, s.row_nb, s.row_m.data(), s.row_id.data(), s.row_rh.data(),
                          s.row_r.data(), s.row_h0.data(), s.cfg.outer_radius),
          "glx_accel_range");
    // clear a fresh encounter flag: the next step detects a lasting one
    std::vector<uint16_t> f(n);
    check(glx_download_particles(s.ctx, begin, n, nullptr, nullptr, nullptr, f.data(), nullptr, nullptr),
          "glx_download_particles");
    std::vector<double> rn(n * 3), vn(n * 3), an(n * 3);
    std::vector<uint32_t> idn(n);
    check(glx_download_particles(s.ctx, begin, n, rn.data(), vn.data(), an.data(), nullptr, idn.data(), nullptr),
          "glx_download_particles");
    for (auto& x : f)
        if ((x & 0xFF) == 0x01) x = 0;
    check(glx_upload_particles(s.ctx, begin, n, rn.data(), vn.data(), an.data(), f.data(), idn.data(), nullptr),
          "glx_upload_particles");
    check(glx_set_counts(s.ctx, s.n_total, s.n_total), "glx_set_counts");
    check(glx_partition(s.ctx, GLX_PART_THREE_WAY, &na, &ne), "glx_partition");
    s.n_alive = s.n_split = na;
    check(glx_set_counts(s.ctx, s.n_alive, s.n_total), "glx_set_counts");
}

ParticleState Executor::download() const {
    const Impl& s = *p_;
    ParticleState out;
    const uint64_t n = s.n_total;
    out.r.resize(3 * n);
    out.v.resize(3 * n);
    out.flags.resize(n);
    out.id.resize(n);
    out.deathtime.assign(n, std::numeric_limits<double>::quiet_NaN());
    check(glx_download_particles(s.ctx, 0, n, out.r.data(), out.v.data(), nullptr, out.flags.data(), out.id.data(),
                                 nullptr),
          "glx_download_particles");
    for (uint64_t k = 0; k < n; k++) out.deathtime[k] = s.get_deathtime(out.id[k]);
    out.n_alive = s.n_alive;
    return out;
}

uint32_t Executor::n_bodies() const { return p_->row_nb; }
std::vector<double> Executor::planet_masses() const { return p_->row_m; }
std::vector<uint32_t> Executor::planet_ids() const { return p_->row_id; }
std::vector<double> Executor::planet_positions() const { return p_->row_r; }

void Executor::last_times(double* kernel_ms, double* partition_ms, double* planet_ms) const {
    float k = 0.0f, q = 0.0f;
    check(glx_last_times(p_->ctx, &k, &q), "glx_last_times");
    if (kernel_ms) *kernel_ms = k;
    if (partition_ms) *partition_ms = q;
    if (planet_ms) *planet_ms = p_->planet_ms;
}

double Executor::kernel_ms_total(uint64_t* n_blocks) const {
    if (n_blocks) *n_blocks = p_->kernel_blocks;
    return p_->kernel_ms_total;
}

}  // namespace glx

// ---------------------------------------------------------------- C ABI ----

struct glx_exec {
    std::unique_ptr<glx::Executor> ex;
};

namespace {

template <typename F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return glx::set_error(GLX_ENOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        // keep the status text of a failed C call; the message names it
        return glx::set_error(GLX_ESTATE, e.what());
    }
}

}  // namespace

extern "C" {

int glx_exec_create(const glx_exec_params* p, uint32_t n_bodies, const double* m, const double* r, const double* v,
                    const uint32_t* planet_id, uint64_t n, const double* pr, const double* pv, const uint32_t* ids,
                    glx_exec** out) {
    if (!p || !out) return glx::set_error(GLX_EINVAL, "glx_exec_create: null argument");
    *out = nullptr;
    return guarded([&]() -> int {
        glx::ExecutorConfig c;
        c.t0 = p->t0;
        c.dt = p->dt;
        c.tbsize = p->tbsize;
        c.outer_radius = p->outer_radius;
        c.cull_radius = p->cull_radius;
        c.rh_factor = p->rh_factor;
        c.resolve_encounters = p->resolve_encounters != 0;
        c.form = p->form;
        c.device = p->device;
        if (p->planet_history) c.planet_history = p->planet_history;
        c.history_table_size = p->history_table_size ? p->history_table_size : 16;
        c.resync_every = p->resync_every ? p->resync_every : 1;
        c.reference_deathtime = p->reference_deathtime != 0;
        if (p->encounter_log) c.encounter_log = p->encounter_log;
        auto e = std::make_unique<glx_exec>();
        e->ex = std::make_unique<glx::Executor>(c, n_bodies, m, r, v, planet_id, n, pr, pv, ids);
        *out = e.release();
        glx::clear_error();
        return GLX_OK;
    });
}

int glx_exec_destroy(glx_exec* e) {
    delete e;
    return GLX_OK;
}

int glx_exec_step(glx_exec* e, uint32_t n_blocks) {
    if (!e) return glx::set_error(GLX_EINVAL, "null executor");
    return guarded([&]() -> int {
        e->ex->run(n_blocks);
        return GLX_OK;
    });
}

int glx_exec_status(glx_exec* e, double* t, uint64_t* n_alive, uint64_t* n_total, uint64_t* n_queued) {
    if (!e) return glx::set_error(GLX_EINVAL, "null executor");
    if (t) *t = e->ex->time();
    if (n_alive) *n_alive = e->ex->n_alive();
    if (n_total) *n_total = e->ex->n_total();
    if (n_queued) *n_queued = e->ex->n_queued();
    return GLX_OK;
}

int glx_exec_take_encounters(glx_exec* e, uint64_t capacity, glx_encounter* out, uint64_t* n_out) {
    if (!e || !n_out) return glx::set_error(GLX_EINVAL, "null argument");
    return guarded([&]() -> int {
        const uint64_t n = e->ex->n_queued();
        if (n > capacity || (n && !out)) {
            *n_out = n;
            return glx::set_error(GLX_EINVAL, "encounter queue holds " + std::to_string(n) + " records, capacity " +
                                                  std::to_string(capacity));
        }
        *n_out = e->ex->take_encounters(out, capacity);
        return GLX_OK;
    });
}

int glx_exec_return(glx_exec* e, uint64_t n, const uint32_t* ids, const double* r, const double* v) {
    if (!e) return glx::set_error(GLX_EINVAL, "null executor");
    return guarded([&]() -> int {
        e->ex->return_particles(n, ids, r, v);
        return GLX_OK;
    });
}

int glx_exec_download(glx_exec* e, double* r, double* v, uint16_t* flags, uint32_t* id, double* deathtime) {
    if (!e) return glx::set_error(GLX_EINVAL, "null executor");
    return guarded([&]() -> int {
        const glx::ParticleState s = e->ex->download();
        const size_t n = s.id.size();
        if (r) std::memcpy(r, s.r.data(), 3 * n * sizeof(double));
        if (v) std::memcpy(v, s.v.data(), 3 * n * sizeof(double));
        if (flags) std::memcpy(flags, s.flags.data(), n * sizeof(uint16_t));
        if (id) std::memcpy(id, s.id.data(), n * sizeof(uint32_t));
        if (deathtime) std::memcpy(deathtime, s.deathtime.data(), n * sizeof(double));
        return GLX_OK;
    });
}

int glx_exec_planets(glx_exec* e, double* r) {
    if (!e || !r) return glx::set_error(GLX_EINVAL, "null argument");
    const auto p = e->ex->planet_positions();
    std::memcpy(r, p.data(), p.size() * sizeof(double));
    return GLX_OK;
}

int glx_exec_bodies(glx_exec* e, uint32_t* n_bodies, double* m, uint32_t* id) {
    if (!e) return glx::set_error(GLX_EINVAL, "null executor");
    const uint32_t nb = e->ex->n_bodies();
    if (n_bodies) *n_bodies = nb;
    if (m) {
        const auto v = e->ex->planet_masses();
        std::memcpy(m, v.data(), v.size() * sizeof(double));
    }
    if (id) {
        const auto v = e->ex->planet_ids();
        std::memcpy(id, v.data(), v.size() * sizeof(uint32_t));
    }
    return GLX_OK;
}

int glx_exec_kernel_total(glx_exec* e, double* kernel_ms, uint64_t* n_blocks) {
    if (!e || !kernel_ms) return glx::set_error(GLX_EINVAL, "null argument");
    *kernel_ms = e->ex->kernel_ms_total(n_blocks);
    return GLX_OK;
}

int glx_exec_times(glx_exec* e, double* kernel_ms, double* partition_ms, double* planet_ms) {
    if (!e) return glx::set_error(GLX_EINVAL, "null executor");
    return guarded([&]() -> int {
        e->ex->last_times(kernel_ms, partition_ms, planet_ms);
        return GLX_OK;
    });
}

}  // extern "C"
