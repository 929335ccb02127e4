// Driver around the reference's own CPU sources (src/wh.cpp, src/data.cpp,
// src/convert.cpp, src/util.cpp of GLISSER), built by oracle/Makefile into
// oracle/_ref/libglisser_ref.so in the container that holds the reference.
//
// TEST INFRASTRUCTURE ONLY.  It exposes the same C entry points as the
// oracle restatement (prefix ref_ instead of orc_) so the tests can check the
// restatement bit-for-bit against the reference's compiled code.  It follows
// the reference's call stack for a CPU time-block: WHIntegrator ctor
// (src/wh.cpp:498-577), integrate_planets_timeblock (src/wh.cpp:584-595),
// integrate_particles_timeblock (src/wh.cpp:597-609), load_h0
// (src/wh.cpp:993-1025) and HostParticlePhaseSpace::stable_partition_*
// (src/data.cpp:105-123); for planet history files the Interpolator
// (src/interp.cpp) driven the way Executor::init / update_planets drive it
// (src/executor.cu:93-130,235-292,700; that file is CUDA, so its few lines
// of glue are repeated here) and the writer save_swift_plhist.

#include "wh.h"
#include "convert.h"
#include "data.h"
#include "interp.h"

#include <cstdint>
#include <cstring>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

// Defined in the reference's src/data.cpp (external linkage) without a
// declaration in its headers.
namespace sr {
namespace data {
bool load_data_nohybrid(HostPlanetPhaseSpace& pl, HostParticlePhaseSpace& pa, const Configuration& config,
                        std::istream& plin, std::istream& icsin);
bool load_data_hybrid(HostPlanetPhaseSpace& pl, HostParticlePhaseSpace& pa, const Configuration& config,
                      std::istream& in);
void save_data_hybrid_binary(const HostPlanetSnapshot& pl, const HostParticlePhaseSpace& pa,
                             const Configuration& config, std::ostream& out);
void save_data_hybrid(const HostPlanetSnapshot& pl, const HostParticlePhaseSpace& pa, const Configuration& config,
                      std::ostream& out);
void save_data_nohybrid(const HostPlanetSnapshot& pl, const HostParticlePhaseSpace& pa, const Configuration& config,
                        std::ostream& plout, std::ostream& icsout);
}  // namespace data
}  // namespace sr

using namespace sr::data;
using namespace sr::wh;

namespace {

thread_local char g_err[256];

f64_3 ld(const double* p, size_t i) { return f64_3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
void st(double* p, size_t i, const f64_3& v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

int classify(const std::runtime_error& e) {
    std::snprintf(g_err, sizeof g_err, "%s", e.what());
    std::string s = e.what();
    if (s.find("unbound") != std::string::npos) return 1;
    if (s.find("kepler") != std::string::npos) return 2;
    return 3;
}

struct RefPlanets {
    HostPlanetPhaseSpace pl;
    HostParticlePhaseSpace pa;
    Configuration cfg;
    WHIntegrator integ;
};

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err; }

void ref_particle_accels(uint32_t n_bodies, const double* m, const double* rh, double outer, const double* planet_r,
                         const double* h0, double time, uint32_t n, const double* r, double* a, uint16_t* flags,
                         const uint32_t* id, double* deathtime) {
    HostPlanetPhaseSpace pl(n_bodies, 1);
    for (uint32_t i = 0; i < n_bodies; i++) pl.m()[i] = m[i];
    for (uint32_t j = 0; j + 1 < n_bodies; j++) pl.r_log().old[j] = ld(planet_r, j);
    HostParticlePhaseSpace pa(n);
    for (uint32_t i = 0; i < n; i++) {
        pa.r()[i] = ld(r, i);
        pa.deathflags()[i] = flags[i];
        pa.id()[i] = id[i];
    }
    WHIntegrator integ;
    integ.planet_rh.assign(rh, rh + n_bodies);
    integ.outer_radius = outer;
    integ.particle_a = Vf64_3(n);
    integ.planet_h0_log = sr::util::History<Vf64_3>(1);
    integ.planet_h0_log.old[0] = ld(h0, 0);
    integ.helio_acc_particles(pl, pa, 0, n, time, 0, true);
    for (uint32_t i = 0; i < n; i++) {
        st(a, i, integ.particle_a[i]);
        flags[i] = pa.deathflags()[i];
    }
    if (deathtime)
        for (auto& kv : pa.deathtime_map()) deathtime[kv.first] = kv.second;
}

int ref_particle_block(uint32_t n_bodies, const double* m, const double* rh, double outer, const double* r_log,
                       const double* h0_log, uint32_t n_steps, double t, double dt, uint32_t n, double* r, double* v,
                       double* a, uint16_t* flags, const uint32_t* id, uint32_t* deathtime_index, double* deathtime,
                       uint64_t* bad) {
    (void)bad;
    HostPlanetPhaseSpace pl(n_bodies, n_steps);
    for (uint32_t i = 0; i < n_bodies; i++) {
        pl.m()[i] = m[i];
        pl.id()[i] = i;
    }
    const size_t rows = static_cast<size_t>(n_steps) * (n_bodies - 1);
    for (size_t k = 0; k < rows; k++) pl.r_log().log[k] = ld(r_log, k);
    pl.r_log().len = n_steps;
    uint32_t max_id = 0;
    for (uint32_t i = 0; i < n; i++) max_id = std::max(max_id, id[i]);
    HostParticlePhaseSpace pa(n);
    pa.deathtime_index() = Vu32(std::max<size_t>(n, max_id + 1));
    for (uint32_t i = 0; i < n; i++) {
        pa.r()[i] = ld(r, i);
        pa.v()[i] = ld(v, i);
        pa.deathflags()[i] = flags[i];
        pa.id()[i] = id[i];
    }
    for (size_t k = 0; k < pa.deathtime_index().size(); k++) pa.deathtime_index()[k] = deathtime_index[k];
    WHIntegrator integ;
    integ.planet_rh.assign(rh, rh + n_bodies);
    integ.outer_radius = outer;
    integ.particle_a = Vf64_3(n);
    for (uint32_t i = 0; i < n; i++) integ.particle_a[i] = ld(a, i);
    integ.particle_mu = integ.particle_dist = integ.particle_energy = integ.particle_vdotr = Vf64(n);
    integ.particle_mask = Vu8(n);
    integ.planet_h0_log = sr::util::History<Vf64_3>(n_steps);
    for (uint32_t s = 0; s < n_steps; s++) integ.planet_h0_log.log[s] = ld(h0_log, s);
    int rc = 0;
    try {
        integ.integrate_particles_timeblock(pl, pa, 0, n, t, dt);
    } catch (const std::runtime_error& e) {
        rc = classify(e);
    }
    for (uint32_t i = 0; i < n; i++) {
        st(r, i, pa.r()[i]);
        st(v, i, pa.v()[i]);
        st(a, i, integ.particle_a[i]);
        flags[i] = pa.deathflags()[i];
    }
    for (size_t k = 0; k < pa.deathtime_index().size(); k++) deathtime_index[k] = pa.deathtime_index()[k];
    if (deathtime)
        for (auto& kv : pa.deathtime_map()) deathtime[kv.first] = kv.second;
    return rc;
}

void* ref_planets_new(uint32_t n, const double* m, const double* r, const double* v, const uint32_t* id,
                      double cull_radius, double rh_factor, double outer, double* h0_0) {
    RefPlanets* p = new RefPlanets();
    p->cfg.tbsize = 1;
    p->cfg.cull_radius = cull_radius;
    p->cfg.encounter_sphere_factor = rh_factor;
    p->cfg.outer_radius = outer;
    p->pl = HostPlanetPhaseSpace(n, 1);
    for (uint32_t i = 0; i < n; i++) {
        p->pl.m()[i] = m[i];
        p->pl.r()[i] = ld(r, i);
        p->pl.v()[i] = ld(v, i);
        p->pl.id()[i] = id[i];
    }
    p->pa = HostParticlePhaseSpace(0);
    p->integ = WHIntegrator(p->pl, p->pa, p->cfg);
    if (h0_0) st(h0_0, 0, p->integ.planet_h0_log.old[0]);
    return p;
}

void ref_planets_free(void* h) { delete static_cast<RefPlanets*>(h); }

int ref_planets_block(void* h, uint32_t n_steps, double dt, double* r_log, double* v_log, double* h0_log) {
    RefPlanets& p = *static_cast<RefPlanets*>(h);
    const size_t n = p.pl.n();
    if (p.pl.r_log().old.size() < (n - 1) * n_steps) {
        p.pl.r_log() = sr::util::History<Vf64_3>((n - 1) * n_steps);
        p.pl.v_log() = sr::util::History<Vf64_3>((n - 1) * n_steps);
    }
    if (p.integ.planet_h0_log.old.size() < n_steps) {
        sr::util::History<Vf64_3> h0(n_steps);
        p.integ.planet_h0_log = h0;
    }
    try {
        p.integ.integrate_planets_timeblock(p.pl, n_steps, 0.0, dt);
    } catch (const std::runtime_error& e) {
        return classify(e);
    }
    for (size_t k = 0; k < (n - 1) * n_steps; k++) {
        st(r_log, k, p.pl.r_log().old[k]);
        st(v_log, k, p.pl.v_log().old[k]);
    }
    for (uint32_t s = 0; s < n_steps; s++) st(h0_log, s, p.integ.planet_h0_log.old[s]);
    return 0;
}

void ref_planets_get(void* h, double* r, double* v, double* rh) {
    RefPlanets& p = *static_cast<RefPlanets*>(h);
    for (size_t i = 0; i < p.pl.n(); i++) {
        if (r) st(r, i, p.pl.r()[i]);
        if (v) st(v, i, p.pl.v()[i]);
        if (rh) rh[i] = p.integ.planet_rh[i];
    }
}

void ref_planets_recalculate_rh(void* h) {
    RefPlanets& p = *static_cast<RefPlanets*>(h);
    p.integ.recalculate_rh(p.pl);
}

void ref_load_h0(uint32_t n_bodies, const double* m, const double* r_log, uint32_t n_steps, double* h0_log) {
    HostPlanetPhaseSpace pl(n_bodies, n_steps);
    for (uint32_t i = 0; i < n_bodies; i++) pl.m()[i] = m[i];
    for (size_t k = 0; k < static_cast<size_t>(n_steps) * (n_bodies - 1); k++) pl.r_log().old[k] = ld(r_log, k);
    pl.r_log().len_old = n_steps;
    WHIntegrator integ;
    integ.planet_h0_log = sr::util::History<Vf64_3>(n_steps);
    integ.load_h0(pl);
    for (uint32_t s = 0; s < n_steps; s++) st(h0_log, s, integ.planet_h0_log.old[s]);
}

void ref_planet_metrics(uint32_t n, const double* m, const double* r, const double* v, double* energy, double* l) {
    HostPlanetPhaseSpace pl(n);
    for (uint32_t i = 0; i < n; i++) {
        pl.m()[i] = m[i];
        pl.r()[i] = ld(r, i);
        pl.v()[i] = ld(v, i);
    }
    f64_3 L;
    calculate_planet_metrics(pl, energy, l ? &L : nullptr);
    if (l) st(l, 0, L);
}

uint64_t ref_stable_partition(const uint16_t* flags, uint64_t begin, uint64_t len, int mode, uint64_t* perm) {
    HostParticlePhaseSpace pa(begin + len);
    for (uint64_t i = 0; i < begin + len; i++) {
        pa.deathflags()[i] = flags[i];
        pa.id()[i] = static_cast<uint32_t>(i);
    }
    auto idx = mode == 0 ? pa.stable_partition_unflagged(begin, len) : pa.stable_partition_alive(begin, len);
    for (uint64_t k = 0; k < len; k++) perm[k] = (*idx)[k];
    return pa.n_alive();
}

void ref_to_elements(double mu, const double* r, const double* v, double* out6) {
    sr::convert::to_elements(mu, ld(r, 0), ld(v, 0), nullptr, &out6[0], &out6[1], &out6[2], &out6[3], &out6[4],
                             &out6[5]);
}

void ref_from_elements(double mu, const double* el6, double* r, double* v) {
    f64_3 R, V;
    sr::convert::from_elements(mu, el6[0], el6[1], el6[2], el6[3], el6[4], el6[5], &R, &V);
    st(r, 0, R);
    st(v, 0, V);
}

// ---- file formats (src/data.cpp:146-815) ----------------------------------

// read_configuration + write_configuration of a configuration text; returns
// 0, or 1 with the exception text in ref_last_error.
int ref_config_normalize(const char* text, char* out, uint64_t cap, uint64_t* len) {
    try {
        Configuration c;
        std::istringstream in(text);
        read_configuration(in, &c);
        std::ostringstream o;
        write_configuration(o, c);
        const std::string s = o.str();
        *len = s.size();
        if (cap < s.size() + 1) return 2;
        std::memcpy(out, s.c_str(), s.size() + 1);
        return 0;
    } catch (const std::exception& e) {
        std::snprintf(g_err, sizeof g_err, "%s", e.what());
        return 1;
    }
}

namespace {

void fill_planets(HostPlanetSnapshot& pl, uint32_t np, const double* pm, const double* pr, const double* pv,
                  const uint32_t* pid) {
    pl = HostPlanetSnapshot(np);
    for (uint32_t i = 0; i < np; i++) {
        pl.m[i] = pm[i];
        pl.r[i] = ld(pr, i);
        pl.v[i] = ld(pv, i);
        pl.id[i] = pid[i];
    }
}

}  // namespace

// save_data_hybrid / save_data_hybrid_binary / save_data_nohybrid
// (format 0 / 1 / 2; format 2 writes path and ppath).
int ref_state_save(const char* path, const char* ppath, int format, int write_momenta, uint32_t np,
                   const double* pm, const double* pr, const double* pv, const uint32_t* pid, uint64_t n,
                   const double* r, const double* v, const uint32_t* id, const uint16_t* flags,
                   const double* deathtime) {
    HostPlanetSnapshot pl;
    fill_planets(pl, np, pm, pr, pv, pid);
    HostParticlePhaseSpace pa(n);
    for (uint64_t i = 0; i < n; i++) {
        pa.r()[i] = ld(r, i);
        pa.v()[i] = ld(v, i);
        pa.id()[i] = id[i];
        pa.deathflags()[i] = flags[i];
        pa.deathtime_map()[id[i]] = deathtime[i];
    }
    Configuration cfg;
    cfg.writemomenta = write_momenta != 0;
    if (format == 2) {
        std::ofstream plo(path), pao(ppath);
        save_data_nohybrid(pl, pa, cfg, plo, pao);
    } else if (format == 1) {
        std::ofstream o(path, std::ios_base::binary);
        save_data_hybrid_binary(pl, pa, cfg, o);
    } else {
        std::ofstream o(path);
        save_data_hybrid(pl, pa, cfg, o);
    }
    return 0;
}

// load_data_hybrid / load_data_nohybrid (format 0 / 2) into caller arrays
// sized for at most cap_pl planets and cap particles; *np, *n receive the
// counts (deathtime by particle order).
int ref_state_load(const char* path, const char* ppath, int format, uint32_t cap_pl, uint64_t cap, uint32_t* np,
                   double* pm, double* pr, double* pv, uint32_t* pid, uint64_t* n, double* r, double* v,
                   uint32_t* id, uint16_t* flags, double* deathtime) {
    HostPlanetPhaseSpace pl;
    HostParticlePhaseSpace pa;
    Configuration cfg = Configuration::create_dummy();
    if (format == 2) {
        std::ifstream pli(path), pai(ppath);
        load_data_nohybrid(pl, pa, cfg, pli, pai);
    } else {
        std::ifstream in(path);
        load_data_hybrid(pl, pa, cfg, in);
    }
    if (pl.n() > cap_pl || pa.n() > cap) return 2;
    *np = (uint32_t)pl.n();
    for (size_t i = 0; i < pl.n(); i++) {
        pm[i] = pl.m()[i];
        st(pr, i, pl.r()[i]);
        st(pv, i, pl.v()[i]);
        pid[i] = pl.id()[i];
    }
    *n = pa.n();
    for (size_t i = 0; i < pa.n(); i++) {
        st(r, i, pa.r()[i]);
        st(v, i, pa.v()[i]);
        id[i] = pa.id()[i];
        flags[i] = pa.deathflags()[i];
        deathtime[i] = format == 2 ? pa.deathtime_map()[(uint32_t)i] : pa.deathtime_map()[pa.id()[i]];
    }
    return 0;
}

// save_binary_track of one snapshot (particles in the given order), appended.
int ref_track_write(const char* path, int append, double time, uint32_t np, const double* pm, const double* pr,
                    const double* pv, const uint32_t* pid, uint64_t n, const double* r, const double* v,
                    const uint32_t* id, int to_elements, int barycentric) {
    HostPlanetSnapshot pl;
    fill_planets(pl, np, pm, pr, pv, pid);
    HostParticleSnapshot pa(n);
    for (uint64_t i = 0; i < n; i++) {
        pa.r[i] = ld(r, i);
        pa.v[i] = ld(v, i);
        pa.id[i] = id[i];
    }
    std::ofstream o(path, std::ios_base::binary | (append ? std::ios_base::app : std::ios_base::trunc));
    save_binary_track(o, pl, pa, time, to_elements != 0, barycentric != 0);
    return 0;
}

// ---- planet history (src/interp.cpp, src/data.cpp:817-873) ------------------

int ref_plhist_write(const char* path, int append, double time, uint32_t n, const double* m, const double* r,
                     const double* v, const uint32_t* id) {
    HostPlanetSnapshot pl;
    fill_planets(pl, n, m, r, v, id);
    std::ofstream o(path, std::ios_base::binary | (append ? std::ios_base::app : std::ios_base::trunc));
    if (!append) begin_swift_plhist(o, pl);
    save_swift_plhist(o, pl, time);
    return o ? 0 : 3;
}

namespace {
struct RefHistory {
    Configuration cfg;
    HostPlanetPhaseSpace pl;
    sr::interp::Interpolator* interp = nullptr;
    WHIntegrator integ;
    double t = 0;
    ~RefHistory() { delete interp; }
};
}  // namespace

void* ref_plhist_open(const char* path, double t0, double dt, uint32_t maxpl) {
    RefHistory* h = new RefHistory();
    try {
        h->cfg.t_0 = t0;
        h->cfg.dt = dt;
        h->cfg.interp_maxpl = maxpl;
        h->cfg.tbsize = 1;
        h->interp = new sr::interp::Interpolator(h->cfg, h->pl, path);
        sr::interp::Interpolator& in = *h->interp;
        h->t = t0;
        while (in.t1 <= h->t) in.next(h->pl);
        if (std::isinf(in.t0)) throw std::runtime_error("lookup file starts past the given Initial-Time");
        in.cur_ts = 0;
        in.n_ts = std::max<size_t>(1, static_cast<size_t>(std::round((in.t1 - h->t) / dt)));
        in.eff_dt = (in.t1 - h->t) / static_cast<double>(in.n_ts);
        in.rel_t = h->t - in.t0;
        in.fill_one(h->pl, in.rel_t);
    } catch (const std::exception& e) {
        std::snprintf(g_err, sizeof g_err, "%s", e.what());
        delete h;
        return nullptr;
    }
    return h;
}

void ref_plhist_free(void* h) { delete static_cast<RefHistory*>(h); }

void ref_plhist_state(void* hp, uint32_t* n_bodies, double* m, uint32_t* id, double* r, double* v) {
    RefHistory& h = *static_cast<RefHistory*>(hp);
    if (n_bodies) *n_bodies = static_cast<uint32_t>(h.pl.n_alive());
    for (size_t k = 0; k < h.pl.n(); k++) {
        if (m) m[k] = h.pl.m()[k];
        if (id) id[k] = h.pl.id()[k];
        if (r) st(r, k, h.pl.r()[k]);
        if (v) st(v, k, h.pl.v()[k]);
    }
}

int ref_plhist_block(void* hp, uint32_t tbsize, double* t_start, double* dt, uint32_t* n_steps, uint32_t* n_bodies,
                     double* m, uint32_t* id, double* r_log, double* h0_log) {
    RefHistory& h = *static_cast<RefHistory*>(hp);
    sr::interp::Interpolator& in = *h.interp;
    try {
        if (in.cur_ts == in.n_ts) {
            double diff = in.rel_t - (in.t1 - in.t0);
            if (!(std::abs(diff) < 1e-8)) throw std::runtime_error("sanity check failed - adjusting t by too much");
            in.next(h.pl);
            in.cur_ts = 0;
            h.t = in.t0;
        }
        const double cur_dt = in.eff_dt;
        const uint32_t n = std::min(tbsize, static_cast<uint32_t>(in.n_ts - in.cur_ts));
        if (h.pl.r_log().old.size() < (h.pl.n() - 1) * n) {
            h.pl.r_log() = sr::util::History<Vf64_3>((h.pl.n() - 1) * n);
            h.pl.v_log() = sr::util::History<Vf64_3>((h.pl.n() - 1) * n);
        }
        in.fill(h.pl, n, in.rel_t, cur_dt);
        in.rel_t += static_cast<double>(n) * cur_dt;
        in.cur_ts += n;
        if (h.integ.planet_h0_log.old.size() < n) h.integ.planet_h0_log = sr::util::History<Vf64_3>(n);
        h.integ.load_h0(h.pl);
        *t_start = h.t;
        *dt = cur_dt;
        *n_steps = n;
        *n_bodies = static_cast<uint32_t>(h.pl.n_alive());
        for (size_t k = 0; k < h.pl.n_alive(); k++) {
            m[k] = h.pl.m()[k];
            id[k] = h.pl.id()[k];
        }
        for (size_t k = 0; k < static_cast<size_t>(n) * (h.pl.n_alive() - 1); k++) st(r_log, k, h.pl.r_log().old[k]);
        if (h0_log)
            for (uint32_t k = 0; k < n; k++) st(h0_log, k, h.integ.planet_h0_log.old[k]);
        h.t += cur_dt * static_cast<double>(n);
    } catch (const std::exception& e) {
        std::snprintf(g_err, sizeof g_err, "%s", e.what());
        return 3;
    }
    return 0;
}

void ref_from_elements_mean(double mu, const double* el6, double* r, double* v) {
    f64_3 R, V;
    sr::convert::from_elements_M(mu, el6[0], el6[1], el6[2], el6[3], el6[4], el6[5], &R, &V);
    st(r, 0, R);
    st(v, 0, V);
}

}  // extern "C"
