// C entry points of the file formats (glx_config_*, glx_state_*, glx_track_*).
#include <cstring>
#include <exception>
#include <fstream>
#include <new>
#include <sstream>
#include <string>

#include "../errors.h"
#include "formats.h"
#include "glisser_amd.h"

using glx::host::State;

namespace {

template <typename F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return glx::set_error(GLX_ENOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return glx::set_error(GLX_EINVAL, e.what());
    }
}

State from_c(const glx_state* c) {
    State s;
    s.pm.assign(c->pm, c->pm + c->n_planets);
    s.pr.assign(c->pr, c->pr + 3 * (size_t)c->n_planets);
    s.pv.assign(c->pv, c->pv + 3 * (size_t)c->n_planets);
    if (c->pid)
        s.pid.assign(c->pid, c->pid + c->n_planets);
    else
        for (uint32_t i = 0; i < c->n_planets; i++) s.pid.push_back(i);
    const size_t n = (size_t)c->n_particles;
    if (n) {
        s.r.assign(c->r, c->r + 3 * n);
        s.v.assign(c->v, c->v + 3 * n);
        s.id.assign(c->id, c->id + n);
        s.flags = c->flags ? std::vector<uint16_t>(c->flags, c->flags + n) : std::vector<uint16_t>(n, 0);
        s.deathtime = c->deathtime ? std::vector<double>(c->deathtime, c->deathtime + n) : std::vector<double>(n, 0.0);
    }
    return s;
}

int check_state(const glx_state* s) {
    if (!s) return glx::set_error(GLX_EINVAL, "null state");
    if (s->n_planets && (!s->pm || !s->pr || !s->pv)) return glx::set_error(GLX_EINVAL, "null planet array");
    if (s->n_particles && (!s->r || !s->v || !s->id)) return glx::set_error(GLX_EINVAL, "null particle array");
    return GLX_OK;
}

State load(const char* path, const char* ppath, int format, uint64_t max_particles) {
    if (!path) throw std::invalid_argument("null path");
    if (format == GLX_STATE_SPLIT) {
        if (!ppath) throw std::invalid_argument("the split format needs a particle file");
        std::ifstream pl(path), pa(ppath);
        if (!pl) throw std::runtime_error(std::string("Planet input file ") + path + " does not exist");
        if (!pa) throw std::runtime_error(std::string("Particle input file ") + ppath + " does not exist");
        return glx::host::load_state_split(pl, pa, max_particles);
    }
    if (format != GLX_STATE_ASCII && format != GLX_STATE_BINARY) throw std::invalid_argument("unknown state format");
    std::ifstream in(path, format == GLX_STATE_BINARY ? std::ios::binary : std::ios::in);
    if (!in) throw std::runtime_error(std::string("Input file ") + path + " does not exist");
    return format == GLX_STATE_BINARY ? glx::host::load_state_binary(in, max_particles)
                                      : glx::host::load_state_ascii(in, max_particles);
}

}  // namespace

extern "C" {

int glx_config_normalize(const char* text, char* out, uint64_t capacity, uint64_t* length) {
    if (!text || !length) return glx::set_error(GLX_EINVAL, "null argument");
    return guarded([&]() -> int {
        glx::host::Configuration c;
        std::istringstream in(text);
        glx::host::read_configuration(in, &c);
        std::ostringstream o;
        glx::host::write_configuration(o, c);
        const std::string s = o.str();
        *length = s.size();
        if (!out || capacity < s.size() + 1) return glx::set_error(GLX_EINVAL, "output buffer too small");
        std::memcpy(out, s.c_str(), s.size() + 1);
        return GLX_OK;
    });
}

int glx_state_save(const char* path, const char* particle_path, int format, int write_momenta, const glx_state* s) {
    if (int rc = check_state(s)) return rc;
    if (!path) return glx::set_error(GLX_EINVAL, "null path");
    return guarded([&]() -> int {
        const State st = from_c(s);
        if (format == GLX_STATE_SPLIT) {
            if (!particle_path) return glx::set_error(GLX_EINVAL, "the split format needs a particle file");
            std::ofstream pl(path), pa(particle_path);
            glx::host::save_state_split(pl, pa, st, write_momenta != 0);
            return pl && pa ? GLX_OK : glx::set_error(GLX_EINVAL, "writing the state files failed");
        }
        if (format != GLX_STATE_ASCII && format != GLX_STATE_BINARY)
            return glx::set_error(GLX_EINVAL, "unknown state format");
        std::ofstream out(path, format == GLX_STATE_BINARY ? std::ios::binary : std::ios::out);
        if (format == GLX_STATE_BINARY)
            glx::host::save_state_binary(out, st, write_momenta != 0);
        else
            glx::host::save_state_ascii(out, st, write_momenta != 0);
        return out ? GLX_OK : glx::set_error(GLX_EINVAL, std::string("writing ") + path + " failed");
    });
}

int glx_state_sizes(const char* path, const char* particle_path, int format, uint64_t max_particles,
                    uint32_t* n_planets, uint64_t* n_particles) {
    if (!n_planets || !n_particles) return glx::set_error(GLX_EINVAL, "null argument");
    return guarded([&]() -> int {
        const State st = load(path, particle_path, format, max_particles);
        *n_planets = (uint32_t)st.n_planets();
        *n_particles = st.n_particles();
        return GLX_OK;
    });
}

int glx_state_load(const char* path, const char* particle_path, int format, uint64_t max_particles, int read_momenta,
                   double big_g, glx_state* s) {
    if (!s) return glx::set_error(GLX_EINVAL, "null state");
    return guarded([&]() -> int {
        State st = load(path, particle_path, format, max_particles);
        glx::host::apply_input_units(st, read_momenta != 0, big_g);
        if (st.n_planets() != s->n_planets || st.n_particles() != s->n_particles)
            return glx::set_error(GLX_EINVAL, "state sizes differ from the arrays given (call glx_state_sizes)");
        if (int rc = check_state(s)) return rc;
        std::memcpy(s->pm, st.pm.data(), st.pm.size() * sizeof(double));
        std::memcpy(s->pr, st.pr.data(), st.pr.size() * sizeof(double));
        std::memcpy(s->pv, st.pv.data(), st.pv.size() * sizeof(double));
        if (s->pid) std::memcpy(s->pid, st.pid.data(), st.pid.size() * sizeof(uint32_t));
        if (st.n_particles()) {
            std::memcpy(s->r, st.r.data(), st.r.size() * sizeof(double));
            std::memcpy(s->v, st.v.data(), st.v.size() * sizeof(double));
            std::memcpy(s->id, st.id.data(), st.id.size() * sizeof(uint32_t));
            if (s->flags) std::memcpy(s->flags, st.flags.data(), st.flags.size() * sizeof(uint16_t));
            if (s->deathtime) std::memcpy(s->deathtime, st.deathtime.data(), st.deathtime.size() * sizeof(double));
        }
        return GLX_OK;
    });
}

int glx_track_write(const char* path, int append, double time, const glx_state* s, int to_elements, int barycentric) {
    if (int rc = check_state(s)) return rc;
    if (!path) return glx::set_error(GLX_EINVAL, "null path");
    return guarded([&]() -> int {
        std::ofstream out(path, std::ios::binary | (append ? std::ios::app : std::ios::trunc));
        glx::host::write_track_snapshot(out, time, from_c(s), to_elements != 0, barycentric != 0);
        return out ? GLX_OK : glx::set_error(GLX_EINVAL, std::string("writing ") + path + " failed");
    });
}

}  // extern "C"

extern "C" {

int glx_track_write_records(const char* path, int append, double time, const glx_state* s, int to_elements,
                            int barycentric, const void* records, uint64_t n) {
    if (!s || !path || (n && !records)) return glx::set_error(GLX_EINVAL, "null argument");
    if (s->n_planets && (!s->pm || !s->pr || !s->pv)) return glx::set_error(GLX_EINVAL, "null planet array");
    return guarded([&]() -> int {
        glx_state planets = *s;
        planets.n_particles = 0;
        std::ofstream out(path, std::ios::binary | (append ? std::ios::app : std::ios::trunc));
        glx::host::write_track_records(out, time, from_c(&planets), to_elements != 0, barycentric != 0,
                                       static_cast<const glx::host::TrackRecord*>(records), n);
        return out ? GLX_OK : glx::set_error(GLX_EINVAL, std::string("writing ") + path + " failed");
    });
}

}  // extern "C"
