// C entry points of the planet history files (glx_plhist_*).
#include <cstring>
#include <exception>
#include <fstream>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../errors.h"
#include "glisser_amd.h"
#include "history.h"

using glx::host::HistoryBlock;
using glx::host::PlanetHistory;
using glx::host::Vec3;

struct glx_plhist {
    std::unique_ptr<PlanetHistory> h;
    HistoryBlock block;
};

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

}  // namespace

extern "C" {

int glx_plhist_write(const char* path, int append, double time, uint32_t n_bodies, const double* m, const double* r,
                     const double* v, const uint32_t* id) {
    if (!path || !m || !r || !v) return glx::set_error(GLX_EINVAL, "glx_plhist_write: null argument");
    if (n_bodies < 1) return glx::set_error(GLX_EINVAL, "glx_plhist_write: n_bodies must include the Sun");
    return guarded([&]() -> int {
        std::vector<Vec3> rr(n_bodies), vv(n_bodies);
        std::vector<uint32_t> ids(n_bodies);
        for (uint32_t i = 0; i < n_bodies; i++) {
            rr[i] = Vec3::load(r, i);
            vv[i] = Vec3::load(v, i);
            ids[i] = id ? id[i] : i;
        }
        std::ofstream out(path, append ? (std::ios_base::binary | std::ios_base::app)
                                       : (std::ios_base::binary | std::ios_base::trunc));
        if (!out) return glx::set_error(GLX_EINVAL, std::string("cannot open ") + path);
        glx::host::write_history_record(out, !append, time, n_bodies, m, rr.data(), vv.data(), ids.data());
        out.flush();
        if (!out) return glx::set_error(GLX_EINVAL, std::string("write failed: ") + path);
        return GLX_OK;
    });
}

int glx_plhist_open(const char* path, double t0, double dt, uint32_t table_size, glx_plhist** out) {
    if (!path || !out) return glx::set_error(GLX_EINVAL, "glx_plhist_open: null argument");
    *out = nullptr;
    return guarded([&]() -> int {
        auto h = std::make_unique<glx_plhist>();
        h->h = std::make_unique<PlanetHistory>(path, t0, dt, table_size);
        *out = h.release();
        glx::clear_error();
        return GLX_OK;
    });
}

int glx_plhist_close(glx_plhist* h) {
    delete h;
    return GLX_OK;
}

int glx_plhist_state(glx_plhist* h, uint32_t* n_bodies, double* m, uint32_t* id, double* r, double* v) {
    if (!h) return glx::set_error(GLX_EINVAL, "glx_plhist_state: null handle");
    const PlanetHistory& p = *h->h;
    if (n_bodies) *n_bodies = p.n_bodies();
    for (size_t k = 0; k < p.table_size(); k++) {
        if (m) m[k] = p.m()[k];
        if (id) id[k] = p.id()[k];
        if (r) p.r()[k].store(r, k);
        if (v) p.v()[k].store(v, k);
    }
    return GLX_OK;
}

int glx_plhist_block(glx_plhist* h, uint32_t tbsize, double* t_start, double* dt, uint32_t* n_steps,
                     uint32_t* n_bodies, double* m, uint32_t* id, double* r_log, double* h0_log) {
    if (!h || !t_start || !dt || !n_steps || !n_bodies || !m || !id || !r_log)
        return glx::set_error(GLX_EINVAL, "glx_plhist_block: null argument");
    return guarded([&]() -> int {
        HistoryBlock& b = h->block;
        h->h->next_block(tbsize, &b);
        *t_start = b.t_start;
        *dt = b.dt;
        *n_steps = b.n_steps;
        *n_bodies = b.n_bodies;
        std::memcpy(m, b.m.data(), b.m.size() * sizeof(double));
        std::memcpy(id, b.id.data(), b.id.size() * sizeof(uint32_t));
        std::memcpy(r_log, b.r_log.data(), b.r_log.size() * sizeof(double));
        if (h0_log) std::memcpy(h0_log, b.h0_log.data(), b.h0_log.size() * sizeof(double));
        return GLX_OK;
    });
}

}  // extern "C"
