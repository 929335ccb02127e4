// C entry points of the host-side producer (planet Jacobi WH integrator,
// indirect-term rows) and the orbital-element conversions.  Exceptions from
// the C++ layer become status codes here.
#include <cmath>
#include <exception>
#include <new>

#include "../errors.h"
#include "convert.h"
#include "glisser_amd.h"
#include "planets.h"

using glx::host::PlanetIntegrator;
using glx::host::Vec3;

struct glx_planets {
    PlanetIntegrator integ;
};

namespace {

template <typename F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const glx::host::PlanetError& e) {
        return glx::set_error(GLX_ESTATE, e.what());
    } catch (const std::bad_alloc&) {
        return glx::set_error(GLX_ENOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return glx::set_error(GLX_EINVAL, e.what());
    }
}

}  // namespace

extern "C" {

int glx_planets_create(uint32_t n, const double* m, const double* r, const double* v, const uint32_t* id,
                       double cull_radius, double rh_factor, double outer_radius, glx_planets** out) {
    if (!out || !m || !r || !v) return glx::set_error(GLX_EINVAL, "glx_planets_create: null argument");
    *out = nullptr;
    if (n < 1 || n > GLX_MAX_BODIES) return glx::set_error(GLX_EINVAL, "glx_planets_create: n outside [1, 16]");
    if (!(m[0] > 0.0)) return glx::set_error(GLX_EINVAL, "glx_planets_create: m[0] must be positive");
    return guarded([&] {
        std::vector<double> mm(m, m + n);
        std::vector<Vec3> rr(n), vv(n);
        std::vector<uint32_t> ids(n);
        for (uint32_t i = 0; i < n; i++) {
            rr[i] = Vec3::load(r, i);
            vv[i] = Vec3::load(v, i);
            ids[i] = id ? id[i] : i;
        }
        *out = new glx_planets{PlanetIntegrator(std::move(mm), std::move(rr), std::move(vv), std::move(ids),
                                                cull_radius, rh_factor, outer_radius)};
        glx::clear_error();
        return GLX_OK;
    });
}

int glx_planets_destroy(glx_planets* p) {
    delete p;
    return GLX_OK;
}

int glx_planets_h0_initial(glx_planets* p, double* h0) {
    if (!p || !h0) return glx::set_error(GLX_EINVAL, "glx_planets_h0_initial: null argument");
    p->integ.h0_initial().store(h0, 0);
    return GLX_OK;
}

int glx_planets_block(glx_planets* p, uint32_t n_steps, double dt, double* r_log, double* v_log, double* h0_log) {
    if (!p) return glx::set_error(GLX_EINVAL, "glx_planets_block: null handle");
    if (!(dt > 0.0) || !std::isfinite(dt)) return glx::set_error(GLX_EINVAL, "glx_planets_block: bad dt");
    return guarded([&] {
        p->integ.integrate_block(n_steps, dt, r_log, v_log, h0_log);
        return GLX_OK;
    });
}

int glx_planets_state(glx_planets* p, double* r, double* v, double* rh) {
    if (!p) return glx::set_error(GLX_EINVAL, "glx_planets_state: null handle");
    for (size_t i = 0; i < p->integ.n(); i++) {
        if (r) p->integ.r()[i].store(r, i);
        if (v) p->integ.v()[i].store(v, i);
        if (rh) rh[i] = p->integ.rh()[i];
    }
    return GLX_OK;
}

int glx_planets_recalculate_rh(glx_planets* p) {
    if (!p) return glx::set_error(GLX_EINVAL, "glx_planets_recalculate_rh: null handle");
    return guarded([&] {
        p->integ.recalculate_rh();
        return GLX_OK;
    });
}

int glx_planets_metrics(glx_planets* p, double* energy, double* l) {
    if (!p) return glx::set_error(GLX_EINVAL, "glx_planets_metrics: null handle");
    Vec3 L;
    p->integ.metrics(energy, l ? &L : nullptr);
    if (l) L.store(l, 0);
    return GLX_OK;
}

int glx_load_h0(uint32_t n_bodies, const double* m, const double* r_log, uint32_t n_steps, double* h0_log) {
    if (!m || !h0_log || (n_bodies > 1 && !r_log)) return glx::set_error(GLX_EINVAL, "glx_load_h0: null argument");
    if (n_bodies < 1 || n_bodies > GLX_MAX_BODIES) return glx::set_error(GLX_EINVAL, "glx_load_h0: bad n_bodies");
    glx::host::h0_from_log(n_bodies, m, r_log, n_steps, h0_log);
    return GLX_OK;
}

int glx_to_elements(uint64_t n, const double* mu, const double* r, const double* v, double* el) {
    if (!mu || !r || !v || !el) return glx::set_error(GLX_EINVAL, "glx_to_elements: null argument");
    for (uint64_t k = 0; k < n; k++) {
        const glx::host::Elements e = glx::host::to_elements(mu[k], Vec3::load(r, k), Vec3::load(v, k));
        double* o = el + 6 * k;
        o[0] = e.a;
        o[1] = e.e;
        o[2] = e.i;
        o[3] = e.capom;
        o[4] = e.om;
        o[5] = e.f;
    }
    return GLX_OK;
}

int glx_from_elements(uint64_t n, const double* mu, const double* el, int anomaly_is_mean, double* r, double* v) {
    if (!mu || !r || !v || !el) return glx::set_error(GLX_EINVAL, "glx_from_elements: null argument");
    for (uint64_t k = 0; k < n; k++) {
        const double* e = el + 6 * k;
        Vec3 R, V;
        if (anomaly_is_mean)
            glx::host::from_elements_M(mu[k], e[0], e[1], e[2], e[3], e[4], e[5], &R, &V);
        else
            glx::host::from_elements(mu[k], e[0], e[1], e[2], e[3], e[4], e[5], &R, &V);
        R.store(r, k);
        V.store(v, k);
    }
    return GLX_OK;
}

}  // extern "C"
