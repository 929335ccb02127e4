// C ABI of the batched form: device memory of many systems, the tile list,
// the captured launch (HIP graph) and the autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "batch.h"
#include "errors.h"
#include "glisser_amd.h"
#include "launch.h"

using glx::BatchArgs;
using glx::SystemDev;

namespace {

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return glx::set_error(code, buf);
}

#define GLX_HIP(call)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return fail(GLX_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

constexpr uint32_t kSlotAlign = 256;            // system ranges start on multiples of this
constexpr size_t kLdsPerCu = 160 * 1024;        // gfx950 LDS per compute unit
constexpr size_t kLdsMaxPerGroup = 63 * 1024;  // dynamic LDS cap (the kernel adds a few static bytes)
constexpr size_t kLdsPer64Lanes = 8 * 1024;     // default staging per wave of a workgroup
constexpr double kHistSlack = 1e-8;             // allowed overrun of a history's last sample

struct SystemHost {
    uint32_t nb, n;
    double t_first, t_last;
    // device slot k holds the caller's particle order[k] (set by upload)
    std::vector<uint32_t> order;
};

// Device order of a system's particles: by 1/a = 2/|r| - v^2/mu, largest
// first (shortest periods, hence largest mean-anomaly steps, first), ties
// and orbits that are not bound (1/a <= 0 or NaN) in the caller's order at
// the end.  A wave then holds particles of similar step size, so the waves
// whose lanes need the long-step Kepler branch are few.  Results do not
// depend on the order: every lane's arithmetic is its own.
std::vector<uint32_t> frequency_order(uint32_t n, const double* r, const double* v, double mu) {
    std::vector<double> key(n);
    for (uint32_t k = 0; k < n; k++) {
        const double* x = r + 3 * (size_t)k;
        const double* u = v + 3 * (size_t)k;
        const double ia = 2.0 / std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) -
                          (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / mu;
        key[k] = ia > 0.0 ? ia : 0.0;  // NaN and unbound: 0
    }
    std::vector<uint32_t> order(n);
    for (uint32_t k = 0; k < n; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return key[a] > key[b]; });
    return order;
}

}  // namespace

struct glx_batch {
    int device = 0;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    uint32_t n_sys = 0, steps_max = 0;
    uint64_t n_slots = 0;
    int n_cu = 0;
    std::vector<SystemHost> host;
    std::vector<SystemDev> sys_h;
    SystemDev* sys_d = nullptr;
    double* d[9] = {};
    uint16_t* flags = nullptr;
    uint32_t* dti = nullptr;
    uint32_t* id = nullptr;
    double* bak_d[9] = {};  // autotuner backup
    uint16_t* bak_flags = nullptr;
    uint32_t* bak_dti = nullptr;
    double* stage = nullptr;  // host <-> device AoS staging, 3 x max particles
    uint32_t* tile_sys = nullptr;
    uint32_t* tile_first = nullptr;
    uint32_t n_tiles = 0;
    uint32_t* clock = nullptr;
    uint32_t* queue = nullptr;
    unsigned long long* sum = nullptr;
    uint64_t clock_h = 0;  // steps done (host copy of the device step clock)
    glx_batch_config cfg{};
    size_t lds_bytes = 0;
    uint32_t grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_form = -1;
    std::vector<void*> allocs;
};

namespace {

template <typename T>
int dalloc(glx_batch* b, T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, (count ? count : 1) * sizeof(T));
    if (e != hipSuccess)
        return fail(GLX_ENOMEM, "hipMalloc(%zu) failed: %s", count * sizeof(T), hipGetErrorString(e));
    b->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return GLX_OK;
}

void drop_graph(glx_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_form = -1;
}

void release(glx_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->s) (void)hipStreamSynchronize(b->s);
    drop_graph(b);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    if (b->s) (void)hipStreamDestroy(b->s);
    delete b;
}

BatchArgs batch_args(const glx_batch* b) {
    BatchArgs B{};
    for (int k = 0; k < 9; k++) B.d[k] = b->d[k];
    B.flags = b->flags;
    B.dti = b->dti;
    B.sys = b->sys_d;
    B.tile_sys = b->tile_sys;
    B.tile_first = b->tile_first;
    B.n_tiles = b->n_tiles;
    B.steps = b->cfg.steps_per_launch;
    B.clock = b->clock;
    B.queue = b->queue;
    B.persistent = b->cfg.persistent;
    return B;
}

int check_config(const glx_batch* b, const glx_batch_config* c) {
    if (c->steps_per_launch < 1 || c->steps_per_launch > b->steps_max)
        return fail(GLX_EINVAL, "steps_per_launch %u outside [1, %u]", c->steps_per_launch, b->steps_max);
    if (c->lanes != 64 && c->lanes != 128 && c->lanes != 256)
        return fail(GLX_EINVAL, "lanes %u: use 64, 128 or 256", c->lanes);
    if (c->persistent > 1) return fail(GLX_EINVAL, "persistent must be 0 or 1");
    if (c->waves_per_simd > 8) return fail(GLX_EINVAL, "waves_per_simd %u outside [0, 8]", c->waves_per_simd);
    if (c->persistent && c->waves_per_simd == 0)
        return fail(GLX_EINVAL, "a persistent launch needs waves_per_simd >= 1 (it sizes the grid)");
    return GLX_OK;
}

// Tiles of `lanes` particles, heaviest systems first (per-step work grows
// with the planet count), each system's tiles in slot order.
int build_tiles(glx_batch* b, uint32_t lanes) {
    std::vector<uint32_t> order(b->n_sys);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return b->host[x].nb > b->host[y].nb; });
    std::vector<uint32_t> ts, tf;
    for (uint32_t s : order) {
        const uint32_t n = b->host[s].n;
        for (uint32_t k = 0; k < (n + lanes - 1) / lanes; k++) {
            ts.push_back(s);
            tf.push_back(b->sys_h[s].first + k * lanes);
        }
    }
    b->n_tiles = (uint32_t)ts.size();
    if (b->n_tiles) {
        GLX_HIP(hipMemcpy(b->tile_sys, ts.data(), ts.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        GLX_HIP(hipMemcpy(b->tile_first, tf.data(), tf.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    return GLX_OK;
}

int apply_config(glx_batch* b, const glx_batch_config* c) {
    GLX_HIP(hipStreamSynchronize(b->s));
    drop_graph(b);
    b->cfg = *c;
    const uint32_t waves_per_group = c->lanes / 64;
    if (c->waves_per_simd) {
        // A workgroup holds waves_per_group waves spread over the CU's four
        // SIMDs; asking for 1/groups of the CU's LDS caps residency.
        uint32_t groups = std::max(1u, 4u * c->waves_per_simd / waves_per_group);
        b->lds_bytes = std::min(kLdsMaxPerGroup, (kLdsPerCu / groups) & ~(size_t)1023);
        b->grid = c->persistent ? (uint32_t)b->n_cu * groups : 0;
    } else {
        b->lds_bytes = kLdsPer64Lanes * waves_per_group;
        b->grid = 0;
    }
    for (uint32_t s = 0; s < b->n_sys; s++) {
        SystemDev& S = b->sys_h[s];
        uint32_t chunk = (uint32_t)(b->lds_bytes / ((size_t)S.nb * 3 * sizeof(double)));
        S.chunk = std::max(1u, std::min(chunk, c->steps_per_launch));
    }
    GLX_HIP(hipMemcpy(b->sys_d, b->sys_h.data(), b->n_sys * sizeof(SystemDev), hipMemcpyHostToDevice));
    if (int rc = build_tiles(b, c->lanes)) return rc;
    if (!c->persistent) b->grid = b->n_tiles;
    return GLX_OK;
}

int capture(glx_batch* b, int form) {
    drop_graph(b);
    const BatchArgs B = batch_args(b);
    GLX_HIP(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
    hipError_t e = glx::launch_batch_interp(b->sys_d, b->n_sys, b->cfg.steps_per_launch, b->clock, b->s);
    if (e == hipSuccess) e = glx::launch_batch_block(B, form, b->grid, b->cfg.lanes, b->lds_bytes, b->s);
    if (e == hipSuccess) e = glx::launch_batch_finish(b->clock, b->queue, b->cfg.steps_per_launch, b->s);
    hipGraph_t g = nullptr;
    const hipError_t e_end = hipStreamEndCapture(b->s, &g);
    if (e != hipSuccess || e_end != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(GLX_EHIP, "capturing the batched launch failed: %s",
                    hipGetErrorString(e != hipSuccess ? e : e_end));
    }
    b->graph = g;
    GLX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_form = form;
    return GLX_OK;
}

int check_form(int form) {
    if (form != GLX_FORM_FP64 && form != GLX_FORM_FAST && form != GLX_FORM_MIXED)
        return fail(GLX_EINVAL, "kernel form %d is not available", form);
    return GLX_OK;
}

int run(glx_batch* b, uint32_t n, int form, float* ms) {
    for (uint32_t s = 0; s < b->n_sys; s++) {
        const SystemDev& S = b->sys_h[s];
        const double last = S.t0 + (double)(b->clock_h + (uint64_t)n * b->cfg.steps_per_launch) * S.dt;
        if (!(last <= b->host[s].t_last + kHistSlack))
            return fail(GLX_EINVAL, "system %u: step time %.17g runs past its history (last sample %.17g)", s,
                        last, b->host[s].t_last);
    }
    if (b->graph_form != form || !b->exec)
        if (int rc = capture(b, form)) return rc;
    GLX_HIP(hipEventRecord(b->e0, b->s));
    for (uint32_t k = 0; k < n; k++) GLX_HIP(hipGraphLaunch(b->exec, b->s));
    GLX_HIP(hipEventRecord(b->e1, b->s));
    GLX_HIP(hipStreamSynchronize(b->s));
    b->clock_h += (uint64_t)n * b->cfg.steps_per_launch;
    if (ms) GLX_HIP(hipEventElapsedTime(ms, b->e0, b->e1));
    return GLX_OK;
}

int sum_steps(glx_batch* b, uint64_t* total) {
    GLX_HIP(glx::launch_sum_steps(b->dti, (uint32_t)b->n_slots, b->sum, b->s));
    unsigned long long h = 0;
    GLX_HIP(hipMemcpyAsync(&h, b->sum, sizeof h, hipMemcpyDeviceToHost, b->s));
    GLX_HIP(hipStreamSynchronize(b->s));
    *total = h;
    return GLX_OK;
}

// Copies of the mutable particle state (autotuner).
int save_state(glx_batch* b, bool restore) {
    if (!b->bak_flags) {
        for (int k = 0; k < 9; k++)
            if (int rc = dalloc(b, &b->bak_d[k], b->n_slots)) return rc;
        if (int rc = dalloc(b, &b->bak_flags, b->n_slots)) return rc;
        if (int rc = dalloc(b, &b->bak_dti, b->n_slots)) return rc;
    }
    const size_t n = b->n_slots;
    for (int k = 0; k < 9; k++)
        GLX_HIP(hipMemcpyAsync(restore ? b->d[k] : b->bak_d[k], restore ? b->bak_d[k] : b->d[k], n * sizeof(double),
                               hipMemcpyDeviceToDevice, b->s));
    GLX_HIP(hipMemcpyAsync(restore ? b->flags : b->bak_flags, restore ? b->bak_flags : b->flags,
                           n * sizeof(uint16_t), hipMemcpyDeviceToDevice, b->s));
    GLX_HIP(hipMemcpyAsync(restore ? b->dti : b->bak_dti, restore ? b->bak_dti : b->dti, n * sizeof(uint32_t),
                           hipMemcpyDeviceToDevice, b->s));
    GLX_HIP(hipStreamSynchronize(b->s));
    return GLX_OK;
}

int set_clock(glx_batch* b, uint64_t value) {
    const uint32_t v[2] = {(uint32_t)value, 0};
    GLX_HIP(hipMemcpyAsync(b->clock, &v[0], sizeof(uint32_t), hipMemcpyHostToDevice, b->s));
    GLX_HIP(hipMemcpyAsync(b->queue, &v[1], sizeof(uint32_t), hipMemcpyHostToDevice, b->s));
    GLX_HIP(hipStreamSynchronize(b->s));
    b->clock_h = value;
    return GLX_OK;
}

}  // namespace

extern "C" {

int glx_batch_create(int device, const glx_system* systems, uint32_t n_systems, uint32_t steps_max,
                     glx_batch** out) {
    if (!out) return fail(GLX_EINVAL, "glx_batch_create: null output");
    *out = nullptr;
    if (n_systems == 0 || !systems) return fail(GLX_EINVAL, "a batch needs at least one system");
    if (steps_max < 1) return fail(GLX_EINVAL, "steps_max must be >= 1");
    uint64_t slots = 0, hist_doubles = 0, log_doubles = 0;
    uint32_t n_max = 1;
    for (uint32_t s = 0; s < n_systems; s++) {
        const glx_system& y = systems[s];
        if (y.n_bodies < 2 || y.n_bodies > GLX_BATCH_MAX_BODIES)
            return fail(GLX_EINVAL, "system %u: n_bodies %u outside [2, %d]", s, y.n_bodies, GLX_BATCH_MAX_BODIES);
        if (!y.m || !y.rh || !y.t || !y.r || !y.v) return fail(GLX_EINVAL, "system %u: null array", s);
        if (!(y.m[0] > 0.0)) return fail(GLX_EINVAL, "system %u: central mass must be positive", s);
        if (!(y.dt > 0.0)) return fail(GLX_EINVAL, "system %u: dt must be positive", s);
        if (y.n_samples < 2) return fail(GLX_EINVAL, "system %u: a history needs at least 2 samples", s);
        for (uint64_t k = 0; k + 1 < y.n_samples; k++)
            if (!(y.t[k + 1] > y.t[k])) return fail(GLX_EINVAL, "system %u: history times must increase", s);
        if (!(y.t0 >= y.t[0] - kHistSlack)) return fail(GLX_EINVAL, "system %u: t0 precedes its history", s);
        slots += ((uint64_t)y.n_particles + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
        hist_doubles += y.n_samples * (1 + 6 * (uint64_t)(y.n_bodies - 1));
        log_doubles += (uint64_t)steps_max * y.n_bodies * 3;
        n_max = std::max(n_max, y.n_particles);
    }
    if (slots > 0xFFFFFFFFull - kSlotAlign) return fail(GLX_EINVAL, "batch too large (%llu particle slots)",
                                                        (unsigned long long)slots);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GLX_EHIP, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(GLX_EINVAL, "device %d out of range", device);

    glx_batch* b = new glx_batch();
    b->device = device;
    b->n_sys = n_systems;
    b->steps_max = steps_max;
    b->n_slots = slots;
    auto bail = [&](int code) {
        release(b);
        return code;
    };
    int rc;
    if (hipSetDevice(device) != hipSuccess) return bail(fail(GLX_EHIP, "hipSetDevice failed"));
    if (hipStreamCreate(&b->s) != hipSuccess || hipEventCreate(&b->e0) != hipSuccess ||
        hipEventCreate(&b->e1) != hipSuccess)
        return bail(fail(GLX_EHIP, "stream / event creation failed"));
    if (hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return bail(fail(GLX_EHIP, "hipDeviceGetAttribute failed"));
    for (int k = 0; k < 9; k++)
        if ((rc = dalloc(b, &b->d[k], slots))) return bail(rc);
    double* hist = nullptr;
    double* logs = nullptr;
    if ((rc = dalloc(b, &b->flags, slots)) || (rc = dalloc(b, &b->dti, slots)) || (rc = dalloc(b, &b->id, slots)) ||
        (rc = dalloc(b, &b->stage, 3 * (size_t)n_max)) || (rc = dalloc(b, &b->sys_d, n_systems)) ||
        (rc = dalloc(b, &b->tile_sys, slots / 64 + n_systems)) ||
        (rc = dalloc(b, &b->tile_first, slots / 64 + n_systems)) || (rc = dalloc(b, &b->clock, 1)) ||
        (rc = dalloc(b, &b->queue, 1)) || (rc = dalloc(b, &b->sum, 1 + glx::kSumParts)) || (rc = dalloc(b, &hist, hist_doubles)) ||
        (rc = dalloc(b, &logs, log_doubles)))
        return bail(rc);
    // Zeroed particle state: padding slots stay flagged-free zeros and never step.
    for (int k = 0; k < 9; k++)
        if (hipMemset(b->d[k], 0, slots * sizeof(double)) != hipSuccess) return bail(fail(GLX_EHIP, "hipMemset failed"));
    if (hipMemset(b->flags, 0, slots * sizeof(uint16_t)) != hipSuccess ||
        hipMemset(b->dti, 0, slots * sizeof(uint32_t)) != hipSuccess ||
        hipMemset(b->id, 0, slots * sizeof(uint32_t)) != hipSuccess)
        return bail(fail(GLX_EHIP, "hipMemset failed"));

    b->host.resize(n_systems);
    b->sys_h.resize(n_systems);
    uint64_t slot = 0, hoff = 0, loff = 0;
    for (uint32_t s = 0; s < n_systems; s++) {
        const glx_system& y = systems[s];
        SystemDev& S = b->sys_h[s];
        std::memset(&S, 0, sizeof S);
        const uint32_t np = y.n_bodies - 1;
        for (uint32_t j = 0; j < y.n_bodies; j++) {
            S.pc.m[j] = y.m[j];
            S.pc.rh2[j] = y.rh[j] * y.rh[j];
            S.pc.id[j] = y.id ? y.id[j] : j;
        }
        double* ht = hist + hoff;
        double* hr = ht + y.n_samples;
        double* hv = hr + y.n_samples * np * 3;
        hoff += y.n_samples * (1 + 6 * (uint64_t)np);
        if (hipMemcpy(ht, y.t, y.n_samples * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(hr, y.r, y.n_samples * np * 3 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(hv, y.v, y.n_samples * np * 3 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(GLX_EHIP, "history upload failed"));
        S.log = logs + loff;
        loff += (uint64_t)steps_max * y.n_bodies * 3;
        S.ht = ht;
        S.hr = hr;
        S.hv = hv;
        S.n_samples = y.n_samples;
        S.t0 = y.t0;
        S.dt = y.dt;
        S.half_dt = y.dt / 2;
        S.mu = y.m[0];
        S.inv_mu = 1.0 / y.m[0];
        S.dt_sqrt_mu = y.dt * std::sqrt(y.m[0]);
        S.inv_sqrt_mu = 1.0 / std::sqrt(y.m[0]);
        S.inv_dt = 1.0 / y.dt;
        S.outer2 = y.outer_radius * y.outer_radius;
        S.fp64_core = glx::fp64_core_constants(y.m[0], y.dt) ? 1u : 0u;
        S.sun_vote = std::sqrt(S.pc.rh2[0]) * (1.0 + glx::kVoteMargin);
        S.outer_vote = y.outer_radius * (1.0 - glx::kVoteMargin);
        S.nb = y.n_bodies;
        S.first = (uint32_t)slot;
        S.n = y.n_particles;
        slot += ((uint64_t)y.n_particles + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
        b->host[s] = SystemHost{y.n_bodies, y.n_particles, y.t[0], y.t[y.n_samples - 1]};
    }
    if ((rc = set_clock(b, 0))) return bail(rc);
    const glx_batch_config def{std::min(steps_max, 2048u), 256, 0, 0};
    if ((rc = apply_config(b, &def))) return bail(rc);
    *out = b;
    glx::clear_error();
    return GLX_OK;
}

int glx_batch_destroy(glx_batch* b) {
    release(b);
    return GLX_OK;
}

int glx_batch_upload(glx_batch* b, uint32_t sys, const double* r, const double* v, const uint16_t* flags,
                     const uint32_t* id) {
    if (!b) return fail(GLX_EINVAL, "null batch");
    if (sys >= b->n_sys) return fail(GLX_EINVAL, "system %u outside [0, %u)", sys, b->n_sys);
    const uint32_t n = b->host[sys].n;
    if (n == 0) return GLX_OK;
    if (!r || !v || !flags || !id) return fail(GLX_EINVAL, "glx_batch_upload: r, v, flags, id are required");
    GLX_HIP(hipSetDevice(b->device));
    const uint32_t f = b->sys_h[sys].first;
    std::vector<uint32_t>& order = b->host[sys].order;
    order = frequency_order(n, r, v, b->sys_h[sys].mu);
    std::vector<double> rv(3 * (size_t)n);
    std::vector<uint16_t> fl(n);
    std::vector<uint32_t> ids(n);
    for (int k = 0; k < 2; k++) {
        const double* src = k ? v : r;
        for (uint32_t j = 0; j < n; j++)
            for (int c = 0; c < 3; c++) rv[3 * (size_t)j + c] = src[3 * (size_t)order[j] + c];
        GLX_HIP(hipMemcpyAsync(b->stage, rv.data(), (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice, b->s));
        GLX_HIP(glx::launch_aos_to_soa(b->stage, b->d[3 * k] + f, b->d[3 * k + 1] + f, b->d[3 * k + 2] + f, n, b->s));
        GLX_HIP(hipStreamSynchronize(b->s));  // rv is reused
    }
    for (uint32_t j = 0; j < n; j++) {
        fl[j] = flags[order[j]];
        ids[j] = id[order[j]];
    }
    GLX_HIP(hipMemcpyAsync(b->flags + f, fl.data(), (size_t)n * sizeof(uint16_t), hipMemcpyHostToDevice, b->s));
    GLX_HIP(hipMemcpyAsync(b->id + f, ids.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, b->s));
    GLX_HIP(hipMemsetAsync(b->dti + f, 0, (size_t)n * sizeof(uint32_t), b->s));
    GLX_HIP(hipStreamSynchronize(b->s));
    return GLX_OK;
}

int glx_batch_download(glx_batch* b, uint32_t sys, double* r, double* v, uint16_t* flags, uint32_t* id,
                       uint32_t* dti) {
    if (!b) return fail(GLX_EINVAL, "null batch");
    if (sys >= b->n_sys) return fail(GLX_EINVAL, "system %u outside [0, %u)", sys, b->n_sys);
    const uint32_t n = b->host[sys].n;
    if (n == 0) return GLX_OK;
    GLX_HIP(hipSetDevice(b->device));
    const uint32_t f = b->sys_h[sys].first;
    // slots back to the uploaded order (identity if nothing was uploaded)
    const std::vector<uint32_t>& order = b->host[sys].order;
    auto dst = [&](uint32_t j) { return order.empty() ? j : order[j]; };
    double* out[2] = {r, v};
    std::vector<double> rv(3 * (size_t)n);
    for (int k = 0; k < 2; k++) {
        if (!out[k]) continue;
        GLX_HIP(glx::launch_soa_to_aos(b->d[3 * k] + f, b->d[3 * k + 1] + f, b->d[3 * k + 2] + f, b->stage, n, b->s));
        GLX_HIP(hipMemcpyAsync(rv.data(), b->stage, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, b->s));
        GLX_HIP(hipStreamSynchronize(b->s));
        for (uint32_t j = 0; j < n; j++)
            for (int c = 0; c < 3; c++) out[k][3 * (size_t)dst(j) + c] = rv[3 * (size_t)j + c];
    }
    std::vector<uint16_t> fl(flags ? n : 0);
    std::vector<uint32_t> ids(id ? n : 0), dt(dti ? n : 0);
    if (flags) GLX_HIP(hipMemcpyAsync(fl.data(), b->flags + f, (size_t)n * sizeof(uint16_t), hipMemcpyDeviceToHost, b->s));
    if (id) GLX_HIP(hipMemcpyAsync(ids.data(), b->id + f, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, b->s));
    if (dti) GLX_HIP(hipMemcpyAsync(dt.data(), b->dti + f, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, b->s));
    GLX_HIP(hipStreamSynchronize(b->s));
    for (uint32_t j = 0; j < n; j++) {
        if (flags) flags[dst(j)] = fl[j];
        if (id) id[dst(j)] = ids[j];
        if (dti) dti[dst(j)] = dt[j];
    }
    return GLX_OK;
}

int glx_batch_init_accel(glx_batch* b) {
    if (!b) return fail(GLX_EINVAL, "null batch");
    GLX_HIP(hipSetDevice(b->device));
    GLX_HIP(glx::launch_batch_init(batch_args(b), b->cfg.lanes, b->s));
    GLX_HIP(hipStreamSynchronize(b->s));
    return GLX_OK;
}

int glx_batch_set_config(glx_batch* b, const glx_batch_config* cfg) {
    if (!b || !cfg) return fail(GLX_EINVAL, "null argument");
    if (int rc = check_config(b, cfg)) return rc;
    GLX_HIP(hipSetDevice(b->device));
    return apply_config(b, cfg);
}

int glx_batch_get_config(glx_batch* b, glx_batch_config* cfg) {
    if (!b || !cfg) return fail(GLX_EINVAL, "null argument");
    *cfg = b->cfg;
    return GLX_OK;
}

int glx_batch_run(glx_batch* b, uint32_t n_launches, int form, float* elapsed_ms) {
    if (!b) return fail(GLX_EINVAL, "null batch");
    if (int rc = check_form(form)) return rc;
    if (elapsed_ms) *elapsed_ms = 0.0f;
    if (n_launches == 0) return GLX_OK;
    GLX_HIP(hipSetDevice(b->device));
    return run(b, n_launches, form, elapsed_ms);
}

int glx_batch_sum_steps(glx_batch* b, uint64_t* total) {
    if (!b || !total) return fail(GLX_EINVAL, "null argument");
    GLX_HIP(hipSetDevice(b->device));
    return sum_steps(b, total);
}

int glx_batch_tune(glx_batch* b, int form, const glx_batch_config* candidates, uint32_t n_candidates,
                   glx_batch_config* best, double* best_rate) {
    if (!b) return fail(GLX_EINVAL, "null batch");
    if (int rc = check_form(form)) return rc;
    std::vector<glx_batch_config> cands;
    if (candidates && n_candidates) {
        cands.assign(candidates, candidates + n_candidates);
    } else {
        for (uint32_t steps : {512u, 1024u, 2048u}) {
            if (steps > b->steps_max) continue;
            for (uint32_t lanes : {64u, 128u, 256u}) {
                cands.push_back({steps, lanes, 0, 0});
                cands.push_back({steps, lanes, 1, 4});
            }
        }
        if (cands.empty()) {
            cands.push_back({b->steps_max, 256, 0, 0});
            cands.push_back({b->steps_max, 256, 1, 4});
        }
    }
    for (const auto& c : cands)
        if (int rc = check_config(b, &c)) return rc;
    GLX_HIP(hipSetDevice(b->device));
    const glx_batch_config before = b->cfg;
    const uint64_t clock0 = b->clock_h;
    if (int rc = save_state(b, false)) return rc;
    uint64_t steps0 = 0;
    if (int rc = sum_steps(b, &steps0)) return rc;
    double best_r = -1.0;
    glx_batch_config best_c = before;
    int rc = GLX_OK;
    for (const auto& c : cands) {
        if ((rc = apply_config(b, &c))) break;
        // one untimed launch (graph instantiation and upload), one timed
        if ((rc = save_state(b, true)) || (rc = set_clock(b, clock0)) || (rc = run(b, 1, form, nullptr))) break;
        if ((rc = save_state(b, true)) || (rc = set_clock(b, clock0))) break;
        float ms = 0.0f;
        if ((rc = run(b, 1, form, &ms))) break;
        uint64_t steps1 = 0;
        if ((rc = sum_steps(b, &steps1))) break;
        const double rate = ms > 0.0f ? (double)(steps1 - steps0) / (ms * 1e-3) : 0.0;
        if (rate > best_r) {
            best_r = rate;
            best_c = c;
        }
    }
    // Back to the state before tuning, with the fastest shape.
    int rc2 = save_state(b, true);
    if (!rc2) rc2 = set_clock(b, clock0);
    if (!rc2) rc2 = apply_config(b, rc ? &before : &best_c);
    if (rc) return rc;
    if (rc2) return rc2;
    if (best) *best = best_c;
    if (best_rate) *best_rate = best_r;
    return GLX_OK;
}

}  // extern "C"
