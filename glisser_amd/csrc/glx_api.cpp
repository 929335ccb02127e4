// C ABI of the gfx950 particle integrator: device memory, streams, events
// and the host<->device copies around the time-block kernel.
//
// Ownership mirrors the reference's split (DeviceData + WHCudaIntegrator +
// Executor, src/data.cuh:55-67, src/wh.cu:272-287, src/executor.cu:151-175)
// but in one context: particle SoA arrays, a scratch set for the stable
// partition, an optional rollback copy, and two planet-block buffers so the
// upload of block k+1 overlaps the kernel of block k (the planet_data_id
// parity of src/executor.cu:442-461).  Ordering between the upload and the
// main stream uses events instead of host synchronisation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "errors.h"
#include "glisser_amd.h"
#include "staging.h"
#include "launch.h"

using glx::BlockArgs;
using glx::ParticleSet;
using glx::PlanetConsts;
using glx::RollbackSet;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define GLX_HIP(call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(GLX_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                            \
    } while (0)

constexpr uint32_t kMaxShards = 8;       // streams of a multi-block launch
constexpr size_t kLdsPerCu = 160 * 1024;  // gfx950 LDS per compute unit

struct PlanetBuf {
    double* dev_log = nullptr;  // [tbsize_max][n_bodies_max][3]
    double* pin_log = nullptr;  // pinned staging of the same size
    hipEvent_t uploaded = nullptr;
    hipEvent_t consumed = nullptr;
    uint32_t n_bodies = 0;
    uint32_t n_steps = 0;
    double mu = 0;
    PlanetConsts pc{};
    bool valid = false;
};

constexpr uint32_t kQueryMax = 1024;  // times per glx_interp_at batch

constexpr uint32_t kQueueSegment = 256;             // default steps per task of a queued launch
constexpr uint32_t kQueueSegmentMax = 4096;         // longest task: a slot wait lasts at most one task
constexpr uint64_t kQueueWaitTicks = 3000000000ull; // 30 s of the 100 MHz counter: bound of a FIFO slot wait
constexpr uint32_t kQueueErrWord = 8;               // pin_small slot of a queued launch's error word

// Sampled planet history resident on the device (glx_upload_history).
struct History {
    double* t = nullptr;
    double* r = nullptr;
    double* v = nullptr;
    double* q_times = nullptr;  // [kQueryMax]
    double* q_rows = nullptr;   // [kQueryMax][GLX_MAX_BODIES][3]
    uint64_t n_samples = 0;
    uint32_t n_bodies = 0;
    double m[GLX_MAX_BODIES] = {};
    uint32_t id[GLX_MAX_BODIES] = {};
    bool valid = false;
};

}  // namespace

namespace glx {
int set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
void clear_error() { g_err.clear(); }
}  // namespace glx

struct glx_ctx {
    int device = 0;
    uint64_t cap = 0;
    uint32_t nb_max = 0, tb_max = 0;
    bool keep_rb = false;
    hipStream_t main_s = nullptr, htd_s = nullptr;
    ParticleSet cur{}, tmp{};
    RollbackSet rb{}, tmp_rb{};
    uint32_t* counts = nullptr;
    uint32_t* offsets = nullptr;
    uint32_t* totals = nullptr;
    unsigned long long* sum_dev = nullptr;
    double* dev_aos = nullptr;       // device staging, cap x 3 doubles
    uint32_t* pin_small = nullptr;   // pinned: totals (2) + sum (2 words)
    std::vector<PlanetBuf> pb;
    hipStream_t shard_s[kMaxShards] = {};
    hipEvent_t shard_done[kMaxShards] = {};
    hipEvent_t k0 = nullptr, k1 = nullptr, q0 = nullptr, q1 = nullptr;
    bool kernel_timed = false, part_timed = false;
    uint64_t n_alive = 0, n_total = 0;
    History hist;
    double hist_t0 = 0, hist_t1 = 0;  // first and last history times (host copy)
    glx::TrackRecordDev* trk_dev = nullptr;  // track records (grown on demand)
    uint64_t trk_cap = 0;
    uint32_t* rank_dev = nullptr;  // id ranks + scan tiles (grown on demand)
    uint64_t rank_cap = 0;
    glx_encounter* enc_dev = nullptr;  // encounter-record staging (grown on demand)
    glx_encounter* enc_pin = nullptr;
    uint64_t enc_cap = 0;
    hipStream_t dth_s = nullptr;      // device -> host copies of staged records
    hipEvent_t enc_packed = nullptr;  // records packed on the main stream
    hipEvent_t enc_copied = nullptr;  // records in the pinned staging buffer
    // Queued launches (glx_run_blocks_queued), allocated on first use:
    // segment and block tables, counters + error word, the task FIFO, and
    // the tables' pinned staging.
    glx::QueueSeg* q_segs = nullptr;
    glx::QueueBlock* q_blocks = nullptr;
    uint32_t* q_ctl = nullptr;
    uint32_t* q_ring = nullptr;
    unsigned char* q_pin = nullptr;
    hipEvent_t q_staged = nullptr;  // the staging's last copy has completed
    bool q_pending = false;         // a queued launch's error word is bound for pin_small[kQueueErrWord]
    uint32_t n_cus = 0;             // compute units of the device
    std::vector<void*> allocs;
};

namespace {

template <typename T>
int dalloc(glx_ctx* c, T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) + 64);
    if (e != hipSuccess) return fail(GLX_ENOMEM, "hipMalloc(%zu) failed: %s", count * sizeof(T), hipGetErrorString(e));
    c->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return GLX_OK;
}

int alloc_set(glx_ctx* c, ParticleSet* s) {
    for (int k = 0; k < 9; k++)
        if (int rc = dalloc(c, &s->d[k], c->cap)) return rc;
    if (int rc = dalloc(c, &s->flags, c->cap)) return rc;
    if (int rc = dalloc(c, &s->dti, c->cap)) return rc;
    return dalloc(c, &s->id, c->cap);
}

int alloc_rb(glx_ctx* c, RollbackSet* s) {
    for (int k = 0; k < 6; k++)
        if (int rc = dalloc(c, &s->d[k], c->cap)) return rc;
    if (int rc = dalloc(c, &s->flags, c->cap)) return rc;
    if (int rc = dalloc(c, &s->dti, c->cap)) return rc;
    return dalloc(c, &s->id, c->cap);
}

void release(glx_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->main_s) (void)hipStreamSynchronize(c->main_s);
    if (c->htd_s) (void)hipStreamSynchronize(c->htd_s);
    if (c->dth_s) (void)hipStreamSynchronize(c->dth_s);
    for (uint32_t k = 0; k < kMaxShards; k++)
        if (c->shard_s[k]) (void)hipStreamSynchronize(c->shard_s[k]);
    for (void* p : c->allocs) (void)hipFree(p);
    for (double* p : {c->hist.t, c->hist.r, c->hist.v, c->hist.q_times, c->hist.q_rows})
        if (p) (void)hipFree(p);
    if (c->enc_dev) (void)hipFree(c->enc_dev);
    if (c->enc_pin) (void)hipHostFree(c->enc_pin);
    if (c->trk_dev) (void)hipFree(c->trk_dev);
    if (c->rank_dev) (void)hipFree(c->rank_dev);
    for (auto& b : c->pb) {
        if (b.pin_log) (void)hipHostFree(b.pin_log);
        if (b.uploaded) (void)hipEventDestroy(b.uploaded);
        if (b.consumed) (void)hipEventDestroy(b.consumed);
    }
    if (c->pin_small) (void)hipHostFree(c->pin_small);
    if (c->q_pin) (void)hipHostFree(c->q_pin);
    if (c->q_staged) (void)hipEventDestroy(c->q_staged);
    for (hipEvent_t e : {c->k0, c->k1, c->q0, c->q1})
        if (e) (void)hipEventDestroy(e);
    for (uint32_t k = 0; k < kMaxShards; k++) {
        if (c->shard_done[k]) (void)hipEventDestroy(c->shard_done[k]);
        if (c->shard_s[k]) (void)hipStreamDestroy(c->shard_s[k]);
    }
    for (hipEvent_t e : {c->enc_packed, c->enc_copied})
        if (e) (void)hipEventDestroy(e);
    if (c->dth_s) (void)hipStreamDestroy(c->dth_s);
    if (c->main_s) (void)hipStreamDestroy(c->main_s);
    if (c->htd_s) (void)hipStreamDestroy(c->htd_s);
    delete c;
}

int check_range(const glx_ctx* c, uint64_t begin, uint64_t n) {
    if (begin > c->cap || n > c->cap - begin)
        return fail(GLX_EINVAL, "particle range [%llu, %llu) exceeds capacity %llu", (unsigned long long)begin,
                    (unsigned long long)(begin + n), (unsigned long long)c->cap);
    return GLX_OK;
}

// Host AoS vec3 array -> three device SoA arrays at offset `begin`.
int put_vec3(glx_ctx* c, const double* host, double* const* soa, uint64_t begin, uint64_t n) {
    GLX_HIP(hipMemcpyAsync(c->dev_aos, host, n * 3 * sizeof(double), hipMemcpyHostToDevice, c->main_s));
    GLX_HIP(glx::launch_aos_to_soa(c->dev_aos, soa[0] + begin, soa[1] + begin, soa[2] + begin, (uint32_t)n,
                                   c->main_s));
    GLX_HIP(hipStreamSynchronize(c->main_s));
    return GLX_OK;
}

int get_vec3(glx_ctx* c, double* host, double* const* soa, uint64_t begin, uint64_t n) {
    GLX_HIP(glx::launch_soa_to_aos(soa[0] + begin, soa[1] + begin, soa[2] + begin, c->dev_aos, (uint32_t)n,
                                   c->main_s));
    GLX_HIP(hipMemcpyAsync(host, c->dev_aos, n * 3 * sizeof(double), hipMemcpyDeviceToHost, c->main_s));
    GLX_HIP(hipStreamSynchronize(c->main_s));
    return GLX_OK;
}

template <typename T>
int put_plain(glx_ctx* c, const T* host, T* dev, uint64_t n) {
    GLX_HIP(hipMemcpyAsync(dev, host, n * sizeof(T), hipMemcpyHostToDevice, c->main_s));
    GLX_HIP(hipStreamSynchronize(c->main_s));
    return GLX_OK;
}

template <typename T>
int get_plain(glx_ctx* c, T* host, const T* dev, uint64_t n) {
    GLX_HIP(hipMemcpyAsync(host, dev, n * sizeof(T), hipMemcpyDeviceToHost, c->main_s));
    GLX_HIP(hipStreamSynchronize(c->main_s));
    return GLX_OK;
}

int fill_consts(uint32_t n_bodies, const double* m, const uint32_t* id, const double* rh, PlanetConsts* pc) {
    std::memset(pc, 0, sizeof *pc);
    for (uint32_t j = 0; j < n_bodies; j++) {
        pc->m[j] = m[j];
        pc->rh2[j] = rh[j] * rh[j];
        pc->id[j] = id ? id[j] : j;
    }
    return GLX_OK;
}

BlockArgs base_args(glx_ctx* c) {
    BlockArgs A{};
    A.rx = c->cur.d[0];
    A.ry = c->cur.d[1];
    A.rz = c->cur.d[2];
    A.vx = c->cur.d[3];
    A.vy = c->cur.d[4];
    A.vz = c->cur.d[5];
    A.ax = c->cur.d[6];
    A.ay = c->cur.d[7];
    A.az = c->cur.d[8];
    A.flags = c->cur.flags;
    A.dti = c->cur.dti;
    return A;
}

constexpr size_t kLdsBudget = 32768;  // bytes of planet rows staged per workgroup

}  // namespace

extern "C" {

const char* glx_last_error(void) { return g_err.c_str(); }

int glx_create(const glx_params* p, glx_ctx** out) {
    if (!p || !out) return fail(GLX_EINVAL, "glx_create: null argument");
    *out = nullptr;
    if (p->n_bodies_max < 1 || p->n_bodies_max > GLX_MAX_BODIES)
        return fail(GLX_EINVAL, "n_bodies_max %u outside [1, %d]", p->n_bodies_max, GLX_MAX_BODIES);
    if (p->tbsize_max < 1) return fail(GLX_EINVAL, "tbsize_max must be >= 1");
    if (p->n_particles > 0xFFFFFFFFull - 2048)
        return fail(GLX_EINVAL, "n_particles %llu too large", (unsigned long long)p->n_particles);
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(GLX_EHIP, "no HIP device available");
    if (p->device < 0 || p->device >= ndev) return fail(GLX_EINVAL, "device %d out of range", p->device);

    glx_ctx* c = new glx_ctx();
    c->device = p->device;
    c->cap = p->n_particles == 0 ? 1 : p->n_particles;
    c->nb_max = p->n_bodies_max;
    c->tb_max = p->tbsize_max;
    c->keep_rb = p->keep_rollback != 0;
    int rc = GLX_OK;
    auto bail = [&](int code) {
        std::string msg = g_err;
        release(c);
        g_err = msg;
        return code;
    };
    if (hipSetDevice(c->device) != hipSuccess) return bail(fail(GLX_EHIP, "hipSetDevice failed"));
    if (hipStreamCreate(&c->main_s) != hipSuccess || hipStreamCreate(&c->htd_s) != hipSuccess ||
        hipStreamCreate(&c->dth_s) != hipSuccess)
        return bail(fail(GLX_EHIP, "hipStreamCreate failed"));
    if (hipEventCreateWithFlags(&c->enc_packed, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->enc_copied, hipEventDisableTiming) != hipSuccess)
        return bail(fail(GLX_EHIP, "hipEventCreate failed"));
    for (uint32_t k = 0; k < kMaxShards; k++)
        if (hipEventCreateWithFlags(&c->shard_done[k], hipEventDisableTiming) != hipSuccess)
            return bail(fail(GLX_EHIP, "shard event creation failed"));
    c->pb.resize(p->n_planet_buffers < 2 ? 2 : p->n_planet_buffers);
    if ((rc = alloc_set(c, &c->cur)) || (rc = alloc_set(c, &c->tmp))) return bail(rc);
    if (c->keep_rb && ((rc = alloc_rb(c, &c->rb)) || (rc = alloc_rb(c, &c->tmp_rb)))) return bail(rc);
    const size_t tiles = glx::part_tiles(c->cap);
    if ((rc = dalloc(c, &c->counts, 2 * tiles)) || (rc = dalloc(c, &c->offsets, 2 * tiles)) ||
        (rc = dalloc(c, &c->totals, 2)) || (rc = dalloc(c, &c->sum_dev, 1 + glx::kSumParts)) ||
        (rc = dalloc(c, &c->dev_aos, 3 * (c->cap > GLX_MAX_BODIES ? c->cap : GLX_MAX_BODIES))))
        return bail(rc);
    if (hipHostMalloc((void**)&c->pin_small, 64, hipHostMallocDefault) != hipSuccess)
        return bail(fail(GLX_ENOMEM, "hipHostMalloc failed"));
    {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus < 1)
            return bail(fail(GLX_EHIP, "hipDeviceGetAttribute(multiprocessor count) failed"));
        c->n_cus = (uint32_t)cus;
    }
    const size_t log_doubles = (size_t)c->tb_max * c->nb_max * 3;
    for (auto& b : c->pb) {
        if ((rc = dalloc(c, &b.dev_log, log_doubles))) return bail(rc);
        if (hipHostMalloc((void**)&b.pin_log, log_doubles * sizeof(double), hipHostMallocDefault) != hipSuccess)
            return bail(fail(GLX_ENOMEM, "hipHostMalloc(planet staging) failed"));
        if (hipEventCreateWithFlags(&b.uploaded, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b.consumed, hipEventDisableTiming) != hipSuccess)
            return bail(fail(GLX_EHIP, "hipEventCreate failed"));
        (void)hipEventRecord(b.uploaded, c->htd_s);
        (void)hipEventRecord(b.consumed, c->main_s);
    }
    for (hipEvent_t* ev : {&c->k0, &c->k1, &c->q0, &c->q1})
        if (hipEventCreate(ev) != hipSuccess) return bail(fail(GLX_EHIP, "hipEventCreate failed"));
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(GLX_EHIP, "hipDeviceSynchronize failed"));
    *out = c;
    g_err.clear();
    return GLX_OK;
}

int glx_destroy(glx_ctx* c) {
    release(c);
    return GLX_OK;
}

int glx_upload_particles(glx_ctx* c, uint64_t begin, uint64_t n, const double* r, const double* v, const double* a,
                         const uint16_t* flags, const uint32_t* id, const uint32_t* dti) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (int rc = check_range(c, begin, n)) return rc;
    if (n == 0) return GLX_OK;
    if (!r || !v || !flags || !id) return fail(GLX_EINVAL, "glx_upload_particles: r, v, flags, id are required");
    GLX_HIP(hipSetDevice(c->device));
    int rc;
    if ((rc = put_vec3(c, r, c->cur.d, begin, n)) || (rc = put_vec3(c, v, c->cur.d + 3, begin, n))) return rc;
    if (a && (rc = put_vec3(c, a, c->cur.d + 6, begin, n))) return rc;
    if ((rc = put_plain(c, flags, c->cur.flags + begin, n)) || (rc = put_plain(c, id, c->cur.id + begin, n)))
        return rc;
    if (dti && (rc = put_plain(c, dti, c->cur.dti + begin, n))) return rc;
    return GLX_OK;
}

int glx_download_particles(glx_ctx* c, uint64_t begin, uint64_t n, double* r, double* v, double* a,
                           uint16_t* flags, uint32_t* id, uint32_t* dti) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (int rc = check_range(c, begin, n)) return rc;
    if (n == 0) return GLX_OK;
    GLX_HIP(hipSetDevice(c->device));
    int rc;
    if (r && (rc = get_vec3(c, r, c->cur.d, begin, n))) return rc;
    if (v && (rc = get_vec3(c, v, c->cur.d + 3, begin, n))) return rc;
    if (a && (rc = get_vec3(c, a, c->cur.d + 6, begin, n))) return rc;
    if (flags && (rc = get_plain(c, flags, c->cur.flags + begin, n))) return rc;
    if (id && (rc = get_plain(c, id, c->cur.id + begin, n))) return rc;
    if (dti && (rc = get_plain(c, dti, c->cur.dti + begin, n))) return rc;
    return GLX_OK;
}

int glx_set_counts(glx_ctx* c, uint64_t n_alive, uint64_t n_total) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (n_total > c->cap || n_alive > n_total)
        return fail(GLX_EINVAL, "counts n_alive=%llu n_total=%llu invalid for capacity %llu",
                    (unsigned long long)n_alive, (unsigned long long)n_total, (unsigned long long)c->cap);
    c->n_alive = n_alive;
    c->n_total = n_total;
    return GLX_OK;
}

int glx_get_counts(glx_ctx* c, uint64_t* n_alive, uint64_t* n_total) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (n_alive) *n_alive = c->n_alive;
    if (n_total) *n_total = c->n_total;
    return GLX_OK;
}

int glx_upload_planet_block(glx_ctx* c, int buf, uint32_t n_bodies, const double* m, const uint32_t* id,
                            const double* rh, const double* r_log, const double* h0_log, uint32_t n_steps) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (buf < 0 || buf >= (int)c->pb.size())
        return fail(GLX_EINVAL, "planet buffer %d outside [0, %zu)", buf, c->pb.size());
    if (n_bodies < 1 || n_bodies > c->nb_max)
        return fail(GLX_EINVAL, "n_bodies %u outside [1, %u]", n_bodies, c->nb_max);
    if (n_steps < 1 || n_steps > c->tb_max) return fail(GLX_EINVAL, "n_steps %u outside [1, %u]", n_steps, c->tb_max);
    if (!m || !rh || !h0_log || (n_bodies > 1 && !r_log)) return fail(GLX_EINVAL, "null planet array");
    if (!(m[0] > 0.0)) return fail(GLX_EINVAL, "central mass m[0] must be positive");
    GLX_HIP(hipSetDevice(c->device));
    PlanetBuf& b = c->pb[buf];
    // The pinned staging may still feed this buffer's previous copy.
    GLX_HIP(hipEventSynchronize(b.uploaded));
    const uint32_t row = n_bodies * 3, prow = (n_bodies - 1) * 3;
    for (uint32_t s = 0; s < n_steps; s++) {
        double* dst = b.pin_log + (size_t)s * row;
        dst[0] = h0_log[3 * (size_t)s];
        dst[1] = h0_log[3 * (size_t)s + 1];
        dst[2] = h0_log[3 * (size_t)s + 2];
        if (prow) std::memcpy(dst + 3, r_log + (size_t)s * prow, prow * sizeof(double));
    }
    // Do not overwrite the device buffer while a kernel still reads it.
    GLX_HIP(hipStreamWaitEvent(c->htd_s, b.consumed, 0));
    GLX_HIP(hipMemcpyAsync(b.dev_log, b.pin_log, (size_t)n_steps * row * sizeof(double), hipMemcpyHostToDevice,
                           c->htd_s));
    GLX_HIP(hipEventRecord(b.uploaded, c->htd_s));
    fill_consts(n_bodies, m, id, rh, &b.pc);
    b.mu = m[0];
    b.n_bodies = n_bodies;
    b.n_steps = n_steps;
    b.valid = true;
    return GLX_OK;
}

}  // extern "C"

namespace {

int check_form(int form) {
    if (form != GLX_FORM_FP64 && form != GLX_FORM_FAST && form != GLX_FORM_MIXED)
        return fail(GLX_EINVAL, "kernel form %d is not available", form);
    return GLX_OK;
}

// Kernel arguments for one block from planet buffer b; `lds_budget` bytes
// of planet rows are staged per workgroup.
BlockArgs block_args(glx_ctx* c, const PlanetBuf& b, double dt, double outer_radius, size_t lds_budget) {
    BlockArgs A = base_args(c);
    A.log = b.dev_log;
    A.first = 0;
    A.n = (uint32_t)c->n_alive;
    A.step_offset = 0;
    A.n_steps = b.n_steps;
    A.n_bodies = b.n_bodies;
    const size_t row_bytes = (size_t)b.n_bodies * 3 * sizeof(double);
    uint32_t chunk = (uint32_t)(lds_budget / row_bytes);
    if (chunk < 1) chunk = 1;
    if (chunk > b.n_steps) chunk = b.n_steps;
    A.chunk = chunk;
    A.dt = dt;
    A.half_dt = dt / 2;
    A.mu = b.mu;
    A.inv_mu = 1.0 / b.mu;
    A.dt_sqrt_mu = dt * std::sqrt(b.mu);
    A.inv_sqrt_mu = 1.0 / std::sqrt(b.mu);
    A.inv_dt = 1.0 / dt;
    A.outer2 = outer_radius * outer_radius;
    A.fp64_core = glx::fp64_core_constants(b.mu, dt) ? 1u : 0u;
    A.sun_vote = std::sqrt(b.pc.rh2[0]) * (1.0 + glx::kVoteMargin);
    A.outer_vote = outer_radius * (1.0 - glx::kVoteMargin);
    A.block_step0 = 0;
    A.pc = b.pc;
    return A;
}

int check_buf(glx_ctx* c, int buf) {
    if (buf < 0 || buf >= (int)c->pb.size())
        return fail(GLX_EINVAL, "planet buffer %d outside [0, %zu)", buf, c->pb.size());
    if (!c->pb[buf].valid) return fail(GLX_ESTATE, "planet buffer %d holds no block", buf);
    return GLX_OK;
}

int queue_run(glx_ctx* c, uint32_t n_blocks, int first_buf, double dt, double outer_radius, int form,
              uint32_t segment_steps);

}  // namespace

extern "C" {

int glx_run_block(glx_ctx* c, int buf, double dt, double outer_radius, int form) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (int rc = check_buf(c, buf)) return rc;
    if (int rc = check_form(form)) return rc;
    if (!(dt > 0.0)) return fail(GLX_EINVAL, "dt must be positive");
    GLX_HIP(hipSetDevice(c->device));
    // With more 256-particle tiles than compute units, the launch's last
    // workgroups would leave SIMDs on one wave at its end (the oldest wave
    // issues first): the queued launch of this one block balances them and
    // gives the same bits.
    if ((c->n_alive + 255) / 256 > c->n_cus) return queue_run(c, 1, buf, dt, outer_radius, form, kQueueSegment);
    PlanetBuf& b = c->pb[buf];
    const BlockArgs A = block_args(c, b, dt, outer_radius, kLdsBudget);
    GLX_HIP(hipStreamWaitEvent(c->main_s, b.uploaded, 0));
    GLX_HIP(hipEventRecord(c->k0, c->main_s));
    GLX_HIP(glx::launch_block(A, form, 0, c->main_s));
    GLX_HIP(hipEventRecord(c->k1, c->main_s));
    GLX_HIP(hipEventRecord(b.consumed, c->main_s));
    c->kernel_timed = true;
    return GLX_OK;
}

namespace {

// Creates the shard streams [0, n) not created yet, on the first launch
// that uses them.  The shards of a launch only overlap if their streams sit
// on different hardware queues: plain streams share the process's small
// queue pool (GPU_MAX_HW_QUEUES, 4 by default) by usage count, so two
// shards can land on one queue, where their launches run one after the
// other (measured on the 200k fp64 workload: 15.0 instead of 12.8 ms per
// block).  A stream with a CU mask gets a hardware queue of its own; the
// mask here holds every CU, so nothing else changes.  Plain streams are
// the fallback where the masked form is refused.
int ensure_shards(glx_ctx* c, uint32_t n) {
    std::vector<uint32_t> mask;
    for (uint32_t k = 0; k < n; k++) {
        if (c->shard_s[k]) continue;
        if (mask.empty()) {
            hipDeviceProp_t prop;
            GLX_HIP(hipGetDeviceProperties(&prop, c->device));
            const uint32_t cus = (uint32_t)prop.multiProcessorCount;
            mask.assign((cus + 31) / 32, 0xFFFFFFFFu);
            if (cus % 32) mask.back() = (1u << (cus % 32)) - 1;
        }
        if (hipExtStreamCreateWithCUMask(&c->shard_s[k], (uint32_t)mask.size(), mask.data()) != hipSuccess) {
            (void)hipGetLastError();
            c->shard_s[k] = nullptr;
            if (hipStreamCreate(&c->shard_s[k]) != hipSuccess) return fail(GLX_EHIP, "shard stream creation failed");
        }
    }
    return GLX_OK;
}

}  // namespace

int glx_run_blocks(glx_ctx* c, uint32_t n_blocks, int first_buf, double dt, double outer_radius, int form,
                   uint32_t n_shards, uint32_t waves_per_simd, uint32_t launches_per_block) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (int rc = check_form(form)) return rc;
    if (!(dt > 0.0)) return fail(GLX_EINVAL, "dt must be positive");
    const int nbuf = (int)c->pb.size();
    if (n_blocks < 1 || (int)n_blocks > nbuf)
        return fail(GLX_EINVAL, "n_blocks %u outside [1, %d planet buffers]", n_blocks, nbuf);
    if (n_shards < 1 || n_shards > kMaxShards) return fail(GLX_EINVAL, "n_shards %u outside [1, %d]", n_shards, kMaxShards);
    if (waves_per_simd != 0 && (waves_per_simd < 3 || waves_per_simd > 8))
        return fail(GLX_EINVAL, "waves_per_simd %u: use 0 (no cap) or 3..8", waves_per_simd);
    if (launches_per_block == 0) launches_per_block = 1;
    if (launches_per_block > 64) return fail(GLX_EINVAL, "launches_per_block %u above 64", launches_per_block);
    for (uint32_t k = 0; k < n_blocks; k++) {
        const int buf = (first_buf + (int)k) % nbuf;
        if (int rc = check_buf(c, buf)) return rc;
        if (c->pb[buf].n_bodies != c->pb[first_buf % nbuf].n_bodies)
            return fail(GLX_EINVAL, "planet buffers of one launch must hold the same bodies");
    }
    GLX_HIP(hipSetDevice(c->device));
    if (int rc = ensure_shards(c, n_shards)) return rc;
    // One workgroup holds one wave per SIMD, so asking for 1/waves of the
    // CU's LDS caps the resident waves per SIMD; freed slots then go to
    // whichever shard's next block is ready.
    const size_t lds_request = waves_per_simd ? (kLdsPerCu / waves_per_simd) & ~(size_t)1023 : 0;
    const size_t budget = lds_request ? lds_request : kLdsBudget;
    const uint32_t n = (uint32_t)c->n_alive;
    uint32_t per = (n + n_shards - 1) / n_shards;
    per = (per + 255) / 256 * 256;
    GLX_HIP(hipEventRecord(c->k0, c->main_s));
    for (uint32_t s = 0; s < n_shards; s++) GLX_HIP(hipStreamWaitEvent(c->shard_s[s], c->k0, 0));
    uint32_t offset = 0;
    for (uint32_t k = 0; k < n_blocks; k++) {
        PlanetBuf& b = c->pb[(first_buf + (int)k) % nbuf];
        const BlockArgs whole = block_args(c, b, dt, outer_radius, budget);
        const size_t row = (size_t)b.n_bodies * 3;
        // The block's rows run as launches_per_block consecutive launches per
        // slice (the same steps, so the same results): the launch that ends
        // the chain is that much shorter, and so is the time the last
        // partly filled round of workgroups leaves the other SIMDs idle.
        // Launches start on multiples of kRadiusPeriod steps (the carried
        // radius is recomputed there anyway), so at most one per period.
        const uint32_t period = glx::kRadiusPeriod;
        const uint32_t segs = (b.n_steps + period - 1) / period;
        const uint32_t parts = std::min(launches_per_block, segs);
        for (uint32_t j = 0; j < parts; j++) {
            const uint32_t r0 = std::min(b.n_steps, (uint32_t)((uint64_t)segs * j / parts) * period);
            const uint32_t r1 = std::min(b.n_steps, (uint32_t)((uint64_t)segs * (j + 1) / parts) * period);
            BlockArgs A = whole;
            A.log = b.dev_log + r0 * row;
            A.n_steps = r1 - r0;
            A.chunk = std::min(A.chunk, A.n_steps);
            A.step_offset = offset + r0;
            A.block_step0 = r0;
            for (uint32_t s = 0; s < n_shards; s++) {
                A.first = s * per;
                A.n = std::min<uint32_t>(n, (s + 1) * per);
                if (A.first >= A.n) continue;
                if (j == 0) GLX_HIP(hipStreamWaitEvent(c->shard_s[s], b.uploaded, 0));
                GLX_HIP(glx::launch_block(A, form, lds_request, c->shard_s[s]));
            }
        }
        offset += b.n_steps;
    }
    for (uint32_t s = 0; s < n_shards; s++) {
        GLX_HIP(hipEventRecord(c->shard_done[s], c->shard_s[s]));
        GLX_HIP(hipStreamWaitEvent(c->main_s, c->shard_done[s], 0));
    }
    GLX_HIP(hipEventRecord(c->k1, c->main_s));
    for (uint32_t k = 0; k < n_blocks; k++) GLX_HIP(hipEventRecord(c->pb[(first_buf + (int)k) % nbuf].consumed, c->main_s));
    c->kernel_timed = true;
    return GLX_OK;
}

}  // extern "C"

namespace {

// Segment-table capacity of a queued launch: every planet buffer in one
// launch, cut into segments of at least kRadiusPeriod steps.
size_t queue_seg_cap(const glx_ctx* c) {
    return c->pb.size() * (size_t)((c->tb_max + glx::kRadiusPeriod - 1) / glx::kRadiusPeriod);
}

int ensure_queue(glx_ctx* c) {
    if (c->q_segs) return GLX_OK;
    const size_t segs = queue_seg_cap(c), blocks = c->pb.size();
    const size_t tiles = (c->cap + 255) / 256 + 1;
    int rc;
    if ((rc = dalloc(c, &c->q_segs, segs)) || (rc = dalloc(c, &c->q_blocks, blocks)) ||
        (rc = dalloc(c, &c->q_ctl, 3)) || (rc = dalloc(c, &c->q_ring, tiles * segs)))
        return rc;
    const size_t pin = segs * sizeof(glx::QueueSeg) + blocks * sizeof(glx::QueueBlock);
    if (hipHostMalloc((void**)&c->q_pin, pin, hipHostMallocDefault) != hipSuccess) {
        c->q_pin = nullptr;
        return fail(GLX_ENOMEM, "hipHostMalloc(queue tables) failed");
    }
    if (hipEventCreateWithFlags(&c->q_staged, hipEventDisableTiming) != hipSuccess) {
        c->q_staged = nullptr;
        return fail(GLX_EHIP, "hipEventCreate failed");
    }
    return GLX_OK;
}

// After the main stream has been synchronized: the error word of the last
// queued launch, if one is outstanding.
int check_queue(glx_ctx* c) {
    if (!c->q_pending) return GLX_OK;
    c->q_pending = false;
    if (c->pin_small[kQueueErrWord] != 0)
        return fail(GLX_EHIP, "queued launch: a task-queue slot was not filled within the wait bound; "
                              "the particles' state is incomplete");
    return GLX_OK;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// One queued launch of n_blocks planet buffers from first_buf (arguments
// checked by the caller: same bodies and central mass, segment_steps a
// multiple of kRadiusPeriod).
int queue_run(glx_ctx* c, uint32_t n_blocks, int first_buf, double dt, double outer_radius, int form,
              uint32_t segment_steps) {
    const int nbuf = (int)c->pb.size();
    const PlanetBuf& b0 = c->pb[first_buf % nbuf];
    GLX_HIP(hipSetDevice(c->device));
    if (int rc = ensure_queue(c)) return rc;
    // The staging may still feed the previous launch's table copy.
    GLX_HIP(hipEventSynchronize(c->q_staged));
    glx::QueueSeg* segs = reinterpret_cast<glx::QueueSeg*>(c->q_pin);
    glx::QueueBlock* blocks = reinterpret_cast<glx::QueueBlock*>(c->q_pin + queue_seg_cap(c) * sizeof(glx::QueueSeg));
    uint32_t n_segs = 0, offset = 0;
    for (uint32_t k = 0; k < n_blocks; k++) {
        const PlanetBuf& b = c->pb[(first_buf + (int)k) % nbuf];
        blocks[k] = glx::QueueBlock{};
        blocks[k].log = b.dev_log;
        blocks[k].step_offset = offset;
        blocks[k].sun_vote = std::sqrt(b.pc.rh2[0]) * (1.0 + glx::kVoteMargin);
        blocks[k].pc = b.pc;
        for (uint32_t r0 = 0; r0 < b.n_steps; r0 += segment_steps)
            segs[n_segs++] = glx::QueueSeg{k, r0, std::min(segment_steps, b.n_steps - r0), 0};
        offset += b.n_steps;
    }
    const uint32_t n_tiles = (uint32_t)((c->n_alive + 255) / 256);
    if ((uint64_t)n_segs * n_tiles + n_tiles >= 0xFFFFFFFFull)
        return fail(GLX_EINVAL, "queued launch of %u segments x %u tiles: too many tasks for one launch", n_segs,
                    n_tiles);
    glx::QueueArgs Q{};
    Q.base = block_args(c, b0, dt, outer_radius, kLdsBudget);
    Q.base.chunk = std::min<uint32_t>(Q.base.chunk, segment_steps);
    Q.segs = c->q_segs;
    Q.blocks = c->q_blocks;
    Q.ctl = c->q_ctl;
    Q.ring = c->q_ring;
    Q.n_segs = n_segs;
    Q.n_tiles = n_tiles;
    Q.wait_ticks = kQueueWaitTicks;
    const size_t seg_bytes = (size_t)n_segs * sizeof(glx::QueueSeg);
    const size_t blk_bytes = (size_t)n_blocks * sizeof(glx::QueueBlock);
    GLX_HIP(hipMemcpyAsync(c->q_segs, segs, seg_bytes, hipMemcpyHostToDevice, c->main_s));
    GLX_HIP(hipMemcpyAsync(c->q_blocks, blocks, blk_bytes, hipMemcpyHostToDevice, c->main_s));
    GLX_HIP(hipEventRecord(c->q_staged, c->main_s));
    for (uint32_t k = 0; k < n_blocks; k++)
        GLX_HIP(hipStreamWaitEvent(c->main_s, c->pb[(first_buf + (int)k) % nbuf].uploaded, 0));
    GLX_HIP(hipEventRecord(c->k0, c->main_s));
    GLX_HIP(glx::launch_queue(Q, form, 0, c->main_s));
    GLX_HIP(hipEventRecord(c->k1, c->main_s));
    c->pin_small[kQueueErrWord] = 0;
    if (Q.n_tiles > 0) {
        GLX_HIP(hipMemcpyAsync(c->pin_small + kQueueErrWord, c->q_ctl + 2, sizeof(uint32_t), hipMemcpyDeviceToHost,
                               c->main_s));
        c->q_pending = true;
    }
    for (uint32_t k = 0; k < n_blocks; k++) GLX_HIP(hipEventRecord(c->pb[(first_buf + (int)k) % nbuf].consumed, c->main_s));
    c->kernel_timed = true;
    return GLX_OK;
}

}  // namespace

extern "C" {

int glx_run_blocks_queued(glx_ctx* c, uint32_t n_blocks, int first_buf, double dt, double outer_radius, int form,
                          uint32_t segment_steps) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (int rc = check_form(form)) return rc;
    if (!(dt > 0.0)) return fail(GLX_EINVAL, "dt must be positive");
    const int nbuf = (int)c->pb.size();
    if (n_blocks < 1 || (int)n_blocks > nbuf)
        return fail(GLX_EINVAL, "n_blocks %u outside [1, %d planet buffers]", n_blocks, nbuf);
    if (segment_steps == 0) segment_steps = kQueueSegment;
    if (segment_steps % glx::kRadiusPeriod != 0 || segment_steps > kQueueSegmentMax)
        return fail(GLX_EINVAL, "segment_steps %u is not a multiple of %u up to %u", segment_steps,
                    glx::kRadiusPeriod, kQueueSegmentMax);
    if (first_buf < 0) return fail(GLX_EINVAL, "first_buf %d is negative", first_buf);
    const PlanetBuf& b0 = c->pb[first_buf % nbuf];
    for (uint32_t k = 0; k < n_blocks; k++) {
        const int buf = (first_buf + (int)k) % nbuf;
        if (int rc = check_buf(c, buf)) return rc;
        if (c->pb[buf].n_bodies != b0.n_bodies)
            return fail(GLX_EINVAL, "planet buffers of one launch must hold the same bodies");
        if (c->pb[buf].mu != b0.mu)
            return fail(GLX_EINVAL, "planet buffers of one launch must hold the same central mass");
    }
    return queue_run(c, n_blocks, first_buf, dt, outer_radius, form, segment_steps);
}

int glx_upload_history(glx_ctx* c, uint32_t n_bodies, const double* m, const uint32_t* id, uint64_t n_samples,
                       const double* t, const double* r, const double* v) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (n_bodies < 1 || n_bodies > c->nb_max)
        return fail(GLX_EINVAL, "n_bodies %u outside [1, %u]", n_bodies, c->nb_max);
    if (n_samples < 2) return fail(GLX_EINVAL, "a history needs at least 2 samples");
    if (!m || !t || (n_bodies > 1 && (!r || !v))) return fail(GLX_EINVAL, "null history array");
    if (!(m[0] > 0.0)) return fail(GLX_EINVAL, "central mass m[0] must be positive");
    for (uint64_t k = 0; k + 1 < n_samples; k++)
        if (!(t[k + 1] > t[k]))
            return fail(GLX_EINVAL, "history times must increase strictly (sample %llu)", (unsigned long long)k + 1);
    GLX_HIP(hipSetDevice(c->device));
    // Kernels queued earlier may still read the old history.
    GLX_HIP(hipStreamSynchronize(c->main_s));
    GLX_HIP(hipStreamSynchronize(c->htd_s));
    History& H = c->hist;
    H.valid = false;
    for (double** p : {&H.t, &H.r, &H.v}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    const size_t pv = (size_t)n_samples * (n_bodies - 1) * 3;
    const size_t pv_alloc = pv ? pv : 1;
    if (hipMalloc((void**)&H.t, n_samples * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&H.r, pv_alloc * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&H.v, pv_alloc * sizeof(double)) != hipSuccess)
        return fail(GLX_ENOMEM, "hipMalloc of a %llu-sample history failed", (unsigned long long)n_samples);
    if (!H.q_times) {
        if (hipMalloc((void**)&H.q_times, kQueryMax * sizeof(double)) != hipSuccess ||
            hipMalloc((void**)&H.q_rows, (size_t)kQueryMax * GLX_MAX_BODIES * 3 * sizeof(double)) != hipSuccess)
            return fail(GLX_ENOMEM, "hipMalloc of the interpolation query space failed");
    }
    GLX_HIP(hipMemcpy(H.t, t, n_samples * sizeof(double), hipMemcpyHostToDevice));
    if (pv) {
        GLX_HIP(hipMemcpy(H.r, r, pv * sizeof(double), hipMemcpyHostToDevice));
        GLX_HIP(hipMemcpy(H.v, v, pv * sizeof(double), hipMemcpyHostToDevice));
    }
    H.n_samples = n_samples;
    H.n_bodies = n_bodies;
    std::memset(H.m, 0, sizeof H.m);
    for (uint32_t j = 0; j < n_bodies; j++) {
        H.m[j] = m[j];
        H.id[j] = id ? id[j] : j;
    }
    H.valid = true;
    c->hist_t0 = t[0];
    c->hist_t1 = t[n_samples - 1];
    return GLX_OK;
}

}  // extern "C"

namespace {

glx::HistoryView history_view(const glx_ctx* c) {
    glx::HistoryView V{};
    V.t = c->hist.t;
    V.r = c->hist.r;
    V.v = c->hist.v;
    V.n_samples = c->hist.n_samples;
    V.n_bodies = c->hist.n_bodies;
    std::memcpy(V.m, c->hist.m, sizeof V.m);
    return V;
}

// Tolerance on leaving the sampled range, as the reference's interpolator
// allows (src/interp.cpp:131).
constexpr double kHistSlack = 1e-8;

}  // namespace

extern "C" {

int glx_interp_block(glx_ctx* c, int buf, double t_start, double dt, uint32_t n_steps, const double* rh) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (!c->hist.valid) return fail(GLX_ESTATE, "no planet history uploaded");
    if (buf < 0 || buf >= (int)c->pb.size())
        return fail(GLX_EINVAL, "planet buffer %d outside [0, %zu)", buf, c->pb.size());
    if (n_steps < 1 || n_steps > c->tb_max) return fail(GLX_EINVAL, "n_steps %u outside [1, %u]", n_steps, c->tb_max);
    if (!(dt > 0.0)) return fail(GLX_EINVAL, "dt must be positive");
    if (!rh) return fail(GLX_EINVAL, "null rh");
    const double first = t_start + dt, last = t_start + (double)n_steps * dt;
    if (!(first >= c->hist_t0 - kHistSlack) || !(last <= c->hist_t1 + kHistSlack))
        return fail(GLX_EINVAL, "block times [%.17g, %.17g] leave the history [%.17g, %.17g]", first, last,
                    c->hist_t0, c->hist_t1);
    GLX_HIP(hipSetDevice(c->device));
    PlanetBuf& b = c->pb[buf];
    GLX_HIP(hipStreamWaitEvent(c->htd_s, b.consumed, 0));
    GLX_HIP(glx::launch_hermite_rows(history_view(c), nullptr, t_start, dt, n_steps, b.dev_log, c->htd_s));
    GLX_HIP(hipEventRecord(b.uploaded, c->htd_s));
    fill_consts(c->hist.n_bodies, c->hist.m, c->hist.id, rh, &b.pc);
    b.mu = c->hist.m[0];
    b.n_bodies = c->hist.n_bodies;
    b.n_steps = n_steps;
    b.valid = true;
    return GLX_OK;
}

int glx_interp_blocks(glx_ctx* c, int first_buf, uint32_t n_blocks, const double* t_start, double dt, uint32_t n_steps,
                      const double* rh) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (!c->hist.valid) return fail(GLX_ESTATE, "no planet history uploaded");
    const int nbuf = (int)c->pb.size();
    if (first_buf < 0) return fail(GLX_EINVAL, "first_buf %d is negative", first_buf);
    if (n_blocks < 1 || (int)n_blocks > nbuf)
        return fail(GLX_EINVAL, "n_blocks %u outside [1, %d planet buffers]", n_blocks, nbuf);
    if (n_steps < 1 || n_steps > c->tb_max) return fail(GLX_EINVAL, "n_steps %u outside [1, %u]", n_steps, c->tb_max);
    if (!(dt > 0.0)) return fail(GLX_EINVAL, "dt must be positive");
    if (!rh || !t_start) return fail(GLX_EINVAL, "null rh or t_start");
    for (uint32_t k = 0; k < n_blocks; k++) {
        const double first = t_start[k] + dt, last = t_start[k] + (double)n_steps * dt;
        if (!(first >= c->hist_t0 - kHistSlack) || !(last <= c->hist_t1 + kHistSlack))
            return fail(GLX_EINVAL, "block %u times [%.17g, %.17g] leave the history [%.17g, %.17g]", k, first, last,
                        c->hist_t0, c->hist_t1);
    }
    GLX_HIP(hipSetDevice(c->device));
    const glx::HistoryView H = history_view(c);
    for (uint32_t k0 = 0; k0 < n_blocks; k0 += glx::kRowBlocksMax) {
        const uint32_t nk = std::min<uint32_t>(glx::kRowBlocksMax, n_blocks - k0);
        glx::RowBlocks R{};
        for (uint32_t k = 0; k < nk; k++) {
            PlanetBuf& b = c->pb[(first_buf + (int)(k0 + k)) % nbuf];
            GLX_HIP(hipStreamWaitEvent(c->htd_s, b.consumed, 0));
            R.t_start[k] = t_start[k0 + k];
            R.out[k] = b.dev_log;
        }
        GLX_HIP(glx::launch_hermite_rows_multi(H, R, nk, dt, n_steps, c->htd_s));
        for (uint32_t k = 0; k < nk; k++) {
            PlanetBuf& b = c->pb[(first_buf + (int)(k0 + k)) % nbuf];
            GLX_HIP(hipEventRecord(b.uploaded, c->htd_s));
            fill_consts(c->hist.n_bodies, c->hist.m, c->hist.id, rh, &b.pc);
            b.mu = c->hist.m[0];
            b.n_bodies = c->hist.n_bodies;
            b.n_steps = n_steps;
            b.valid = true;
        }
    }
    return GLX_OK;
}

int glx_interp_at(glx_ctx* c, uint32_t n_times, const double* times, double* r, double* h0) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (!c->hist.valid) return fail(GLX_ESTATE, "no planet history uploaded");
    if (n_times == 0) return GLX_OK;
    if (!times) return fail(GLX_EINVAL, "null times");
    for (uint32_t k = 0; k < n_times; k++)
        if (!(times[k] >= c->hist_t0 - kHistSlack) || !(times[k] <= c->hist_t1 + kHistSlack))
            return fail(GLX_EINVAL, "time %.17g leaves the history [%.17g, %.17g]", times[k], c->hist_t0, c->hist_t1);
    GLX_HIP(hipSetDevice(c->device));
    const uint32_t nb = c->hist.n_bodies, np = nb - 1;
    std::vector<double> rows((size_t)kQueryMax * nb * 3);
    for (uint32_t k0 = 0; k0 < n_times; k0 += kQueryMax) {
        const uint32_t cnt = std::min(kQueryMax, n_times - k0);
        GLX_HIP(hipMemcpyAsync(c->hist.q_times, times + k0, cnt * sizeof(double), hipMemcpyHostToDevice, c->htd_s));
        GLX_HIP(glx::launch_hermite_rows(history_view(c), c->hist.q_times, 0.0, 0.0, cnt, c->hist.q_rows, c->htd_s));
        GLX_HIP(hipMemcpyAsync(rows.data(), c->hist.q_rows, (size_t)cnt * nb * 3 * sizeof(double),
                               hipMemcpyDeviceToHost, c->htd_s));
        GLX_HIP(hipStreamSynchronize(c->htd_s));
        for (uint32_t k = 0; k < cnt; k++) {
            const double* row = rows.data() + (size_t)k * nb * 3;
            if (h0) std::memcpy(h0 + (size_t)(k0 + k) * 3, row, 3 * sizeof(double));
            if (r && np) std::memcpy(r + (size_t)(k0 + k) * np * 3, row + 3, (size_t)np * 3 * sizeof(double));
        }
    }
    return GLX_OK;
}

int glx_download_planet_block(glx_ctx* c, int buf, uint32_t n_bodies, uint32_t n_steps, double* r_log,
                              double* h0_log) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (int rc = check_buf(c, buf)) return rc;
    const PlanetBuf& b = c->pb[buf];
    if (n_bodies != b.n_bodies || n_steps != b.n_steps)
        return fail(GLX_EINVAL, "buffer %d holds %u bodies x %u steps, not %u x %u", buf, b.n_bodies, b.n_steps,
                    n_bodies, n_steps);
    GLX_HIP(hipSetDevice(c->device));
    GLX_HIP(hipEventSynchronize(b.uploaded));
    const uint32_t nb = b.n_bodies, np = nb - 1;
    std::vector<double> rows((size_t)b.n_steps * nb * 3);
    GLX_HIP(hipMemcpy(rows.data(), b.dev_log, rows.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (uint32_t s = 0; s < b.n_steps; s++) {
        const double* row = rows.data() + (size_t)s * nb * 3;
        if (h0_log) std::memcpy(h0_log + (size_t)s * 3, row, 3 * sizeof(double));
        if (r_log && np) std::memcpy(r_log + (size_t)s * np * 3, row + 3, (size_t)np * 3 * sizeof(double));
    }
    return GLX_OK;
}

}  // extern "C"

namespace {

// Accelerations and flags of particles [begin, end) from one planet row.
int accel_range(glx_ctx* c, uint64_t begin, uint64_t end, uint32_t n_bodies, const double* m, const uint32_t* id,
                const double* rh, const double* planet_r, const double* h0, double outer_radius) {
    if (n_bodies < 1 || n_bodies > c->nb_max)
        return fail(GLX_EINVAL, "n_bodies %u outside [1, %u]", n_bodies, c->nb_max);
    if (!m || !rh || !h0 || (n_bodies > 1 && !planet_r)) return fail(GLX_EINVAL, "null planet array");
    GLX_HIP(hipSetDevice(c->device));
    // Stage one planet row through planet buffer 0's staging space.
    PlanetBuf& b = c->pb[0];
    GLX_HIP(hipEventSynchronize(b.uploaded));
    GLX_HIP(hipStreamSynchronize(c->main_s));
    b.pin_log[0] = h0[0];
    b.pin_log[1] = h0[1];
    b.pin_log[2] = h0[2];
    if (n_bodies > 1) std::memcpy(b.pin_log + 3, planet_r, (size_t)(n_bodies - 1) * 3 * sizeof(double));
    double* row_dev = c->dev_aos;  // free between copies
    GLX_HIP(hipMemcpyAsync(row_dev, b.pin_log, (size_t)n_bodies * 3 * sizeof(double), hipMemcpyHostToDevice,
                           c->main_s));
    BlockArgs A = base_args(c);
    A.log = row_dev;
    A.first = (uint32_t)begin;
    A.n = (uint32_t)end;
    A.n_steps = 1;
    A.n_bodies = n_bodies;
    A.outer2 = outer_radius * outer_radius;
    fill_consts(n_bodies, m, id, rh, &A.pc);
    GLX_HIP(glx::launch_init_accel(A, c->main_s));
    GLX_HIP(hipStreamSynchronize(c->main_s));
    return GLX_OK;
}

}  // namespace

extern "C" {

int glx_init_accel(glx_ctx* c, uint32_t n_bodies, const double* m, const uint32_t* id, const double* rh,
                   const double* planet_r, const double* h0, double outer_radius) {
    if (!c) return fail(GLX_EINVAL, "null context");
    return accel_range(c, 0, c->n_alive, n_bodies, m, id, rh, planet_r, h0, outer_radius);
}

int glx_accel_range(glx_ctx* c, uint64_t begin, uint64_t n, uint32_t n_bodies, const double* m, const uint32_t* id,
                    const double* rh, const double* planet_r, const double* h0, double outer_radius) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (int rc = check_range(c, begin, n)) return rc;
    return accel_range(c, begin, begin + n, n_bodies, m, id, rh, planet_r, h0, outer_radius);
}

int glx_pack_encounters(glx_ctx* c, uint64_t begin, uint64_t n, double t_block, glx_encounter* out) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (n > 0 && !out) return fail(GLX_EINVAL, "null output");
    const glx_encounter* staged = nullptr;
    if (int rc = glx::pack_encounters_staged(c, begin, n, t_block, &staged)) return rc;
    if (n > 0) std::memcpy(out, staged, n * sizeof(glx_encounter));
    return GLX_OK;
}

int glx_kill_encounters(glx_ctx* c) {
    if (!c) return fail(GLX_EINVAL, "null context");
    GLX_HIP(hipSetDevice(c->device));
    GLX_HIP(glx::launch_kill_encounters(c->cur.flags, (uint32_t)c->n_alive, c->main_s));
    return GLX_OK;
}

int glx_partition(glx_ctx* c, int mode, uint64_t* n_unflagged, uint64_t* n_encounter) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (mode != GLX_PART_UNFLAGGED && mode != GLX_PART_THREE_WAY)
        return fail(GLX_EINVAL, "partition mode %d unknown", mode);
    GLX_HIP(hipSetDevice(c->device));
    const uint32_t n = (uint32_t)c->n_alive;
    uint64_t n0 = 0, n1 = 0;
    GLX_HIP(hipEventRecord(c->q0, c->main_s));
    if (n > 0) {
        GLX_HIP(glx::launch_partition(c->cur, c->tmp, c->keep_rb ? &c->rb : nullptr,
                                      c->keep_rb ? &c->tmp_rb : nullptr, n, mode == GLX_PART_THREE_WAY, c->counts,
                                      c->offsets, c->totals, c->main_s));
        GLX_HIP(hipMemcpyAsync(c->pin_small, c->totals, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->main_s));
    }
    GLX_HIP(hipEventRecord(c->q1, c->main_s));
    GLX_HIP(hipStreamSynchronize(c->main_s));
    if (int rc = check_queue(c)) return rc;
    if (n > 0) {
        n0 = c->pin_small[0];
        n1 = mode == GLX_PART_THREE_WAY ? c->pin_small[1] : 0;
    }
    c->n_alive = n0;
    c->part_timed = true;
    if (n_unflagged) *n_unflagged = n0;
    if (n_encounter) *n_encounter = n1;
    return GLX_OK;
}

int glx_snapshot_rollback(glx_ctx* c) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (!c->keep_rb) return fail(GLX_ESTATE, "context was created without a rollback copy");
    GLX_HIP(hipSetDevice(c->device));
    GLX_HIP(glx::launch_snapshot(c->cur, c->rb, (uint32_t)c->n_total, c->main_s));
    return GLX_OK;
}

int glx_download_rollback(glx_ctx* c, uint64_t begin, uint64_t n, double* r, double* v) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (!c->keep_rb) return fail(GLX_ESTATE, "context was created without a rollback copy");
    if (int rc = check_range(c, begin, n)) return rc;
    if (n == 0) return GLX_OK;
    GLX_HIP(hipSetDevice(c->device));
    int rc;
    if (r && (rc = get_vec3(c, r, c->rb.d, begin, n))) return rc;
    if (v && (rc = get_vec3(c, v, c->rb.d + 3, begin, n))) return rc;
    return GLX_OK;
}

int glx_sync(glx_ctx* c) {
    if (!c) return fail(GLX_EINVAL, "null context");
    GLX_HIP(hipSetDevice(c->device));
    GLX_HIP(hipStreamSynchronize(c->htd_s));
    GLX_HIP(hipStreamSynchronize(c->main_s));
    return check_queue(c);
}

int glx_last_times(glx_ctx* c, float* kernel_ms, float* partition_ms) {
    if (!c) return fail(GLX_EINVAL, "null context");
    GLX_HIP(hipSetDevice(c->device));
    if (kernel_ms) {
        *kernel_ms = 0.f;
        if (c->kernel_timed) {
            GLX_HIP(hipEventSynchronize(c->k1));
            GLX_HIP(hipEventElapsedTime(kernel_ms, c->k0, c->k1));
        }
    }
    if (partition_ms) {
        *partition_ms = 0.f;
        if (c->part_timed) {
            GLX_HIP(hipEventSynchronize(c->q1));
            GLX_HIP(hipEventElapsedTime(partition_ms, c->q0, c->q1));
        }
    }
    return GLX_OK;
}

int glx_sum_steps(glx_ctx* c, uint64_t n, uint64_t* total) {
    if (!c || !total) return fail(GLX_EINVAL, "null argument");
    if (int rc = check_range(c, 0, n)) return rc;
    GLX_HIP(hipSetDevice(c->device));
    GLX_HIP(glx::launch_sum_steps(c->cur.dti, (uint32_t)n, c->sum_dev, c->main_s));
    GLX_HIP(hipMemcpyAsync(c->pin_small + 4, c->sum_dev, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           c->main_s));
    GLX_HIP(hipStreamSynchronize(c->main_s));
    unsigned long long v;
    std::memcpy(&v, c->pin_small + 4, sizeof v);
    *total = v;
    return GLX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- tracks ----

#include "host/formats.h"

namespace {

template <typename T>
int grow(T** dev, uint64_t* cap, uint64_t n) {
    if (n <= *cap) return GLX_OK;
    if (*dev) (void)hipFree(*dev);
    *dev = nullptr;
    *cap = 0;
    if (hipMalloc((void**)dev, std::max<uint64_t>(n, 4096) * sizeof(T)) != hipSuccess)
        return fail(GLX_ENOMEM, "hipMalloc of %llu track entries failed", (unsigned long long)n);
    *cap = std::max<uint64_t>(n, 4096);
    return GLX_OK;
}

}  // namespace

extern "C" {

int glx_track_pack(glx_ctx* c, uint32_t n_planets, const double* pm, const double* pr, const double* pv,
                   int to_elements, int barycentric, void* records, uint64_t capacity, uint64_t* n_out) {
    if (!c || !n_out) return fail(GLX_EINVAL, "null argument");
    const uint64_t n = c->n_alive;
    *n_out = n;
    if (n == 0) return GLX_OK;
    if (!records || capacity < n) return fail(GLX_EINVAL, "record buffer holds %llu, %llu needed",
                                              (unsigned long long)capacity, (unsigned long long)n);
    if (n_planets < 1 || !pm || !pr || !pv) return fail(GLX_EINVAL, "the planets (Sun first) are required");
    glx::host::State pl;
    pl.pm.assign(pm, pm + n_planets);
    pl.pr.assign(pr, pr + 3 * (size_t)n_planets);
    pl.pv.assign(pv, pv + 3 * (size_t)n_planets);
    glx::TrackCenter center{};
    glx::host::track_center(pl, to_elements && barycentric, &center.mass, center.r, center.v);
    GLX_HIP(hipSetDevice(c->device));
    if (int rc = grow(&c->trk_dev, &c->trk_cap, n)) return rc;
    // ids ranked on the device when they are dense (max id < 4 x capacity)
    GLX_HIP(glx::launch_max_id(c->cur.id, (uint32_t)n, c->totals, c->main_s));
    GLX_HIP(hipMemcpyAsync(c->pin_small, c->totals, sizeof(uint32_t), hipMemcpyDeviceToHost, c->main_s));
    GLX_HIP(hipStreamSynchronize(c->main_s));
    const uint32_t max_id = c->pin_small[0];
    const bool dense = (uint64_t)max_id < 4 * c->cap;
    uint32_t* rank = nullptr;
    if (dense) {
        const uint32_t n_ids = max_id + 1;
        if (int rc = grow(&c->rank_dev, &c->rank_cap, n_ids + glx::rank_scratch_tiles(n_ids))) return rc;
        rank = c->rank_dev;
        GLX_HIP(glx::launch_rank_ids(c->cur.id, (uint32_t)n, n_ids, rank, rank + n_ids, c->main_s));
        // unique ids: the largest has rank n - 1
        GLX_HIP(hipMemcpyAsync(c->pin_small, rank + max_id, sizeof(uint32_t), hipMemcpyDeviceToHost, c->main_s));
        GLX_HIP(hipStreamSynchronize(c->main_s));
        if (c->pin_small[0] != n - 1) return fail(GLX_ESTATE, "particle ids are not unique");
    }
    GLX_HIP(glx::launch_pack_track(c->cur, (uint32_t)n, rank, center, to_elements, c->trk_dev, c->main_s));
    GLX_HIP(hipMemcpyAsync(records, c->trk_dev, n * sizeof(glx::TrackRecordDev), hipMemcpyDeviceToHost, c->main_s));
    GLX_HIP(hipStreamSynchronize(c->main_s));
    if (!dense) {
        auto* r = static_cast<glx::TrackRecordDev*>(records);
        std::sort(r, r + n, [](const glx::TrackRecordDev& a, const glx::TrackRecordDev& b) { return a.id < b.id; });
    }
    return GLX_OK;
}

}  // extern "C"

namespace glx {

// Grows the record staging to hold n records (waits for the context's
// queued work first: the buffers may be in use).
int grow_encounter_staging(glx_ctx* c, uint64_t n) {
    if (n <= c->enc_cap) return GLX_OK;
    GLX_HIP(hipStreamSynchronize(c->main_s));
    GLX_HIP(hipStreamSynchronize(c->dth_s));
    if (c->enc_dev) (void)hipFree(c->enc_dev);
    if (c->enc_pin) (void)hipHostFree(c->enc_pin);
    c->enc_dev = nullptr;
    c->enc_pin = nullptr;
    c->enc_cap = 0;
    const uint64_t cap = std::max<uint64_t>(n, 4096);
    if (hipMalloc((void**)&c->enc_dev, cap * sizeof(glx_encounter)) != hipSuccess ||
        hipHostMalloc((void**)&c->enc_pin, cap * sizeof(glx_encounter), hipHostMallocDefault) != hipSuccess)
        return fail(GLX_ENOMEM, "encounter staging of %llu records failed", (unsigned long long)cap);
    c->enc_cap = cap;
    return GLX_OK;
}

int pack_encounters_begin(glx_ctx* c, uint64_t begin, uint64_t n, double t_block) {
    if (!c) return fail(GLX_EINVAL, "null context");
    if (int rc = check_range(c, begin, n)) return rc;
    GLX_HIP(hipSetDevice(c->device));
    if (int rc = grow_encounter_staging(c, n)) return rc;
    if (n > 0) {
        GLX_HIP(glx::launch_pack_encounters(c->cur, c->keep_rb ? &c->rb : nullptr, (uint32_t)begin, (uint32_t)n,
                                            t_block, c->enc_dev, c->main_s));
        GLX_HIP(hipEventRecord(c->enc_packed, c->main_s));
        GLX_HIP(hipStreamWaitEvent(c->dth_s, c->enc_packed, 0));
        GLX_HIP(hipMemcpyAsync(c->enc_pin, c->enc_dev, n * sizeof(glx_encounter), hipMemcpyDeviceToHost,
                               c->dth_s));
    }
    GLX_HIP(hipEventRecord(c->enc_copied, c->dth_s));
    return GLX_OK;
}

int pack_encounters_wait(glx_ctx* c, const glx_encounter** out) {
    if (!c || !out) return fail(GLX_EINVAL, "null argument");
    GLX_HIP(hipSetDevice(c->device));
    GLX_HIP(hipEventSynchronize(c->enc_copied));
    *out = c->enc_pin;
    return GLX_OK;
}

int pack_encounters_staged(glx_ctx* c, uint64_t begin, uint64_t n, double t_block, const glx_encounter** out) {
    if (!out) return fail(GLX_EINVAL, "null output");
    *out = nullptr;
    if (int rc = pack_encounters_begin(c, begin, n, t_block)) return rc;
    if (n == 0) return GLX_OK;
    return pack_encounters_wait(c, out);
}

}  // namespace glx
