/*
 * glisser_amd.h -- C ABI of the MI355X (gfx950) test-particle time-block
 * integrator.
 *
 * The entry points replace the device half of GLISSER's particle path:
 * the per-particle Wisdom-Holman time-block kernel (src/wh.cu), the
 * encounter-kill transform and the partition that follow it in the executor
 * (src/executor.cu), and the host<->device copies that feed them.  Plain
 * pointers and sizes only; every call returns a status (0 = success) and
 * glx_last_error() describes the last failure on the calling thread.  No
 * C++ exception crosses this boundary.
 *
 * Host-side layout of vector arrays is the reference's f64_3 layout
 * (src/types.h:78-213): x, y, z doubles with a 24-byte stride.  On the
 * device the context keeps particles as structure-of-arrays.
 *
 * Threading: one host thread per context.  Calls marked "async" only enqueue
 * work on the context's streams; glx_sync() waits for all of it.
 */
#ifndef GLISSER_AMD_H
#define GLISSER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glx_ctx glx_ctx;

enum glx_status {
    GLX_OK = 0,
    GLX_EINVAL = 1,  /* bad argument (size, index, null pointer) */
    GLX_EHIP = 2,    /* a HIP runtime call failed */
    GLX_ENOMEM = 3,  /* device or pinned allocation failed */
    GLX_ESTATE = 4   /* call not valid in the context's current state */
};

/* Arithmetic forms of the time-block kernel. */
enum glx_form {
    GLX_FORM_FP64 = 0,  /* reference arithmetic: fixed 5 Newton iterations, IEEE div/sqrt */
    GLX_FORM_FAST = 1,  /* FP64 throughout, fewer and cheaper operations (early-exit Newton) */
    GLX_FORM_MIXED = 2  /* FAST drift; planet r^-3 from the FP32 rsqrt, corrected in FP64; FP64 sums */
};

/* Partition modes of glx_partition(). */
enum glx_partition_mode {
    GLX_PART_UNFLAGGED = 0, /* flags == 0 first, everything else after (stable) */
    GLX_PART_THREE_WAY = 1  /* flags == 0, then (flags & 0xFE) == 0, then the rest (stable) */
};

#define GLX_MAX_BODIES 16

typedef struct glx_params {
    uint64_t n_particles;  /* capacity; DeviceParticlePhaseSpace(n) at src/executor.cu:167 */
    uint32_t n_bodies_max; /* Sun + planets, <= GLX_MAX_BODIES */
    uint32_t tbsize_max;   /* Time-Block-Size: rows of one planet log */
    int32_t device;        /* HIP device ordinal */
    int32_t keep_rollback; /* allocate the rollback copy (Resolve-Encounters 1) */
    uint32_t n_planet_buffers; /* planet-block buffers (0 or < 2 means 2: the reference's double buffer) */
} glx_params;

/* One record of the host close-encounter queue (glx_pack_encounters,
 * glx_exec_take_encounters). */
typedef struct glx_encounter {
    uint32_t id;
    uint16_t flags;     /* 0x01, planet id in the high byte (0: the Sun) */
    uint32_t step;      /* steps of the block completed when flagged */
    double t_block;     /* time at the start of that block */
    double r0[3], v0[3];  /* state at the start of the block (rollback copy) */
    double r[3], v[3];    /* state at the flagging step */
} glx_encounter;

/* Create / destroy a context and all its device memory.
 * Replaces the allocations of Executor::init (src/executor.cu:151-175) and
 * WHCudaIntegrator's constructor (src/wh.cu:272-282). */
int glx_create(const glx_params* params, glx_ctx** out);
int glx_destroy(glx_ctx* ctx);

/* Text of the last error on this thread ("" if none). */
const char* glx_last_error(void);

/* Copy particles [begin, begin+n) host -> device (synchronous).
 * r, v, a: n x 3 doubles; flags, id, deathtime_index: n entries.  a and
 * deathtime_index may be NULL (then left unchanged).
 * Replaces Executor::upload_data (src/executor.cu:322-337) together with
 * WHCudaIntegrator::upload_data_cuda (src/wh.cu:338-342). */
int glx_upload_particles(glx_ctx* ctx, uint64_t begin, uint64_t n, const double* r, const double* v,
                         const double* a, const uint16_t* flags, const uint32_t* id,
                         const uint32_t* deathtime_index);

/* Copy particles [begin, begin+n) device -> host (synchronous); any output
 * pointer may be NULL.  Replaces Executor::download_data
 * (src/executor.cu:358-440). */
int glx_download_particles(glx_ctx* ctx, uint64_t begin, uint64_t n, double* r, double* v, double* a,
                           uint16_t* flags, uint32_t* id, uint32_t* deathtime_index);

/* Number of particles the next glx_run_block advances ([0, n_alive)) and the
 * total count held (n_total >= n_alive).  DeviceParticlePhaseSpace::n_alive /
 * n_total (src/data.cuh:8-30). */
int glx_set_counts(glx_ctx* ctx, uint64_t n_alive, uint64_t n_total);
int glx_get_counts(glx_ctx* ctx, uint64_t* n_alive, uint64_t* n_total);

/* Upload one planet block into buffer buf (0 .. n_planet_buffers-1), async on the upload
 * stream.  m, rh: n_bodies doubles (GM units; m[0] is the Sun); id: n_bodies
 * planet ids; r_log: n_steps x (n_bodies-1) x 3 heliocentric planet
 * positions, row s at t0 + (s+1) dt; h0_log: n_steps x 3 indirect terms.
 * The host arrays may be reused as soon as the call returns.
 * Replaces Executor::upload_planet_log (src/executor.cu:442-461) and
 * WHCudaIntegrator::upload_planet_log_cuda (src/wh.cu:306-311). */
int glx_upload_planet_block(glx_ctx* ctx, int buf, uint32_t n_bodies, const double* m, const uint32_t* id,
                            const double* rh, const double* r_log, const double* h0_log, uint32_t n_steps);

/* Upload a sampled planet history, synchronous: n_samples >= 2 times t
 * (strictly increasing) with heliocentric r and v of bodies 1..n_bodies-1
 * (n_samples x (n_bodies-1) x 3 each); m: n_bodies GM values (m[0] is the
 * Sun); id: n_bodies planet ids (NULL means 0..n_bodies-1).  Replaces the
 * history reader of the Interpolator (src/interp.cpp:14-96); the history
 * here is Cartesian and interpolated with cubic Hermite polynomials. */
int glx_upload_history(glx_ctx* ctx, uint32_t n_bodies, const double* m, const uint32_t* id, uint64_t n_samples,
                       const double* t, const double* r, const double* v);

/* Fill planet buffer buf with n_steps rows interpolated from the history at
 * t_start + (s+1) dt, async on the upload stream: no host work and no copy.
 * rh: n_bodies encounter radii.  GLX_EINVAL if a time leaves the history
 * (by more than 1e-8).  Replaces Interpolator::fill, WHIntegrator::load_h0
 * and the planet-log upload (src/interp.cpp:113-162, src/wh.cpp:993-1025,
 * src/executor.cu:442-461). */
int glx_interp_block(glx_ctx* ctx, int buf, double t_start, double dt, uint32_t n_steps, const double* rh);

/* glx_interp_block for n_blocks buffers first_buf, first_buf+1, ... (modulo
 * the buffer count), block k starting at t_start[k], in one launch per 32
 * blocks: the same rows, without a host round of API calls per block ahead
 * of a multi-block run.  No reference counterpart (the reference fills one
 * block per host iteration). */
int glx_interp_blocks(glx_ctx* ctx, int first_buf, uint32_t n_blocks, const double* t_start, double dt,
                      uint32_t n_steps, const double* rh);

/* Interpolated planet positions (n_times x (n_bodies-1) x 3) and h0
 * (n_times x 3) at arbitrary times, synchronous; either output may be NULL.
 * Interpolator::fill_one (src/interp.cpp:99-111) plus load_h0. */
int glx_interp_at(glx_ctx* ctx, uint32_t n_times, const double* times, double* r, double* h0);

/* Copy the rows of planet buffer buf back, synchronous: r_log
 * n_steps x (n_bodies-1) x 3 and h0_log n_steps x 3 (either may be NULL);
 * n_bodies and n_steps must be those of the block the buffer holds. */
int glx_download_planet_block(glx_ctx* ctx, int buf, uint32_t n_bodies, uint32_t n_steps, double* r_log,
                              double* h0_log);

/* Advance particles [0, n_alive) through every step of planet buffer buf
 * (async on the main stream; waits for that buffer's upload).  With more
 * 256-particle tiles than the device has compute units it runs as the
 * queued launch of glx_run_blocks_queued (one block, 256-step segments;
 * same results), otherwise as one launch.
 * Replaces WHCudaIntegrator::integrate_particles_timeblock_cuda
 * (src/wh.cu:344-395) and the MVSKernel it launches (src/wh.cu:15-207). */
int glx_run_block(glx_ctx* ctx, int buf, double dt, double outer_radius, int form);

/* Advance particles [0, n_alive) through n_blocks consecutive planet blocks
 * held in buffers first_buf, first_buf+1, ... (modulo the buffer count),
 * async.  The particles are split into n_shards slices, each run on its own
 * stream as a chain of launches, with no compaction between blocks (flagged
 * particles simply stop).  waves_per_simd > 0 caps the resident waves per
 * SIMD (3..8) through the LDS request, so that a SIMD freed early picks up
 * the next ready launch of another slice; 0 leaves occupancy uncapped.
 * launches_per_block (1..64, 0 means 1) runs each block's steps as that
 * many consecutive launches, split on multiples of 64 steps (so at most
 * ceil(Time-Block-Size / 64) of them; same results; a shorter last launch
 * shortens the partly idle end of the chain).  deathtime_index becomes the step
 * count within the launch (k * Time-Block-Size + step + 1).  With n_blocks
 * = 1, n_shards = 1 and one launch per block it is glx_run_block.  No
 * reference counterpart: the reference launches one block per host
 * iteration (src/executor.cu:640-646). */
int glx_run_blocks(glx_ctx* ctx, uint32_t n_blocks, int first_buf, double dt, double outer_radius, int form,
                   uint32_t n_shards, uint32_t waves_per_simd, uint32_t launches_per_block);

/* The same integration as glx_run_blocks (same buffers, same results bit
 * for bit, same deathtime_index convention) as ONE persistent launch: the
 * particles are cut into tiles of 256 and every block into segments of
 * segment_steps steps (a multiple of 64 up to 4096; 0 means 256).  The
 * workgroups (at most one per tile) take tasks from a device FIFO that
 * starts with every tile's first segment; finishing a tile's segment
 * appends its next one.  A workgroup starts its next task as soon as it
 * finishes one, so the SIMDs keep their resident waves until the last
 * segments, where a chain of launches idles part of the device at the end
 * of each one.  Async on the main stream (waits for the blocks' uploads).
 * A wait for a FIFO entry is bounded in real time; should it ever expire
 * (it cannot, short of a device fault) the next glx_sync or glx_partition
 * returns GLX_EHIP.  The blocks must hold the same number of bodies and the
 * same central mass.  No reference counterpart (see glx_run_blocks). */
int glx_run_blocks_queued(glx_ctx* ctx, uint32_t n_blocks, int first_buf, double dt, double outer_radius, int form,
                          uint32_t segment_steps);

/* Initial accelerations and flags of particles [0, n_alive) from planet row
 * 0 of `planet_r` ((n_bodies-1) x 3) and h0 (3 doubles), async.  The
 * device-side counterpart of the helio_acc_particles<true> call in the
 * WHIntegrator constructor (src/wh.cpp:531-543). */
int glx_init_accel(glx_ctx* ctx, uint32_t n_bodies, const double* m, const uint32_t* id, const double* rh,
                   const double* planet_r, const double* h0, double outer_radius);

/* glx_init_accel over particles [begin, begin+n) instead of [0, n_alive):
 * accelerations and flags of particles handed back after an encounter was
 * resolved on the host, from the planet row of the current time (the
 * helio_acc_particles call of Executor::handle_encounters,
 * src/executor.cu:558-567), synchronous. */
int glx_accel_range(glx_ctx* ctx, uint64_t begin, uint64_t n, uint32_t n_bodies, const double* m,
                    const uint32_t* id, const double* rh, const double* planet_r, const double* h0,
                    double outer_radius);

/* flags & 0x01 -> (flags & 0xFF00) | 0x80 over [0, n_alive), async.
 * Replaces KillEncounterKernel (src/executor.cu:63-78,656-660). */
int glx_kill_encounters(glx_ctx* ctx);

/* Device arrays of one particle set (structure of arrays: x, y, z
 * components), owned by the caller. */
typedef struct glx_device_particles {
    double* r[3];
    double* v[3];
    double* a[3];
    uint16_t* flags;
    uint32_t* deathtime_index;
    uint32_t* id;
} glx_device_particles;

/* Most particles one glx_partition_static call takes. */
#define GLX_PARTITION_STATIC_MAX (1u << 25)

/* glx_partition without any allocation, for callers that cannot allocate
 * device or host memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.
 * Stable partition of the n <= GLX_PARTITION_STATIC_MAX particles of `in`
 * into `out` (device arrays, distinct from in's) on the current device,
 * async on `stream` (a hipStream_t; NULL: the null stream): the same
 * classes and order as glx_partition (mode as there); class_sizes (device
 * memory, 2 x u32) receives the number of unflagged particles and of the
 * encounter class (THREE_WAY; else 0).  The tile counts live in a fixed
 * 512 KiB device array per GPU, kept until the process ends: calls that
 * overlap on one GPU race for it and must be ordered by the caller (one
 * stream, or events).  GLX_EINVAL for larger n. */
int glx_partition_static(const glx_device_particles* in, const glx_device_particles* out, uint64_t n, int mode,
                         uint32_t* class_sizes, void* stream);

/* Stable partition of [0, n_alive) moving every particle array (and the
 * rollback copy when kept); synchronous.  *n_unflagged receives the new
 * n_alive; *n_encounter the size of the encounter class (THREE_WAY) or 0.
 * Replaces the thrust::partition calls of src/executor.cu:663-665,780-797
 * with an order-preserving (std::stable_partition-equal) compaction. */
int glx_partition(glx_ctx* ctx, int mode, uint64_t* n_unflagged, uint64_t* n_encounter);

/* Copy r, v, flags, deathtime_index and id of [0, n_total) into the rollback
 * copy, async.  Replaces the memcpy_dtd block of src/executor.cu:624-635. */
int glx_snapshot_rollback(glx_ctx* ctx);

/* Particles [begin, begin+n) as encounter records, synchronous: packed on
 * the device (state at the block start from the rollback copy when one is
 * kept, else the current state; current state; flags; id; deathtime index
 * as `step`), one copy to the host.  The download half of resync2 for the
 * encounter range (src/executor.cu:812-849). */
int glx_pack_encounters(glx_ctx* ctx, uint64_t begin, uint64_t n, double t_block, glx_encounter* out);

/* Download the rollback copy of [begin, begin+n) (r, v may be NULL). */
int glx_download_rollback(glx_ctx* ctx, uint64_t begin, uint64_t n, double* r, double* v);

/* Wait for all work queued on the context. */
int glx_sync(glx_ctx* ctx);

/* Device time of the last glx_run_block kernel and of the last
 * glx_partition, in milliseconds (HIP events on the launching stream);
 * the cudaEventElapsedTime pairs of src/executor.cu:637-648,776-806. */
int glx_last_times(glx_ctx* ctx, float* kernel_ms, float* partition_ms);

/* Sum of deathtime_index over [0, n) -- the particle-steps done by the last
 * block -- computed on the device (synchronous). */
int glx_sum_steps(glx_ctx* ctx, uint64_t n, uint64_t* total);

/* sin and cos of n doubles as the GLX_FORM_FP64 Kepler drift evaluates
 * them: correctly rounded (round to nearest) for |x| < 2^20, the device math
 * library beyond.  The reference's CPU drift calls glibc sin/cos
 * (src/wh.cpp:46-65, :928), which is correctly rounded on the Kepler
 * solve's arguments in all but a few calls per million.
 * glx_sincos_device runs the device routine on GPU `device` (synchronous;
 * at most 2^28 values); glx_sincos_host runs the same code on the host. */
int glx_sincos_device(int device, uint64_t n, const double* x, double* s, double* c);
int glx_sincos_host(uint64_t n, const double* x, double* s, double* c);

/* ---------------------------------------------------------------------
 * Batched runs: many independent planetary systems (different bodies,
 * particle counts and timesteps) advanced together, one launch per
 * time-block for all of them.  No reference counterpart: the reference
 * integrates one system per process (src/executor.cu:598-764); a batch is
 * what a caller would otherwise run as many single-system executors.
 * ------------------------------------------------------------------- */

typedef struct glx_batch glx_batch;

#define GLX_BATCH_MAX_BODIES 9

/* One system of a batch.  m, rh: n_bodies values (GM units, m[0] the Sun;
 * rh[0] the Sun's cull radius); id: n_bodies planet ids or NULL.  The
 * planets come from a sampled history (n_samples >= 2 strictly increasing
 * times t, heliocentric r and v of bodies 1..n_bodies-1, n_samples x
 * (n_bodies-1) x 3 each) interpolated on the device with cubic Hermite
 * polynomials; step s of the batch (counted from 0) is at t0 + (s+1) dt. */
typedef struct glx_system {
    uint32_t n_bodies;     /* 2 .. GLX_BATCH_MAX_BODIES */
    uint32_t n_particles;
    double t0;
    double dt;
    double outer_radius;
    const double* m;
    const uint32_t* id;
    const double* rh;
    uint64_t n_samples;
    const double* t;
    const double* r;
    const double* v;
} glx_system;

/* Launch shape of a batch: steps per launch (<= the batch's maximum),
 * lanes per workgroup (64, 128 or 256), persistent (0: one workgroup per
 * tile of `lanes` particles; 1: a fixed grid of workgroups takes tiles from
 * a device queue), waves_per_simd (0: no cap; 1..8: resident waves per SIMD,
 * which also sizes the persistent grid). */
typedef struct glx_batch_config {
    uint32_t steps_per_launch;
    uint32_t lanes;
    uint32_t persistent;
    uint32_t waves_per_simd;
} glx_batch_config;

/* Create a batch on device `device` holding the systems (histories and
 * constants are copied; the caller's arrays may be freed after the call).
 * steps_max bounds steps_per_launch.  Particles start zeroed: upload them
 * with glx_batch_upload, then glx_batch_init_accel. */
int glx_batch_create(int device, const glx_system* systems, uint32_t n_systems, uint32_t steps_max,
                     glx_batch** out);
int glx_batch_destroy(glx_batch* b);

/* Particles of system `sys` (its n_particles entries; r, v: n x 3), synchronous.
 * On download any output may be NULL; deathtime_index counts the steps since
 * the batch started (flagged particles keep the step count of their last step).
 * The device keeps a system's particles ordered by orbital frequency (so the
 * lanes of a wave take similar Kepler steps); download returns them in the
 * order of the upload. */
int glx_batch_upload(glx_batch* b, uint32_t sys, const double* r, const double* v, const uint16_t* flags,
                     const uint32_t* id);
int glx_batch_download(glx_batch* b, uint32_t sys, double* r, double* v, uint16_t* flags, uint32_t* id,
                       uint32_t* deathtime_index);

/* Initial accelerations and flags of every system at its t0 (planet rows
 * from the history), synchronous: the batched counterpart of the
 * helio_acc_particles<true> call in the WHIntegrator constructor
 * (src/wh.cpp:531-543). */
int glx_batch_init_accel(glx_batch* b);

/* Set / read the launch shape.  Setting it captures the launch
 * (interpolation, particle kernel, clock) into a HIP graph on the next run. */
int glx_batch_set_config(glx_batch* b, const glx_batch_config* cfg);
int glx_batch_get_config(glx_batch* b, glx_batch_config* cfg);

/* Replay the captured launch n_launches times (steps_per_launch steps of
 * every system each), then wait; *elapsed_ms (may be NULL) receives the
 * device time between events around the replays.  GLX_EINVAL if a system's
 * history does not cover the steps. */
int glx_batch_run(glx_batch* b, uint32_t n_launches, int form, float* elapsed_ms);

/* Autotuner: time one launch of each candidate shape (n_candidates entries;
 * NULL / 0 = the built-in set over steps {512, 1024, 2048}, lanes {64, 128,
 * 256}, dispatch and persistent modes) from the current state, restore the
 * state, keep the fastest shape and write it to *best (particle-steps per
 * second to *best_rate; either may be NULL). */
int glx_batch_tune(glx_batch* b, int form, const glx_batch_config* candidates, uint32_t n_candidates,
                   glx_batch_config* best, double* best_rate);

/* Sum of deathtime_index over every particle of the batch (synchronous):
 * the particle-steps done since the batch started. */
int glx_batch_sum_steps(glx_batch* b, uint64_t* total);

/* ---------------------------------------------------------------------
 * Host time-block loop (C entry points of the C++ class glx::Executor in
 * glisser_amd_executor.hpp): Executor::init / loop / resync2 / the
 * encounter hand-off of the reference (src/executor.cu:128-877), driving the
 * entry points above.  Planets come from the host Jacobi WH producer.
 * ------------------------------------------------------------------- */

typedef struct glx_exec glx_exec;

typedef struct glx_exec_params {
    double t0;
    double dt;                  /* Time-Step */
    uint32_t tbsize;            /* Time-Block-Size */
    double outer_radius;        /* Outer-Limit */
    double cull_radius;         /* Cull-Radius */
    double rh_factor;           /* Encounter-RH-Factor */
    int32_t resolve_encounters; /* Resolve-Encounters: hand encounter particles to the host queue */
    int32_t form;               /* glx_form */
    int32_t device;
    const char* planet_history; /* Read-Planet-History file, or NULL: integrate the planets given */
    uint32_t history_table_size; /* Planet-History-Max-Planet-Count (0 means 16) */
    uint32_t resync_every;      /* Resync-Interval (0 means 1); 1 with resolve_encounters */
    int32_t reference_deathtime; /* death time = end of the resync block cast to float, as the reference;
                                    0: the time of the flagging step */
    const char* encounter_log;  /* NULL, or a file for the reference's encounter.out lines */
} glx_exec_params;

/* Create the loop: planets (n_bodies GM values, heliocentric r, v; unused
 * and may be NULL with a planet history), n particles (heliocentric pr,
 * pv, ids).  Uploads, computes the initial accelerations and flags,
 * compacts, produces the first planet block.  With a planet history the
 * blocks follow its records (Executor::update_planets,
 * src/executor.cu:226-299): each record interval of length T holds
 * max(1, round(T / dt)) equal steps, at most tbsize per block. */
int glx_exec_create(const glx_exec_params* p, uint32_t n_bodies, const double* m, const double* r, const double* v,
                    const uint32_t* planet_id, uint64_t n, const double* pr, const double* pv, const uint32_t* ids,
                    glx_exec** out);
int glx_exec_destroy(glx_exec* e);
/* n_blocks time-blocks (Executor::loop + resync2 each). */
int glx_exec_step(glx_exec* e, uint32_t n_blocks);
int glx_exec_status(glx_exec* e, double* t, uint64_t* n_alive, uint64_t* n_total, uint64_t* n_queued);
/* Move every queued encounter record into out (capacity entries); GLX_EINVAL
 * with *n_out = queue length if it does not fit. */
int glx_exec_take_encounters(glx_exec* e, uint64_t capacity, glx_encounter* out, uint64_t* n_out);
/* Hand resolved particles back at the loop's current time (r, v: n x 3). */
int glx_exec_return(glx_exec* e, uint64_t n, const uint32_t* ids, const double* r, const double* v);
/* Every particle slot (n_total): r, v (n x 3), flags, id, deathtime (time of
 * the flagging step, or with reference_deathtime the end time of the block
 * whose resync removed the particle, cast to float; NaN if not set); any
 * output may be NULL. */
int glx_exec_download(glx_exec* e, double* r, double* v, uint16_t* flags, uint32_t* id, double* deathtime);
/* Heliocentric planet positions ((n_bodies - 1) x 3, n_bodies from
 * glx_exec_bodies, at most GLX_MAX_BODIES) at the current time. */
int glx_exec_planets(glx_exec* e, double* r);
/* The bodies at the current time: count (Sun included), GM values and ids
 * (any output may be NULL).  With a planet history the set can change
 * between blocks. */
int glx_exec_bodies(glx_exec* e, uint32_t* n_bodies, double* m, uint32_t* id);
/* Device time of the last kernel and compaction, host time of the last planet block (ms). */
int glx_exec_times(glx_exec* e, double* kernel_ms, double* partition_ms, double* planet_ms);
/* Device time of every block kernel since glx_exec_create (ms) and the
 * number of blocks that ran one (n_blocks may be NULL).  The reference
 * executor keeps no such total; it is for callers that time the host loop
 * against its kernels. */
int glx_exec_kernel_total(glx_exec* e, double* kernel_ms, uint64_t* n_blocks);

/* ---------------------------------------------------------------------
 * File formats of the reference (configuration, state, binary track):
 * src/data.cpp:146-815, README.md.  The C++ API is in
 * glisser_amd/csrc/host/formats.h; these are its C entry points.
 * ------------------------------------------------------------------- */

/* Parse a configuration text (the reference's "Key value" lines and
 * consistency rules, src/data.cpp:195-333) and write its canonical form
 * (every key, src/data.cpp:335-371) into out (capacity bytes incl. the NUL);
 * *length receives the text length.  GLX_EINVAL with the reference's message
 * on a bad line or an inconsistent configuration. */
int glx_config_normalize(const char* text, char* out, uint64_t capacity, uint64_t* length);

enum glx_state_format {
    GLX_STATE_ASCII = 0,  /* one file, text: src/data.cpp:427-483,620-643 */
    GLX_STATE_BINARY = 1, /* one file, binary: src/data.cpp:485-525,586-618 */
    GLX_STATE_SPLIT = 2   /* planet file + particle file, text: src/data.cpp:374-425,645-668 */
};

/* Caller-owned arrays of a state: planets (n_planets GM values, r, v as
 * n x 3, ids) and particles (r, v, ids, flags, deathtime). */
typedef struct glx_state {
    uint32_t n_planets;
    uint64_t n_particles;
    double* pm;
    double* pr;
    double* pv;
    uint32_t* pid;     /* NULL on save: 0..n-1 */
    double* r;
    double* v;
    uint32_t* id;
    uint16_t* flags;   /* NULL on save: zeros */
    double* deathtime; /* NULL on save: zeros */
} glx_state;

/* Write a state file (split: path = planets, particle_path = particles);
 * write_momenta multiplies planet velocities by GM. */
int glx_state_save(const char* path, const char* particle_path, int format, int write_momenta, const glx_state* s);
/* Sizes of a state file (at most max_particles particles). */
int glx_state_sizes(const char* path, const char* particle_path, int format, uint64_t max_particles,
                    uint32_t* n_planets, uint64_t* n_particles);
/* Read a state file into arrays sized by glx_state_sizes, then divide
 * planet momenta by GM (read_momenta) and scale GM by big_g (load_data,
 * src/data.cpp:527-584).  The binary deathtime is read as the double the
 * writer stores (the reference reads a float there). */
int glx_state_load(const char* path, const char* particle_path, int format, uint64_t max_particles, int read_momenta,
                   double big_g, glx_state* s);
/* Append (or start) a binary track file with one snapshot of the state's
 * planets and particles, in the given particle order (save_binary_track,
 * src/data.cpp:718-815). */
int glx_track_write(const char* path, int append, double time, const glx_state* s, int to_elements, int barycentric);

/* Track records of the context's alive particles [0, n_alive) formed on the
 * device, synchronous: elements about the Sun or the barycentre of the
 * planets given (pm, pr, pv: n_planets bodies, Sun first), or the Cartesian
 * state, as 28-byte records (u32 id + 6 float32) ordered by id -- the
 * sort_by_id + save_binary_track particle loop of the reference
 * (targets/main.cpp:148-156, src/data.cpp:770-812).  records holds capacity
 * records; *n_out = n_alive.  GLX_ESTATE if particle ids repeat. */
int glx_track_pack(glx_ctx* ctx, uint32_t n_planets, const double* pm, const double* pr, const double* pv,
                   int to_elements, int barycentric, void* records, uint64_t capacity, uint64_t* n_out);
/* Append (or start) a track file with one snapshot: the planets of s (its
 * particles are not used) and n records from glx_track_pack. */
int glx_track_write_records(const char* path, int append, double time, const glx_state* s, int to_elements,
                            int barycentric, const void* records, uint64_t n);

/* The element conversion glx_track_pack runs on the device, as doubles:
 * n Cartesian rows (x, y, z, vx, vy, vz) about central GM mu into n rows
 * (a, e, i, Omega, omega, f), in the operation order of the reference's
 * to_elements (src/convert.cpp:310-463) with correctly rounded atan2, acos,
 * asin, sin and cos, the values glibc returns in all but rare cases.
 * glx_track_elements_device runs the device code on GPU `device`
 * (synchronous; at most 2^26 rows); glx_track_elements_host runs the same
 * code on the host. */
int glx_track_elements_device(int device, uint64_t n, double mu, const double* rv, double* el);
int glx_track_elements_host(uint64_t n, double mu, const double* rv, double* el);

/* The correctly rounded functions of that conversion over arrays:
 * out[k] = atan2(a[k], b[k]), acos(a[k]) or asin(a[k]) (b unused), on GPU
 * `device` (synchronous; at most 2^26 values) or on the host. */
enum glx_cr_fn { GLX_CR_ATAN2 = 0, GLX_CR_ACOS = 1, GLX_CR_ASIN = 2 };
int glx_cr_math_device(int device, int fn, uint64_t n, const double* a, const double* b, double* out);
int glx_cr_math_host(int fn, uint64_t n, const double* a, const double* b, double* out);

/* ---------------------------------------------------------------------
 * Planet history files (README.md:154-171): the reference's "planetary
 * history" of float32 heliocentric elements per record, written by
 * save_swift_plhist (src/data.cpp:817-873) and read back by its
 * Interpolator (src/interp.cpp), which moves a, e, i, Omega, omega
 * linearly between records and the mean anomaly at a corrected constant
 * rate.  Blocks never cross a record: each interval of length T holds
 * max(1, round(T / dt)) steps of equal size.  Rows equal the reference CPU
 * build's bit for bit.
 * ------------------------------------------------------------------- */

typedef struct glx_plhist glx_plhist;

/* Start (append = 0: header with the Sun's GM first) or extend a history
 * file with one record of n_bodies heliocentric bodies (entry 0 the Sun;
 * id NULL means 0..n_bodies-1): begin_swift_plhist + save_swift_plhist. */
int glx_plhist_write(const char* path, int append, double time, uint32_t n_bodies, const double* m, const double* r,
                     const double* v, const uint32_t* id);

/* Open a history for a run starting at t0 with Time-Step dt and a planet
 * table of table_size slots (Planet-History-Max-Planet-Count, the Sun
 * included): the Interpolator constructor and the history set-up of
 * Executor::init (src/interp.cpp:13-91, src/executor.cu:93-130).
 * GLX_EINVAL with the reference's message if the history starts after t0 or
 * ends before it. */
int glx_plhist_open(const char* path, double t0, double dt, uint32_t table_size, glx_plhist** out);
int glx_plhist_close(glx_plhist* h);

/* The planet table (table_size slots, the first *n_bodies in use): GM, ids
 * and heliocentric r, v of the last interpolated time (at open: t0).  Any
 * output may be NULL. */
int glx_plhist_state(glx_plhist* h, uint32_t* n_bodies, double* m, uint32_t* id, double* r, double* v);

/* The next planet block of at most tbsize steps: *t_start, step *dt,
 * *n_steps, bodies (*n_bodies GM values and ids, Sun first), r_log
 * (n_steps x (n_bodies-1) x 3, row s at t_start + (s+1) dt; room for
 * tbsize x (table_size-1) x 3) and h0_log (n_steps x 3, may be NULL).
 * Executor::update_planets with Interpolator::fill and load_h0
 * (src/executor.cu:235-292, src/interp.cpp:113-165, src/wh.cpp:993-1025).
 * GLX_EINVAL at the end of the file ("EOS reached for lookup file."). */
int glx_plhist_block(glx_plhist* h, uint32_t tbsize, double* t_start, double* dt, uint32_t* n_steps,
                     uint32_t* n_bodies, double* m, uint32_t* id, double* r_log, double* h0_log);

/* ---------------------------------------------------------------------
 * Host-side producer of the kernel's planet input and orbital conversions.
 * Built without FP contraction so it rounds like the reference CPU build.
 * ------------------------------------------------------------------- */

typedef struct glx_planets glx_planets;

/* Planet Jacobi Wisdom-Holman integrator: the planet half of the
 * WHIntegrator constructor (src/wh.cpp:498-531).  m in GM units (m[0] is
 * the Sun), heliocentric r, v (n x 3); id may be NULL (0..n-1). */
int glx_planets_create(uint32_t n, const double* m, const double* r, const double* v, const uint32_t* id,
                       double cull_radius, double rh_factor, double outer_radius, glx_planets** out);
int glx_planets_destroy(glx_planets* p);

/* Indirect term h0 at the initial time (the constructor's h0_log.old[0],
 * src/wh.cpp:531-540). */
int glx_planets_h0_initial(glx_planets* p, double* h0);

/* n_steps planet steps: WHIntegrator::integrate_planets_timeblock
 * (src/wh.cpp:584-595) with step_planets (src/wh.cpp:1027-1067).
 * r_log, v_log: n_steps x (n-1) x 3; h0_log: n_steps x 3 (any may be NULL).
 * GLX_ESTATE on an unbound or unconverged planet orbit (where the
 * reference throws). */
int glx_planets_block(glx_planets* p, uint32_t n_steps, double dt, double* r_log, double* v_log, double* h0_log);

/* Current heliocentric state (n x 3) and encounter radii (n); any output may
 * be NULL. */
int glx_planets_state(glx_planets* p, double* r, double* v, double* rh);

/* WHIntegrator::recalculate_rh (src/wh.cpp:472-495). */
int glx_planets_recalculate_rh(glx_planets* p);

/* calculate_planet_metrics (src/wh.cpp:422-470): energy and L (3). */
int glx_planets_metrics(glx_planets* p, double* energy, double* l);

/* WHIntegrator::load_h0 (src/wh.cpp:993-1025): h0 rows from a position log
 * of bodies 1..n_bodies-1. */
int glx_load_h0(uint32_t n_bodies, const double* m, const double* r_log, uint32_t n_steps, double* h0_log);

/* to_elements (src/convert.cpp:310-463) for n states about mu[k];
 * el: n x 6 (a, e, i, Omega, omega, f). */
int glx_to_elements(uint64_t n, const double* mu, const double* r, const double* v, double* el);

/* from_elements / from_elements_M (src/convert.cpp:230-308) for n element
 * sets el (n x 6: a, e, i, Omega, omega, f-or-M). */
int glx_from_elements(uint64_t n, const double* mu, const double* el, int anomaly_is_mean, double* r, double* v);

#ifdef __cplusplus
}
#endif

#endif /* GLISSER_AMD_H */
