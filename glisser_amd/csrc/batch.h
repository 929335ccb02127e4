// Batched form: many independent planetary systems, each with its own
// bodies, particle count, timestep and planet history, advanced by one
// kernel launch per time-block (structures shared by batch.hip and the
// C ABI in batch_api.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "wh_device.h"

namespace glx {

// Bodies per system in the batched kernel (Sun + up to 8 planets): the planet
// loop is unrolled to this bound with a wave-uniform guard per planet.
constexpr uint32_t kBatchMaxBodies = 9;

// One system on the device.  Particles [first, first + n) of the batch's
// particle arrays belong to it; first is a multiple of 256, so no workgroup
// of 64, 128 or 256 lanes spans two systems.
struct SystemDev {
    PlanetConsts pc;     // GM, squared encounter radii, planet ids
    double* log;         // planet rows of the current launch: [steps][nb][3] (h0, then planets)
    const double* ht;    // history: sample times [n_samples]
    const double* hr;    // heliocentric planet positions [n_samples][nb-1][3]
    const double* hv;    // and velocities
    uint64_t n_samples;
    double t0;           // time at the start of the batch
    double dt;
    double half_dt;
    double mu;
    double inv_mu;
    double dt_sqrt_mu;   // dt sqrt(mu)
    double inv_sqrt_mu;
    double inv_dt;
    double outer2;
    uint32_t fp64_core;  // fp64 form: fp64_core_constants(mu, dt)
    double sun_vote;     // carried |r| below this or above outer_vote
    double outer_vote;   // triggers the exact Sun / outer tests
    uint32_t nb;         // bodies (Sun + planets)
    uint32_t first;      // first particle slot
    uint32_t n;          // particles
    uint32_t chunk;      // planet rows staged in LDS at a time
};

// Arguments of one batched launch.  A tile is one workgroup's worth of
// particles of one system; tiles are listed heaviest first.
struct BatchArgs {
    double* d[9];                 // rx ry rz vx vy vz ax ay az
    uint16_t* flags;
    uint32_t* dti;
    const SystemDev* sys;
    const uint32_t* tile_sys;     // system of each tile
    const uint32_t* tile_first;   // first particle slot of each tile
    uint32_t n_tiles;
    uint32_t steps;               // steps per launch
    const uint32_t* clock;        // steps completed since the batch start
    uint32_t* queue;              // persistent mode: next tile to take
    uint32_t persistent;          // 0: one workgroup per tile; 1: workgroups loop over a tile queue
};

// Planet rows of the coming launch for every system (grid: steps x systems).
hipError_t launch_batch_interp(const SystemDev* sys, uint32_t n_sys, uint32_t steps, const uint32_t* clock,
                               hipStream_t stream);
// The particle kernel: `grid` workgroups of `lanes` lanes, `lds_bytes` of
// dynamic LDS each.
hipError_t launch_batch_block(const BatchArgs& B, int form, uint32_t grid, uint32_t lanes, size_t lds_bytes,
                              hipStream_t stream);
// clock += steps and queue = 0, at the end of every launch.
hipError_t launch_batch_finish(uint32_t* clock, uint32_t* queue, uint32_t steps, hipStream_t stream);
// Initial accelerations and flags at each system's t0 (rows from the history).
hipError_t launch_batch_init(const BatchArgs& B, uint32_t lanes, hipStream_t stream);

}  // namespace glx
