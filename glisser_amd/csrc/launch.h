// Host-side launchers of the device kernels (one translation unit each).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "wh_device.h"

namespace glx {

// Particle arrays of one particle set on the device (SoA).  The `d` slots
// hold rx, ry, rz, vx, vy, vz, ax, ay, az.
struct ParticleSet {
    double* d[9];
    uint16_t* flags;
    uint32_t* dti;
    uint32_t* id;
};

// Rollback copy: r, v (6 slots), flags, dti, id.
struct RollbackSet {
    double* d[6];
    uint16_t* flags;
    uint32_t* dti;
    uint32_t* id;
};

// lds_bytes: dynamic LDS to request per workgroup (at least the staged
// chunk); a larger request caps the workgroups resident per CU.
hipError_t launch_block(const BlockArgs& A, int form, size_t lds_bytes, hipStream_t stream);
hipError_t launch_init_accel(const BlockArgs& A, hipStream_t stream);

// Stable partition scratch sizes for n elements.
constexpr uint32_t kPartTile = 1024;  // elements per workgroup (256 lanes x 4)
inline uint32_t part_tiles(uint64_t n) { return (uint32_t)((n + kPartTile - 1) / kPartTile); }

// counts: 2 * tiles u32; offsets: 2 * tiles u32; totals: 2 u32 (device);
// `tmp`/`tmp_rb` are scratch particle sets with capacity >= n.
hipError_t launch_partition(ParticleSet cur, ParticleSet tmp, const RollbackSet* rb, const RollbackSet* tmp_rb,
                            uint32_t n, int three_way, uint32_t* counts, uint32_t* offsets, uint32_t* totals,
                            hipStream_t stream);

hipError_t launch_kill_encounters(uint16_t* flags, uint32_t n, hipStream_t stream);
hipError_t launch_snapshot(ParticleSet cur, RollbackSet rb, uint32_t n, hipStream_t stream);
hipError_t launch_sum_steps(const uint32_t* dti, uint32_t n, unsigned long long* out, hipStream_t stream);

// AoS (x,y,z) <-> SoA transposes for host copies staged through device memory.
hipError_t launch_aos_to_soa(const double* aos, double* x, double* y, double* z, uint32_t n, hipStream_t stream);
hipError_t launch_soa_to_aos(const double* x, const double* y, const double* z, double* aos, uint32_t n,
                             hipStream_t stream);

}  // namespace glx
