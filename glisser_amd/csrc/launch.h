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
// Resets Q's FIFO and counters, then runs every task of Q in one persistent
// grid (as many workgroups as fit on the device, at most one per tile).
// lds_bytes as launch_block.
hipError_t launch_queue(const QueueArgs& Q, int form, size_t lds_bytes, hipStream_t stream);

// Stable partition scratch sizes for n elements.
constexpr uint32_t kPartTile = 1024;  // elements per workgroup (256 lanes x 4)
inline uint32_t part_tiles(uint64_t n) { return (uint32_t)((n + kPartTile - 1) / kPartTile); }

// counts: 2 * tiles u32; offsets: 2 * tiles u32; totals: 2 u32 (device);
// `tmp`/`tmp_rb` are scratch particle sets with capacity >= n.
hipError_t launch_partition(ParticleSet cur, ParticleSet tmp, const RollbackSet* rb, const RollbackSet* tmp_rb,
                            uint32_t n, int three_way, uint32_t* counts, uint32_t* offsets, uint32_t* totals,
                            hipStream_t stream);

// The same stable partition from src into dst (distinct arrays, nothing
// copied back): counts and offsets hold 2 * part_tiles(n) entries each,
// totals receives the two class sizes.  Writes every scratch entry it
// reads, so nothing needs clearing first.
hipError_t launch_partition_into(ParticleSet src, ParticleSet dst, uint32_t n, int three_way, uint32_t* counts,
                                 uint32_t* offsets, uint32_t* totals, hipStream_t stream);

hipError_t launch_kill_encounters(uint16_t* flags, uint32_t n, hipStream_t stream);
hipError_t launch_snapshot(ParticleSet cur, RollbackSet rb, uint32_t n, hipStream_t stream);
// Sum of dti[0, n) into out[0]; out holds 1 + kSumParts entries (the rest
// are the workgroups' partial sums).
constexpr uint32_t kSumParts = 256;
hipError_t launch_sum_steps(const uint32_t* dti, uint32_t n, unsigned long long* out, hipStream_t stream);
// Host close-encounter records of particles [begin, begin + n); rb may be null.
hipError_t launch_pack_encounters(ParticleSet cur, const RollbackSet* rb, uint32_t begin, uint32_t n, double t_block,
                                  glx_encounter* out, hipStream_t stream);

// A sampled planet history on the device: times t[n_samples] (strictly
// increasing) and heliocentric r, v of bodies 1..n_bodies-1 as
// [n_samples][n_bodies-1][3]; m holds the bodies' GM (m[0] = Sun).
struct HistoryView {
    const double* t;
    const double* r;
    const double* v;
    uint64_t n_samples;
    uint32_t n_bodies;
    double m[GLX_MAX_BODIES];
};

// Start times and output buffers of up to kRowBlocksMax blocks of rows.
constexpr uint32_t kRowBlocksMax = 32;
struct RowBlocks {
    double t_start[kRowBlocksMax];
    double* out[kRowBlocksMax];
};
// The rows of launch_hermite_rows (times null) for n_blocks blocks at once:
// block k's n_steps rows at R.t_start[k] + (s+1) dt into R.out[k].
hipError_t launch_hermite_rows_multi(const HistoryView& H, const RowBlocks& R, uint32_t n_blocks, double dt,
                                     uint32_t n_steps, hipStream_t stream);

// Planet-log rows (h0, then planet positions; n_bodies x 3 doubles each) at
// times[s], or at t_start + (s+1) dt when times is null, for s < n_steps.
hipError_t launch_hermite_rows(const HistoryView& H, const double* times, double t_start, double dt, uint32_t n_steps,
                               double* out, hipStream_t stream);

// Track snapshots on the device: the centre the elements refer to, and the
// 28-byte record (u32 id + a, e, i, Omega, omega, f or x, y, z, vx, vy, vz as
// float32).
struct TrackCenter {
    double mass;
    double r[3];
    double v[3];
};
struct TrackRecordDev {
    uint32_t id;
    float x[6];
};
static_assert(sizeof(TrackRecordDev) == 28, "track record is 28 bytes");

// max of id[0, n) into *out (device).
hipError_t launch_max_id(const uint32_t* id, uint32_t n, uint32_t* out, hipStream_t stream);
// Tiles of the rank scan for ids [0, n_ids): size of the tile_sums scratch.
uint32_t rank_scratch_tiles(uint32_t n_ids);
// present[0, n_ids) becomes rank[id] = number of ids in id[0, n) below id
// (ids unique and < n_ids).
hipError_t launch_rank_ids(const uint32_t* id, uint32_t n, uint32_t n_ids, uint32_t* present, uint32_t* tile_sums,
                           hipStream_t stream);
// Records of particles [0, n) written at out[rank[id]] (rank null: slot order).
hipError_t launch_pack_track(ParticleSet cur, uint32_t n, const uint32_t* rank, const TrackCenter& c, int to_elements,
                             TrackRecordDev* out, hipStream_t stream);

// AoS (x,y,z) <-> SoA transposes for host copies staged through device memory.
hipError_t launch_aos_to_soa(const double* aos, double* x, double* y, double* z, uint32_t n, hipStream_t stream);
hipError_t launch_soa_to_aos(const double* x, const double* y, const double* z, double* aos, uint32_t n,
                             hipStream_t stream);

}  // namespace glx
