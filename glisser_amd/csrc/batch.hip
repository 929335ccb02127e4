// Batched time-block kernels: independent planetary systems with different
// bodies, particle counts and timesteps in one launch.
//
// The reference runs one system per process and one kernel per block
// (src/executor.cu:640-646).  Here a launch covers every system: its
// workgroups are tiles of one system's particles (listed heaviest system
// first, so the expensive tiles are dispatched early and the launch ends
// on cheap ones), each tile stages its own system's planet rows through LDS
// and advances its particles `steps` steps with the same per-particle code
// as the single-system kernel (planet loop unrolled to 9 bodies with a
// wave-uniform guard).  In persistent mode a fixed grid of workgroups takes
// tiles from a device counter instead of one workgroup per tile.  A launch
// is interp (planet rows of every system) -> particles -> finish (step
// clock + steps, queue reset), captured once into a HIP graph and replayed,
// so no host work or host-to-device copy sits between launches.
#include <hip/hip_runtime.h>

#include "batch.h"
#include "interp_device.h"
#include "wh_fast.h"

namespace glx {

namespace {

__global__ __launch_bounds__(64) void batch_interp_kernel(const SystemDev* __restrict__ sys, uint32_t steps,
                                                          const uint32_t* __restrict__ clock) {
    const uint32_t s = blockIdx.x * 64 + threadIdx.x;
    if (s >= steps) return;
    const SystemDev& S = sys[blockIdx.y];
    const double t = step_time(S.t0, (uint64_t)clock[0] + s, S.dt);
    hermite_row(S.ht, S.hr, S.hv, S.n_samples, S.nb, S.pc.m, t, S.log + (size_t)s * S.nb * 3);
}

__global__ void batch_finish_kernel(uint32_t* clock, uint32_t* queue, uint32_t steps) {
    if (threadIdx.x == 0) {
        clock[0] = clock[0] + steps;
        queue[0] = 0;
    }
}

// Bodies up to which a specialised tile keeps the GM and squared radii in
// vector registers (more would cost the kernel its third wave per SIMD).
constexpr uint32_t kBatchRegBodies = 4;

// One tile.  NB > 0: the system's body count as a compile-time constant
// (planet loop unrolled without guards; up to kBatchRegBodies bodies the GM
// and squared radii in vector registers); NB == 0: any count up to
// kBatchMaxBodies.
template <int FORM, int NB>
__device__ __forceinline__ void run_tile_nb(const BatchArgs& B, uint32_t tile, double* lds) {
    const uint32_t si = B.tile_sys[tile];
    const SystemDev& S = B.sys[si];
    const uint32_t i = B.tile_first[tile] + threadIdx.x;
    const bool valid = i < S.first + S.n;
    const uint32_t nb = S.nb;
    const uint32_t row = nb * 3;
    const uint32_t base = B.clock[0];  // steps done since the batch started

    Particle p;
    uint32_t dti = 0;
    if (valid) {
        p.rx = B.d[0][i];
        p.ry = B.d[1][i];
        p.rz = B.d[2][i];
        p.vx = B.d[3][i];
        p.vy = B.d[4][i];
        p.vz = B.d[5][i];
        p.ax = B.d[6][i];
        p.ay = B.d[7][i];
        p.az = B.d[8][i];
        p.flags = B.flags[i];
        // particles flagged before this launch keep their death step; the
        // others end the launch at its last step unless flagged on the way
        dti = p.flags == 0 ? base + B.steps : B.dti[i];
    } else {
        p.rx = p.ry = p.rz = p.vx = p.vy = p.vz = p.ax = p.ay = p.az = 0.0;
        p.flags = 0xFFFFu;  // padding lanes never step
    }
    if constexpr (FORM != GLX_FORM_FP64) radius_from_position(p);

    constexpr bool kInRegs = NB > 0 && NB <= (int)kBatchRegBodies;
    constexpr int kH = kInRegs ? NB : 1;
    double hm[kH], hrh2[kH];
    const double* pm = S.pc.m;
    const double* prh2 = S.pc.rh2;
    if constexpr (kInRegs) {
#pragma unroll
        for (int j = 0; j < NB; j++) {
            hm[j] = S.pc.m[j];
            hrh2[j] = S.pc.rh2[j];
            pin(hm[j]);
            pin(hrh2[j]);
        }
        pm = hm;
        prh2 = hrh2;
    }
    TaylorLead lead = taylor_lead();
    pin_lead<FORM>(lead);
    bool pending = false;  // FAST / MIXED: closing half kick still due
    for (uint32_t cs = 0; cs < B.steps; cs += S.chunk) {
        const uint32_t cn = min(S.chunk, B.steps - cs);
        __syncthreads();  // every wave is done with the previous chunk (or tile)
        const double* src = S.log + (size_t)cs * row;
        const uint32_t total = cn * row;
        for (uint32_t k = threadIdx.x; k < total; k += blockDim.x) lds[k] = src[k];
        __syncthreads();
        chunk_steps<NB, FORM, (NB > 0 ? NB : (int)kBatchMaxBodies)>(p, dti, S, pm, prh2, lds, row, cn, nb, base + cs,
                                                             base + cs, pending, lead);
    }
    if constexpr (FORM != GLX_FORM_FP64) close_kick(p, S.half_dt, pending);

    if (valid) {
        B.d[0][i] = p.rx;
        B.d[1][i] = p.ry;
        B.d[2][i] = p.rz;
        B.d[3][i] = p.vx;
        B.d[4][i] = p.vy;
        B.d[5][i] = p.vz;
        B.d[6][i] = p.ax;
        B.d[7][i] = p.ay;
        B.d[8][i] = p.az;
        B.flags[i] = (uint16_t)p.flags;
        B.dti[i] = dti;
    }
}

// FAST / MIXED tiles of systems with 2 to 4 bodies run a copy specialised
// for their body count (the tile's system, hence the branch, is the same
// for the whole workgroup); more bodies, and the fp64 form, run the generic
// copy.  A copy for 5 or more bodies would take the kernel past the
// registers of three waves per SIMD, whichever tiles it serves.
template <int FORM>
__device__ __forceinline__ void run_tile(const BatchArgs& B, uint32_t tile, double* lds) {
    if constexpr (FORM != GLX_FORM_FP64) {
        switch (B.sys[B.tile_sys[tile]].nb) {
            case 2: return run_tile_nb<FORM, 2>(B, tile, lds);
            case 3: return run_tile_nb<FORM, 3>(B, tile, lds);
            case 4: return run_tile_nb<FORM, 4>(B, tile, lds);
            default: break;
        }
    }
    run_tile_nb<FORM, 0>(B, tile, lds);
}

template <int FORM>
__global__ __launch_bounds__(256) void batch_block_kernel(BatchArgs B) {
    extern __shared__ double lds[];
    __shared__ uint32_t next_tile;
    if (!B.persistent) {
        if (blockIdx.x < B.n_tiles) run_tile<FORM>(B, blockIdx.x, lds);
        return;
    }
    // Persistent: take tiles until the queue runs past the last one.  The
    // tile index is the same for the whole workgroup, so every barrier
    // below is reached by all of its waves, and every workgroup exits once
    // the counter passes n_tiles.
    for (;;) {
        __syncthreads();  // next_tile of the previous round has been read
        if (threadIdx.x == 0) next_tile = atomicAdd(B.queue, 1u);
        __syncthreads();
        const uint32_t tile = __builtin_amdgcn_readfirstlane(next_tile);
        if (tile >= B.n_tiles) break;
        run_tile<FORM>(B, tile, lds);
    }
}

__global__ __launch_bounds__(256) void batch_init_kernel(BatchArgs B) {
    __shared__ double row[kBatchMaxBodies * 3];
    const uint32_t tile = blockIdx.x;
    const SystemDev& S = B.sys[B.tile_sys[tile]];
    if (threadIdx.x == 0) hermite_row(S.ht, S.hr, S.hv, S.n_samples, S.nb, S.pc.m, S.t0, row);
    __syncthreads();
    const uint32_t i = B.tile_first[tile] + threadIdx.x;
    if (i >= S.first + S.n) return;
    Particle p;
    p.rx = B.d[0][i];
    p.ry = B.d[1][i];
    p.rz = B.d[2][i];
    p.flags = B.flags[i];
    accel_fp64<0, kBatchMaxBodies>(p, row, S.pc, S.outer2, S.nb);
    B.d[6][i] = p.ax;
    B.d[7][i] = p.ay;
    B.d[8][i] = p.az;
    B.flags[i] = (uint16_t)p.flags;
}

}  // namespace

hipError_t launch_batch_interp(const SystemDev* sys, uint32_t n_sys, uint32_t steps, const uint32_t* clock,
                               hipStream_t stream) {
    if (n_sys == 0 || steps == 0) return hipSuccess;
    batch_interp_kernel<<<dim3((steps + 63) / 64, n_sys), dim3(64), 0, stream>>>(sys, steps, clock);
    return hipGetLastError();
}

hipError_t launch_batch_block(const BatchArgs& B, int form, uint32_t grid, uint32_t lanes, size_t lds_bytes,
                              hipStream_t stream) {
    if (grid == 0 || B.n_tiles == 0) return hipSuccess;
    switch (form) {
        case GLX_FORM_FP64:
            batch_block_kernel<GLX_FORM_FP64><<<dim3(grid), dim3(lanes), lds_bytes, stream>>>(B);
            break;
        case GLX_FORM_FAST:
            batch_block_kernel<GLX_FORM_FAST><<<dim3(grid), dim3(lanes), lds_bytes, stream>>>(B);
            break;
        case GLX_FORM_MIXED:
            batch_block_kernel<GLX_FORM_MIXED><<<dim3(grid), dim3(lanes), lds_bytes, stream>>>(B);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_finish(uint32_t* clock, uint32_t* queue, uint32_t steps, hipStream_t stream) {
    batch_finish_kernel<<<dim3(1), dim3(64), 0, stream>>>(clock, queue, steps);
    return hipGetLastError();
}

hipError_t launch_batch_init(const BatchArgs& B, uint32_t lanes, hipStream_t stream) {
    if (B.n_tiles == 0) return hipSuccess;
    batch_init_kernel<<<dim3(B.n_tiles), dim3(lanes), 0, stream>>>(B);
    return hipGetLastError();
}

}  // namespace glx
