// Device-side planet log from a sampled planet history (cubic Hermite).
//
// The reference fills each block's planet log on the host (planet WH step,
// src/wh.cpp:584-595, or element-linear interpolation, src/interp.cpp:
// 113-162) and uploads it before every kernel (src/executor.cu:442-461).
// Here the history -- heliocentric r and v of every planet at sample times
// t_k -- is uploaded once and each block's rows are produced on the device,
// next to the particles, so no per-block host work or host-to-device copy
// remains.  Row s of a block is the state at t_start + (s+1) dt (the log
// convention of src/wh.cpp:1060), laid out as the particle kernel reads it:
// the indirect term h0 first (load_h0, src/wh.cpp:993-1025), then the planet
// positions.
//
// On [t_k, t_k+1], h = t_k+1 - t_k, u = (t - t_k) / h:
//   r(t) = h00 r_k + h10 h v_k + h01 r_k+1 + h11 h v_k+1
//   h00 = 2u^3 - 3u^2 + 1, h10 = u^3 - 2u^2 + u, h01 = 3u^2 - 2u^3, h11 = u^3 - u^2
// with the sums evaluated left to right and no FMA contraction, so each row
// is bit-identical to the CPU restatement of the same formula.  The work is
// one lane per step (n_bodies - 1 planets each): 2048 steps x 4 planets is a
// few microseconds, so it stays FP64 in every kernel form.
#include <hip/hip_runtime.h>

#include "interp_device.h"
#include "launch.h"

namespace glx {

namespace {

__global__ __launch_bounds__(64) void hermite_rows_kernel(HistoryView H, const double* times, double t_start,
                                                          double dt, uint32_t n_steps, double* out) {
    const uint32_t s = blockIdx.x * 64 + threadIdx.x;
    if (s >= n_steps) return;
    const double t = times ? times[s] : step_time(t_start, s, dt);
    hermite_row(H.t, H.r, H.v, H.n_samples, H.n_bodies, H.m, t, out + (size_t)s * H.n_bodies * 3);
}

// Several blocks' rows in one launch: block y of the grid writes n_steps
// rows at R.t_start[y] + (s+1) dt into R.out[y].
__global__ __launch_bounds__(64) void hermite_rows_multi_kernel(HistoryView H, RowBlocks R, double dt,
                                                                uint32_t n_steps) {
    const uint32_t s = blockIdx.x * 64 + threadIdx.x;
    if (s >= n_steps) return;
    const double t = step_time(R.t_start[blockIdx.y], s, dt);
    hermite_row(H.t, H.r, H.v, H.n_samples, H.n_bodies, H.m, t, R.out[blockIdx.y] + (size_t)s * H.n_bodies * 3);
}

}  // namespace

hipError_t launch_hermite_rows_multi(const HistoryView& H, const RowBlocks& R, uint32_t n_blocks, double dt,
                                     uint32_t n_steps, hipStream_t stream) {
    if (n_steps == 0 || n_blocks == 0) return hipSuccess;
    if (n_blocks > kRowBlocksMax) return hipErrorInvalidValue;
    hermite_rows_multi_kernel<<<dim3((n_steps + 63) / 64, n_blocks), dim3(64), 0, stream>>>(H, R, dt, n_steps);
    return hipGetLastError();
}

hipError_t launch_hermite_rows(const HistoryView& H, const double* times, double t_start, double dt, uint32_t n_steps,
                               double* out, hipStream_t stream) {
    if (n_steps == 0) return hipSuccess;
    hermite_rows_kernel<<<dim3((n_steps + 63) / 64), dim3(64), 0, stream>>>(H, times, t_start, dt, n_steps, out);
    return hipGetLastError();
}

}  // namespace glx
