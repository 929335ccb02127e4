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

#include "launch.h"

namespace glx {

namespace {

__global__ __launch_bounds__(64) void hermite_rows_kernel(HistoryView H, const double* times, double t_start,
                                                          double dt, uint32_t n_steps, double* out) {
#pragma clang fp contract(off)
    const uint32_t s = blockIdx.x * 64 + threadIdx.x;
    if (s >= n_steps) return;
    const double t = times ? times[s] : t_start + (double)(s + 1) * dt;
    // Last sample with t_k <= t, within [0, n_samples - 2].
    uint64_t lo = 0, hi = H.n_samples - 1;
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (H.t[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    const double h = H.t[lo + 1] - H.t[lo];
    const double u = (t - H.t[lo]) / h;
    const double u2 = u * u, u3 = u2 * u;
    const double h00 = 2.0 * u3 - 3.0 * u2 + 1.0;
    const double h10 = u3 - 2.0 * u2 + u;
    const double h01 = 3.0 * u2 - 2.0 * u3;
    const double h11 = u3 - u2;
    const double w10 = h10 * h, w11 = h11 * h;

    const uint32_t np = H.n_bodies - 1;
    double* row = out + (size_t)s * H.n_bodies * 3;
    const double* r0 = H.r + lo * np * 3;
    const double* v0 = H.v + lo * np * 3;
    const double* r1 = r0 + (size_t)np * 3;
    const double* v1 = v0 + (size_t)np * 3;
    double hx = 0, hy = 0, hz = 0, cx = 0, cy = 0, cz = 0;
    for (uint32_t j = 0; j < np; j++) {
        const uint32_t o = 3 * j;
        const double x = h00 * r0[o] + w10 * v0[o] + h01 * r1[o] + w11 * v1[o];
        const double y = h00 * r0[o + 1] + w10 * v0[o + 1] + h01 * r1[o + 1] + w11 * v1[o + 1];
        const double z = h00 * r0[o + 2] + w10 * v0[o + 2] + h01 * r1[o + 2] + w11 * v1[o + 2];
        row[3 + o] = x;
        row[4 + o] = y;
        row[5 + o] = z;
        // load_h0: the first planet's term starts h0, the others accumulate
        // separately and are added at the end.
        const double r2 = x * x + y * y + z * z;
        const double inv_r3 = 1.0 / (sqrt(r2) * r2);
        if (j == 0) {
            hx = -x * H.m[1] * inv_r3;
            hy = -y * H.m[1] * inv_r3;
            hz = -z * H.m[1] * inv_r3;
        } else {
            const double mfac = H.m[j + 1] * inv_r3;
            cx -= x * mfac;
            cy -= y * mfac;
            cz -= z * mfac;
        }
    }
    row[0] = hx + cx;
    row[1] = hy + cy;
    row[2] = hz + cz;
}

}  // namespace

hipError_t launch_hermite_rows(const HistoryView& H, const double* times, double t_start, double dt, uint32_t n_steps,
                               double* out, hipStream_t stream) {
    if (n_steps == 0) return hipSuccess;
    hermite_rows_kernel<<<dim3((n_steps + 63) / 64), dim3(64), 0, stream>>>(H, times, t_start, dt, n_steps, out);
    return hipGetLastError();
}

}  // namespace glx
