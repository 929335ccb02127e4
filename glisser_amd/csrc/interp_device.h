// Cubic Hermite evaluation of a sampled planet history into one planet-log
// row (device code shared by the single-system and the batched kernels).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace glx {

// Time of step s of a block starting at t_start: t_start + (s + 1) dt.
__device__ __forceinline__ double step_time(double t_start, uint64_t s, double dt) {
#pragma clang fp contract(off)
    return t_start + (double)(s + 1) * dt;
}

// Row at time t: the indirect term h0 (load_h0, src/wh.cpp:993-1025) then
// the positions of bodies 1..nb-1, from samples at times ht[n_samples]
// with heliocentric hr, hv ([n_samples][nb-1][3]).  On [t_k, t_k+1],
// h = t_k+1 - t_k, u = (t - t_k) / h:
//   r(t) = h00 r_k + h10 h v_k + h01 r_k+1 + h11 h v_k+1
//   h00 = 2u^3 - 3u^2 + 1, h10 = u^3 - 2u^2 + u, h01 = 3u^2 - 2u^3, h11 = u^3 - u^2
// with the sums evaluated left to right and no FMA contraction, so each row
// is bit-identical to the CPU restatement of the same formula.  k is the
// last sample with t_k <= t, clamped to [0, n_samples - 2].
__device__ __forceinline__ void hermite_row(const double* __restrict__ ht, const double* __restrict__ hr,
                                            const double* __restrict__ hv, uint64_t n_samples, uint32_t nb,
                                            const double* __restrict__ m, double t, double* __restrict__ row) {
#pragma clang fp contract(off)
    uint64_t lo = 0, hi = n_samples - 1;
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (ht[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    const double h = ht[lo + 1] - ht[lo];
    const double u = (t - ht[lo]) / h;
    const double u2 = u * u, u3 = u2 * u;
    const double h00 = 2.0 * u3 - 3.0 * u2 + 1.0;
    const double h10 = u3 - 2.0 * u2 + u;
    const double h01 = 3.0 * u2 - 2.0 * u3;
    const double h11 = u3 - u2;
    const double w10 = h10 * h, w11 = h11 * h;

    const uint32_t np = nb - 1;
    const double* r0 = hr + lo * np * 3;
    const double* v0 = hv + lo * np * 3;
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
            hx = -x * m[1] * inv_r3;
            hy = -y * m[1] * inv_r3;
            hz = -z * m[1] * inv_r3;
        } else {
            const double mfac = m[j + 1] * inv_r3;
            cx -= x * mfac;
            cy -= y * mfac;
            cz -= z * mfac;
        }
    }
    row[0] = hx + cx;
    row[1] = hy + cy;
    row[2] = hz + cz;
}

}  // namespace glx
