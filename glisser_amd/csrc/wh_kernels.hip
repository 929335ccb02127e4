// Time-block kernels: every active particle advances through all steps of
// one planet block (replaces MVSKernel / MVSKernel_, src/wh.cu:15-267).
//
// Launch shape: 1-D, 256 lanes per workgroup, one particle per lane.  The
// planet block is staged into LDS `chunk` steps at a time (the full block
// at Time-Block-Size 2048 with 5 bodies is 240 KB, above the 160 KB of LDS
// per CU).  Every lane of a wave reads the same LDS address (broadcast).
#include <hip/hip_runtime.h>

#include "wh_device.h"
#include "wh_fast.h"
#include "launch.h"

namespace glx {

// Bodies whose per-step constants the specialised FAST / MIXED kernels keep
// in vector registers.
constexpr int kRegBodies = 5;

template <int NB, int FORM>
__global__ __launch_bounds__(256) void wh_block_kernel(BlockArgs A) {
    extern __shared__ double lds[];
    const uint32_t i = A.first + blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < A.n;
    const uint32_t nb = NB > 0 ? (uint32_t)NB : A.n_bodies;
    const uint32_t row = nb * 3;

    Particle p;
    uint32_t dti = 0;
    if (valid) {
        p.rx = A.rx[i];
        p.ry = A.ry[i];
        p.rz = A.rz[i];
        p.vx = A.vx[i];
        p.vy = A.vy[i];
        p.vz = A.vz[i];
        p.ax = A.ax[i];
        p.ay = A.ay[i];
        p.az = A.az[i];
        p.flags = A.flags[i];
        // particles flagged before this block keep their death step; the
        // others end the block at its last step unless they are flagged
        // on the way
        dti = p.flags == 0 ? A.step_offset + A.n_steps : A.dti[i];
    } else {
        p.rx = p.ry = p.rz = p.vx = p.vy = p.vz = p.ax = p.ay = p.az = 0.0;
        p.flags = 0xFFFFu;  // padding lanes never step
    }

    if constexpr (FORM != GLX_FORM_FP64) radius_from_position(p);

    // With a compile-time body count of at most kRegBodies, the FAST / MIXED
    // forms hold the bodies' GM and squared radii in vector registers: the
    // scalar file is full with the launch's other constants, and a value
    // that overflows it is read back every step.  (More bodies would push
    // the kernel past the registers of three waves per SIMD.)
    constexpr bool kInRegs = NB > 0 && NB <= kRegBodies && FORM != GLX_FORM_FP64;
    constexpr int kH = kInRegs ? NB : 1;
    double hm[kH], hrh2[kH];
    const double* pm = A.pc.m;
    const double* prh2 = A.pc.rh2;
    if constexpr (kInRegs) {
#pragma unroll
        for (int j = 0; j < NB; j++) {
            hm[j] = A.pc.m[j];
            hrh2[j] = A.pc.rh2[j];
            pin(hm[j]);
            pin(hrh2[j]);
        }
        pm = hm;
        prh2 = hrh2;
    }

    // FAST / MIXED: the start polynomials' highest coefficients in vector
    // registers
    TaylorLead lead = taylor_lead();
    pin_lead<FORM>(lead);
    bool pending = false;  // FAST / MIXED: closing half kick still due
    for (uint32_t cs = 0; cs < A.n_steps; cs += A.chunk) {
        const uint32_t cn = min(A.chunk, A.n_steps - cs);
        __syncthreads();  // every wave is done with the previous chunk
        const double* src = A.log + (size_t)cs * row;
        const uint32_t total = cn * row;
        for (uint32_t k = threadIdx.x; k < total; k += blockDim.x) lds[k] = src[k];
        __syncthreads();
        chunk_steps<NB, FORM>(p, dti, A, pm, prh2, lds, row, cn, nb, A.step_offset + cs, A.block_step0 + cs, pending,
                              lead);
    }
    if constexpr (FORM != GLX_FORM_FP64) close_kick(p, A.half_dt, pending);

    if (valid) {
        A.rx[i] = p.rx;
        A.ry[i] = p.ry;
        A.rz[i] = p.rz;
        A.vx[i] = p.vx;
        A.vy[i] = p.vy;
        A.vz[i] = p.vz;
        A.ax[i] = p.ax;
        A.ay[i] = p.ay;
        A.az[i] = p.az;
        A.flags[i] = (uint16_t)p.flags;
        A.dti[i] = dti;
    }
}

// Initial accelerations and flags from one planet row (t0 positions + h0).
template <int NB>
__global__ __launch_bounds__(256) void init_accel_kernel(BlockArgs A) {
    extern __shared__ double lds[];
    const uint32_t nb = NB > 0 ? (uint32_t)NB : A.n_bodies;
    for (uint32_t k = threadIdx.x; k < nb * 3; k += blockDim.x) lds[k] = A.log[k];
    __syncthreads();
    const uint32_t i = A.first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    Particle p;
    p.rx = A.rx[i];
    p.ry = A.ry[i];
    p.rz = A.rz[i];
    p.flags = A.flags[i];
    accel_fp64<NB>(p, lds, A.pc, A.outer2, nb);
    A.ax[i] = p.ax;
    A.ay[i] = p.ay;
    A.az[i] = p.az;
    A.flags[i] = (uint16_t)p.flags;
}

template <int FORM>
static hipError_t launch_block_form(const BlockArgs& A, size_t lds_bytes, hipStream_t stream) {
    const uint32_t threads = 256;
    const uint32_t blocks = (A.n - A.first + threads - 1) / threads;
    switch (A.n_bodies) {
#define GLX_CASE(NBV)                                                                          \
    case NBV:                                                                                  \
        hipLaunchKernelGGL((wh_block_kernel<NBV, FORM>), dim3(blocks), dim3(threads), lds_bytes, \
                           stream, A);                                                         \
        break;
        GLX_CASE(2)
        GLX_CASE(3)
        GLX_CASE(4)
        GLX_CASE(5)
        GLX_CASE(6)
        GLX_CASE(7)
        GLX_CASE(8)
        GLX_CASE(9)
#undef GLX_CASE
        default:
            hipLaunchKernelGGL((wh_block_kernel<0, FORM>), dim3(blocks), dim3(threads), lds_bytes, stream, A);
    }
    return hipGetLastError();
}

hipError_t launch_block(const BlockArgs& A, int form, size_t lds_bytes, hipStream_t stream) {
    if (A.n <= A.first) return hipSuccess;
    const size_t need = (size_t)A.chunk * A.n_bodies * 3 * sizeof(double);
    if (lds_bytes < need) lds_bytes = need;
    switch (form) {
        case GLX_FORM_FP64:
            return launch_block_form<GLX_FORM_FP64>(A, lds_bytes, stream);
        case GLX_FORM_FAST:
            return launch_block_form<GLX_FORM_FAST>(A, lds_bytes, stream);
        case GLX_FORM_MIXED:
            return launch_block_form<GLX_FORM_MIXED>(A, lds_bytes, stream);
        default:
            return hipErrorInvalidValue;
    }
}

hipError_t launch_init_accel(const BlockArgs& A, hipStream_t stream) {
    if (A.n <= A.first) return hipSuccess;
    const uint32_t threads = 256;
    const uint32_t blocks = (A.n - A.first + threads - 1) / threads;
    const size_t lds_bytes = (size_t)A.n_bodies * 3 * sizeof(double);
    hipLaunchKernelGGL((init_accel_kernel<0>), dim3(blocks), dim3(threads), lds_bytes, stream, A);
    return hipGetLastError();
}

}  // namespace glx
