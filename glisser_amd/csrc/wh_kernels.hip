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

// Particle-state loads and stores of block_tile.  COH: system-coherent
// accesses (no cache keeps them), for state handed from one workgroup to
// another (possibly on another XCD) inside a queued launch; otherwise plain.
template <bool COH, class T>
__device__ __forceinline__ T ld(const T* p) {
    if constexpr (COH) return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else return *p;
}
template <bool COH, class T>
__device__ __forceinline__ void st(T* p, T v) {
    if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else *p = v;
}

// Particles [tile, tile + 256) of A (those below A.n) through A's steps,
// staged in `lds`.  Every lane of the workgroup calls it (barriers inside).
// C: BlockArgs, or a queued launch's TaskArgs (the same members); COH: the
// state accesses (ld / st).
template <int NB, int FORM, class C, bool COH = false>
__device__ __forceinline__ void block_tile(const C& A, uint32_t tile, double* lds) {
    const uint32_t i = tile + threadIdx.x;
    const bool valid = i < A.n;
    const uint32_t nb = NB > 0 ? (uint32_t)NB : A.n_bodies;
    const uint32_t row = nb * 3;

    Particle p;
    uint32_t dti = 0;
    if (valid) {
        p.rx = ld<COH>(A.rx + i);
        p.ry = ld<COH>(A.ry + i);
        p.rz = ld<COH>(A.rz + i);
        p.vx = ld<COH>(A.vx + i);
        p.vy = ld<COH>(A.vy + i);
        p.vz = ld<COH>(A.vz + i);
        p.ax = ld<COH>(A.ax + i);
        p.ay = ld<COH>(A.ay + i);
        p.az = ld<COH>(A.az + i);
        p.flags = ld<COH>(A.flags + i);
        // particles flagged before this block keep their death step; the
        // others end the block at its last step unless they are flagged
        // on the way
        dti = p.flags == 0 ? A.step_offset + A.n_steps : ld<COH>(A.dti + i);
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
        st<COH>(A.rx + i, p.rx);
        st<COH>(A.ry + i, p.ry);
        st<COH>(A.rz + i, p.rz);
        st<COH>(A.vx + i, p.vx);
        st<COH>(A.vy + i, p.vy);
        st<COH>(A.vz + i, p.vz);
        st<COH>(A.ax + i, p.ax);
        st<COH>(A.ay + i, p.ay);
        st<COH>(A.az + i, p.az);
        st<COH>(A.flags + i, (uint16_t)p.flags);
        st<COH>(A.dti + i, dti);
    }
}

template <int NB, int FORM>
__global__ __launch_bounds__(256) void wh_block_kernel(BlockArgs A) {
    extern __shared__ double lds[];
    block_tile<NB, FORM>(A, A.first + blockIdx.x * blockDim.x, lds);
}

template <class T>
using ConstPtr = const __attribute__((address_space(4))) T*;

// One task of a queued launch: the members block_tile and the step code read
// (those of BlockArgs), from the launch's shared constants, the task's
// segment and its block's planet constants (referenced where the workgroup
// keeps them, so indexing them needs no private copy).
struct TaskArgs {
    double *rx, *ry, *rz, *vx, *vy, *vz, *ax, *ay, *az;
    uint16_t* flags;
    uint32_t* dti;
    const double* log;
    uint32_t n, n_steps, chunk, step_offset, block_step0, n_bodies, fp64_core;
    double dt, half_dt, mu, outer2, inv_mu, dt_sqrt_mu, inv_sqrt_mu, inv_dt, sun_vote, outer_vote;
    const PlanetConsts& pc;
};

// Several consecutive blocks in one persistent grid.  Each workgroup takes
// the next slot of the task FIFO, runs that task (one tile through one
// segment) and appends the tile's next segment, until every slot is taken.
// The FIFO holds runnable tasks only.  While more tiles are unfinished than
// there are workgroups (every workgroup holds at most one task, and there
// are no more workgroups than tiles), the slot a workgroup takes, at most
// n_tiles + (tasks finished) - 1, has been claimed by the workgroup that
// appends it, so the wait for its entry is the few instructions between that
// claim and the write.  Only at the end of the run, once fewer tiles are
// unfinished than there are workgroups, does a taken slot wait for a running
// task to finish (at most one task).  Every wait ends, and it is bounded in
// real time all the same (wait_ticks): past it the launch records the
// failure in ctl[2], every workgroup stops at its next task, and the host
// reports the error.  FIFO order runs the tiles that are furthest behind
// first, so the tiles finish together.
//
// Versus one launch per segment and slice, a workgroup starts its next task
// at once instead of at the end of its launch, so every SIMD keeps its
// resident waves until the last segments.  The per-particle arithmetic is
// block_tile's and segments start on multiples of kRadiusPeriod steps, so
// results are bit for bit those of any other launch split.
//
// Hand-off between workgroups (possibly on different XCDs, whose L2 caches
// are not coherent with each other): the tiles' state, the FIFO and its
// counters are read and written with system-coherent accesses, which no
// cache keeps, so no cache needs writing back or invalidating (the
// agent-scope release / acquire that would otherwise be needed writes back
// and invalidates a whole L2 at every task).  Every lane waits for its state
// stores to complete before the barrier after which one lane appends the
// tile's next segment; a workgroup that takes an entry issues its state
// loads only after the barrier that follows reading it.
template <int NB, int FORM>
__global__ __launch_bounds__(256) void wh_queue_kernel(QueueArgs Q) {
    extern __shared__ double lds[];
    __shared__ uint32_t grab, lag;
    const uint32_t n_tasks = Q.n_segs * Q.n_tiles;
    const uint32_t row = (NB > 0 ? (uint32_t)NB : Q.base.n_bodies) * 3;
    for (;;) {
        if (threadIdx.x == 0) {
            uint32_t v = 0;  // 0: no task, stop
            if (__hip_atomic_load(&Q.ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) == 0) {
                const uint32_t h = __hip_atomic_fetch_add(&Q.ctl[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                lag = 0;
                if (h < n_tasks) {
                    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
                    while ((v = __hip_atomic_load(&Q.ring[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) == 0) {
                        if (__builtin_amdgcn_s_memrealtime() - t0 > Q.wait_ticks) {
                            __hip_atomic_store(&Q.ctl[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                    }
                    // rounds of n_tiles slots this tile's segment is behind
                    // the slot it was taken from
                    const uint32_t g = (v - 1) / Q.n_tiles, round = h / Q.n_tiles;
                    lag = round > g ? round - g : 0;
                }
            }
            grab = v;
        }
        __syncthreads();
        const uint32_t v = __builtin_amdgcn_readfirstlane(grab);
        if (v == 0) break;
        // A SIMD issues from its highest-priority ready wave, the oldest one
        // first among equals, and a persistent grid's waves keep their ages:
        // left alone, the oldest workgroups run their tasks several times
        // faster than the youngest, tiles taken by slow workgroups fall
        // behind, and the last tiles finish milliseconds after the rest.
        // A tile that is behind raises its waves' priority for its task.
        switch (__builtin_amdgcn_readfirstlane(lag)) {
            case 0: __builtin_amdgcn_s_setprio(0); break;
            case 1: __builtin_amdgcn_s_setprio(1); break;
            case 2: __builtin_amdgcn_s_setprio(2); break;
            default: __builtin_amdgcn_s_setprio(3); break;
        }
        const uint32_t task = v - 1, g = task / Q.n_tiles, tile = task - g * Q.n_tiles;
        // The tables do not change during the launch: read through the
        // constant address space, the task's values (wave-uniform) load
        // into scalar registers as kernel arguments do.
        const ConstPtr<QueueSeg> sg = (ConstPtr<QueueSeg>)(Q.segs + g);
        const uint32_t r0 = sg->r0, seg_steps = sg->n_steps;
        const ConstPtr<QueueBlock> B = (ConstPtr<QueueBlock>)(Q.blocks + sg->block);
        const PlanetConsts& pcs = *(const PlanetConsts*)(&B->pc);
        const BlockArgs& b = Q.base;
        const TaskArgs A{b.rx, b.ry, b.rz, b.vx, b.vy, b.vz, b.ax, b.ay, b.az, b.flags, b.dti,
                         B->log + (size_t)r0 * row, b.n, seg_steps, min(b.chunk, seg_steps),
                         B->step_offset + r0, r0, b.n_bodies, b.fp64_core,
                         b.dt, b.half_dt, b.mu, b.outer2, b.inv_mu, b.dt_sqrt_mu, b.inv_sqrt_mu, b.inv_dt,
                         B->sun_vote, b.outer_vote, pcs};
        // block_tile starts with a barrier, after which thread 0 may write
        // `grab` again: every lane has read it by then
        block_tile<NB, FORM, TaskArgs, true>(A, b.first + tile * 256u, lds);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);  // this lane's state stores have completed
        __syncthreads();
        if (threadIdx.x == 0 && g + 1 < Q.n_segs) {
            const uint32_t slot = __hip_atomic_fetch_add(&Q.ctl[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&Q.ring[slot], v + Q.n_tiles, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Empties a queued launch's FIFO but for every tile's first segment
// (slots [0, n_tiles)) and clears its counters.
__global__ void queue_reset_kernel(uint32_t* ctl, uint32_t* ring, uint32_t n_tiles, uint32_t n_tasks) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ctl[0] = 0;
        ctl[1] = n_tiles;
        ctl[2] = 0;
    }
    if (i < n_tasks) ring[i] = i < n_tiles ? i + 1 : 0;
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

// One persistent grid of wh_queue_kernel<NB, FORM>: as many workgroups as
// fit on the device at once (fewer if there are fewer tasks).
template <int NB, int FORM>
static hipError_t launch_queue_nb(const QueueArgs& Q, size_t lds_bytes, hipStream_t stream) {
    const void* kern = reinterpret_cast<const void*>(&wh_queue_kernel<NB, FORM>);
    int per_cu = 0, cus = 0, dev = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, lds_bytes);
    if (e != hipSuccess) return e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    // no more workgroups than tiles: at most n_tiles tasks are ever runnable
    const uint64_t fit = (uint64_t)(per_cu > 0 ? per_cu : 1) * (uint64_t)(cus > 0 ? cus : 1);
    const uint32_t groups = (uint32_t)(Q.n_tiles < fit ? Q.n_tiles : fit);
    hipLaunchKernelGGL((wh_queue_kernel<NB, FORM>), dim3(groups), dim3(256), lds_bytes, stream, Q);
    return hipGetLastError();
}

template <int FORM>
static hipError_t launch_queue_form(const QueueArgs& Q, size_t lds_bytes, hipStream_t stream) {
    switch (Q.base.n_bodies) {
#define GLX_CASE(NBV) \
    case NBV:         \
        return launch_queue_nb<NBV, FORM>(Q, lds_bytes, stream);
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
            return launch_queue_nb<0, FORM>(Q, lds_bytes, stream);
    }
}

hipError_t launch_queue(const QueueArgs& Q, int form, size_t lds_bytes, hipStream_t stream) {
    if (Q.n_segs == 0 || Q.n_tiles == 0) return hipSuccess;
    const size_t need = (size_t)Q.base.chunk * Q.base.n_bodies * 3 * sizeof(double);
    if (lds_bytes < need) lds_bytes = need;
    const uint32_t n_tasks = Q.n_segs * Q.n_tiles;
    hipLaunchKernelGGL(queue_reset_kernel, dim3((n_tasks + 255) / 256), dim3(256), 0, stream, Q.ctl, Q.ring,
                       Q.n_tiles, n_tasks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (form) {
        case GLX_FORM_FP64:
            return launch_queue_form<GLX_FORM_FP64>(Q, lds_bytes, stream);
        case GLX_FORM_FAST:
            return launch_queue_form<GLX_FORM_FAST>(Q, lds_bytes, stream);
        case GLX_FORM_MIXED:
            return launch_queue_form<GLX_FORM_MIXED>(Q, lds_bytes, stream);
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
