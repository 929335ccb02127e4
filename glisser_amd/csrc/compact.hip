// Order-preserving particle compaction and the small companion kernels of
// the executor's time-block loop.
//
// The reference partitions with thrust::partition (src/executor.cu:41-61,
// 663-665,780-797), which leaves the order inside each class unspecified.
// Here the partition is stable -- byte-identical to std::stable_partition
// over the same flags (src/data.cpp:17-39) -- in three passes:
//   1. per-tile class counts (1024 elements per workgroup),
//   2. one-workgroup exclusive scan of the tile counts,
//   3. per-tile scatter, ranks from wave64 ballots + mbcnt-style popcounts
//      and a 4-wave prefix in LDS.
// The particle set is first copied to a scratch set, then scattered back,
// so the arrays keep their device addresses.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "launch.h"

namespace glx {

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kItems = kPartTile / kThreads;  // 4

__device__ __forceinline__ uint32_t part_class(uint32_t f, int three_way) {
    if (f == 0) return 0;
    if (three_way && (f & 0xFEu) == 0) return 1;
    return 2;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__global__ __launch_bounds__(256) void part_count_kernel(const uint16_t* __restrict__ flags, uint32_t n,
                                                         int three_way, uint32_t* __restrict__ counts,
                                                         uint32_t tiles) {
    __shared__ uint32_t red[2][kThreads / 64];
    const uint32_t base = blockIdx.x * kPartTile;
    uint32_t c0 = 0, c1 = 0;
#pragma unroll
    for (uint32_t k = 0; k < kItems; k++) {
        const uint32_t idx = base + k * kThreads + threadIdx.x;
        if (idx < n) {
            const uint32_t c = part_class(flags[idx], three_way);
            c0 += c == 0;
            c1 += c == 1;
        }
    }
    c0 = wave_sum(c0);
    c1 = wave_sum(c1);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wave] = c0;
        red[1][wave] = c1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s0 = 0, s1 = 0;
        for (uint32_t w = 0; w < kThreads / 64; w++) {
            s0 += red[0][w];
            s1 += red[1][w];
        }
        counts[blockIdx.x] = s0;
        counts[tiles + blockIdx.x] = s1;
    }
}

// Exclusive scan of the two tile-count rows; one workgroup of 1024 lanes.
__global__ __launch_bounds__(1024) void part_scan_kernel(const uint32_t* __restrict__ counts,
                                                         uint32_t* __restrict__ offsets,
                                                         uint32_t* __restrict__ totals, uint32_t tiles) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry_s;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t cls = 0; cls < 2; cls++) {
        if (threadIdx.x == 0) carry_s = 0;
        __syncthreads();
        for (uint32_t base = 0; base < tiles; base += 1024) {
            const uint32_t idx = base + threadIdx.x;
            const uint32_t v = idx < tiles ? counts[cls * tiles + idx] : 0;
            uint32_t incl = v;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t t = __shfl_up(incl, off, 64);
                if (lane >= (uint32_t)off) incl += t;
            }
            if (lane == 63) wsum[wave] = incl;
            __syncthreads();
            uint32_t before = 0, total = 0;
            for (uint32_t w = 0; w < 16; w++) {
                const uint32_t s = wsum[w];
                if (w < wave) before += s;
                total += s;
            }
            const uint32_t carry = carry_s;
            if (idx < tiles) offsets[cls * tiles + idx] = carry + before + incl - v;
            __syncthreads();
            if (threadIdx.x == 0) carry_s = carry + total;
            __syncthreads();
        }
        if (threadIdx.x == 0) totals[cls] = carry_s;
        __syncthreads();
    }
}

struct CopyArgs {
    ParticleSet src, dst;
    RollbackSet rsrc, rdst;
    int has_rb;
};

__device__ __forceinline__ void copy_one(const CopyArgs& c, uint32_t from, uint32_t to) {
#pragma unroll
    for (int k = 0; k < 9; k++) c.dst.d[k][to] = c.src.d[k][from];
    c.dst.flags[to] = c.src.flags[from];
    c.dst.dti[to] = c.src.dti[from];
    c.dst.id[to] = c.src.id[from];
    if (c.has_rb) {
#pragma unroll
        for (int k = 0; k < 6; k++) c.rdst.d[k][to] = c.rsrc.d[k][from];
        c.rdst.flags[to] = c.rsrc.flags[from];
        c.rdst.dti[to] = c.rsrc.dti[from];
        c.rdst.id[to] = c.rsrc.id[from];
    }
}

__global__ __launch_bounds__(256) void copy_kernel(CopyArgs c, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) copy_one(c, i, i);
}

__global__ __launch_bounds__(256) void part_scatter_kernel(CopyArgs c, uint32_t n, int three_way,
                                                           const uint32_t* __restrict__ offsets,
                                                           const uint32_t* __restrict__ totals, uint32_t tiles) {
    __shared__ uint32_t wcount[kThreads / 64][3];
    const uint32_t base = blockIdx.x * kPartTile;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t n0 = totals[0];
    const uint32_t n1 = three_way ? totals[1] : 0u;
    const uint32_t o0 = offsets[blockIdx.x];
    const uint32_t o1 = three_way ? offsets[tiles + blockIdx.x] : 0u;
    uint32_t run[3] = {o0, n0 + o1, n0 + n1 + (base - o0 - o1)};
#pragma unroll
    for (uint32_t k = 0; k < kItems; k++) {
        const uint32_t idx = base + k * kThreads + threadIdx.x;
        const bool valid = idx < n;
        const uint32_t cl = valid ? part_class(c.src.flags[idx], three_way) : 3u;
        const uint64_t m0 = __ballot(cl == 0);
        const uint64_t m1 = __ballot(cl == 1);
        const uint64_t m2 = __ballot(cl == 2);
        if (lane == 0) {
            wcount[wave][0] = __popcll(m0);
            wcount[wave][1] = __popcll(m1);
            wcount[wave][2] = __popcll(m2);
        }
        __syncthreads();
        uint32_t pre[3] = {0, 0, 0}, tot[3] = {0, 0, 0};
#pragma unroll
        for (uint32_t w = 0; w < kThreads / 64; w++) {
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const uint32_t v = wcount[w][q];
                if (w < wave) pre[q] += v;
                tot[q] += v;
            }
        }
        if (valid) {
            const uint64_t mine = cl == 0 ? m0 : (cl == 1 ? m1 : m2);
            const uint32_t dest = run[cl] + pre[cl] + (uint32_t)__popcll(mine & below);
            copy_one(c, idx, dest);
        }
#pragma unroll
        for (int q = 0; q < 3; q++) run[q] += tot[q];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void kill_encounters_kernel(uint16_t* flags, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t f = flags[i];
    if ((f & 0x01u) == 0x01u) flags[i] = (uint16_t)((f & 0xFF00u) | 0x80u);
}

__global__ __launch_bounds__(256) void snapshot_kernel(ParticleSet s, RollbackSet r, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < 6; k++) r.d[k][i] = s.d[k][i];
        r.flags[i] = s.flags[i];
        r.dti[i] = s.dti[i];
        r.id[i] = s.id[i];
    }
}

// Particles [begin, begin + n) as host close-encounter records: state at the
// start of the block from the rollback copy (or the current state when no
// copy is kept), state now, flags, id and step count.
__global__ __launch_bounds__(256) void pack_encounters_kernel(ParticleSet s, RollbackSet r, int has_rb,
                                                              uint32_t begin, uint32_t n, double t_block,
                                                              glx_encounter* out) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t i = begin + k;
        glx_encounter e;
        e.id = s.id[i];
        e.flags = s.flags[i];
        e.step = s.dti[i];
        e.t_block = t_block;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            e.r[c] = s.d[c][i];
            e.v[c] = s.d[3 + c][i];
            e.r0[c] = has_rb ? r.d[c][i] : e.r[c];
            e.v0[c] = has_rb ? r.d[3 + c][i] : e.v[c];
        }
        out[k] = e;
    }
}

// One workgroup sums dti[0, n) into *out (no atomics, nothing to clear).
// Sum of dti over a grid-strided share of [0, n) per workgroup into
// part[blockIdx.x]; then one workgroup adds the parts into *out (no atomics,
// nothing to clear).
__device__ __forceinline__ unsigned long long block_sum_1024(unsigned long long s) {
    __shared__ unsigned long long part[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < 16; w++) t += part[w];
    return t;
}

__global__ __launch_bounds__(1024) void sum_steps_kernel(const uint32_t* __restrict__ dti, uint32_t n,
                                                         unsigned long long* part) {
    unsigned long long s = 0;
    for (uint32_t i = blockIdx.x * 1024 + threadIdx.x; i < n; i += gridDim.x * 1024) s += dti[i];
    const unsigned long long t = block_sum_1024(s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(1024) void sum_parts_kernel(const unsigned long long* __restrict__ part, uint32_t n,
                                                         unsigned long long* out) {
    const unsigned long long t = block_sum_1024(threadIdx.x < n ? part[threadIdx.x] : 0);
    if (threadIdx.x == 0) *out = t;
}

__global__ __launch_bounds__(256) void aos_to_soa_kernel(const double* __restrict__ aos, double* x, double* y,
                                                         double* z, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = aos[3 * (size_t)i];
    y[i] = aos[3 * (size_t)i + 1];
    z[i] = aos[3 * (size_t)i + 2];
}

__global__ __launch_bounds__(256) void soa_to_aos_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                         const double* __restrict__ z, double* aos, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    aos[3 * (size_t)i] = x[i];
    aos[3 * (size_t)i + 1] = y[i];
    aos[3 * (size_t)i + 2] = z[i];
}

inline uint32_t grid_for(uint32_t n) {
    uint32_t g = (n + kThreads - 1) / kThreads;
    return g == 0 ? 1 : g;
}

}  // namespace

hipError_t launch_partition(ParticleSet cur, ParticleSet tmp, const RollbackSet* rb, const RollbackSet* tmp_rb,
                            uint32_t n, int three_way, uint32_t* counts, uint32_t* offsets, uint32_t* totals,
                            hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t tiles = part_tiles(n);
    CopyArgs to_tmp{cur, tmp, {}, {}, 0};
    if (rb && tmp_rb) {
        to_tmp.rsrc = *rb;
        to_tmp.rdst = *tmp_rb;
        to_tmp.has_rb = 1;
    }
    hipLaunchKernelGGL(part_count_kernel, dim3(tiles), dim3(kThreads), 0, stream, cur.flags, n, three_way, counts,
                       tiles);
    hipLaunchKernelGGL(part_scan_kernel, dim3(1), dim3(1024), 0, stream, counts, offsets, totals, tiles);
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(kThreads), 0, stream, to_tmp, n);
    CopyArgs back{tmp, cur, to_tmp.rdst, to_tmp.rsrc, to_tmp.has_rb};
    hipLaunchKernelGGL(part_scatter_kernel, dim3(tiles), dim3(kThreads), 0, stream, back, n, three_way, offsets,
                       totals, tiles);
    return hipGetLastError();
}

hipError_t launch_partition_into(ParticleSet src, ParticleSet dst, uint32_t n, int three_way, uint32_t* counts,
                                 uint32_t* offsets, uint32_t* totals, hipStream_t stream) {
    const uint32_t tiles = part_tiles(n);
    if (tiles > 0)
        hipLaunchKernelGGL(part_count_kernel, dim3(tiles), dim3(kThreads), 0, stream, src.flags, n, three_way, counts,
                           tiles);
    // with no tile the scan only writes the two class sizes (zero)
    hipLaunchKernelGGL(part_scan_kernel, dim3(1), dim3(1024), 0, stream, counts, offsets, totals, tiles);
    if (tiles > 0) {
        CopyArgs c{src, dst, {}, {}, 0};
        hipLaunchKernelGGL(part_scatter_kernel, dim3(tiles), dim3(kThreads), 0, stream, c, n, three_way, offsets,
                           totals, tiles);
    }
    return hipGetLastError();
}

hipError_t launch_kill_encounters(uint16_t* flags, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kill_encounters_kernel, dim3(grid_for(n)), dim3(kThreads), 0, stream, flags, n);
    return hipGetLastError();
}

hipError_t launch_snapshot(ParticleSet cur, RollbackSet rb, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(snapshot_kernel, dim3(grid_for(n)), dim3(kThreads), 0, stream, cur, rb, n);
    return hipGetLastError();
}

hipError_t launch_pack_encounters(ParticleSet cur, const RollbackSet* rb, uint32_t begin, uint32_t n, double t_block,
                                  glx_encounter* out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const RollbackSet none{};
    hipLaunchKernelGGL(pack_encounters_kernel, dim3(grid_for(n)), dim3(kThreads), 0, stream, cur, rb ? *rb : none,
                       rb ? 1 : 0, begin, n, t_block, out);
    return hipGetLastError();
}

hipError_t launch_sum_steps(const uint32_t* dti, uint32_t n, unsigned long long* out, hipStream_t stream) {
    const uint32_t groups = std::max(1u, std::min(kSumParts, (n + 8191) / 8192));
    hipLaunchKernelGGL(sum_steps_kernel, dim3(groups), dim3(1024), 0, stream, dti, n, out + 1);
    hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(1024), 0, stream, out + 1, groups, out);
    return hipGetLastError();
}

hipError_t launch_aos_to_soa(const double* aos, double* x, double* y, double* z, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(aos_to_soa_kernel, dim3(grid_for(n)), dim3(kThreads), 0, stream, aos, x, y, z, n);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(const double* x, const double* y, const double* z, double* aos, uint32_t n,
                             hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(soa_to_aos_kernel, dim3(grid_for(n)), dim3(kThreads), 0, stream, x, y, z, aos, n);
    return hipGetLastError();
}

}  // namespace glx
