// Track snapshots formed on the device: every alive particle's orbital
// elements (or Cartesian state) as float32, ordered by particle id, packed
// into the 28-byte records of the binary track (u32 id + 6 float32).
//
// The reference forms each track snapshot on the host: a copy of the alive
// particles, sort_by_id, then to_elements per particle in save_binary_track
// (targets/main.cpp:148-156,231-246, src/data.cpp:718-815).  Here one
// launch computes the elements next to the particles and writes each
// record straight to its place in id order: ids are ranked by an
// exclusive scan over a presence table (ids are unique and, on this path,
// below 4x the particle capacity), so no sort runs at all.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "elem_cr.h"
#include "errors.h"
#include "glisser_amd.h"
#include "launch.h"

namespace glx {

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kScanTile = 1024;  // ids per workgroup of the rank scan

uint32_t grid_for(uint64_t n) {
    const uint64_t g = (n + kThreads - 1) / kThreads;
    return (uint32_t)(g < 65535 ? g : 65535);
}

// Osculating elements about central GM mu, in the operation order of the
// host conversion (src/convert.cpp:310-463); written to a[6] as
// a, e, i, Omega, omega, f.  Its atan2, acos, asin, sin and cos are the
// correctly rounded ones (elem_cr.h, sincos_cr.h), which glibc's results
// equal in all but rare cases; pi is 2 asin(1) rounded, as the host has it.
// The same code runs on the host (glx_track_elements_host).
__host__ __device__ void elements_of(double mu, double x, double y, double z, double vx, double vy, double vz,
                                     double* out) {
#pragma clang fp contract(off)
    constexpr double pi = 0x1.921fb54442d18p+1;
    const double prec = 1.0e-13;
    const double hx = y * vz - z * vy;
    const double hy = z * vx - x * vz;
    const double hz = x * vy - y * vx;
    const double hsq = hx * hx + hy * hy + hz * hz;
    if (!(hsq > prec)) {  // radial orbit
        out[0] = sqrt(x * x + y * y + z * z);
        out[1] = HUGE_VAL;
        out[2] = cr::asin_cr(z / sqrt(x * x + y * y + z * z));
        out[3] = cr::atan2_cr(y, x);
        out[4] = HUGE_VAL;
        out[5] = HUGE_VAL;
        return;
    }
    const double i = cr::acos_cr(hz / sqrt(hsq));
    const double capom = (fabs(i) < prec || fabs(pi - fabs(i)) < prec) ? 0.0 : cr::atan2_cr(hx, -hy);
    const double vsq = vx * vx + vy * vy + vz * vz;
    const double vdotr = x * vx + y * vy + z * vz;
    const double rr = sqrt(x * x + y * y + z * z);
    const double xhat = x / rr, yhat = y / rr, zhat = z / rr;
    double nx, ny;
    cr::sincos_cr(capom, &ny, &nx);
    double fac = vsq * rr - mu;
    const double Px = fac * xhat - vdotr * vx;
    const double Py = fac * yhat - vdotr * vy;
    const double Pz = fac * zhat - vdotr * vz;
    const double modP = sqrt(Px * Px + Py * Py + Pz * Pz);
    const double e = modP / mu;
    double om;
    if (fabs(e) < prec) {
        om = 0.0;
    } else if ((i < prec) || (pi - i < prec)) {
        om = cr::atan2_cr(Py, Px);
    } else {
        const double ecosw = (nx * Px + ny * Py) / mu;
        om = cr::acos_cr(ecosw / e);
        if (fabs(Pz) > prec) om *= fabs(Pz) / Pz;
    }
    const double energy = vsq / 2.0 - mu / rr;
    double a, f;
    if (fabs(energy) < prec) {
        a = 0.5 * hsq / mu;
        f = fabs(vdotr) < prec ? 0.0 : 2.0 * cr::acos_cr(sqrt(a / rr)) * vdotr / fabs(vdotr);
    } else if (energy > 0.0) {
        a = -0.5 * mu / energy;
        if (fabs(vdotr) < prec) {
            f = 0.0;
        } else {
            fac = a * (1.0 - e * e) / rr - 1.0;
            f = cr::acos_cr(fac / e) * vdotr / fabs(vdotr);
        }
    } else {
        a = -0.5 * mu / energy;
        if (fabs(e) > prec) {
            if (fabs(vdotr) < prec) {
                f = rr < a ? 0.0 : pi;
            } else {
                fac = a * (1.0 - e * e) / rr - 1.0;
                f = cr::acos_cr(fac / e) * vdotr / fabs(vdotr);
                if (f != f) f = rr < a ? 0.0 : pi;  // NaN
            }
        } else {
            fac = (x * nx + y * ny) / rr;
            f = cr::acos_cr(fac);
            if (fabs(z) > prec) {
                f *= fabs(z) / z;
            } else if ((i < prec) || (pi - i < prec)) {
                double si, ci;
                cr::sincos_cr(i, &si, &ci);
                f = cr::atan2_cr(y, x) * ci;
            }
        }
    }
    out[0] = a;
    out[1] = e;
    out[2] = i;
    out[3] = capom;
    out[4] = om;
    out[5] = f;
}

__global__ __launch_bounds__(kThreads) void clear_u32_kernel(uint32_t* p, uint32_t n) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = 0;
}

__global__ __launch_bounds__(kThreads) void mark_ids_kernel(const uint32_t* __restrict__ id, uint32_t n,
                                                            uint32_t* present) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) present[id[k]] = 1;
}

// Sum of each kScanTile-id tile of the presence table.
__global__ __launch_bounds__(kThreads) void tile_sums_kernel(const uint32_t* __restrict__ present, uint32_t n_ids,
                                                             uint32_t* sums) {
    __shared__ uint32_t part[kThreads / 64];
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (uint32_t k = threadIdx.x; k < kScanTile; k += kThreads)
        if (base + k < n_ids) s += present[base + k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kThreads / 64; w++) t += part[w];
        sums[blockIdx.x] = t;
    }
}

// Exclusive scan of the tile sums in one workgroup of 1024 lanes: each lane
// sums its slice serially, the lanes' sums are scanned by wave shuffles
// (inclusive within each wave, then over the 16 wave totals), and each lane
// writes its slice's exclusive prefixes.
__global__ __launch_bounds__(1024) void scan_tiles_kernel(uint32_t* sums, uint32_t n_tiles) {
    __shared__ uint32_t wave_tot[16];
    const uint32_t per = (n_tiles + 1023) / 1024;
    const uint32_t b = threadIdx.x * per;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t s = 0;
    for (uint32_t k = b; k < b + per && k < n_tiles; k++) s += sums[k];
    uint32_t inc = s;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
    }
    if (lane == 63) wave_tot[w] = inc;
    __syncthreads();
    if (w == 0) {
        uint32_t t = lane < 16 ? wave_tot[lane] : 0;
        for (uint32_t off = 1; off < 16; off <<= 1) {
            const uint32_t y = __shfl_up(t, off, 64);
            if (lane >= off) t += y;
        }
        if (lane < 16) wave_tot[lane] = t;
    }
    __syncthreads();
    uint32_t run = inc - s + (w ? wave_tot[w - 1] : 0);
    for (uint32_t k = b; k < b + per && k < n_tiles; k++) {
        const uint32_t x = sums[k];
        sums[k] = run;
        run += x;
    }
}

// rank[id] = number of present ids below id (exclusive scan within each
// tile plus the tile's offset).  One workgroup of kThreads lanes per tile,
// four ids per lane.
__global__ __launch_bounds__(kThreads) void rank_ids_kernel(uint32_t* present, uint32_t n_ids,
                                                            const uint32_t* __restrict__ offsets) {
    __shared__ uint32_t part[kThreads / 64];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4;
    uint32_t v[4];
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        v[q] = base + q < n_ids ? present[base + q] : 0;
        s += v[q];
    }
    // inclusive scan of s across the wave, then across the waves
    const uint32_t lane = threadIdx.x & 63;
    uint32_t inc = s;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off, 64);
        if (lane >= (uint32_t)off) inc += y;
    }
    if (lane == 63) part[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint32_t wave_off = 0;
    for (uint32_t w = 0; w < (threadIdx.x >> 6); w++) wave_off += part[w];
    uint32_t run = offsets[blockIdx.x] + wave_off + inc - s;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (base + q < n_ids) present[base + q] = run;
        run += v[q];
    }
}

__global__ __launch_bounds__(kThreads) void pack_track_kernel(ParticleSet s, uint32_t n, const uint32_t* __restrict__ rank,
                                                              TrackCenter c, int to_elements, TrackRecordDev* out) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
#pragma clang fp contract(off)
        const uint32_t id = s.id[k];
        const double x = s.d[0][k], y = s.d[1][k], z = s.d[2][k];
        const double vx = s.d[3][k], vy = s.d[4][k], vz = s.d[5][k];
        TrackRecordDev r;
        r.id = id;
        if (to_elements) {
            double el[6];
            elements_of(c.mass, x - c.r[0], y - c.r[1], z - c.r[2], vx - c.v[0], vy - c.v[1], vz - c.v[2], el);
#pragma unroll
            for (int q = 0; q < 6; q++) r.x[q] = (float)el[q];
        } else {
            r.x[0] = (float)x;
            r.x[1] = (float)y;
            r.x[2] = (float)z;
            r.x[3] = (float)vx;
            r.x[4] = (float)vy;
            r.x[5] = (float)vz;
        }
        out[rank ? rank[id] : k] = r;
    }
}

// Elements of n Cartesian rows (x, y, z, vx, vy, vz) into n rows of six.
__global__ __launch_bounds__(kThreads) void elements_kernel(double mu, const double* __restrict__ rv, double* el,
                                                            uint64_t n) {
    const uint64_t k = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (k >= n) return;
    const double* q = rv + 6 * k;
    elements_of(mu, q[0], q[1], q[2], q[3], q[4], q[5], el + 6 * k);
}

__host__ __device__ inline double cr_math(int fn, double a, double b) {
    return fn == GLX_CR_ATAN2 ? cr::atan2_cr(a, b) : fn == GLX_CR_ACOS ? cr::acos_cr(a) : cr::asin_cr(a);
}

__global__ __launch_bounds__(kThreads) void cr_math_kernel(int fn, const double* __restrict__ a,
                                                           const double* __restrict__ b, double* out, uint64_t n) {
    const uint64_t k = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (k >= n) return;
    out[k] = cr_math(fn, a[k], fn == GLX_CR_ATAN2 ? b[k] : 0.0);
}

// max over id[0, n) folded into *out (cleared beforehand): each workgroup
// reduces its grid-stride share, then one atomic max per workgroup.
__global__ __launch_bounds__(kThreads) void max_id_kernel(const uint32_t* __restrict__ id, uint32_t n, uint32_t* out) {
    __shared__ uint32_t part[kThreads / 64];
    uint32_t m = 0;
    for (uint32_t k = blockIdx.x * kThreads + threadIdx.x; k < n; k += gridDim.x * kThreads) m = max(m, id[k]);
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kThreads / 64; w++) t = max(t, part[w]);
        atomicMax(out, t);
    }
}

}  // namespace

hipError_t launch_max_id(const uint32_t* id, uint32_t n, uint32_t* out, hipStream_t stream) {
    const uint32_t g = grid_for(n);
    hipLaunchKernelGGL(clear_u32_kernel, dim3(1), dim3(kThreads), 0, stream, out, 1u);
    hipLaunchKernelGGL(max_id_kernel, dim3(g < 1 ? 1 : g > 256 ? 256 : g), dim3(kThreads), 0, stream, id, n, out);
    return hipGetLastError();
}

uint32_t rank_scratch_tiles(uint32_t n_ids) { return (n_ids + kScanTile - 1) / kScanTile; }

hipError_t launch_rank_ids(const uint32_t* id, uint32_t n, uint32_t n_ids, uint32_t* present, uint32_t* tile_sums,
                           hipStream_t stream) {
    const uint32_t tiles = rank_scratch_tiles(n_ids);
    hipLaunchKernelGGL(clear_u32_kernel, dim3(grid_for(n_ids)), dim3(kThreads), 0, stream, present, n_ids);
    if (n) hipLaunchKernelGGL(mark_ids_kernel, dim3(grid_for(n)), dim3(kThreads), 0, stream, id, n, present);
    hipLaunchKernelGGL(tile_sums_kernel, dim3(tiles), dim3(kThreads), 0, stream, present, n_ids, tile_sums);
    hipLaunchKernelGGL(scan_tiles_kernel, dim3(1), dim3(1024), 0, stream, tile_sums, tiles);
    hipLaunchKernelGGL(rank_ids_kernel, dim3(tiles), dim3(kThreads), 0, stream, present, n_ids, tile_sums);
    return hipGetLastError();
}

hipError_t launch_pack_track(ParticleSet cur, uint32_t n, const uint32_t* rank, const TrackCenter& c, int to_elements,
                             TrackRecordDev* out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pack_track_kernel, dim3(grid_for(n)), dim3(kThreads), 0, stream, cur, n, rank, c, to_elements,
                       out);
    return hipGetLastError();
}

}  // namespace glx

namespace {

int hip_fail(const char* what, hipError_t e) {
    return glx::set_error(GLX_EHIP, std::string(what) + " failed: " + hipGetErrorString(e));
}

// Runs `launch(device_in, device_out)` on GPU `device` over n_in input and
// n_out output doubles copied in and out synchronously.
template <class F>
int run_on_device(int device, const double* in, uint64_t n_in, double* out, uint64_t n_out, F launch) {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail("hipSetDevice", e);
    double* d = nullptr;
    e = hipMalloc((void**)&d, (n_in + n_out) * sizeof(double));
    if (e != hipSuccess) return glx::set_error(GLX_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    int rc = GLX_OK;
    if ((e = hipMemcpy(d, in, n_in * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) {
        rc = hip_fail("hipMemcpy", e);
    } else {
        launch(d, d + n_in);
        if ((e = hipGetLastError()) != hipSuccess || (e = hipDeviceSynchronize()) != hipSuccess)
            rc = hip_fail("kernel", e);
        else if ((e = hipMemcpy(out, d + n_in, n_out * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
            rc = hip_fail("hipMemcpy", e);
    }
    (void)hipFree(d);
    return rc;
}

constexpr uint64_t kMaxRows = (uint64_t)1 << 26;

int check_cr_args(int fn, uint64_t n, const double* a, const double* b, const double* out) {
    if (fn != GLX_CR_ATAN2 && fn != GLX_CR_ACOS && fn != GLX_CR_ASIN)
        return glx::set_error(GLX_EINVAL, "fn must be GLX_CR_ATAN2, GLX_CR_ACOS or GLX_CR_ASIN");
    if (n > 0 && (!a || !out || (fn == GLX_CR_ATAN2 && !b))) return glx::set_error(GLX_EINVAL, "null argument");
    return GLX_OK;
}

}  // namespace

extern "C" {

int glx_track_elements_host(uint64_t n, double mu, const double* rv, double* el) {
    if (n > 0 && (!rv || !el)) return glx::set_error(GLX_EINVAL, "null argument");
    for (uint64_t k = 0; k < n; k++)
        glx::elements_of(mu, rv[6 * k], rv[6 * k + 1], rv[6 * k + 2], rv[6 * k + 3], rv[6 * k + 4], rv[6 * k + 5],
                         el + 6 * k);
    return GLX_OK;
}

int glx_track_elements_device(int device, uint64_t n, double mu, const double* rv, double* el) {
    if (n > 0 && (!rv || !el)) return glx::set_error(GLX_EINVAL, "null argument");
    if (n == 0) return GLX_OK;
    if (n > kMaxRows) return glx::set_error(GLX_EINVAL, "at most 2^26 states per call");
    return run_on_device(device, rv, 6 * n, el, 6 * n, [&](const double* din, double* dout) {
        hipLaunchKernelGGL(glx::elements_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0, mu, din, dout,
                           n);
    });
}

int glx_cr_math_host(int fn, uint64_t n, const double* a, const double* b, double* out) {
    if (int rc = check_cr_args(fn, n, a, b, out)) return rc;
    for (uint64_t k = 0; k < n; k++) out[k] = glx::cr_math(fn, a[k], fn == GLX_CR_ATAN2 ? b[k] : 0.0);
    return GLX_OK;
}

int glx_cr_math_device(int device, int fn, uint64_t n, const double* a, const double* b, double* out) {
    if (int rc = check_cr_args(fn, n, a, b, out)) return rc;
    if (n == 0) return GLX_OK;
    if (n > kMaxRows) return glx::set_error(GLX_EINVAL, "at most 2^26 values per call");
    // one device buffer: a, then b for atan2
    const bool two = fn == GLX_CR_ATAN2;
    std::vector<double> host_in(a, a + n);
    if (two) host_in.insert(host_in.end(), b, b + n);
    return run_on_device(device, host_in.data(), host_in.size(), out, n, [&](const double* din, double* dout) {
        hipLaunchKernelGGL(glx::cr_math_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0, fn, din,
                           two ? din + n : din, dout, n);
    });
}

}  // extern "C"
