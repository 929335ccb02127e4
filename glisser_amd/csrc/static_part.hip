// glx_partition_static: the stable partition of glx_partition on arrays the
// caller owns, with no device allocation -- its tile counts and offsets live
// in a file-scope device array of fixed size (one copy per GPU, held until
// the process ends).
//
// Host code must not pass `scratch` itself to a kernel: in host code the
// name is a host-side stand-in, not a device address.  The device address
// is looked up once per device with hipGetSymbolAddress (with that device
// current) and every launch gets that pointer.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "errors.h"
#include "glisser_amd.h"
#include "launch.h"

namespace {

constexpr uint32_t kStaticTiles = GLX_PARTITION_STATIC_MAX / glx::kPartTile;
static_assert((uint64_t)kStaticTiles * glx::kPartTile == GLX_PARTITION_STATIC_MAX, "whole tiles");

// tile counts [2 * tiles] then tile offsets [2 * tiles]: 512 KiB
static __device__ __align__(256) uint32_t scratch[4 * kStaticTiles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t* g_ptr[kMaxDevices] = {};

int hip_fail(const char* what, hipError_t e) {
    return glx::set_error(GLX_EHIP, std::string(what) + " failed: " + hipGetErrorString(e));
}

// Device address of this device's `scratch` (looked up on the first call
// made with the device current).
int scratch_ptr(uint32_t** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail("hipGetDevice", e);
    if (dev < 0 || dev >= kMaxDevices) return glx::set_error(GLX_EINVAL, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail("hipGetSymbolAddress", e);
        g_ptr[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_ptr[dev];
    return GLX_OK;
}

bool complete(const glx_device_particles* p) {
    for (int k = 0; k < 3; k++)
        if (!p->r[k] || !p->v[k] || !p->a[k]) return false;
    return p->flags && p->deathtime_index && p->id;
}

glx::ParticleSet as_set(const glx_device_particles* p) {
    glx::ParticleSet s{};
    for (int k = 0; k < 3; k++) {
        s.d[k] = p->r[k];
        s.d[3 + k] = p->v[k];
        s.d[6 + k] = p->a[k];
    }
    s.flags = p->flags;
    s.dti = p->deathtime_index;
    s.id = p->id;
    return s;
}

}  // namespace

extern "C" int glx_partition_static(const glx_device_particles* in, const glx_device_particles* out, uint64_t n,
                                    int mode, uint32_t* class_sizes, void* stream) {
    if (!in || !out || !class_sizes) return glx::set_error(GLX_EINVAL, "null argument");
    if (mode != GLX_PART_UNFLAGGED && mode != GLX_PART_THREE_WAY)
        return glx::set_error(GLX_EINVAL, "partition mode " + std::to_string(mode) + " unknown");
    if (n > GLX_PARTITION_STATIC_MAX)
        return glx::set_error(GLX_EINVAL, std::to_string(n) + " particles exceed the " +
                                              std::to_string(GLX_PARTITION_STATIC_MAX) +
                                              " a glx_partition_static call can hold");
    if (!complete(in) || !complete(out)) return glx::set_error(GLX_EINVAL, "a particle array pointer is null");
    if (in->flags == out->flags || in->r[0] == out->r[0])
        return glx::set_error(GLX_EINVAL, "input and output arrays must be distinct");
    uint32_t* ptr = nullptr;
    if (int rc = scratch_ptr(&ptr)) return rc;
    const uint32_t tiles = glx::part_tiles(n);
    hipError_t e = glx::launch_partition_into(as_set(in), as_set(out), (uint32_t)n, mode == GLX_PART_THREE_WAY, ptr,
                                              ptr + 2 * tiles, class_sizes, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail("glx_partition_static launch", e);
    return GLX_OK;
}
