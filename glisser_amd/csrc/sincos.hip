// The fp64 form's sin/cos (correctly rounded, sincos_cr.h) evaluated over
// arrays on the device and, with the same code, on the host.  The kernel
// path is what the Kepler drift of the fp64 form runs; the host twin lets a
// caller (and the CPU tests) reproduce its values without a GPU.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "errors.h"
#include "glisser_amd.h"
#include "sincos_cr.h"

namespace {

__global__ __launch_bounds__(256) void sincos_kernel(const double* __restrict__ x, double* __restrict__ s,
                                                     double* __restrict__ c, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    glx::cr::sincos_cr(x[i], &s[i], &c[i]);
}

int hip_fail(const char* what, hipError_t e) {
    return glx::set_error(GLX_EHIP, std::string(what) + " failed: " + hipGetErrorString(e));
}

}  // namespace

extern "C" {

int glx_sincos_host(uint64_t n, const double* x, double* s, double* c) {
    if (n > 0 && (!x || !s || !c)) return glx::set_error(GLX_EINVAL, "null argument");
    for (uint64_t i = 0; i < n; i++) glx::cr::sincos_cr(x[i], &s[i], &c[i]);
    return GLX_OK;
}

int glx_sincos_device(int device, uint64_t n, const double* x, double* s, double* c) {
    if (n > 0 && (!x || !s || !c)) return glx::set_error(GLX_EINVAL, "null argument");
    if (n == 0) return GLX_OK;
    if (n > (uint64_t)1 << 28) return glx::set_error(GLX_EINVAL, "at most 2^28 arguments per call");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail("hipSetDevice", e);
    double* d = nullptr;
    e = hipMalloc((void**)&d, 3 * n * sizeof(double));
    if (e != hipSuccess) return glx::set_error(GLX_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    int rc = GLX_OK;
    const uint32_t blocks = (uint32_t)((n + 255) / 256);
    if ((e = hipMemcpy(d, x, n * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) {
        rc = hip_fail("hipMemcpy", e);
    } else {
        hipLaunchKernelGGL(sincos_kernel, dim3(blocks), dim3(256), 0, 0, d, d + n, d + 2 * n, n);
        if ((e = hipGetLastError()) != hipSuccess || (e = hipDeviceSynchronize()) != hipSuccess)
            rc = hip_fail("sincos_kernel", e);
        else if ((e = hipMemcpy(s, d + n, n * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess ||
                 (e = hipMemcpy(c, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
            rc = hip_fail("hipMemcpy", e);
    }
    (void)hipFree(d);
    return rc;
}

}  // extern "C"
