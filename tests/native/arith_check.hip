// arith_check.hip -- device check of the division helpers the traversal
// uses in place of IEEE division (rcp_rn, div_rn in sg_device.hpp).  Built
// with the library's flags into a test-only shared library; the GPU test
// calls sg_arith_check through ctypes.
//
// Every mode walks all 2^32 bit patterns x once (grid-stride loop, 2048
// blocks x 256 threads, no buffer indexed by x) and counts results whose
// bits differ from IEEE division (two NaNs count as equal):
//   mode 0: rcp_rn(x)                       vs 1.f / x
//   mode 1: div_rn(a, x, 1/x, range(x))     vs a / x, a pseudo-random with a
//           binary exponent in [-70, 70] and random sign (seed picks a)
//   mode 2: the same with a = +0,  mode 3: with a = -0
#include <hip/hip_runtime.h>

#include "sg_device.hpp"

namespace {

__device__ __forceinline__ bool same_bits(float a, float b) {
    if (a != a && b != b) return true;
    return __float_as_uint(a) == __float_as_uint(b);
}

__global__ __launch_bounds__(256) void k_arith_check(uint32_t mode, uint32_t seed, unsigned long long* count,
                                                     uint32_t* example) {
    unsigned long long bad = 0;
    uint32_t ex = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (1ull << 32); i += stride) {
        const float x = __uint_as_float((uint32_t)i);
        float got, want;
        if (mode == 0) {
            got = sg::rcp_rn(x);
            want = 1.f / x;
        } else {
            float a;
            if (mode == 1) {
                const uint64_t h = sg::hash3(seed, i, 0x5eedull);
                const uint32_t e = (uint32_t)((h >> 8) % 141u) + 127u - 70u;
                a = __uint_as_float((uint32_t)(h >> 32) & 0x807FFFFFu | e << 23);
            } else {
                a = mode == 2 ? 0.f : -0.f;
            }
            got = sg::div_rn(a, x, 1.f / x, sg::div_fast_range(x));
            want = a / x;
        }
        if (!same_bits(got, want)) {
            ++bad;
            ex = (uint32_t)i;
        }
    }
    if (bad) {
        atomicAdd(count, bad);
        atomicExch(example, ex);
    }
}

}  // namespace

extern "C" int sg_arith_check(int device, uint32_t mode, uint32_t seed, unsigned long long* mismatches,
                              uint32_t* example) {
    if (mode > 3 || !mismatches || !example) return 1;
    if (hipSetDevice(device) != hipSuccess) return 2;
    unsigned long long* d_count = nullptr;
    uint32_t* d_ex = nullptr;
    if (hipMalloc(&d_count, sizeof *d_count) != hipSuccess) return 3;
    if (hipMalloc(&d_ex, sizeof *d_ex) != hipSuccess) {
        (void)hipFree(d_count);
        return 3;
    }
    int rc = 0;
    if (hipMemset(d_count, 0, sizeof *d_count) != hipSuccess || hipMemset(d_ex, 0, sizeof *d_ex) != hipSuccess)
        rc = 4;
    if (!rc) {
        hipLaunchKernelGGL(k_arith_check, dim3(2048), dim3(256), 0, 0, mode, seed, d_count, d_ex);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = 5;
    }
    if (!rc && (hipMemcpy(mismatches, d_count, sizeof *d_count, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(example, d_ex, sizeof *d_ex, hipMemcpyDeviceToHost) != hipSuccess))
        rc = 6;
    (void)hipFree(d_count);
    (void)hipFree(d_ex);
    return rc;
}
