/* GPU check of the short forms of the correctly rounded sqrt and 1/x that the sweeps use on
 * their hit paths (sqrt_rn_short, rcp_rn_short in morphogenesis_amd/csrc/mg_device.h): every
 * float of their ranges against the compiler's own sqrtf and 1.0f / x, built with the library's
 * flags (IEEE division and square root). Test infrastructure, loaded by
 * tests/test_gpu_rounding.py. */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../morphogenesis_amd/csrc/mg_device.h"

__global__ void k_sqrt_check(uint32_t lo, uint64_t n, unsigned long long* bad, uint32_t* first)
{
    unsigned long long cnt = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t u = lo + (uint32_t)i;
        volatile float xv = __uint_as_float(u);
        const float x = xv;
        if (__float_as_uint(sqrt_rn_short(x)) != __float_as_uint(sqrtf(x))) {
            cnt++;
            atomicMin(first, u);
        }
    }
    if (cnt) atomicAdd(bad, cnt);
}

__global__ void k_rcp_check(uint32_t lo, uint64_t n, unsigned long long* bad, uint32_t* first)
{
    unsigned long long cnt = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t u = lo + (uint32_t)i;
        for (uint32_t sign = 0; sign < 2; sign++) {
            volatile float xv = __uint_as_float(u | (sign << 31));
            const float x = xv;
            if (__float_as_uint(rcp_rn_short(x)) != __float_as_uint(1.0f / x)) {
                cnt++;
                atomicMin(first, u);
            }
        }
    }
    if (cnt) atomicAdd(bad, cnt);
}

/* which = 0: sqrt over the bit patterns [lo, hi]; 1: 1/x over +-[lo, hi].
 * out[0] = inputs compared, out[1] = inputs that differ, out[2] = lowest differing pattern
 * (0xFFFFFFFF if none). Returns a hipError_t code. */
extern "C" int mg_rounding_check(int which, uint32_t lo, uint32_t hi, unsigned long long* out)
{
    if (hi < lo) return (int)hipErrorInvalidValue;
    unsigned long long* bad = nullptr;
    uint32_t* first = nullptr;
    hipError_t e = hipMalloc(&bad, sizeof(*bad));
    if (e == hipSuccess) e = hipMalloc(&first, sizeof(*first));
    const unsigned long long zero = 0;
    const uint32_t none = 0xFFFFFFFFu;
    if (e == hipSuccess) e = hipMemcpy(bad, &zero, sizeof(zero), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(first, &none, sizeof(none), hipMemcpyHostToDevice);
    const uint64_t n = (uint64_t)hi - lo + 1u;
    if (e == hipSuccess) {
        if (which == 0)
            hipLaunchKernelGGL(k_sqrt_check, dim3(8192), dim3(256), 0, 0, lo, n, bad, first);
        else
            hipLaunchKernelGGL(k_rcp_check, dim3(8192), dim3(256), 0, 0, lo, n, bad, first);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    unsigned long long h = 0;
    uint32_t f = none;
    if (e == hipSuccess) e = hipMemcpy(&h, bad, sizeof(h), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&f, first, sizeof(f), hipMemcpyDeviceToHost);
    out[0] = n * (which == 0 ? 1u : 2u);
    out[1] = h;
    out[2] = f;
    if (bad) (void)hipFree(bad);
    if (first) (void)hipFree(first);
    return (int)e;
}
