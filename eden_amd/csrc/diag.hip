// diag.hip -- self-test of device arithmetic the kernels rely on for bit-exactness:
// div_rn against the compiler's division, over hashed operand pairs and a list of special
// values. Reached through include/edenhip_diag.h; not used by the simulation path.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/edenhip_diag.h"
#include "device_model.hpp"

using namespace edenhip;

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// operand i of pair k: raw random bits, or (every other pair) a random significand and sign
// with an exponent drawn from [80, 175], around the fast path's range [87, 167]
__device__ __forceinline__ float operand(uint64_t seed, uint64_t k, int i) {
    const uint64_t h = mix64(seed ^ mix64(2 * k + (uint64_t)i));
    uint32_t bits = (uint32_t)h;
    if (k & 1) {
        const uint32_t e = 80u + (uint32_t)((h >> 32) % 96u);
        bits = (bits & 0x807FFFFFu) | (e << 23);
    }
    return __uint_as_float(bits);
}

__device__ __forceinline__ bool same(float a, float b) {
    const bool na = a != a, nb = b != b;
    return (na && nb) || __float_as_uint(a) == __float_as_uint(b);
}

__global__ void divide_check(uint64_t seed, int64_t count, const float *__restrict__ special, int32_t n_special,
                             unsigned long long *__restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned long long bad = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
        const float n = operand(seed, (uint64_t)k, 0), d = operand(seed, (uint64_t)k, 1);
        if (!same(div_rn(n, d), n / d)) bad++;
    }
    const int64_t ns = (int64_t)n_special * n_special;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < ns; k += stride) {
        const float n = special[k / n_special], d = special[k % n_special];
        if (!same(div_rn(n, d), n / d)) bad++;
    }
    if (bad) atomicAdd(out, bad);
}

}  // namespace

extern "C" int edenhip_diag_divide(uint64_t seed, int64_t count, const float *special, int32_t n_special,
                                   int64_t *mismatches) {
    if (!mismatches || count < 0 || n_special < 0 || (n_special > 0 && !special)) return -1;
    float *d_special = nullptr;
    unsigned long long *d_out = nullptr;
    const size_t sb = sizeof(float) * (size_t)(n_special > 0 ? n_special : 1);
    if (hipMalloc(&d_special, sb) != hipSuccess) return -2;
    if (hipMalloc(&d_out, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipFree(d_special);
        return -2;
    }
    int rc = 0;
    unsigned long long h_out = 0;
    if ((n_special > 0 && hipMemcpy(d_special, special, sizeof(float) * (size_t)n_special, hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemset(d_out, 0, sizeof(unsigned long long)) != hipSuccess)
        rc = -2;
    if (rc == 0) {
        hipLaunchKernelGGL(divide_check, dim3(4096), dim3(256), 0, 0, seed, count, d_special, n_special, d_out);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(&h_out, d_out, sizeof(h_out), hipMemcpyDeviceToHost) != hipSuccess)
            rc = -2;
    }
    (void)hipFree(d_special);
    (void)hipFree(d_out);
    *mismatches = (int64_t)h_out;
    return rc;
}
