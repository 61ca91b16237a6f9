// Synthetic-input generator on the device: Philox4x32-10 (Salmon, Moraes,
// Dror, Shaw, SC'11) keyed by (seed, global index), the same stream layout as
// the host generator in a2afft/inputs.py.  Uniform values are exact
// integer-derived doubles, so device and host agree bit for bit; normal
// values use the device's log/sin/cos and agree to rounding.
#include <algorithm>

#include "common.h"

namespace a2a {

__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
        const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}

__device__ __forceinline__ uint64_t bits53(uint32_t hi, uint32_t lo) {
    return ((uint64_t(hi) << 32) | lo) >> 11;
}

template <typename V, int KIND>
__device__ __forceinline__ void fill_one(V* __restrict__ out, uint64_t i, uint64_t seed, uint64_t g0) {
    const uint64_t g = g0 + i;
    uint32_t c[4] = {uint32_t(g), uint32_t(g >> 32), uint32_t(KIND), 0u};
    philox(c, uint32_t(seed), uint32_t(seed >> 32));
    double re, im;
    if (KIND == 0) {
        re = double(bits53(c[0], c[1])) * (1.0 / 4503599627370496.0) - 1.0;
        im = double(bits53(c[2], c[3])) * (1.0 / 4503599627370496.0) - 1.0;
    } else {
        const double u1 = double(bits53(c[0], c[1]) + 1) * (1.0 / 9007199254740992.0);
        const double u2 = double(bits53(c[2], c[3])) * (1.0 / 9007199254740992.0);
        const double r = sqrt(-2.0 * log(u1));
        double s, co;
        sincospi(2.0 * u2, &s, &co);
        re = r * co;
        im = r * s;
    }
    V v;
    v.x = decltype(v.x)(re);
    v.y = decltype(v.y)(im);
    out[i] = v;
}

// Grid-stride: a dispatch holds fewer than 2^32 work-items, so grids of more
// elements loop (each element's value depends on its index only).
template <typename V, int KIND>
__global__ __launch_bounds__(256) void fill_kernel(V* __restrict__ out, uint64_t n, uint64_t seed, uint64_t g0) {
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256)
        fill_one<V, KIND>(out, i, seed, g0);
}

}  // namespace a2a

using namespace a2a;

extern "C" int a2afft_fill_random(void* out, a2afft_dtype dtype, int64_t n, uint64_t seed, int kind, int64_t g0,
                                  void* stream) {
    if (!out || n < 0 || g0 < 0 || (kind != 0 && kind != 1) || (dtype != A2AFFT_C128 && dtype != A2AFFT_C64))
        return set_error(A2AFFT_E_INVALID_ARG, "fill_random: bad arguments");
    if (n == 0) return A2AFFT_OK;
    const uint64_t blocks = std::min<uint64_t>((uint64_t(n) + 255) / 256, kStrideBlocks);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == A2AFFT_C128) {
        if (kind == 0) hipLaunchKernelGGL((fill_kernel<double2, 0>), dim3(blocks), dim3(256), 0, s, (double2*)out, uint64_t(n), seed, uint64_t(g0));
        else hipLaunchKernelGGL((fill_kernel<double2, 1>), dim3(blocks), dim3(256), 0, s, (double2*)out, uint64_t(n), seed, uint64_t(g0));
    } else {
        if (kind == 0) hipLaunchKernelGGL((fill_kernel<float2, 0>), dim3(blocks), dim3(256), 0, s, (float2*)out, uint64_t(n), seed, uint64_t(g0));
        else hipLaunchKernelGGL((fill_kernel<float2, 1>), dim3(blocks), dim3(256), 0, s, (float2*)out, uint64_t(n), seed, uint64_t(g0));
    }
    return finish_call(stream, hip_check(int(hipGetLastError()), "fill_random launch"));
}
