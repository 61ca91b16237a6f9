// Remap kernels of the a2afft hot path on gfx950: the six pack/unpack index
// maps (src/a2afft/distribution/kernels/a2apencils.cu), the three fast-axis
// transposes (a2achangefastaxis.cu), the device-local all-to-all that
// replaces MPI_Alltoall (a2adistribution.cpp:117-125), and the copy / scale
// kernels of dfft/kernels/a2affts.cu:78-126.
//
// All remaps are pure permutations of complex values: the index arithmetic is
// the reference's, so results are bit-identical by construction.  Pack/unpack
// maps move runs of l_axis contiguous elements (idx % l_axis is preserved), so
// one thread moves a 16-byte unit and both the load and the store coalesce.
// The transposes stage 32x32 tiles through LDS.
#include <algorithm>

#include "common.h"
#include "remap.h"
#include "remap_index.h"

#include <stdio.h>

namespace a2a {

// One thread moves one 16-byte unit (VEC complex values of one run).
template <int KIND, typename U, int VEC>
__global__ __launch_bounds__(256) void remap_runs_kernel(const U* __restrict__ src, U* __restrict__ dst,
                                                         RemapParams p, uint32_t units_per_rank) {
    const uint32_t u = blockIdx.x * 256u + threadIdx.x;
    if (u >= units_per_rank) return;
    const size_t rbase = size_t(blockIdx.y) * units_per_rank;
    const uint32_t nd = remap_dest<KIND>(u * VEC, p) / VEC;
    dst[rbase + nd] = src[rbase + u];
}

// recv[s][q*ns + i] = send[q][s*ns + i]; grid.y = source rank q.
template <typename U>
__global__ __launch_bounds__(256) void alltoall_kernel(const U* __restrict__ send, U* __restrict__ recv,
                                                       uint32_t units_per_rank, uint32_t ns_units,
                                                       FastDiv dns) {
    const uint32_t e = blockIdx.x * 256u + threadIdx.x;
    if (e >= units_per_rank) return;
    const uint32_t q = blockIdx.y;
    const uint32_t s = dns.div(e);
    const uint32_t i = e - s * ns_units;
    recv[size_t(s) * units_per_rank + size_t(q) * ns_units + i] = send[size_t(q) * units_per_rank + e];
}

// Batched 3D transpose with a fixed middle axis, per rank buffer:
//   dst[((b*C + c)*M + m)*A + a] = src[((b*A + a)*M + m)*C + c]
// fast_z_to_x: (B,A,M,C) = (1, lx, ly, lz)          a2achangefastaxis.cu:14-15
// fast_x_to_y: (B,A,M,C) = (1, lz*ly, 1, lx)        a2achangefastaxis.cu:45-46
// fast_y_to_z: (B,A,M,C) = (lx, lz, 1, ly)          a2achangefastaxis.cu:76-77
// grid.x: tiles of 32x32 over (a, c); grid.y: b*M + m; grid.z: rank.
template <typename V>
__global__ __launch_bounds__(256) void transpose_kernel(const V* __restrict__ src, V* __restrict__ dst,
                                                        uint32_t A, uint32_t M, uint32_t C,
                                                        uint32_t tiles_c, uint64_t rank_elems) {
    __shared__ V tile[32][33];
    const uint32_t ta = blockIdx.x / tiles_c, tc = blockIdx.x - ta * tiles_c;
    const uint32_t bm = blockIdx.y;
    const uint32_t b = bm / M, m = bm - b * M;
    const size_t rbase = size_t(blockIdx.z) * rank_elems;
    const size_t bbase = size_t(b) * A * M * C;
    const uint32_t tx = threadIdx.x & 31u, ty = threadIdx.x >> 5;   // 32 x 8
    const uint32_t a0 = ta * 32u, c0 = tc * 32u;
#pragma unroll
    for (uint32_t r = 0; r < 32; r += 8) {
        uint32_t a = a0 + ty + r, c = c0 + tx;
        if (a < A && c < C) tile[ty + r][tx] = src[rbase + bbase + (size_t(a) * M + m) * C + c];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < 32; r += 8) {
        uint32_t c = c0 + ty + r, a = a0 + tx;
        if (a < A && c < C) dst[rbase + bbase + (size_t(c) * M + m) * A + a] = tile[tx][ty + r];
    }
}

// The same transpose with 16-byte lanes on both sides: tiles of TA x TC
// elements (512-byte rows: 64 x 64 complex64, 32 x 32 complex128); a lane
// loads EW = 16 / sizeof(V) consecutive c of one row and stores EW
// consecutive a of one output row, both non-temporal (each value is read
// and written once).  Needs A and C multiples of EW; partial tiles are
// bounds-checked per vector.
template <typename V, int TA, int TC>
__global__ __launch_bounds__(256) void transpose16_kernel(const V* __restrict__ src, V* __restrict__ dst,
                                                          uint32_t A, uint32_t M, uint32_t C,
                                                          uint32_t tiles_c, uint64_t rank_elems) {
    constexpr int EW = 16 / int(sizeof(V));
    constexpr int LR = TC / EW;        // lanes per loaded row
    constexpr int RP = 256 / LR;       // rows per load pass
    constexpr int LW = TA / EW;        // lanes per stored row
    constexpr int WP = 256 / LW;       // rows per store pass
    typedef uint32_t v4 __attribute__((ext_vector_type(4)));
    __shared__ V tile[TA][TC + 1];
    const uint32_t ta = blockIdx.x / tiles_c, tc = blockIdx.x - ta * tiles_c;
    const uint32_t bm = blockIdx.y;
    const uint32_t b = bm / M, m = bm - b * M;
    const size_t base = size_t(blockIdx.z) * rank_elems + size_t(b) * A * M * C;
    const uint32_t a0 = ta * TA, c0 = tc * TC;
    const uint32_t tid = threadIdx.x;
    {
        const uint32_t lc = (tid % LR) * EW, lr = tid / LR;
#pragma unroll
        for (int r = 0; r < TA; r += RP) {
            const uint32_t a = a0 + lr + r, c = c0 + lc;
            if (a < A && c < C) {
                const v4 u = __builtin_nontemporal_load(
                    reinterpret_cast<const v4*>(src + base + (size_t(a) * M + m) * C + c));
                const V* e = reinterpret_cast<const V*>(&u);
#pragma unroll
                for (int k = 0; k < EW; ++k) tile[lr + r][lc + k] = e[k];
            }
        }
    }
    __syncthreads();
    {
        const uint32_t la = (tid % LW) * EW, lr = tid / LW;
#pragma unroll
        for (int r = 0; r < TC; r += WP) {
            const uint32_t c = c0 + lr + r, a = a0 + la;
            if (a < A && c < C) {
                v4 u;
                V* e = reinterpret_cast<V*>(&u);
#pragma unroll
                for (int k = 0; k < EW; ++k) e[k] = tile[la + k][lr + r];
                __builtin_nontemporal_store(u, reinterpret_cast<v4*>(dst + base + (size_t(c) * M + m) * A + a));
            }
        }
    }
}

// One 16-byte (or 8-byte) vector per lane, non-temporal both ways: every value
// is read once and written once, and 16-byte lanes measured 6.5 TB/s on a
// 1 GiB copy against 5.8 TB/s for 8-byte lanes.
typedef uint32_t copy_v4 __attribute__((ext_vector_type(4)));
typedef uint32_t copy_v2 __attribute__((ext_vector_type(2)));
template <typename V>
__global__ __launch_bounds__(256) void copy_kernel(V* __restrict__ out, const V* __restrict__ in, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256u)
        __builtin_nontemporal_store(__builtin_nontemporal_load(in + i), out + i);
}

// out = data / scale, the division of scale_fft (a2affts.cu:83-88) per component,
// W components per lane (one 16-byte vector, or 8 bytes for the tail form).
template <typename R, int W>
__global__ __launch_bounds__(256) void scale_kernel(R* __restrict__ out, const R* __restrict__ in, uint64_t n,
                                                    R scale) {
    typedef R vec __attribute__((ext_vector_type(W)));
    for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256u) {
        vec v = __builtin_nontemporal_load(reinterpret_cast<const vec*>(in) + i);
#pragma unroll
        for (int c = 0; c < W; ++c) v[c] = v[c] / scale;
        __builtin_nontemporal_store(v, reinterpret_cast<vec*>(out) + i);
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

static int check_geometry(int world_size, int64_t nlocal, const int lgs[3], const int dims[3], int nbuffers) {
    if (world_size <= 0 || nlocal <= 0 || !lgs || !dims || nbuffers <= 0)
        return set_error(A2AFFT_E_INVALID_ARG, "remap: bad arguments");
    if (nbuffers > 65535) return set_error(A2AFFT_E_SIZE, "remap: more than 65535 rank buffers");
    if (int64_t(lgs[0]) * lgs[1] * lgs[2] != nlocal)
        return set_error(A2AFFT_E_INVALID_ARG, "remap: nlocal != lx*ly*lz");
    if (nlocal >= (int64_t(1) << 32)) return set_error(A2AFFT_E_SIZE, "remap: nlocal >= 2^32");
    if (int64_t(dims[0]) * dims[1] * dims[2] != world_size)
        return set_error(A2AFFT_E_RANKS, "remap: dims product != world_size");
    for (int a = 0; a < 3; ++a)
        if (lgs[a] <= 0 || dims[a] <= 0) return set_error(A2AFFT_E_INVALID_ARG, "remap: non-positive extent");
    if (nlocal % world_size) return set_error(A2AFFT_E_CHUNK, "remap: nlocal %% world_size != 0");
    if ((int64_t(lgs[0]) * lgs[1]) % world_size || (int64_t(lgs[0]) * lgs[2]) % world_size ||
        (int64_t(lgs[1]) * lgs[2]) % world_size)
        return set_error(A2AFFT_E_CHUNK, "remap: (l_a*l_b) %% world_size != 0 (a2adistribution.cpp:78-80)");
    for (int a = 0; a < 3; ++a)
        if (world_size % dims[a]) return set_error(A2AFFT_E_RANKS, "remap: world_size %% dims != 0");
    return A2AFFT_OK;
}

template <int KIND>
static int launch_kind(const void* src, void* dst, const RemapParams& p, int64_t nlocal, int esize, int nbuffers,
                       hipStream_t s) {
    const bool vec2 = (esize == 8) && (p.L % 2 == 0);
    const uint32_t units = uint32_t(vec2 ? nlocal / 2 : nlocal);
    dim3 grid((units + 255) / 256, nbuffers);
    if (esize == 16)
        hipLaunchKernelGGL((remap_runs_kernel<KIND, uint4, 1>), grid, dim3(256), 0, s, (const uint4*)src,
                           (uint4*)dst, p, units);
    else if (vec2)
        hipLaunchKernelGGL((remap_runs_kernel<KIND, uint4, 2>), grid, dim3(256), 0, s, (const uint4*)src,
                           (uint4*)dst, p, units);
    else
        hipLaunchKernelGGL((remap_runs_kernel<KIND, uint2, 1>), grid, dim3(256), 0, s, (const uint2*)src,
                           (uint2*)dst, p, units);
    return hip_check(int(hipGetLastError()), "remap kernel launch");
}

int launch_remap(int kind, const void* src, void* dst, int world_size, int64_t nlocal, const int lgs[3],
                 const int dims[3], a2afft_dtype dtype, int nbuffers, hipStream_t s) {
    A2A_TRY(check_geometry(world_size, nlocal, lgs, dims, nbuffers));
    if (!src || !dst) return set_error(A2AFFT_E_INVALID_ARG, "remap: null buffer");
    if (src == dst) return set_error(A2AFFT_E_INVALID_ARG, "remap: source and dest alias");
    const int es = esize_of(dtype);
    RemapParams p = make_params(kind, world_size, nlocal, lgs, dims);
    switch (kind) {
    case Z_A2A_TO_PENCILS: return launch_kind<Z_A2A_TO_PENCILS>(src, dst, p, nlocal, es, nbuffers, s);
    case Z_PENCILS_TO_A2A: return launch_kind<Z_PENCILS_TO_A2A>(src, dst, p, nlocal, es, nbuffers, s);
    case X_A2A_TO_PENCILS: return launch_kind<X_A2A_TO_PENCILS>(src, dst, p, nlocal, es, nbuffers, s);
    case X_PENCILS_TO_A2A: return launch_kind<X_PENCILS_TO_A2A>(src, dst, p, nlocal, es, nbuffers, s);
    case Y_A2A_TO_PENCILS: return launch_kind<Y_A2A_TO_PENCILS>(src, dst, p, nlocal, es, nbuffers, s);
    case Y_PENCILS_TO_A2A: {
        int st = launch_kind<Y_PENCILS_TO_A2A>(src, dst, p, nlocal, es, nbuffers, s);
        if (st < 0) return st;
        return (dims[1] == 2 || dims[0] == 1) ? A2AFFT_OK : A2AFFT_W_DIVERGES_FROM_REFERENCE;
    }
    default: return set_error(A2AFFT_E_INVALID_ARG, "remap: unknown kind %d", kind);
    }
}

int launch_transpose(int which, const void* src, void* dst, const int lgs[3], int64_t nlocal, a2afft_dtype dtype,
                     int nbuffers, hipStream_t s) {
    if (!src || !dst || !lgs || nbuffers <= 0 || nlocal <= 0)
        return set_error(A2AFFT_E_INVALID_ARG, "transpose: bad arguments");
    if (src == dst) return set_error(A2AFFT_E_INVALID_ARG, "transpose: source and dest alias");
    if (int64_t(lgs[0]) * lgs[1] * lgs[2] != nlocal)
        return set_error(A2AFFT_E_INVALID_ARG, "transpose: nlocal != lx*ly*lz");
    if (nbuffers > 65535) return set_error(A2AFFT_E_SIZE, "transpose: more than 65535 rank buffers");
    const uint32_t lx = lgs[0], ly = lgs[1], lz = lgs[2];
    uint32_t B, A, M, C;
    if (which == 0) { B = 1; A = lx; M = ly; C = lz; }            // fast_z_to_x
    else if (which == 1) { B = 1; A = lz * ly; M = 1; C = lx; }   // fast_x_to_y
    else { B = lx; A = lz; M = 1; C = ly; }                       // fast_y_to_z
    if (uint64_t(B) * M > 65535) return set_error(A2AFFT_E_SIZE, "transpose: grid too large");
    // 16-byte lanes both ways when the rows split into them (complex64: A and
    // C even, and 16-byte aligned buffers): 64 x 64 (complex64) or 32 x 32
    // (complex128) tiles
    const bool c128 = dtype == A2AFFT_C128;
    const bool aligned = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) % 16) == 0 &&
                         (c128 || (A % 2 == 0 && C % 2 == 0 && uint64_t(nlocal) % 2 == 0));
    if (aligned) {
        const uint32_t T = c128 ? 32 : 64;
        const uint32_t ta = (A + T - 1) / T, tc = (C + T - 1) / T;
        if (uint64_t(ta) * tc >= (uint64_t(1) << 24)) return set_error(A2AFFT_E_SIZE, "transpose: grid too large");
        dim3 grid(ta * tc, B * M, nbuffers);
        if (c128)
            hipLaunchKernelGGL((transpose16_kernel<double2, 32, 32>), grid, dim3(256), 0, s, (const double2*)src,
                               (double2*)dst, A, M, C, tc, uint64_t(nlocal));
        else
            hipLaunchKernelGGL((transpose16_kernel<float2, 64, 64>), grid, dim3(256), 0, s, (const float2*)src,
                               (float2*)dst, A, M, C, tc, uint64_t(nlocal));
        return hip_check(int(hipGetLastError()), "transpose kernel launch");
    }
    const uint32_t ta = (A + 31) / 32, tc = (C + 31) / 32;
    // a dispatch holds fewer than 2^32 work-items along x (256 a workgroup)
    if (uint64_t(ta) * tc >= (uint64_t(1) << 24)) return set_error(A2AFFT_E_SIZE, "transpose: grid too large");
    dim3 grid(ta * tc, B * M, nbuffers);
    if (dtype == A2AFFT_C128)
        hipLaunchKernelGGL(transpose_kernel<double2>, grid, dim3(256), 0, s, (const double2*)src, (double2*)dst, A,
                           M, C, tc, uint64_t(nlocal));
    else
        hipLaunchKernelGGL(transpose_kernel<float2>, grid, dim3(256), 0, s, (const float2*)src, (float2*)dst, A, M,
                           C, tc, uint64_t(nlocal));
    return hip_check(int(hipGetLastError()), "transpose kernel launch");
}

int launch_alltoall(const void* send, void* recv, int world_size, int64_t nlocal, a2afft_dtype dtype,
                    hipStream_t s) {
    if (!send || !recv || world_size <= 0 || nlocal <= 0 || nlocal % world_size)
        return set_error(A2AFFT_E_INVALID_ARG, "alltoall: bad arguments");
    if (send == recv) return set_error(A2AFFT_E_INVALID_ARG, "alltoall: send and recv alias");
    if (world_size > 65535 || nlocal >= (int64_t(1) << 32)) return set_error(A2AFFT_E_SIZE, "alltoall: too large");
    const int es = esize_of(dtype);
    const int64_t ns = nlocal / world_size;
    const bool vec2 = (es == 8) && (ns % 2 == 0);
    const uint32_t upr = uint32_t(vec2 ? nlocal / 2 : nlocal);
    const uint32_t nsu = uint32_t(vec2 ? ns / 2 : ns);
    FastDiv dns(nsu);
    dim3 grid((upr + 255) / 256, world_size);
    if (es == 16 || vec2)
        hipLaunchKernelGGL(alltoall_kernel<uint4>, grid, dim3(256), 0, s, (const uint4*)send, (uint4*)recv, upr,
                           nsu, dns);
    else
        hipLaunchKernelGGL(alltoall_kernel<uint2>, grid, dim3(256), 0, s, (const uint2*)send, (uint2*)recv, upr,
                           nsu, dns);
    return hip_check(int(hipGetLastError()), "alltoall kernel launch");
}

int launch_copy(void* out, const void* in, int64_t n, a2afft_dtype dtype, hipStream_t s) {
    if (!out || !in || n < 0) return set_error(A2AFFT_E_INVALID_ARG, "copy: bad arguments");
    if (n == 0) return A2AFFT_OK;
    // 16-byte lanes when both pointers are 16-byte aligned and the byte count
    // splits into them; otherwise 8-byte lanes (complex64 at an odd offset)
    const uint64_t bytes = uint64_t(n) * (dtype == A2AFFT_C128 ? 16u : 8u);
    const bool wide = bytes % 16 == 0 && (reinterpret_cast<uintptr_t>(out) | reinterpret_cast<uintptr_t>(in)) % 16 == 0;
    const uint64_t units = wide ? bytes / 16 : bytes / 8;
    const uint64_t blocks = std::min<uint64_t>((units + 255) / 256, kStrideBlocks);
    if (wide)
        hipLaunchKernelGGL(copy_kernel<copy_v4>, dim3(blocks), dim3(256), 0, s, (copy_v4*)out, (const copy_v4*)in,
                           units);
    else
        hipLaunchKernelGGL(copy_kernel<copy_v2>, dim3(blocks), dim3(256), 0, s, (copy_v2*)out, (const copy_v2*)in,
                           units);
    return hip_check(int(hipGetLastError()), "copy kernel launch");
}

int launch_scale(void* out, const void* in, int64_t n, double scale, a2afft_dtype dtype, hipStream_t s) {
    if (!out || !in || n < 0) return set_error(A2AFFT_E_INVALID_ARG, "scale: bad arguments");
    if (n == 0) return A2AFFT_OK;
    // 16 bytes per lane (one complex128, or two complex64) when both pointers
    // are 16-byte aligned and the count pairs up; otherwise 8 bytes per lane
    const bool c128 = dtype == A2AFFT_C128;
    const bool aligned = (reinterpret_cast<uintptr_t>(out) | reinterpret_cast<uintptr_t>(in)) % 16 == 0;
    const bool wide = aligned && (c128 || n % 2 == 0);
    const uint64_t units = c128 ? (wide ? uint64_t(n) : uint64_t(n) * 2) : wide ? uint64_t(n) / 2 : uint64_t(n);
    const uint64_t blocks = std::min<uint64_t>((units + 255) / 256, kStrideBlocks);
    if (c128 && wide)
        hipLaunchKernelGGL((scale_kernel<double, 2>), dim3(blocks), dim3(256), 0, s, (double*)out, (const double*)in,
                           units, scale);
    else if (c128)
        hipLaunchKernelGGL((scale_kernel<double, 1>), dim3(blocks), dim3(256), 0, s, (double*)out, (const double*)in,
                           units, scale);
    else if (wide)
        hipLaunchKernelGGL((scale_kernel<float, 4>), dim3(blocks), dim3(256), 0, s, (float*)out, (const float*)in,
                           units, float(scale));
    else
        hipLaunchKernelGGL((scale_kernel<float, 2>), dim3(blocks), dim3(256), 0, s, (float*)out, (const float*)in,
                           units, float(scale));
    return hip_check(int(hipGetLastError()), "scale kernel launch");
}

}  // namespace a2a

// ---------------------------------------------------------------------------
// C ABI: launcher level (distribution/kernels/kernels.h:3-17,
// dfft/kernels/kernels.h:9-11).
// ---------------------------------------------------------------------------
using namespace a2a;

#define A2A_REMAP_ENTRY(NAME, KIND)                                                                         \
    extern "C" int NAME(const void* source, void* dest, int blockSize, int world_size, int64_t nlocal,      \
                        const int local_grid_size[3], const int dims[3], a2afft_dtype dtype, int nbuffers,  \
                        void* stream) {                                                                     \
        if (blockSize <= 0 || blockSize > 1024) return set_error(A2AFFT_E_INVALID_ARG, "bad blockSize");   \
        int st = launch_remap(KIND, source, dest, world_size, nlocal, local_grid_size, dims, dtype, nbuffers, \
                              (hipStream_t)stream);                                                         \
        return finish_call(stream, st);                                                                     \
    }

A2A_REMAP_ENTRY(a2afft_launch_d_z_a2a_to_z_pencils, Z_A2A_TO_PENCILS)
A2A_REMAP_ENTRY(a2afft_launch_d_z_pencils_to_z_a2a, Z_PENCILS_TO_A2A)
A2A_REMAP_ENTRY(a2afft_launch_d_x_a2a_to_x_pencils, X_A2A_TO_PENCILS)
A2A_REMAP_ENTRY(a2afft_launch_d_x_pencils_to_x_a2a, X_PENCILS_TO_A2A)
A2A_REMAP_ENTRY(a2afft_launch_d_y_a2a_to_y_pencils, Y_A2A_TO_PENCILS)
A2A_REMAP_ENTRY(a2afft_launch_d_y_pencils_to_y_a2a, Y_PENCILS_TO_A2A)

#define A2A_TRANSPOSE_ENTRY(NAME, WHICH)                                                                    \
    extern "C" int NAME(const void* source, void* dest, const int local_grid_size[3], int blockSize,        \
                        int64_t nlocal, a2afft_dtype dtype, int nbuffers, void* stream) {                   \
        if (blockSize <= 0 || blockSize > 1024) return set_error(A2AFFT_E_INVALID_ARG, "bad blockSize");   \
        int st = launch_transpose(WHICH, source, dest, local_grid_size, nlocal, dtype, nbuffers,            \
                                  (hipStream_t)stream);                                                     \
        return finish_call(stream, st);                                                                     \
    }

A2A_TRANSPOSE_ENTRY(a2afft_launch_d_fast_z_to_x, 0)
A2A_TRANSPOSE_ENTRY(a2afft_launch_d_fast_x_to_y, 1)
A2A_TRANSPOSE_ENTRY(a2afft_launch_d_fast_y_to_z, 2)

extern "C" int a2afft_alltoall(const void* send, void* recv, int world_size, int64_t nlocal, a2afft_dtype dtype,
                               void* stream) {
    return finish_call(stream, launch_alltoall(send, recv, world_size, nlocal, dtype, (hipStream_t)stream));
}

extern "C" int a2afft_launch_scale_fft(void* out, const void* data, int Ng, int64_t n_elems, int blockSize,
                                       a2afft_dtype dtype, void* stream) {
    if (Ng <= 0 || blockSize <= 0 || blockSize > 1024) return set_error(A2AFFT_E_INVALID_ARG, "scale: bad args");
    const double scale = double(int64_t(Ng) * Ng * Ng);
    return finish_call(stream, launch_scale(out, data, n_elems, scale, dtype, (hipStream_t)stream));
}

extern "C" int a2afft_launch_fast_copy_fft(void* out, const void* data, int64_t n_elems, int blockSize,
                                           a2afft_dtype dtype, void* stream) {
    if (blockSize <= 0 || blockSize > 1024) return set_error(A2AFFT_E_INVALID_ARG, "copy: bad blockSize");
    return finish_call(stream, launch_copy(out, data, n_elems, dtype, (hipStream_t)stream));
}
