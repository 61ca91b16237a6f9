// Pattern copy: the memory traffic of one Stockham pass without its
// arithmetic.  Every thread computes its line's address exactly as
// stockham_unit does (the same line_base / elem_off, workgroup order, tile
// shape and cache policy), loads its E elements and stores them back
// unchanged; the workgroup holds the pass kernel's LDS footprint, so as many
// workgroups share a CU.  The
// pass kernel's time beside this kernel's time shows how much of it the
// arithmetic and LDS exchanges add to the memory traffic of its own access
// pattern.  It is not a strict bound: a pure copy issues all of a
// workgroup's loads, then all its stores, while the pass spreads them over
// its LDS phases, which can move the same bytes faster.
#include "fft.h"
#include "stockham.h"

namespace a2a {

// Workgroup bound: 1024 threads leave 128 VGPRs a lane, too few for 32
// complex128 or 64 complex64 values; 512 threads leave 256.
template <typename V, int E> struct PatternBound {
    static constexpr int threads = sizeof(V) * E >= 512 ? 512 : 1024;
};

template <typename V, int E, bool NTMP>
__global__ __launch_bounds__((PatternBound<V, E>::threads)) void pattern_copy_kernel(V* __restrict__ data, FftGeom g,
                                                                                    int T, int TPL, int strided,
                                                                                    int flags) {
    const uint32_t block = (flags & KF_XCD) ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    bool valid;
    const int64_t base = line_base(g, strided != 0, T, TPL, block, int(threadIdx.x), &valid, (flags & KF_A2MAJ) != 0);
    if (!valid) return;
    V v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = ld_global<NTMP>(data + base + elem_off(g, e * TPL));
    // the stores must not be folded away as rewrites of the loaded values
#pragma unroll
    for (int e = 0; e < E; ++e) asm volatile("" : "+v"(v[e].x), "+v"(v[e].y));
#pragma unroll
    for (int e = 0; e < E; ++e) st_global<NTMP>(data + base + elem_off(g, e * TPL), v[e]);
}

// lds_slots() of stockham.h for run-time (N, T, strided, pad).
static int lds_slots_rt(int N, int T, int strided, int pad) {
    if (npass_of(N) < 2) return 1;
    if (strided) return pad ? N * (T + 1) : N * T;
    return pad ? T * (N + N / radix_of(N, 0)) : T * N;
}

template <typename V, bool NTMP>
static const void* pattern_kernel_for(int E, int threads) {
    const int lim = E == 64 ? PatternBound<V, 64>::threads : E == 32 ? PatternBound<V, 32>::threads : 1024;
    if (threads > lim || (E == 64 && sizeof(V) > 8)) return nullptr;
    switch (E) {
    case 2: return reinterpret_cast<const void*>(&pattern_copy_kernel<V, 2, NTMP>);
    case 4: return reinterpret_cast<const void*>(&pattern_copy_kernel<V, 4, NTMP>);
    case 8: return reinterpret_cast<const void*>(&pattern_copy_kernel<V, 8, NTMP>);
    case 16: return reinterpret_cast<const void*>(&pattern_copy_kernel<V, 16, NTMP>);
    case 32: return reinterpret_cast<const void*>(&pattern_copy_kernel<V, 32, NTMP>);
    case 64:
        if constexpr (sizeof(V) == 8) return reinterpret_cast<const void*>(&pattern_copy_kernel<V, 64, NTMP>);
        return nullptr;
    default: return nullptr;
    }
}

int launch_pattern_copy(const AxisPass& p, void* data, a2afft_dtype dt, int N, const KSel& ks, hipStream_t s) {
    if (!data) return set_error(A2AFFT_E_INVALID_ARG, "pattern copy: null data");
    if (ks.flags & (KF_DFT | KF_MIX | KF_MIXL | KF_COH | KF_CZ))
        return set_error(A2AFFT_E_UNSUPPORTED, "pattern copy: fixed-schedule Stockham passes only");
    FftGeom g;
    uint32_t grid;
    int st = resolve_with(p.fft, p.a1, p.a2, p.strided, N, ks.T, ks.E, &g, &grid, p.a2_start, p.a2_count);
    if (st < 0) return set_error(st, "pattern copy: pass geometry (N=%d, T=%d, E=%d)", N, ks.T, ks.E);
    if (grid == 0) return A2AFFT_OK;
    const bool dbl = dt == A2AFFT_C128, nt = (ks.flags & KF_NT) != 0;
    const int TPL = N / ks.E, threads = ks.T * TPL;
    const void* fn = dbl ? (nt ? pattern_kernel_for<CT<double>::V, true>(ks.E, threads)
                               : pattern_kernel_for<CT<double>::V, false>(ks.E, threads))
                         : (nt ? pattern_kernel_for<CT<float>::V, true>(ks.E, threads)
                               : pattern_kernel_for<CT<float>::V, false>(ks.E, threads));
    if (!fn) return set_error(A2AFFT_E_UNSUPPORTED, "pattern copy: E=%d, %d threads", ks.E, threads);
    // a complex value per slot, or one component with KF_HALFX
    const size_t slot_bytes = (dbl ? 16u : 8u) / ((ks.flags & KF_HALFX) ? 2u : 1u);
    const size_t lds = size_t(lds_slots_rt(N, ks.T, p.strided, (ks.flags & KF_PAD) ? 1 : 0)) * slot_bytes;
    int T = ks.T, strided = p.strided, flags = ks.flags;
    void* args[] = {&data, &g, &T, (void*)&TPL, &strided, &flags};
    A2A_TRY(hip_check(int(hipLaunchKernel(fn, dim3(grid), dim3(threads), args, lds, s)), "pattern copy launch"));
    return hip_check(int(hipGetLastError()), "pattern copy launch");
}

}  // namespace a2a
