// Host-side interface of the Stockham FFT kernels.
#pragma once

#include "common.h"
#include "fft_shape.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

namespace a2a {

// One batched 1D FFT pass over an arbitrary strided layout.  Lines are
// enumerated by (u1 < a1.n, u2 < a2.n); the element n of line (u1, u2) sits at
// a1.off(u1) + a2.off(u2) + fft.off(n).  strided = 1 means the lines along
// a2 are adjacent in memory (a2.s_elem == 1 inside a block) and a workgroup
// takes T of them side by side; strided = 0 means each line's own elements
// are the contiguous ones.
struct AxisPass {
    AxisMap fft, a1, a2;
    int strided;
    int scale_mode;   // 0 none, 1 multiply, 2 divide
    double scale;
    uint32_t a2_start = 0;   // window along a2 (count 0 = whole axis)
    uint32_t a2_count = 0;
};

// Supported transform lengths: every N in 2..4096.  Fixed-schedule Stockham
// kernels cover powers of two and 2^a 3^b 5^c (2^a <= 16) up to 60
// (fft_len_ok; a precision may lack some of the latter), the runtime-schedule
// Stockham kernel (KF_MIX) the other 2^a 3^b 5^c lengths, the chirp-z pass
// (KF_CZ) or the direct DFT (KF_DFT) the rest.
bool fft_length_supported(int N);

// Device twiddle table for (N, dir, dtype); created on first use, cached for
// the life of the process (per device).  Returns null on failure.
const void* twiddle_table(int N, int dir, a2afft_dtype dtype);
// The table a pass with kernel choice `s` reads (Stockham, runtime-schedule
// Stockham or direct DFT).
struct AxisPass;
const void* pass_twiddles(const AxisPass& pass, int N, int dir, a2afft_dtype dtype, const KSel& s);

// Launch one pass.  data: complex of dtype.  dir: -1 forward, +1 inverse.
// sel: the kernel variant (null = the default for this geometry).
int launch_fft_pass(const AxisPass& pass, void* data, a2afft_dtype dtype, int N, int dir, hipStream_t s,
                    const KSel* sel);

// A copy kernel with the access pattern of the Stockham pass `sel` would run
// (pattern.hip): that pass's memory traffic without its arithmetic.
// Fixed-schedule Stockham selections only.
int launch_pattern_copy(const AxisPass& pass, void* data, a2afft_dtype dtype, int N, const KSel& sel, hipStream_t s);

// In-place DFTs of `lines` contiguous lines of N points, any supported N
// (the batched 1D launchers' dispatcher).
int fft_lines(void* data, int N, int64_t lines, a2afft_dtype dtype, int dir, hipStream_t s);

// Batched contiguous lines longer than 4096 points (fft_long.hip): four-step
// over N = N1 * N2 with both factors <= 4096 (long_fft_split), through a
// workspace from the stream-ordered allocator; lengths with no such split up
// to kBluesteinMaxN as Bluestein's chirp-z convolution of power-of-two length.
constexpr int kBluesteinMaxN = 1 << 23;
// The chirp-z form for lines of any N <= kBluesteinMaxN.
int launch_chirp_z(void* data, int N, int64_t lines, a2afft_dtype dtype, int dir, hipStream_t s);
int launch_long_fft(void* data, int N, int64_t lines, a2afft_dtype dtype, int dir, hipStream_t s);

// Chirp-z passes (KF_CZ) over any axis pass geometry, N in 9..kDftMaxN: one
// kernel per pass convolves each line with two power-of-two FFTs of
// M = cz_length(N) <= 8192 points in registers and LDS (fft_long.hip).
bool chirp_pass_ok(int N);
// Builds every table a chirp-z pass of length N reads, both directions, and
// returns the chirp of direction `dir` (the pass's table), or null.
const void* chirp_pass_tables(int N, int dir, a2afft_dtype dtype);
// The pass's kernel for (N, precision, layout), or null.
const void* chirp_pass_kernel(int N, bool dbl, int strided);
// Runs a resolved chirp-z pass: `units` lines of geometry g (strided: adjacent
// lines adjacent in memory) in place on `data`.
int launch_chirp_pass(const FftGeom& g, uint32_t units, int strided, void* data, a2afft_dtype dtype, int N, int dir,
                      hipStream_t s);

// Whether `extra` bytes of twiddles (at most 16 KiB) may sit beside a
// workgroup's `tile` bytes of LDS: where they leave as many workgroups per CU
// (160 KiB of LDS), or at least two.  Measured: 960-point strided tiles gain
// though they drop from 4 to 3 workgroups a CU (0.741 -> 0.647 ms per
// 960x960x128 x pass), 640-point wide tiles lose from 2 to 1 (0.467 -> 0.620).
inline bool lds_room_for_twiddles(int tile, int extra) {
    constexpr int kCuLds = 163840;
    if (extra > 16384 || tile <= 0) return false;
    const int before = kCuLds / tile, after = kCuLds / (tile + extra);
    return after == before || after >= 2;
}

// Stockham FFT with the radix schedule in the kernel arguments (KF_MIX,
// fft_mixed.hip) for 2^a 3^b 5^c lengths the fixed schedules do not cover.
constexpr int kMixMaxPasses = 16;
struct MixedPlan {
    int N, E, npass, T, TPL, twsize, NT;   // NT: threads per workgroup
    int radix[kMixMaxPasses], ns[kMixMaxPasses], toff[kMixMaxPasses];
    FastDiv dns[kMixMaxPasses];   // division by ns[p]
    int twl;   // LDS offset of the twiddles (pass p's as [r][k] at toff[p] + r ns[p] + k), 0 = global table
};
// force_T > 0: that tile (lines per workgroup) if it fits the pass, else false.
bool mixed_plan(const AxisPass& pass, int N, bool dbl, MixedPlan* mp, int force_T = 0);
// Complex values a thread holds in the runtime schedule of a length (0: none).
int mixed_plan_regs(int N, bool dbl);

// Stockham FFT with butterflies mapped to threads pass by pass (KF_MIXL,
// fft_mixl.hip) for 2^a 3^b 5^c lengths the fixed schedules do not cover:
// every pass spreads its T*N/R butterflies over the workgroup, so radices
// 16, 8, 4, 3 and 5 mix freely (480 = 8*4*3*5) instead of being limited to
// divisors of one per-thread element count.
constexpr int kMixlMaxPasses = 8;   // every 2^a 3^b 5^c length up to 4096 takes at most 7
struct MixlPlan {
    int N, npass, T, NT, EM, twsize;   // EM: register budget, complex values per thread
    int radix[kMixlMaxPasses], ns[kMixlMaxPasses], toff[kMixlMaxPasses], pad[kMixlMaxPasses];
    FastDiv dns[kMixlMaxPasses];   // j / ns[p]
    FastDiv dB[kMixlMaxPasses];    // butterfly index / (N / radix[p]) (contiguous tiles)
    FastDiv dT;                   // butterfly index / T (strided tiles)
    // Address steps of contiguous tiles, fixed per pass: the LDS slots of a
    // butterfly's inputs (outputs) lie rds[p] (wrs[p]) apart; in global
    // memory, gin between the first pass's inputs and gout between the last
    // pass's outputs, or 0 where the axis's block extent breaks the step (then
    // each address is computed).
    int rds[kMixlMaxPasses], wrs[kMixlMaxPasses];
    int64_t gin, gout;
    // Where the twiddles sit in LDS (in values, after the data tile), pass p's
    // as [r][k] at toff[p] + r ns[p] + k; 0 = read from the global table.
    int twl;
};
bool mixl_plan(const AxisPass& pass, int N, bool dbl, MixlPlan* lp, int force_T = 0);
const void* mixl_kernel_for(int EM, bool dbl, int strided, int flags, int dir);
int mixl_lds_bytes(const MixlPlan& lp, bool dbl, int strided);
const void* mixl_twiddles(const MixlPlan& lp, int dir, a2afft_dtype dtype);
const void* mixed_kernel_for(int N, int T, int E, bool dbl, int strided, int dir);
bool mixed_is_wide(int N, int T, int E);   // more than 256 threads: strided tiles only
int mixed_lds_bytes(const MixedPlan& mp, bool dbl);
const void* mixed_twiddles(const MixedPlan& mp, int dir, a2afft_dtype dtype);

// Launch shape of a pass (for tests and tuning reports).
struct FftShape {
    int T, E, NT, grid, flags;
};
int fft_pass_shape(const AxisPass& pass, a2afft_dtype dtype, int N, const KSel* sel, FftShape* out);

// The built kernel for (N, precision, strided, shape, direction), or null.
const void* find_kernel(int N, bool dbl, int strided, const KSel& s, int dir);

// The built variants that can run this pass; returns the count.
int fft_candidates(const AxisPass& pass, a2afft_dtype dtype, int N, KSel* out, int max);
KSel fft_default_sel(const AxisPass& pass, a2afft_dtype dtype, int N);

// Split-precision passes (radix-16 stages on MFMA, fft_mfma.hip).
// mfma_pass_check: E_UNSUPPORTED unless the pass can run split.
int mfma_pass_check(const AxisPass& pass, int N, a2afft_precision prec);
// sel: T = tile lines, E = 0 for the interleaved-K split-BF16 kernel (flags =
// its variant), E = N for the split-K-halves kernels; null = the default.
int launch_mfma_pass(const AxisPass& pass, void* data, int N, int dir, a2afft_precision prec, hipStream_t s,
                     const KSel* sel = nullptr);
int mfma_candidates(const AxisPass& pass, int N, a2afft_precision prec, KSel* out, int max);
KSel mfma_default_sel(const AxisPass& pass, int N, a2afft_precision prec, bool windowed);
const void* mfma_twiddle_table(int N, int dir);

// exp(dir*2*pi*i*m/M) in long double (quadrant-reduced), for table building.
void twiddle_ld(int64_t m, int64_t M, int dir, long double* re, long double* im);

}  // namespace a2a
