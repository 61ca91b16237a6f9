// Launch shapes and address arithmetic of the Stockham kernels, shared by
// the device code and by host code (dispatch, and the CPU test that checks
// every address a launch touches).
#pragma once

#include "common.h"

namespace a2a {

// Radix schedules.  Pass p has radix radix_of(N,p); Ns = product of earlier radices.
// Powers of two use radix 2..16 passes.  Mixed lengths 2^a * 3^b * 5^c
// (2^a <= 16, b + c >= 1, N <= 60) use the power-of-two factor as one radix,
// then the 3s, then the 5s, with one thread holding a whole line (E = N).
A2A_HD constexpr bool is_pow2(int N) { return N > 0 && (N & (N - 1)) == 0; }
A2A_HD constexpr int pow2_part(int N) {
    int p = 1;
    while (N > 1 && N % 2 == 0) { N /= 2; p *= 2; }
    return p;
}
A2A_HD constexpr int count_factor(int N, int f) {
    int c = 0;
    while (N > 1 && N % f == 0) { N /= f; ++c; }
    return c;
}
A2A_HD constexpr bool mixed_len_ok(int N) {
    if (N < 3 || N > 60 || is_pow2(N)) return false;
    int m = N / pow2_part(N);
    while (m % 3 == 0) m /= 3;
    while (m % 5 == 0) m /= 5;
    return m == 1 && pow2_part(N) <= 16;
}
A2A_HD constexpr int npass_of(int N) {
    if (!is_pow2(N)) return (pow2_part(N) > 1 ? 1 : 0) + count_factor(N, 3) + count_factor(N, 5);
    return N <= 16 ? 1 : (N <= 256 ? 2 : (N <= 4096 ? 3 : 4));
}
A2A_HD constexpr int radix_of(int N, int p) {
    if (!is_pow2(N)) {
        const int p2 = pow2_part(N);
        if (p2 > 1) {
            if (p == 0) return p2;
            --p;
        }
        return p < count_factor(N, 3) ? 3 : 5;
    }
    // 2:[2] 4:[4] 8:[8] 16:[16] 32:[8,4] 64:[8,8] 128:[16,8] 256:[16,16]
    // 512:[8,8,8] 1024:[16,8,8] 2048:[16,16,8] 4096:[16,16,16]
    // 8192:[16,16,16,2] (the chirp-z convolution rows of 2049..4096-point lines)
    return N <= 16 ? N
         : N == 32 ? (p == 0 ? 8 : 4)
         : N == 64 ? 8
         : N == 128 ? (p == 0 ? 16 : 8)
         : N == 256 ? 16
         : N == 512 ? 8
         : N == 1024 ? (p == 0 ? 16 : 8)
         : N == 2048 ? (p == 2 ? 8 : 16)
         : N == 8192 ? (p == 3 ? 2 : 16)
         : 16;
}
// Granule of E (elements per thread): every radix must divide it.
A2A_HD constexpr int rmax_of(int N) {
    if (!is_pow2(N)) return N;
    return N <= 16 ? N : (N == 32 || N == 64 || N == 512) ? 8 : 16;
}
A2A_HD constexpr int ns_of(int N, int p) {
    int ns = 1;
    for (int i = 0; i < p; ++i) ns *= radix_of(N, i);
    return ns;
}
// Offset of pass p's twiddles in the per-length table: entries [k*R + r] for
// k < Ns, r < R hold exp(DIR*2*pi*i*r*k/(Ns*R)); pass 0 has none.
A2A_HD constexpr int twoff_of(int N, int p) {
    int off = 0;
    for (int i = 1; i < p; ++i) off += ns_of(N, i) * radix_of(N, i);
    return off;
}
A2A_HD constexpr int twsize_of(int N) { return twoff_of(N, npass_of(N)); }
A2A_HD constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }

// Launch geometry of one axis pass (all offsets in elements).
struct FftGeom {
    AxisMap fft;        // the transform axis
    AxisMap a1;         // outer enumeration axis of lines / tiles
    AxisMap a2;         // inner enumeration axis (z columns for strided tiles)
    uint32_t n2;        // units along a2 (lines, or tiles of T columns)
    uint32_t a2_start;  // first a2 coordinate of this launch's window
    uint32_t nunits;    // total lines (contiguous) or tiles (strided)
    FastDiv dn2;
    FastDiv dn1;        // a1.n, for the a2-major unit order
    int64_t tstride;    // strided tiles: element step between the T columns
    int scale_mode;     // 0: none, 1: multiply by scale, 2: divide by scale
    double scale;
};


// Start address of a thread's line, including off(t) of its first element.
// Contiguous: workgroup `block` owns lines block*T .. block*T+T-1.  Strided:
// workgroup `block` owns tile (u1, zt) = T adjacent columns.  The kernels
// call this with compile-time T and TPL.
A2A_HD int64_t line_base(const FftGeom& g, bool strided, int T, int TPL, uint32_t block, int tid, bool* valid,
                         bool a2_major = false) {
    const int t = strided ? tid / T : tid % TPL;
    const int line = strided ? tid % T : tid / TPL;
    int64_t base;
    *valid = true;
    if (strided) {
        // units are numbered a1-major (u1 = block / n2) or a2-major
        const uint32_t u1 = a2_major ? g.dn1.mod(block) : g.dn2.div(block);
        const uint32_t zt = a2_major ? g.dn1.div(block) : block - u1 * g.n2;
        base = g.a1.off(u1) + g.a2.off(g.a2_start + zt * uint32_t(T)) + int64_t(line) * g.tstride;
    } else {
        const uint32_t ln = block * uint32_t(T) + uint32_t(line);
        *valid = ln < g.nunits;
        const uint32_t lv = *valid ? ln : 0u;
        const uint32_t u1 = a2_major ? g.dn1.mod(lv) : g.dn2.div(lv);
        const uint32_t u2 = a2_major ? g.dn1.div(lv) : lv - u1 * g.n2;
        base = g.a1.off(u1) + g.a2.off(g.a2_start + u2);
    }
    // off(n) = off(e*TPL) + off(t) for n = t + e*TPL: the dispatcher only
    // launches when the axis block extent and TPL divide one another.
    return base + g.fft.off(uint32_t(t));
}

// Workgroup renumbering for XCD-aware tile order: hardware hands workgroup b
// to XCD b % 8; this gives XCD x the contiguous logical range
// [x*per + min(x, rem), ...) of the G workgroups.  A bijection on [0, G).
A2A_HD uint32_t xcd_block(uint32_t b, uint32_t G) {
    const uint32_t x = b & 7u, l = b >> 3, per = G >> 3, rem = G & 7u;
    return x < rem ? x * (per + 1) + l : rem * (per + 1) + (x - rem) * per + l;
}

A2A_HD int64_t elem_off(const FftGeom& g, int n0) { return g.fft.off(uint32_t(n0)); }

// ---------------------------------------------------------------------------
// Launch shapes.  Contiguous lines: E elements per thread (at least the
// largest radix), 256 threads per workgroup.  Strided tiles: T adjacent
// columns (128 B rows of complex64 / complex128) and E chosen so a workgroup
// has at most 256 threads; LDS stays at or below 64 KiB.
// ---------------------------------------------------------------------------
A2A_HD constexpr int cont_E(int N) { return N <= 16 ? N : (N <= 64 ? 8 : 16); }
A2A_HD constexpr int cont_T(int N) { return (256 * cont_E(N)) / N > 0 ? (256 * cont_E(N)) / N : 1; }
A2A_HD constexpr int str_T(int N, bool dbl) {
    return !dbl ? (N <= 512 ? 16 : 8192 / N) : (N <= 512 ? 8 : 4096 / N);
}
A2A_HD constexpr int str_E(int N, bool dbl) {
    return (N * str_T(N, dbl)) / 256 > rmax_of(N) ? (N * str_T(N, dbl)) / 256 : rmax_of(N);
}

inline bool fft_len_ok(int N) { return (N >= 2 && N <= 4096 && is_pow2(N)) || mixed_len_ok(N); }

// Kernel variant flags.
enum : int {
    KF_XCD = 1,    // XCD-aware workgroup numbering
    KF_PAD = 2,    // padded LDS layout
    KF_NT = 4,     // non-temporal global loads and stores
    KF_A2MAJ = 8,  // units numbered a2-major (z tiles outermost)
    KF_COH = 16,   // device-coherent (L2-bypassing) global accesses, for in-launch hand-offs
    KF_HALFX = 32, // LDS exchanges move one component at a time: half the LDS, two more barriers
    KF_DFT = 64,   // direct O(N^2) DFT, one line per workgroup: the lengths no Stockham kernel covers
    KF_MIX = 128,  // Stockham FFT with the radix schedule in the arguments: long 2^a 3^b 5^c lengths
    KF_CZ = 2048,  // Bluestein chirp-z convolution over power-of-two FFTs: prime factors above 5
    KF_MIXL = 4096,  // Stockham FFT with per-pass butterfly mapping: long 2^a 3^b 5^c lengths
};
constexpr int kDftMaxN = 4096;   // longest direct-DFT line (N complex values in LDS)

// Chirp-z passes (KF_CZ, fft_long.hip).  Line u of a launch window is
// (u1, u2) = (u / n2, u % n2) (n2 = window count), its element n at
// a1.off(u1) + a2.off(a2_start + u2) + fft.off(n).
A2A_HD int64_t cz_line_base(const FftGeom& g, uint32_t u) {
    const uint32_t u1 = g.dn2.div(u);
    return g.a1.off(u1) + g.a2.off(g.a2_start + (u - u1 * g.n2));
}
// The convolution of M <= 2 kDftMaxN points runs in one kernel
// (cz_fused_kernel): a workgroup of 512 threads holds kCzFusedE elements of each of
// T = 8192 / M lines in registers, thread tid taking element t + e * (M / E)
// of line `line` of its tile as below (the Stockham kernels' lane orders).
constexpr int kCzFusedE = 16, kCzFusedThreads = 512;
A2A_HD constexpr int cz_fused_lines(int M) { return (kCzFusedThreads * kCzFusedE) / M; }
A2A_HD void cz_fused_thread(bool strided, int M, int tid, int* t, int* line) {
    const int T = cz_fused_lines(M), TPL = M / kCzFusedE;
    *t = strided ? tid / T : tid % TPL;
    *line = strided ? tid % T : tid / TPL;
}
// The convolution length: the power of two >= 2N - 1.
A2A_HD int cz_length(int N) {
    int M = 1;
    while (M < 2 * N - 1) M *= 2;
    return M;
}

// Kernel choice for one pass: tile width T, elements per thread E and the
// variant flags.  NT = T * N / E threads per workgroup.
struct KSel {
    int T, E, flags;
};

// True when shape (T, E) can run this pass: E holds whole butterflies of every
// radix, the axis block extent and N/E divide one another (additive offsets),
// and a strided tile of T columns stays inside one rank's z extent.
inline bool shape_fits(const AxisMap& fft, const AxisMap& a2, int strided, int N, int T, int E) {
    if (E < rmax_of(N) || E > N || N % E || E % rmax_of(N)) return false;
    const uint32_t TPL = uint32_t(N / E);
    if (!(fft.l % TPL == 0 || TPL % fft.l == 0)) return false;
    if (T * (N / E) > 1024) return false;
    if (strided && T > 1 && (a2.s_elem != 1 || a2.l % uint32_t(T) || a2.n % uint32_t(T))) return false;
    return true;
}

// Default shape of a pass (before tuning).
inline KSel default_sel(int N, bool dbl, int strided, const AxisMap& fft, const AxisMap& a2) {
    if (!is_pow2(N)) {   // mixed lengths: one thread per line
        if (!strided) return KSel{64, N, 0};
        if (shape_fits(fft, a2, 1, N, 16, N)) return KSel{16, N, 0};
        if (shape_fits(fft, a2, 1, N, 4, N)) return KSel{4, N, 0};
        return KSel{1, N, 0};
    }
    if (!strided) return KSel{cont_T(N), cont_E(N), 0};
    KSel s{str_T(N, dbl), str_E(N, dbl), 0};
    if (!shape_fits(fft, a2, 1, N, s.T, s.E)) s = KSel{1, rmax_of(N), 0};
    return s;
}

// Final geometry of a pass for a chosen shape; returns 0 or a negative status.
// The launch covers a2 coordinates [a2_start, a2_start + a2_count) (count 0 =
// the whole axis); a strided window must start and end on tile boundaries.
// The geometry once the shape is known to fit (any length).
inline int resolve_geom(const AxisMap& fft, const AxisMap& a1, const AxisMap& a2, int strided, int T, FftGeom* g,
                        uint32_t* grid, uint32_t a2_start, uint32_t a2_count) {
    if (a2_count == 0) a2_count = a2.n - a2_start;
    if (uint64_t(a2_start) + a2_count > a2.n) return A2AFFT_E_INVALID_ARG;
    if (strided && (a2_start % uint32_t(T) || a2_count % uint32_t(T))) return A2AFFT_E_INVALID_ARG;
    g->a2_start = a2_start;
    g->fft = fft;
    g->a1 = a1;
    g->a2 = a2;
    g->scale_mode = 0;
    g->scale = 1.0;
    g->tstride = a2.s_elem;
    uint64_t units, blocks;
    if (!strided) {
        g->n2 = a2_count;
        units = uint64_t(a1.n) * a2_count;
        blocks = (units + T - 1) / T;
    } else {
        g->n2 = a2_count / uint32_t(T);
        units = uint64_t(a1.n) * g->n2;
        blocks = units;
    }
    if (units >= (uint64_t(1) << 32) || blocks >= (uint64_t(1) << 31)) return A2AFFT_E_SIZE;
    g->nunits = uint32_t(units);
    g->dn2.set(g->n2 ? g->n2 : 1);
    g->dn1.set(a1.n ? a1.n : 1);
    *grid = uint32_t(blocks);
    return A2AFFT_OK;
}

inline int resolve_with(const AxisMap& fft, const AxisMap& a1, const AxisMap& a2, int strided, int N, int T, int E,
                        FftGeom* g, uint32_t* grid, uint32_t a2_start = 0, uint32_t a2_count = 0) {
    if (!fft_len_ok(N)) return A2AFFT_E_UNSUPPORTED;
    if (fft.n != uint32_t(N)) return A2AFFT_E_INVALID_ARG;
    if (!shape_fits(fft, a2, strided, N, T, E)) return A2AFFT_E_UNSUPPORTED;
    return resolve_geom(fft, a1, a2, strided, T, g, grid, a2_start, a2_count);
}

// Batched lines longer than kDftMaxN (fft_long.hip): N = N1 * N2 with both
// factors in 2..kDftMaxN, N1 >= N2 and N1 the smallest such divisor (as close
// to sqrt(N) as the divisors allow); false when N has no such split (a prime
// above kDftMaxN, or N above kDftMaxN^2).
inline bool long_fft_split(int64_t N, int* n1, int* n2) {
    if (N <= kDftMaxN || N > int64_t(kDftMaxN) * kDftMaxN) return false;
    for (int64_t d = 2; d <= kDftMaxN; ++d)
        if (N % d == 0 && N / d <= kDftMaxN && N / d >= 2 && d >= N / d) {
            *n1 = int(d);
            *n2 = int(N / d);
            return true;
        }
    return false;
}

// The four-step's twiddle-transpose of each line [N1][N2] -> [N2][N1] in
// 32 x 32 tiles: workgroup b covers tile (tn2, tk1) of line `line`, with
// b = (line * tk1s + tk1) * tn2s + tn2, tn2s = ceil(N2 / 32), tk1s = ceil(N1 / 32).
A2A_HD void long_tile(uint32_t b, uint32_t tn2s, uint32_t tk1s, uint32_t* tn2, uint32_t* tk1, uint32_t* line) {
    *tn2 = b % tn2s;
    const uint32_t rest = b / tn2s;
    *tk1 = rest % tk1s;
    *line = rest / tk1s;
}

// The square case N1 = N2 = M transposes in place: workgroup b handles the
// tile pair (a, c), a <= c, of line `line` (rows of the upper triangle of the
// tiles x tiles grid in order), with b = line * pairs + pair.
A2A_HD void long_pair(uint32_t b, uint32_t tiles, uint32_t* a, uint32_t* c, uint32_t* line) {
    const uint32_t pairs = tiles * (tiles + 1) / 2;
    *line = b / pairs;
    uint32_t p = b % pairs, row = 0;
    while (p >= tiles - row) {
        p -= tiles - row;
        ++row;
    }
    *a = row;
    *c = row + p;
}

// Axis maps of the three fused passes over the virtual-rank block layout:
// z pass = lines (x, y), contiguous; x pass = tiles (y, z-columns); y pass =
// tiles (x, z-columns).  The reference's axis order z, x, y (a2adfft.cpp:31-85).
struct PassMaps {
    AxisMap fft, a1, a2;
    int strided;
};

// n: global extents (x, y, z); a single C-order grid is dims [1,1,1].
inline void fused_pass_maps(const int n[3], const int dims[3], const int lgs[3], int64_t nlocal, PassMaps out[3]) {
    const int64_t lx = lgs[0], ly = lgs[1], lz = lgs[2];
    const AxisMap ax = make_axis(uint32_t(lx), uint32_t(n[0]), int64_t(dims[1]) * dims[2] * nlocal, ly * lz);
    const AxisMap ay = make_axis(uint32_t(ly), uint32_t(n[1]), int64_t(dims[2]) * nlocal, lz);
    const AxisMap az = make_axis(uint32_t(lz), uint32_t(n[2]), nlocal, 1);
    out[0] = PassMaps{az, ax, ay, 0};
    out[1] = PassMaps{ax, ay, az, 1};
    out[2] = PassMaps{ay, ax, az, 1};
}

// `lines` contiguous pencils of length N (forward_1d_fft, a2affts.cu:7-40).
inline PassMaps batched_pass_maps(int N, uint32_t lines) {
    return PassMaps{make_axis(uint32_t(N), uint32_t(N), 0, 1), make_axis(lines, lines, 0, N), make_axis(1, 1, 0, 0),
                    0};
}

}  // namespace a2a
