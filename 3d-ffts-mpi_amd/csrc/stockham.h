// Stockham radix-4/8/16 FFT kernels for gfx950.
//
// One workgroup owns T complete lines of length N.  Each thread holds E
// elements of one line in registers: element n = t + e*TPL (TPL = N/E
// threads per line).  A radix-R pass (decimation in time, autosort) reads
// x[j + r*N/R] for butterfly j, which for j = t + b*TPL is exactly the
// thread's own register e = b + r*(E/R); so the first pass runs straight
// from the global load, the last pass leaves its outputs in the registers
// the global store needs (out[j + q*N/R]), and only the passes in between
// exchange through LDS.  Global loads and stores therefore touch the same
// addresses, which makes the kernel in-place safe: a workgroup reads all of
// its lines before it writes any of them.
//
// Two line shapes:
//   contiguous (STRIDED=false): lane order runs along the line, so each
//     wavefront reads whole 256-512 B runs; used for the z axis of the block
//     layout (segments of lz elements, one per rank along z) and for batched
//     contiguous pencils.
//   strided (STRIDED=true): the T lines of a workgroup are T consecutive z
//     columns, lane order runs across the lines, so a wavefront reads rows of
//     T*sizeof(complex) = 128 B; used for the x and y axes.
//
// Arithmetic is written with explicit fma and built with -ffp-contract=off,
// so a line is computed bit-identically whichever shape or launch computes it.
#pragma once

#include <type_traits>

#include "common.h"
#include "fft_shape.h"

namespace a2a {

template <typename real> struct CT;
template <> struct CT<float> { using V = float2; };
template <> struct CT<double> { using V = double2; };

template <typename V> __device__ __forceinline__ V c_add(V a, V b) {
    V r; r.x = a.x + b.x; r.y = a.y + b.y; return r;
}
template <typename V> __device__ __forceinline__ V c_sub(V a, V b) {
    V r; r.x = a.x - b.x; r.y = a.y - b.y; return r;
}
template <typename V> __device__ __forceinline__ V c_mul(V a, V w) {
    V r;
    r.x = fma(a.x, w.x, -(a.y * w.y));
    r.y = fma(a.x, w.y, a.y * w.x);
    return r;
}
// a * exp(DIR * i*pi/2)
template <int DIR, typename V> __device__ __forceinline__ V c_w4(V a) {
    V r;
    if (DIR < 0) { r.x = a.y; r.y = -a.x; } else { r.x = -a.y; r.y = a.x; }
    return r;
}

template <typename real> struct K;
template <> struct K<float> {
    static constexpr float r2 = 0.70710678118654752440f;   // sqrt(2)/2
    static constexpr float c8 = 0.92387953251128675613f;   // cos(pi/8)
    static constexpr float s8 = 0.38268343236508977173f;   // sin(pi/8)
    static constexpr float s3 = 0.86602540378443864676f;   // sin(2pi/3)
    static constexpr float c51 = 0.30901699437494742410f;  // cos(2pi/5)
    static constexpr float c52 = -0.80901699437494742410f; // cos(4pi/5)
    static constexpr float s51 = 0.95105651629515357212f;  // sin(2pi/5)
    static constexpr float s52 = 0.58778525229247312917f;  // sin(4pi/5)
};
template <> struct K<double> {
    static constexpr double r2 = 0.70710678118654752440;
    static constexpr double c8 = 0.92387953251128675613;
    static constexpr double s8 = 0.38268343236508977173;
    static constexpr double s3 = 0.86602540378443864676;
    static constexpr double c51 = 0.30901699437494742410;
    static constexpr double c52 = -0.80901699437494742410;
    static constexpr double s51 = 0.95105651629515357212;
    static constexpr double s52 = 0.58778525229247312917;
};

// a * exp(DIR * i*pi/4)
template <int DIR, typename V> __device__ __forceinline__ V c_w8(V a) {
    using real = decltype(a.x);
    const real c = K<real>::r2;
    V r;
    if (DIR < 0) { r.x = (a.x + a.y) * c; r.y = (a.y - a.x) * c; }
    else { r.x = (a.x - a.y) * c; r.y = (a.x + a.y) * c; }
    return r;
}
// a * exp(DIR * 3i*pi/4)
template <int DIR, typename V> __device__ __forceinline__ V c_w8_3(V a) {
    using real = decltype(a.x);
    const real c = K<real>::r2;
    V r;
    if (DIR < 0) { r.x = (a.y - a.x) * c; r.y = -((a.x + a.y) * c); }
    else { r.x = -((a.x + a.y) * c); r.y = (a.x - a.y) * c; }
    return r;
}
// a * exp(DIR * i*pi/8) and a * exp(DIR * 3i*pi/8)
template <int DIR, typename V> __device__ __forceinline__ V c_w16(V a) {
    using real = decltype(a.x);
    V w; w.x = K<real>::c8; w.y = DIR < 0 ? -K<real>::s8 : K<real>::s8;
    return c_mul(a, w);
}
template <int DIR, typename V> __device__ __forceinline__ V c_w16_3(V a) {
    using real = decltype(a.x);
    V w; w.x = K<real>::s8; w.y = DIR < 0 ? -K<real>::c8 : K<real>::c8;
    return c_mul(a, w);
}

// In-place R-point DFT, natural order in and out: y[q] = sum_r x[r] W_R^{DIR*r*q}.
template <int DIR, typename V> __device__ __forceinline__ void dft2(V& a, V& b) {
    V t = a; a = c_add(t, b); b = c_sub(t, b);
}
template <int DIR, typename V> __device__ __forceinline__ void dft4(V& x0, V& x1, V& x2, V& x3) {
    V t0 = c_add(x0, x2), t1 = c_sub(x0, x2);
    V t2 = c_add(x1, x3), t3 = c_w4<DIR>(c_sub(x1, x3));
    x0 = c_add(t0, t2); x2 = c_sub(t0, t2);
    x1 = c_add(t1, t3); x3 = c_sub(t1, t3);
}
template <int DIR, typename V> __device__ __forceinline__ void dft8(V* x) {
    V e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    V o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    o1 = c_w8<DIR>(o1);
    o2 = c_w4<DIR>(o2);
    o3 = c_w8_3<DIR>(o3);
    x[0] = c_add(e0, o0); x[4] = c_sub(e0, o0);
    x[1] = c_add(e1, o1); x[5] = c_sub(e1, o1);
    x[2] = c_add(e2, o2); x[6] = c_sub(e2, o2);
    x[3] = c_add(e3, o3); x[7] = c_sub(e3, o3);
}
template <int DIR, typename V> __device__ __forceinline__ void dft16(V* x) {
    // n = n1 + 4*n2, k = k1 + 4*k2: inner DFT4 over n2, twiddle W16^{n1*k1},
    // outer DFT4 over n1.
    V a[4][4];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
        a[n1][0] = x[n1]; a[n1][1] = x[n1 + 4]; a[n1][2] = x[n1 + 8]; a[n1][3] = x[n1 + 12];
        dft4<DIR>(a[n1][0], a[n1][1], a[n1][2], a[n1][3]);
    }
    a[1][1] = c_w16<DIR>(a[1][1]);
    a[1][2] = c_w8<DIR>(a[1][2]);
    a[1][3] = c_w16_3<DIR>(a[1][3]);
    a[2][1] = c_w8<DIR>(a[2][1]);
    a[2][2] = c_w4<DIR>(a[2][2]);
    a[2][3] = c_w8_3<DIR>(a[2][3]);
    a[3][1] = c_w16_3<DIR>(a[3][1]);
    a[3][2] = c_w8_3<DIR>(a[3][2]);
    {   // W16^9 = -W16^1
        V t = c_w16<DIR>(a[3][3]);
        a[3][3].x = -t.x; a[3][3].y = -t.y;
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        V y0 = a[0][k1], y1 = a[1][k1], y2 = a[2][k1], y3 = a[3][k1];
        dft4<DIR>(y0, y1, y2, y3);
        x[k1] = y0; x[k1 + 4] = y1; x[k1 + 8] = y2; x[k1 + 12] = y3;
    }
}
// 3- and 5-point DFTs: y_k = sum_n x_n W^{nk}, W = exp(DIR*2*pi*i/R); the
// sums over conjugate pairs share their real parts.
template <int DIR, typename V> __device__ __forceinline__ void dft3(V* x) {
    using real = decltype(x[0].x);
    const real s = DIR < 0 ? -K<real>::s3 : K<real>::s3;
    const V a = c_add(x[1], x[2]), d = c_sub(x[1], x[2]);
    V m;   // x0 - (x1 + x2) / 2
    m.x = x[0].x - a.x * real(0.5);
    m.y = x[0].y - a.y * real(0.5);
    V r;   // i * s * (x1 - x2)
    r.x = -(d.y * s);
    r.y = d.x * s;
    x[0] = c_add(x[0], a);
    x[1] = c_add(m, r);
    x[2] = c_sub(m, r);
}
template <int DIR, typename V> __device__ __forceinline__ void dft5(V* x) {
    using real = decltype(x[0].x);
    const real c1 = K<real>::c51, c2 = K<real>::c52;
    const real s1 = DIR < 0 ? -K<real>::s51 : K<real>::s51;
    const real s2 = DIR < 0 ? -K<real>::s52 : K<real>::s52;
    const V a1 = c_add(x[1], x[4]), d1 = c_sub(x[1], x[4]);
    const V a2 = c_add(x[2], x[3]), d2 = c_sub(x[2], x[3]);
    V m1, m2, r1, r2;
    m1.x = x[0].x + (a1.x * c1 + a2.x * c2);
    m1.y = x[0].y + (a1.y * c1 + a2.y * c2);
    m2.x = x[0].x + (a1.x * c2 + a2.x * c1);
    m2.y = x[0].y + (a1.y * c2 + a2.y * c1);
    // r1 = i * (s1 d1 + s2 d2), r2 = i * (s2 d1 - s1 d2)
    r1.x = -(d1.y * s1 + d2.y * s2);
    r1.y = d1.x * s1 + d2.x * s2;
    r2.x = -(d1.y * s2 - d2.y * s1);
    r2.y = d1.x * s2 - d2.x * s1;
    x[0].x = x[0].x + (a1.x + a2.x);
    x[0].y = x[0].y + (a1.y + a2.y);
    x[1] = c_add(m1, r1);
    x[4] = c_sub(m1, r1);
    x[2] = c_add(m2, r2);
    x[3] = c_sub(m2, r2);
}
// 6- and 10-point DFTs as prime-factor (Good-Thomas) products of 2 and 3 or
// 5 points, no twiddles: input n = (3 n1 + 2 n2) mod 6, resp. (5 n1 + 2 n2)
// mod 10; output k = (3 k1 + 4 k2) mod 6, resp. (5 k1 + 6 k2) mod 10, where
// n k reduces to the sub-DFTs' own exponents.
template <int DIR, typename V> __device__ __forceinline__ void dft6(V* x) {
    V a0[3] = {x[0], x[2], x[4]}, a1[3] = {x[3], x[5], x[1]};
    dft3<DIR>(a0);
    dft3<DIR>(a1);
    x[0] = c_add(a0[0], a1[0]); x[3] = c_sub(a0[0], a1[0]);
    x[4] = c_add(a0[1], a1[1]); x[1] = c_sub(a0[1], a1[1]);
    x[2] = c_add(a0[2], a1[2]); x[5] = c_sub(a0[2], a1[2]);
}
template <int DIR, typename V> __device__ __forceinline__ void dft10(V* x) {
    V a0[5] = {x[0], x[2], x[4], x[6], x[8]}, a1[5] = {x[5], x[7], x[9], x[1], x[3]};
    dft5<DIR>(a0);
    dft5<DIR>(a1);
    x[0] = c_add(a0[0], a1[0]); x[5] = c_sub(a0[0], a1[0]);
    x[6] = c_add(a0[1], a1[1]); x[1] = c_sub(a0[1], a1[1]);
    x[2] = c_add(a0[2], a1[2]); x[7] = c_sub(a0[2], a1[2]);
    x[8] = c_add(a0[3], a1[3]); x[3] = c_sub(a0[3], a1[3]);
    x[4] = c_add(a0[4], a1[4]); x[9] = c_sub(a0[4], a1[4]);
}
// 15 points as the product of 3 and 5 the same way: input n = (5 n1 + 3 n2)
// mod 15 (three 5-point DFTs over n2), output k = (10 k1 + 6 k2) mod 15 (five
// 3-point DFTs over n1); 10 = 1 mod 3 and 6 = 1 mod 5.
template <int DIR, typename V> __device__ __forceinline__ void dft15(V* x) {
    V a0[5] = {x[0], x[3], x[6], x[9], x[12]}, a1[5] = {x[5], x[8], x[11], x[14], x[2]},
      a2[5] = {x[10], x[13], x[1], x[4], x[7]};
    dft5<DIR>(a0);
    dft5<DIR>(a1);
    dft5<DIR>(a2);
#pragma unroll
    for (int k2 = 0; k2 < 5; ++k2) {
        V t[3] = {a0[k2], a1[k2], a2[k2]};
        dft3<DIR>(t);
        x[(6 * k2) % 15] = t[0];
        x[(10 + 6 * k2) % 15] = t[1];
        x[(20 + 6 * k2) % 15] = t[2];
    }
}
template <int R, int DIR, typename V> __device__ __forceinline__ void dft(V* x) {
    if constexpr (R == 3) dft3<DIR>(x);
    else if constexpr (R == 15) dft15<DIR>(x);
    else if constexpr (R == 5) dft5<DIR>(x);
    else if constexpr (R == 6) dft6<DIR>(x);
    else if constexpr (R == 10) dft10<DIR>(x);
    else if constexpr (R == 2) dft2<DIR>(x[0], x[1]);
    else if constexpr (R == 4) dft4<DIR>(x[0], x[1], x[2], x[3]);
    else if constexpr (R == 8) dft8<DIR>(x);
    else dft16<DIR>(x);
}

// LDS slot of element d of line `line`.  Contiguous tiles are line-major,
// strided tiles column-interleaved (the T lines side by side).  With PAD a
// contiguous line gets one spare slot per R0 elements (R0 = first radix):
// the first exchange writes at stride R0, which otherwise lands every lane of
// a wavefront in a few banks.  A padded strided row is T+1 wide.
template <int N, int T, bool STRIDED, int PAD>
__device__ __forceinline__ constexpr int lds_slot(int line, int d) {
    constexpr int R0 = radix_of(N, 0);
    if constexpr (STRIDED) return PAD ? d * (T + 1) + line : d * T + line;
    else return PAD ? line * (N + N / R0) + d + d / R0 : line * N + d;
}
template <int N, int T, bool STRIDED, int PAD>
constexpr int lds_slots() {
    return npass_of(N) < 2 ? 1
                           : (STRIDED ? (PAD ? N * (T + 1) : N * T)
                                      : (PAD ? T * (N + N / radix_of(N, 0)) : T * N));
}

// Global access of one complex value; NTMP marks it non-temporal (a native
// 2-vector, so it stays one 8- or 16-byte access).  COH makes it a relaxed
// device-scope atomic access (complex64 only): coherent across the XCDs'
// separate L2 caches without cache write-back or invalidation, for kernels
// that hand data between workgroups inside one launch.
template <bool NTMP, bool COH = false, typename V> __device__ __forceinline__ V ld_global(const V* p) {
    if constexpr (COH) {
        static_assert(sizeof(V) == 8, "coherent access is for complex64");
        const uint64_t u = __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        return __builtin_bit_cast(V, u);
    } else if constexpr (NTMP) {
        using R = decltype(p->x);
        typedef R nv2 __attribute__((ext_vector_type(2)));
        const nv2 t = __builtin_nontemporal_load(reinterpret_cast<const nv2*>(p));
        V v;
        v.x = t.x;
        v.y = t.y;
        return v;
    } else {
        return *p;
    }
}
template <bool NTMP, bool COH = false, typename V> __device__ __forceinline__ void st_global(V* p, V v) {
    if constexpr (COH) {
        static_assert(sizeof(V) == 8, "coherent access is for complex64");
        __hip_atomic_store(reinterpret_cast<uint64_t*>(p), __builtin_bit_cast(uint64_t, v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    } else if constexpr (NTMP) {
        using R = decltype(p->x);
        typedef R nv2 __attribute__((ext_vector_type(2)));
        nv2 t;
        t.x = v.x;
        t.y = v.y;
        __builtin_nontemporal_store(t, reinterpret_cast<nv2*>(p));
    } else {
        *p = v;
    }
}

// One slot of an LDS exchange: the complex value (C = 2), its real part
// (C = 0) or its imaginary part (C = 1).
template <int C, typename L, typename V> __device__ __forceinline__ void lds_put(L& slot, const V& v) {
    if constexpr (C == 2) slot = v;
    else if constexpr (C == 0) slot = v.x;
    else slot = v.y;
}
template <int C, typename L, typename V> __device__ __forceinline__ void lds_get(const L& slot, V& v) {
    if constexpr (C == 2) v = slot;
    else if constexpr (C == 0) v.x = slot;
    else v.y = slot;
}

// The Stockham write of a pass's outputs (radix R, Ns = product of the
// earlier radices) and the read of the next pass's inputs, through LDS, for
// component C of every value.
template <int C, int N, int E, int T, bool STRIDED, int PAD, int R, int Ns, typename V, typename L>
__device__ __forceinline__ void stockham_exchange(V (&v)[E], L* lds, int t, int line, bool active) {
    constexpr int TPL = N / E;
    constexpr int NB = E / R;
    if (active) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * TPL;
            const int d0 = int(unsigned(j) / unsigned(Ns)) * (Ns * R) + int(unsigned(j) % unsigned(Ns));
#pragma unroll
            for (int q = 0; q < R; ++q) lds_put<C>(lds[lds_slot<N, T, STRIDED, PAD>(line, d0 + q * Ns)], v[b + q * NB]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) lds_get<C>(lds[lds_slot<N, T, STRIDED, PAD>(line, t + e * TPL)], v[e]);
}

// Passes p..NP-1 of the Stockham schedule on the thread's registers.
// `active` = false for threads beyond the unit's NT (a wider workgroup runs
// the unit): they take part in the barriers but write nothing to LDS.  `lds`
// holds complex values (L = V), or one real component per slot (L = real):
// then each exchange moves the real parts, then the imaginary parts, through
// half as much LDS, so more workgroups fit a CU.
template <typename real, int N, int E, int T, bool STRIDED, int DIR, int PAD, int p, typename L>
__device__ __forceinline__ void stockham_passes(typename CT<real>::V (&v)[E], L* lds,
                                                const typename CT<real>::V* __restrict__ tw, int t, int line,
                                                bool active) {
    using V = typename CT<real>::V;
    constexpr int NP = npass_of(N);
    if constexpr (p < NP) {
        constexpr int TPL = N / E;
        constexpr int R = radix_of(N, p);
        constexpr int Ns = ns_of(N, p);
        constexpr int NB = E / R;
        constexpr int toff = twoff_of(N, p);
        static_assert(E % R == 0, "E must be a multiple of every radix");
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            V x[R];
#pragma unroll
            for (int r = 0; r < R; ++r) x[r] = v[b + r * NB];
            if constexpr (p > 0) {
                const int j = t + b * TPL;
                const int k = int(unsigned(j) % unsigned(Ns));
#pragma unroll
                for (int r = 1; r < R; ++r) x[r] = c_mul(x[r], tw[toff + k * R + r]);
            }
            dft<R, DIR>(x);
#pragma unroll
            for (int q = 0; q < R; ++q) v[b + q * NB] = x[q];
        }
        if constexpr (p + 1 < NP) {
            if constexpr (p > 0) __syncthreads();
            if constexpr (sizeof(L) == sizeof(V)) {
                stockham_exchange<2, N, E, T, STRIDED, PAD, R, Ns>(v, lds, t, line, active);
            } else {
                stockham_exchange<0, N, E, T, STRIDED, PAD, R, Ns>(v, lds, t, line, active);
                __syncthreads();
                stockham_exchange<1, N, E, T, STRIDED, PAD, R, Ns>(v, lds, t, line, active);
            }
        }
        stockham_passes<real, N, E, T, STRIDED, DIR, PAD, p + 1>(v, lds, tw, t, line, active);
    }
}

struct NoHook {
    __device__ __forceinline__ void operator()() const {}
};

// One work unit: T lines (contiguous) or one tile of T columns (strided),
// loaded, transformed and stored by the NT = T*N/E threads of a workgroup.
// `block` is the unit index; `lds` holds lds_slots() complex values.  Threads
// with tid >= NT take part in the barriers only.  Every thread calls `hook`
// once, after the unit's loads have been consumed and before its stores are
// issued.
template <typename real, int N, int E, int T, bool STRIDED, int DIR, int FLAGS, typename L, typename Hook = NoHook>
__device__ __forceinline__ void stockham_unit(typename CT<real>::V* __restrict__ data, const FftGeom& g,
                                              const typename CT<real>::V* __restrict__ tw, uint32_t block, L* lds,
                                              int tid, Hook&& hook = Hook{}) {
    using V = typename CT<real>::V;
    constexpr int TPL = N / E;
    constexpr int NT = T * TPL;
    constexpr int PAD = (FLAGS & KF_PAD) ? 1 : 0;
    constexpr bool NTMP = (FLAGS & KF_NT) != 0;
    constexpr bool COH = (FLAGS & KF_COH) != 0;
    constexpr bool A2MAJ = (FLAGS & KF_A2MAJ) != 0;

    const bool active = tid < NT;
    const int tl = active ? tid : 0;
    const int t = STRIDED ? tl / T : tl % TPL;
    const int line = STRIDED ? tl % T : tl / TPL;
    bool valid;
    const int64_t base = line_base(g, STRIDED, T, TPL, block, tl, &valid, A2MAJ);
    valid = valid && active;

    V v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        if (valid) v[e] = ld_global<NTMP, COH>(data + base + elem_off(g, e * TPL));
        else { v[e].x = 0; v[e].y = 0; }
    }

    stockham_passes<real, N, E, T, STRIDED, DIR, PAD, 0>(v, lds, tw, t, line, active);
    hook();

    if (!valid) return;
    if (g.scale_mode == 1) {
        const real s = real(g.scale);
#pragma unroll
        for (int e = 0; e < E; ++e) { v[e].x *= s; v[e].y *= s; }
    } else if (g.scale_mode == 2) {
        const real s = real(g.scale);
#pragma unroll
        for (int e = 0; e < E; ++e) { v[e].x = v[e].x / s; v[e].y = v[e].y / s; }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) st_global<NTMP, COH>(data + base + elem_off(g, e * TPL), v[e]);
}

// One workgroup per unit.  XCD = 1 renumbers workgroups so each of the 8
// XCDs (which receive workgroups round-robin) works through a contiguous
// range of units.
template <typename real, int N, int E, int T, bool STRIDED, int DIR, int FLAGS>
__global__ __launch_bounds__(T * (N / E)) void stockham_kernel(typename CT<real>::V* __restrict__ data, FftGeom g,
                                                               const typename CT<real>::V* __restrict__ tw) {
    using V = typename CT<real>::V;
    constexpr int PAD = (FLAGS & KF_PAD) ? 1 : 0;
    // complex slots, or one real component per slot with KF_HALFX
    using L = typename std::conditional<(FLAGS & KF_HALFX) != 0, real, V>::type;
    __shared__ L lds[lds_slots<N, T, STRIDED, PAD>()];
    const uint32_t block = (FLAGS & KF_XCD) ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    stockham_unit<real, N, E, T, STRIDED, DIR, FLAGS>(data, g, tw, block, lds, int(threadIdx.x));
}

// Registry of built kernel variants (filled by the per-precision files).
struct KEntry {
    int N, dbl, strided, T, E, flags, dir;
    const void* fn;
};

template <typename real, int N, int E, int T, bool S, int FLAGS>
inline void add_variant(KEntry* out, int* n) {
    constexpr int dbl = sizeof(real) == 8;
    out[(*n)++] = KEntry{N, dbl, S, T, E, FLAGS, -1, (const void*)&stockham_kernel<real, N, E, T, S, -1, FLAGS>};
    out[(*n)++] = KEntry{N, dbl, S, T, E, FLAGS, 1, (const void*)&stockham_kernel<real, N, E, T, S, 1, FLAGS>};
}

// Plain, XCD-ordered, padded and both.
template <typename real, int N, int E, int T, bool S>
inline void add_shape(KEntry* out, int* n) {
    add_variant<real, N, E, T, S, 0>(out, n);
    if constexpr (npass_of(N) > 1) {
        add_variant<real, N, E, T, S, KF_XCD>(out, n);
        add_variant<real, N, E, T, S, KF_PAD>(out, n);
        add_variant<real, N, E, T, S, KF_XCD | KF_PAD>(out, n);
    }
}

// Non-temporal and a2-major variants of a shape (benchmark lengths).
template <typename real, int N, int E, int T, bool S>
inline void add_extra_flags(KEntry* out, int* n) {
    add_variant<real, N, E, T, S, KF_XCD | KF_NT>(out, n);
    add_variant<real, N, E, T, S, KF_XCD | KF_PAD | KF_NT>(out, n);
    add_variant<real, N, E, T, S, KF_XCD | KF_A2MAJ>(out, n);
    add_variant<real, N, E, T, S, KF_XCD | KF_PAD | KF_A2MAJ>(out, n);
    add_variant<real, N, E, T, S, KF_A2MAJ>(out, n);
    add_variant<real, N, E, T, S, KF_XCD | KF_NT | KF_A2MAJ>(out, n);
}

// Half-width LDS exchange variants of a shape (KF_HALFX): the same tiles with
// twice the workgroups per CU where LDS limits occupancy.
template <typename real, int N, int E, int T, bool S>
inline void add_halfx(KEntry* out, int* n) {
    add_variant<real, N, E, T, S, KF_XCD | KF_HALFX>(out, n);
    add_variant<real, N, E, T, S, KF_XCD | KF_PAD | KF_HALFX>(out, n);
    add_variant<real, N, E, T, S, KF_XCD | KF_PAD | KF_NT | KF_HALFX>(out, n);
    add_variant<real, N, E, T, S, KF_A2MAJ | KF_HALFX>(out, n);
    add_variant<real, N, E, T, S, KF_XCD | KF_A2MAJ | KF_HALFX>(out, n);
}

// Mixed lengths 2^a 3^b 5^c: one thread per line (E = N), contiguous tiles of
// 64 lines, strided tiles of 16 or 4 columns and the single-column fallback.
template <typename real, int N>
inline void add_mixed(KEntry* out, int* n) {
    static_assert(!is_pow2(N) && mixed_len_ok(N), "mixed length");
    add_variant<real, N, N, 64, false, 0>(out, n);
    add_variant<real, N, N, 16, true, 0>(out, n);
    add_variant<real, N, N, 4, true, 0>(out, n);
    add_variant<real, N, N, 1, true, 0>(out, n);
}
template <typename real, int... Ns>
inline void add_mixed_all(KEntry* out, int* n) {
    (add_mixed<real, Ns>(out, n), ...);
}

// Default contiguous / strided shapes and the single-column fallback.
template <typename real, int N>
inline void add_defaults(KEntry* out, int* n) {
    constexpr bool dbl = sizeof(real) == 8;
    add_shape<real, N, cont_E(N), cont_T(N), false>(out, n);
    add_shape<real, N, str_E(N, dbl), str_T(N, dbl), true>(out, n);
    add_variant<real, N, rmax_of(N), 1, true, 0>(out, n);
}

}  // namespace a2a
