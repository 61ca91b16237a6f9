// Split-precision Stockham passes with the radix-16 DFT stages on MFMA.
//
// A radix-16 stage applies the 16-point DFT matrix F (F[q][r] = W16^(DIR*q*r))
// to every butterfly of the line.  Sixteen butterflies form the 16 columns of
// one matrix product on v_mfma_f32_16x16x32_{bf16,f16}, in the real form
//   [Re y]   [Re F  -Im F] [Re x]
//   [Im y] = [Im F   Re F] [Im x]      (K = 32: real parts, then imaginary)
// computed as two 16x16x32 products (real rows, imaginary rows).  Operands
// are split into hi + lo halves of the 16-bit type (hi = round(x),
// lo = round(x - hi)) and each product is hi*hi + hi*lo + lo*hi with FP32
// accumulation.  The twiddles between stages are FP32 complex products,
// applied to the stage's outputs (for the next stage), from tables built in
// long double and rounded once.  The FP16 split scales each butterfly by a
// power of two before the split (FP16's range) and undoes it exactly after.
//
// Schedule: a radix-16 stages, then one VALU stage of radix R = N / 16^a
// (2, 4 or 8) when N is not a power of 16 -- 512 = 16 * 16 * 2.
//
// Two kernels:
//   mfma_fft_direct (256 and 512): one LDS exchange, operands straight from
//     and results straight to global memory (described at the kernel).
//   mfma_fft_kernel (16..128): 256 threads, a tile of T lines (contiguous
//     lines, or T adjacent columns for strided passes) staged in LDS as a
//     real and an imaginary plane of floats.  Each wavefront owns T/4 whole
//     lines during the MFMA stages, so those need no barriers; the load, the
//     final stage and the store use a coalescing thread map and are
//     separated by two barriers.
//
// The operand layouts used: A (16 x 32) lane l holds row l % 16, B (32 x 16)
// lane l holds column l % 16, each with 8 K values chosen by l / 16 -- the
// same (l / 16, element) -> k map for A and B, so the sum over k does not
// depend on it; the result lane l holds column l % 16, rows 4 * (l / 16) + i.
#include "fft.h"
#include "stockham.h"

#include <cmath>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

namespace a2a {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename H> struct MfTraits;
template <> struct MfTraits<__bf16> {
    using V8 = bf16x8;
    static constexpr bool kScale = false;
    __device__ static f32x4 mma(V8 a, V8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};
template <> struct MfTraits<_Float16> {
    using V8 = f16x8;
    static constexpr bool kScale = true;
    __device__ static f32x4 mma(V8 a, V8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};

constexpr int kMfThreads = 256;

A2A_HD constexpr int mf_stages16(int N) { return N >= 4096 ? 3 : N >= 256 ? 2 : N >= 16 ? 1 : 0; }
A2A_HD constexpr int mf_pow16(int a) { return 1 << (4 * a); }   // not recursive: inlined in kernels
A2A_HD constexpr int mf_rlast(int N) { return N / mf_pow16(mf_stages16(N)); }   // 1, 2, 4 or 8
A2A_HD constexpr int mf_nstages(int N) { return mf_stages16(N) + (mf_rlast(N) > 1 ? 1 : 0); }
A2A_HD constexpr int mf_radix(int N, int p) { return p < mf_stages16(N) ? 16 : mf_rlast(N); }
A2A_HD constexpr int mf_ns(int N, int p) { return mf_pow16(p < mf_stages16(N) ? p : mf_stages16(N)); }
// Table: 16 roots W16^t, then for stage p >= 1 entries [k*R + r], k < Ns, r < R.
A2A_HD constexpr int mf_twoff(int N, int p) {
    int off = 16;
    for (int i = 1; i < p; ++i) off += mf_ns(N, i) * mf_radix(N, i);
    return off;
}
A2A_HD constexpr int mf_twsize(int N) { return mf_twoff(N, mf_nstages(N)); }
// Tile width: 16 lines for 256 and 512 (the direct kernel); otherwise 128 B
// rows for strided tiles and enough lines that each of the 4 wavefronts owns
// whole lines holding a multiple of 16 butterflies.
A2A_HD constexpr int mf_T(int N, bool strided) {
    return N >= 256 ? 16 : ((1024 / N) > (strided ? 16 : 4) ? (1024 / N) : (strided ? 16 : 4));
}
A2A_HD constexpr int mf_ld(int N) { return N + N / 16; }   // padded plane row
A2A_HD constexpr int mf_slot(int d) { return d + (d >> 4); }

bool mf_length_ok(int N) { return N >= 16 && N <= 512 && (N & (N - 1)) == 0; }

template <typename H>
__device__ __forceinline__ void split2(float x, H& hi, H& lo) {
    hi = H(x);
    lo = H(x - float(hi));
}

template <typename H, int N, int T, bool STRIDED, int DIR>
__global__ __launch_bounds__(kMfThreads) void mfma_fft_kernel(float2* __restrict__ data, FftGeom g,
                                                               const float2* __restrict__ tw) {
    using Tr = MfTraits<H>;
    using V8 = typename Tr::V8;
    constexpr int LD = mf_ld(N);
    constexpr int A16 = mf_stages16(N);
    constexpr int RL = mf_rlast(N);
    constexpr int NST = mf_nstages(N);
    constexpr int LW = T / 4;            // lines per wavefront
    constexpr int BPL = N / 16;          // radix-16 butterflies per line
    constexpr int CL = BPL >= 16 ? 1 : 16 / BPL;   // lines per chunk of whole MFMA column groups
    constexpr int GC = CL * BPL / 16;              // column groups per chunk
    constexpr int TPLC = kMfThreads / T;           // contiguous map: threads per line
    static_assert(T % 4 == 0 && LW % CL == 0, "tile does not split into MFMA groups");
    __shared__ float lds[2 * T * LD];
    float* const re = lds;
    float* const im = lds + T * LD;

    const int tid = int(threadIdx.x);
    const uint32_t block = xcd_block(blockIdx.x, gridDim.x);

    // ---- load: global -> LDS planes (coalescing map) ----
    const int mline = STRIDED ? tid % T : tid / TPLC;   // line of this thread in the load / store map
    const int m0 = STRIDED ? tid / T : tid % TPLC;      // first element
    constexpr int MSTEP = STRIDED ? kMfThreads / T : TPLC;
    bool valid;
    const int64_t lbase = line_base(g, STRIDED, T, 1, block, mline, &valid, false);
#pragma unroll 4
    for (int d = m0; d < N; d += MSTEP) {
        float2 v = make_float2(0.f, 0.f);
        if (valid) v = data[lbase + g.fft.off(uint32_t(d))];
        re[mline * LD + mf_slot(d)] = v.x;
        im[mline * LD + mf_slot(d)] = v.y;
    }
    __syncthreads();

    // ---- radix-16 stages on MFMA, each wavefront on its own lines ----
    const int lane = tid & 63;
    const int cn = lane & 15;   // column (butterfly) of this lane; row of A
    const int kg = lane >> 4;   // K group: 0,1 real parts x[0..7], x[8..15]; 2,3 imaginary parts
    const int L0 = (tid >> 6) * LW;
    // A operands: rows of [Re F, -Im F] and [Im F, Re F], split hi + lo.
    V8 arh, arl, aih, ail;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int r = 8 * (kg & 1) + i;
        const float2 w = tw[(cn * r) & 15];
        const float ar = kg < 2 ? w.x : -w.y;
        const float ai = kg < 2 ? w.y : w.x;
        H h, l;
        split2<H>(ar, h, l);
        arh[i] = h;
        arl[i] = l;
        split2<H>(ai, h, l);
        aih[i] = h;
        ail[i] = l;
    }
    const float* const src = kg < 2 ? re : im;
#pragma unroll
    for (int p = 0; p < A16; ++p) {
        const int Ns = mf_pow16(p);
        // One chunk of CL lines (GC column groups) at a time: all of a chunk's
        // operands are read before any of its outputs is written, and chunks
        // (whole lines) do not overlap.
#pragma unroll 1
        for (int c = 0; c < LW / CL; ++c) {
            V8 bh[GC], bl[GC];
            float sc[GC];
#pragma unroll
            for (int G = 0; G < GC; ++G) {
                const int b = G * 16 + cn;
                const int line = L0 + c * CL + b / BPL, j = b % BPL;
                float x[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) x[i] = src[line * LD + mf_slot(j + (8 * (kg & 1) + i) * BPL)];
                float s = 1.f;
                if constexpr (Tr::kScale) {
                    // power of two that brings the butterfly's largest part near 2^14
                    float mx = 0.f;
#pragma unroll
                    for (int i = 0; i < 8; ++i) mx = fmaxf(mx, fabsf(x[i]));
                    mx = fmaxf(mx, __shfl_xor(mx, 16));
                    mx = fmaxf(mx, __shfl_xor(mx, 32));
                    int e = 0;
                    (void)frexpf(mx, &e);
                    const int sh = 14 - e;
                    s = mx > 0.f ? ldexpf(1.f, sh < -126 ? -126 : (sh > 126 ? 126 : sh)) : 1.f;
#pragma unroll
                    for (int i = 0; i < 8; ++i) x[i] *= s;
                }
                sc[G] = s;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    H h, l;
                    split2<H>(x[i], h, l);
                    bh[G][i] = h;
                    bl[G][i] = l;
                }
            }
#pragma unroll
            for (int G = 0; G < GC; ++G) {
                f32x4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
                yr = Tr::mma(arh, bh[G], yr);
                yr = Tr::mma(arh, bl[G], yr);
                yr = Tr::mma(arl, bh[G], yr);
                yi = Tr::mma(aih, bh[G], yi);
                yi = Tr::mma(aih, bl[G], yi);
                yi = Tr::mma(ail, bh[G], yi);
                const int b = G * 16 + cn;
                const int line = L0 + c * CL + b / BPL, j = b % BPL;
                const float inv = 1.f / sc[G];   // exact: a power of two
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int q = 4 * kg + i;
                    const int d = (j / Ns) * (Ns * 16) + (j % Ns) + q * Ns;
                    float2 y = make_float2(yr[i] * inv, yi[i] * inv);
                    if (p + 1 < NST) {   // twiddle of the next stage for the element it will read at d
                        const int R2 = mf_radix(N, p + 1), Ns2 = mf_ns(N, p + 1);
                        const int j2 = d % (N / R2), r2 = d / (N / R2), k2 = j2 % Ns2;
                        y = c_mul(y, tw[mf_twoff(N, p + 1) + k2 * R2 + r2]);
                    }
                    re[line * LD + mf_slot(d)] = y.x;
                    im[line * LD + mf_slot(d)] = y.y;
                }
            }
        }
    }
    __syncthreads();

    // ---- last stage (radix RL on VALU) fused with the store ----
    const int scale_mode = g.scale_mode;
    const float fs = float(g.scale);
    auto put = [&](int d, float2 v) {
        if (scale_mode == 1) { v.x *= fs; v.y *= fs; }
        else if (scale_mode == 2) { v.x = v.x / fs; v.y = v.y / fs; }
        data[lbase + g.fft.off(uint32_t(d))] = v;
    };
    if (!valid) return;
    if constexpr (RL == 1) {
#pragma unroll 4
        for (int d = m0; d < N; d += MSTEP) put(d, make_float2(re[mline * LD + mf_slot(d)], im[mline * LD + mf_slot(d)]));
    } else {
        constexpr int NB = N / RL;   // butterflies of the last stage per line
#pragma unroll 4
        for (int j = m0; j < NB; j += MSTEP) {
            float2 x[RL];
#pragma unroll
            for (int r = 0; r < RL; ++r)
                x[r] = make_float2(re[mline * LD + mf_slot(j + r * NB)], im[mline * LD + mf_slot(j + r * NB)]);
            dft<RL, DIR>(x);
#pragma unroll
            for (int q = 0; q < RL; ++q) put(j + q * NB, x[q]);
        }
    }
}


// Splits 8 floats into hi and lo operand vectors, converting in pairs
// (one v_cvt_pk per pair for each half).
template <typename H> struct Split8;
template <> struct Split8<__bf16> {
    __device__ static void run(const float (&x)[8], bf16x8& hi, bf16x8& lo) {
        typedef __bf16 b2 __attribute__((ext_vector_type(2)));
        typedef float f2 __attribute__((ext_vector_type(2)));
        uint32_t hp[4], lp[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const f2 v = {x[2 * k], x[2 * k + 1]};
            hp[k] = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, b2));
            const f2 r = {x[2 * k] - __uint_as_float(hp[k] << 16), x[2 * k + 1] - __uint_as_float(hp[k] & 0xffff0000u)};
            lp[k] = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, b2));
        }
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        const u4 h = {hp[0], hp[1], hp[2], hp[3]}, l = {lp[0], lp[1], lp[2], lp[3]};
        hi = __builtin_bit_cast(bf16x8, h);
        lo = __builtin_bit_cast(bf16x8, l);
    }
};
template <> struct Split8<_Float16> {
    __device__ static void run(const float (&x)[8], f16x8& hi, f16x8& lo) {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        typedef float f2 __attribute__((ext_vector_type(2)));
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const f2 v = {x[2 * k], x[2 * k + 1]};
            const h2 h = __builtin_convertvector(v, h2);
            const f2 r = {x[2 * k] - float(h[0]), x[2 * k + 1] - float(h[1])};
            const h2 l = __builtin_convertvector(r, h2);
            hi[2 * k] = h[0];
            hi[2 * k + 1] = h[1];
            lo[2 * k] = l[0];
            lo[2 * k + 1] = l[1];
        }
    }
};

// Lines of 256 = 16 * 16 and 512 = 16 * 16 * 2 in tiles of 16 lines, with one
// LDS exchange, like the FP32 kernels: the first radix-16 stage takes its
// operands straight from global memory and the last stage (radix 16, or the
// radix-2 stage of 512) stores straight to it.  A stage-0 column group is
// one butterfly index across the 16 lines of a strided tile (lanes l % 16 =
// the 16 adjacent columns) or 16 consecutive butterflies of one contiguous
// line, so each load and store instruction covers whole 128-byte rows; each
// lane loads the real or the imaginary float its K slot needs.  For 512 the
// two stage-1 groups whose outputs meet in a radix-2 butterfly (butterflies j
// and j + 16) are computed by the same wavefront, so the radix-2 stage runs
// in registers.
//
// Index algebra (BPL = N/16 butterflies per line, q = 4*(l/16) + i):
//   stage 0 butterfly j: inputs j + r*BPL, output q at d = 16 j + q, LDS slot
//     17 j + q; stage-1 twiddle W_{16 BPL}^{(d mod 16) (d / BPL)} = table entry
//     16 + 16 q + j / (BPL/16).
//   stage 1 butterfly j1: inputs at slots slot(j1) + r * 17 BPL/16; output q
//     at d = j1 + 16 q (+ 256 for j1 >= 16); radix-2 twiddle of d: 272 + 2 d + 1.
template <typename H, int N, bool STRIDED, int DIR, int NW, bool O32>
__global__ __launch_bounds__(64 * NW, NW == 8 ? 4 : 2) void mfma_fft_direct(float2* __restrict__ data, FftGeom g,
                                                            const float2* __restrict__ tw) {
    // O32: every byte offset of the pass fits in 32 bits, so addresses are the
    // buffer pointer plus a 32-bit lane offset.
    using Off = typename std::conditional<O32, uint32_t, int64_t>::type;
    using Tr = MfTraits<H>;
    using V8 = typename Tr::V8;
    static_assert(N == 256 || N == 512, "direct kernel lengths");
    static_assert(NW == 4 || NW == 8, "4 or 8 wavefronts");
    constexpr int T = 16;
    constexpr int BPL = N / 16;       // radix-16 butterflies per line
    constexpr int JB = BPL / 16;      // contiguous: column groups per line (1 or 2)
    constexpr int A = 16 / NW;        // contiguous: lines per wavefront; strided: stage-1 pairs per wavefront
    constexpr int GW = BPL / NW;      // column groups per wavefront and stage
    constexpr int LD = N + N / 16 + 1;   // odd row: the 16 lines of a strided group hit distinct banks
    constexpr int RS = 17 * BPL / 16;    // LDS slot step between the inputs of a stage-1 butterfly
    __shared__ float lds[2 * T * LD + 2 * (N == 512 ? 768 : 256)];

    const int tid = int(threadIdx.x);
    const int lane = tid & 63, cn = lane & 15, kg = lane >> 4, kh = kg & 1, ri = kg >> 1, w = tid >> 6;
    // XCD-aware tile order: measured fastest for every pass, the far-strided x
    // pass included (0.51 ms against 0.56-0.70 ms for z-tiles-outermost or
    // hardware order on 512^3 over dims [1,4,4])
    const uint32_t block = xcd_block(blockIdx.x, gridDim.x);
    const char* const datab = reinterpret_cast<const char*>(data);
    char* const outb = reinterpret_cast<char*>(data);

    // line bases (bytes): strided lanes stay on column cn; contiguous wavefront w owns lines A*w .. A*w+A-1
    Off base[A];
    bool valid[A];
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const int line = STRIDED ? cn : A * w + a;
        base[a] = Off(8 * line_base(g, STRIDED, T, 1, block, line, &valid[a], false));
    }

    // ---- stage 0 operands from global memory ----
    // off(j + r*BPL) = off(j) + off(r*BPL): the axis block extent and BPL are
    // powers of two, so one divides the other and j < BPL never carries.
    // Lanes l and l ^ 32 need the real and the imaginary parts of the same 8
    // inputs: each loads 4 of them whole and they trade halves.
    Off offr[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) offr[i] = Off(8 * g.fft.off(uint32_t((8 * kh + 4 * ri + i) * BPL)));
    float2 c[GW][4];
#pragma unroll
    for (int gi = 0; gi < GW; ++gi) {
        const int a = STRIDED ? 0 : gi / JB;
        const int j = STRIDED ? w * GW + gi : (gi % JB) * 16 + cn;
        const Off pl = base[a] + Off(8 * g.fft.off(uint32_t(j)));
#pragma unroll
        for (int i = 0; i < 4; ++i)
            c[gi][i] = valid[a] ? ld_global<true>(reinterpret_cast<const float2*>(datab + (pl + offr[i])))
                                : make_float2(0.f, 0.f);
    }
    // twiddles of stage 1 (and of the radix-2 stage of 512) into LDS
    constexpr int NTW = (N == 512 ? 768 : 256);
    float2* const ltw = reinterpret_cast<float2*>(lds + 2 * T * LD);
    for (int t = tid; t < NTW; t += 64 * NW) ltw[t] = tw[16 + t];
    float x[GW][8];
#pragma unroll
    for (int gi = 0; gi < GW; ++gi) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float give = ri ? c[gi][i].x : c[gi][i].y;
            const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(give), __float_as_uint(give), false, false);
            const float got = __uint_as_float(ri ? sw[0] : sw[1]);   // the partner's value
            x[gi][i] = ri ? got : c[gi][i].x;
            x[gi][4 + i] = ri ? c[gi][i].y : got;
        }
    }

    V8 arh, arl, aih, ail;
    {
        float fr[8], fi[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float2 wv = tw[(cn * (8 * kh + i)) & 15];
            fr[i] = kg < 2 ? wv.x : -wv.y;
            fi[i] = kg < 2 ? wv.y : wv.x;
        }
        Split8<H>::run(fr, arh, arl);
        Split8<H>::run(fi, aih, ail);
    }

    // one radix-16 DFT of the lane's column group: y[i] = output 4*kg + i
    auto dft16_group = [&](float (&xv)[8], float2 (&y)[4]) {
        float s = 1.f;
        if constexpr (Tr::kScale) {
            float mx = 0.f;
#pragma unroll
            for (int i = 0; i < 8; ++i) mx = fmaxf(mx, fabsf(xv[i]));
            mx = fmaxf(mx, __shfl_xor(mx, 16));
            mx = fmaxf(mx, __shfl_xor(mx, 32));
            int e = 0;
            (void)frexpf(mx, &e);
            const int sh = 14 - e;
            s = mx > 0.f ? ldexpf(1.f, sh < -126 ? -126 : (sh > 126 ? 126 : sh)) : 1.f;
#pragma unroll
            for (int i = 0; i < 8; ++i) xv[i] *= s;
        }
        V8 bh, bl;
        Split8<H>::run(xv, bh, bl);
        f32x4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
        yr = Tr::mma(arh, bh, yr);
        yr = Tr::mma(arh, bl, yr);
        yr = Tr::mma(arl, bh, yr);
        yi = Tr::mma(aih, bh, yi);
        yi = Tr::mma(aih, bl, yi);
        yi = Tr::mma(ail, bh, yi);
        if constexpr (Tr::kScale) {
            const float inv = 1.f / s;   // exact: a power of two
#pragma unroll
            for (int i = 0; i < 4; ++i) y[i] = make_float2(yr[i] * inv, yi[i] * inv);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) y[i] = make_float2(yr[i], yi[i]);
        }
    };

    __syncthreads();   // twiddles in LDS

    // ---- stage 0 (Ns = 1): outputs to LDS with the stage-1 twiddles ----
    {
        const int lrow = STRIDED ? cn : A * w;   // + gi / JB for contiguous
        float* const wre = lds + lrow * LD + 4 * kg;
        float* const wim = wre + T * LD;
        const float2* const twq = ltw + 16 * 4 * kg;   // + 16 i + j / JB
#pragma unroll
        for (int gi = 0; gi < GW; ++gi) {
            const int lo = STRIDED ? 0 : (gi / JB) * LD;
            const int j = STRIDED ? w * GW + gi : (gi % JB) * 16 + cn;
            float2 y[4];
            dft16_group(x[gi], y);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float2 v = c_mul(y[i], twq[16 * i + j / JB]);
                wre[lo + 17 * j + i] = v.x;
                wim[lo + 17 * j + i] = v.y;
            }
        }
    }
    __syncthreads();

    // ---- stage 1 (Ns = 16), then the radix-2 stage of 512, to global memory ----
    // Stage 1 only reads LDS, so it runs pair by pair: pair a is butterfly
    // j1 = jlo (and, for 512, its radix-2 partner jlo + 16).
    const float* const rp = lds + ri * T * LD + 8 * kh * RS;
    auto read_group = [&](int line, int j1, float (&xv)[8]) {
        const float* const q = rp + line * LD + j1 + j1 / 16;
#pragma unroll
        for (int i = 0; i < 8; ++i) xv[i] = q[i * RS];
    };
    // the inverse's 1/(nx*ny*nz) as one multiplication (exact for powers of two)
    const bool scaled = g.scale_mode != 0;
    const float fs = g.scale_mode == 2 ? float(1.0 / g.scale) : float(g.scale);
    // outputs d = jlo + 16*q (+ 256): off(d) = off(jlo) + off(16*q) (+ off(256))
    Off offq[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) offq[i] = Off(8 * g.fft.off(uint32_t(16 * (4 * kg + i))));
    const Off off256 = N == 512 ? Off(8 * g.fft.off(256u)) : Off(0);
    auto put = [&](Off o, float2 v) {
        if (scaled) { v.x *= fs; v.y *= fs; }
        st_global<true>(reinterpret_cast<float2*>(outb + o), v);
    };
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const int ab = STRIDED ? 0 : a;
        const int line = STRIDED ? cn : A * w + a;
        const int jlo = STRIDED ? A * w + a : cn;   // j1 % 16
        const Off po = base[ab] + Off(8 * g.fft.off(uint32_t(jlo)));
        float xv[8];
        read_group(line, jlo, xv);
        float2 y0[4];
        dft16_group(xv, y0);
        if constexpr (N == 256) {
            if (valid[ab]) {
#pragma unroll
                for (int i = 0; i < 4; ++i) put(po + offq[i], y0[i]);
            }
        } else {
            read_group(line, jlo + 16, xv);
            float2 y1[4];
            dft16_group(xv, y1);
            const float2* const t2 = ltw + (mf_twoff(N, 2) - 16) + 1 + 2 * jlo + 32 * 4 * kg;   // + 32 i
            if (valid[ab]) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float2 t = c_mul(y1[i], t2[32 * i]);
                    put(po + offq[i], c_add(y0[i], t));
                    put(po + offq[i] + off256, c_sub(y0[i], t));
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Split-BF16 and split-FP16 passes of 256 = 16 * 16 and 512 = 16 * 16 * 2
// point lines with the complex values interleaved along K, one LDS exchange
// of operands that are already split, and buffer addressing.
//
// K order.  The 32 K slots of a radix-16 product hold the 16 inputs as
// (re, im) pairs: lane group g (= lane / 16) holds inputs r = 4g .. 4g+3, so
// a lane's 8 bf16 operand values are 4 whole complex numbers -- the 4 values
// it loads -- and no lane exchange is needed.  Rows of the DFT operand are
// (Re F, -Im F) pairs for the real outputs and (Im F, Re F) pairs for the
// imaginary outputs.  Products are accumulated small to large:
// lo*hi, hi*lo, hi*hi.
//
// Stockham indices (B0 = N/16 radix-16 butterflies per line, R2 = N/256):
//   stage 0 butterfly j reads x[j + B0 r] and writes d = 16 j + q;
//   stage 1 butterfly j1 reads d = j1 + B0 r, i.e. output q = j1 % 16 of
//     stage-0 butterfly j = j1/16 + R2 r, twiddled by W_256^(r q), and writes
//     256 (j1/16) + j1 % 16 + 16 q';
//   stage 2 (R2 = 2) butterfly j2 = j1 % 16 + 16 q' combines the stage-1
//     outputs q' of j1 = j2 % 16 and j2 % 16 + 16, the second twiddled by
//     W_512^j2, and writes j2 and j2 + 256.
// So a stage-0 output (j, q) is K slot r = j / R2 of stage-1 butterfly
// j1 = q + 16 (j % R2).
//
// Strided tiles (16 adjacent columns; MFMA columns = the 16 columns):
//   phase s = 0 .. R2-1: wave w runs the stage-0 butterflies
//   j = s + R2 (4w + m), m = 0..3, so for each of its outputs q a lane holds
//   K slots 4w .. 4w+3 of stage-1 butterfly q + 16 s in one column -- a
//   16-byte hi and a 16-byte lo operand written with one store each; after a
//   barrier wave w runs stage-1 butterflies 4w .. 4w+3 (+16 s) and either
//   stores them (R2 = 1), keeps them (s = 0) or completes the radix-2 stage
//   (s = 1).  The exchange buffer holds one phase: 32 KiB.
// Contiguous lines (16 lines per workgroup): stage 0 runs with the data as
//   the A operand (rows = 16 consecutive butterflies), so a lane gets output
//   q = lane % 16 of butterflies 4g .. 4g+3, which are 4 (R2 = 1) or 2
//   (R2 = 2) consecutive K slots of one stage-1 butterfly; stage 1 runs with
//   columns = 16 consecutive butterflies j1 of a line, so the final stores
//   cover 128-byte rows.  Each wave exchanges only its own lines (8 KiB of
//   LDS per phase), so no workgroup barrier is needed.
//
// Addresses are 32-bit byte offsets from the buffer start (the launcher
// checks the pass's extent fits), the wave-uniform part in the scalar
// offset; loads and stores outside the extent are dropped by the buffer
// range check.
//
// FP16 (F16 = true).  FP16's range needs the operands scaled.  Each line gets
// one power-of-two scale s that brings its largest component magnitude into
// [2^14, 2^15): a strided column's maximum over its 4 lane groups (xor
// shuffles) and the 4 waves (64 floats of LDS, read after the barrier that
// publishes the twiddles), a contiguous line's over its wave.  Stage 0 takes
// x s; the stage-1 operands are its outputs times the twiddle / 16 (the LDS
// twiddle table is stored divided by 16, exactly), so they stay below 2^15
// whatever the stage-0 growth; the stores multiply by 16 / s.  Every product
// is exact in its power of two, so the hi + lo split keeps about 22 bits of
// each value relative to its own line, as the per-butterfly scale of the
// split-K-halves kernels does.
// ---------------------------------------------------------------------------
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
// Strided-tile variant: each load and store instruction covers the rows of 4
// neighbouring butterflies (lane groups transposed in registers).
constexpr int kMfRowAdj = 1 << 8;
typedef int i32x2 __attribute__((ext_vector_type(2)));

// Packed (a, b) rounded to bf16 (a in the low half).
__device__ __forceinline__ uint32_t pk_bf16(float a, float b) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef __bf16 b2 __attribute__((ext_vector_type(2)));
    const f2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, b2));
}
// Packed (a, b) rounded to fp16 (a in the low half).
__device__ __forceinline__ uint32_t pk_f16(float a, float b) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const f2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, h2));
}
// The two halves of a packed fp16 pair as floats.
__device__ __forceinline__ float2 unpk_f16(uint32_t p) {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const h2 h = __builtin_bit_cast(h2, p);
    return make_float2(float(h[0]), float(h[1]));
}
// One complex value as a packed 16-bit pair hi = round(v) and lo = round(v - hi)
// (bf16, or fp16 with F16).
template <bool F16 = false> __device__ __forceinline__ void split_c(float2 v, uint32_t& hi, uint32_t& lo) {
    if constexpr (F16) {
        hi = pk_f16(v.x, v.y);
        const float2 h = unpk_f16(hi);
        lo = pk_f16(v.x - h.x, v.y - h.y);
    } else {
        hi = pk_bf16(v.x, v.y);
        lo = pk_bf16(v.x - __uint_as_float(hi << 16), v.y - __uint_as_float(hi & 0xffff0000u));
    }
}
template <bool F16 = false> __device__ __forceinline__ void split4(const float2 (&x)[4], u32x4& h, u32x4& l) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t a, b;
        split_c<F16>(x[k], a, b);
        h[k] = a;
        l[k] = b;
    }
}
template <bool F16 = false> __device__ __forceinline__ f32x4 mma_h(u32x4 a, u32x4 b, f32x4 c) {
    if constexpr (F16)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0,
                                                      0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                       0, 0, 0);
}

// DFT-16 operand rows (real and imaginary output rows, hi and lo).
struct F16Op {
    u32x4 rh, rl, ih, il;
};
// Normal form: columns = butterflies (data = B operand); lane (c, g) gets
// outputs 4g .. 4g+3 of column c.
template <bool F16 = false>
__device__ __forceinline__ void dft16_cols(const F16Op& F, u32x4 bh, u32x4 bl, f32x4& yr, f32x4& yi) {
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
    yr = mma_h<F16>(F.rl, bh, z);
    yi = mma_h<F16>(F.il, bh, z);
    yr = mma_h<F16>(F.rh, bl, yr);
    yi = mma_h<F16>(F.ih, bl, yi);
    yr = mma_h<F16>(F.rh, bh, yr);
    yi = mma_h<F16>(F.ih, bh, yi);
}
// Transposed form: rows = butterflies (data = A operand); lane (c, g) gets
// output c of rows (butterflies) 4g .. 4g+3.
template <bool F16 = false>
__device__ __forceinline__ void dft16_rows(const F16Op& F, u32x4 ah, u32x4 al, f32x4& yr, f32x4& yi) {
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
    yr = mma_h<F16>(al, F.rh, z);
    yi = mma_h<F16>(al, F.ih, z);
    yr = mma_h<F16>(ah, F.rl, yr);
    yi = mma_h<F16>(ah, F.il, yi);
    yr = mma_h<F16>(ah, F.rh, yr);
    yi = mma_h<F16>(ah, F.ih, yi);
}

// Power-of-two scale for the fp16 operands of a line whose largest component
// magnitude is mx: brings mx into [2^14, 2^15), inside fp16's range with the
// headroom a radix-16 stage needs once the next stage's operands are divided
// by 16.  1 for an all-zero line.
__device__ __forceinline__ float f16_scale(float mx) {
    int e = 0;
    (void)frexpf(mx, &e);
    const int sh = 15 - e;
    return mx > 0.f ? ldexpf(1.f, sh < -126 ? -126 : (sh > 126 ? 126 : sh)) : 1.f;
}
// Largest component magnitude, on the bit patterns: for magnitudes (sign
// cleared) unsigned order is float order, and integer max needs no NaN
// canonicalisation of its inputs (one v_max3_u32 per complex value).
__device__ __forceinline__ uint32_t cmax(uint32_t m, float2 v) {
    const uint32_t a = __float_as_uint(v.x) & 0x7fffffffu, b = __float_as_uint(v.y) & 0x7fffffffu;
    return max(m, max(a, b));
}

template <bool NTMP>
__device__ __forceinline__ float2 bld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    const i32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, int(voff), int(soff), NTMP ? 2 : 0);
    return __builtin_bit_cast(float2, v);
}
template <bool NTMP>
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float2 v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(i32x2, v), r, int(voff), int(soff), NTMP ? 2 : 0);
}

// Workgroup barrier that orders LDS accesses only: global loads and stores
// in flight stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// 4 x 4 transpose of 4-complex blocks across the lane groups (lane / 16) of
// a wave: block b of lane group g moves to block g of lane group b.  Two
// 2 x 2 block transposes: the high bits of group and block through
// v_permlane32_swap (upper 32 lanes of the first operand <-> lower 32 lanes
// of the second), then the low bits through v_permlane16_swap (odd rows of
// the first operand <-> even rows of the second).
__device__ __forceinline__ void swap_lanes32(float& x, float& y) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    x = __uint_as_float(r[0]);
    y = __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap_lanes16(float& x, float& y) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    x = __uint_as_float(r[0]);
    y = __uint_as_float(r[1]);
}
__device__ __forceinline__ void xpose4(float2 (&B)[4][4]) {
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            swap_lanes32(B[k][e].x, B[k + 2][e].x);
            swap_lanes32(B[k][e].y, B[k + 2][e].y);
        }
#pragma unroll
    for (int k = 0; k < 4; k += 2)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            swap_lanes16(B[k][e].x, B[k + 1][e].x);
            swap_lanes16(B[k][e].y, B[k + 1][e].y);
        }
}

template <int N, bool STRIDED, int FLAGS, bool F16 = false>
__global__ __launch_bounds__(256) void mf2_kernel(float2* __restrict__ data, FftGeom g, const float2* __restrict__ tw,
                                                  uint32_t extent) {
    static_assert(N == 256 || N == 512, "mf2 lengths");
    constexpr int R2 = N / 256;
    constexpr int B0 = N / 16;
    constexpr bool NTMP = (FLAGS & KF_NT) != 0;
    constexpr bool A2MAJ = (FLAGS & KF_A2MAJ) != 0;
    constexpr bool ROWADJ = (FLAGS & kMfRowAdj) != 0;
    __shared__ u32x4 xh[1024];
    __shared__ u32x4 xl[1024];
    // W_256^(q r) at [16 q + r]; the table is symmetric (each entry is
    // computed from the product q r), so lanes that vary q read [16 r + q]:
    // consecutive lanes, consecutive banks
    __shared__ float2 tw1[256];
    __shared__ float2 tw2[R2 > 1 ? 256 : 1];    // W_512^k, k < 256
    // fp16: each strided column's largest magnitude, per wave
    __shared__ uint32_t colmax[F16 && STRIDED ? 64 : 1];

    const int tid = int(threadIdx.x);
    const int lane = tid & 63, cn = lane & 15, gq = lane >> 4;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t block = (FLAGS & KF_XCD) ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(data, 0, int(extent), 0x00020000);

    // fp16: the stage-1 operands are the stage-0 outputs divided by 16 (exact),
    // folded into the twiddles, so they stay inside fp16's range
    if constexpr (F16) {
        const float2 t = tw[16 + tid];
        tw1[tid] = make_float2(t.x * 0.0625f, t.y * 0.0625f);
    } else {
        tw1[tid] = tw[16 + tid];
    }
    if constexpr (R2 > 1) tw2[tid] = tw[16 + 256 + 2 * tid + 1];

    F16Op F;
    {
        float2 fr[4], fi[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float2 wv = tw[(cn * (4 * gq + e)) & 15];
            fr[e] = make_float2(wv.x, -wv.y);
            fi[e] = make_float2(wv.y, wv.x);
        }
        split4<F16>(fr, F.rh, F.rl);
        split4<F16>(fi, F.ih, F.il);
    }
    // byte offsets of the K rows n = B0 (4 gq + e), and of the outputs 16 q'
    uint32_t offr[4], offq[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        offr[e] = 8u * uint32_t(g.fft.off(uint32_t(B0 * (4 * gq + e))));
        offq[e] = 8u * uint32_t(g.fft.off(uint32_t(16 * (4 * gq + e))));
    }
    const uint32_t off256 = R2 > 1 ? 8u * uint32_t(g.fft.off(256u)) : 0u;
    // the inverse's 1/(nx*ny*nz) as one multiplication (exact for the
    // power-of-two lengths these kernels take); 1 when the pass does not scale
    const float fs = g.scale_mode == 0 ? 1.f : g.scale_mode == 2 ? float(1.0 / g.scale) : float(g.scale);
    // u: fp16 only, fs * 16 / (the line's operand scale), undoing the scaling
    // with the pass's scale in one multiplication (both factors are powers of
    // two or fs is applied to a value the power of two scales exactly)
    auto sc = [&](float2 v, float u = 1.f) {
        const float f = F16 ? u : fs;
        v.x *= f;
        v.y *= f;
        return v;
    };

    if constexpr (STRIDED) {
        bool valid;
        const uint32_t lb = 8u * uint32_t(line_base(g, true, 16, 1, block, cn, &valid, A2MAJ));
        float2 v[R2][4][4];
        if constexpr (ROWADJ) {
            // lane group gq loads all 16 K rows of butterfly m = gq (rows
            // j + B0 r, so one load covers the rows of 4 neighbouring
            // butterflies), then the 4 x 4 block transpose hands each lane
            // group its K block of every butterfly
#pragma unroll
            for (int s = 0; s < R2; ++s) {
                const uint32_t vj = lb + 8u * uint32_t(g.fft.off(uint32_t(s + R2 * (4 * w + gq))));
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    v[s][r >> 2][r & 3] = bld<NTMP>(rs, vj, 8u * uint32_t(g.fft.off(uint32_t(B0 * r))));
            }
        } else {
#pragma unroll
            for (int s = 0; s < R2; ++s)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const uint32_t oj = 8u * uint32_t(g.fft.off(uint32_t(s + R2 * (4 * w + m))));
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[s][m][e] = bld<NTMP>(rs, lb + offr[e], oj);
                }
        }
        // fp16: one power-of-two scale per column (line), from its largest
        // magnitude over the 4 lane groups and the 4 waves
        float us = 1.f;
        if constexpr (F16) {
            uint32_t mx = 0;
#pragma unroll
            for (int s = 0; s < R2; ++s)
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int e = 0; e < 4; ++e) mx = cmax(mx, v[s][m][e]);
            mx = max(mx, uint32_t(__shfl_xor(int(mx), 16)));
            mx = max(mx, uint32_t(__shfl_xor(int(mx), 32)));
            if (gq == 0) colmax[w * 16 + cn] = mx;
        }
        lds_barrier();   // twiddles (and column maxima) in LDS
        if constexpr (F16) {
            const uint32_t mx = max(max(colmax[cn], colmax[16 + cn]), max(colmax[32 + cn], colmax[48 + cn]));
            const float s0 = f16_scale(__uint_as_float(mx));
            us = (16.f / s0) * fs;
#pragma unroll
            for (int s = 0; s < R2; ++s)
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        v[s][m][e].x *= s0;
                        v[s][m][e].y *= s0;
                    }
        }
        // row-adjacent stores: lane group g' stores butterfly 4w + g'
        const uint32_t vq = ROWADJ ? lb + 8u * uint32_t(g.fft.off(uint32_t(4 * w + gq))) : 0u;
        float2 y0[4][4];
#pragma unroll
        for (int s = 0; s < R2; ++s) {
            if constexpr (ROWADJ) xpose4(v[s]);
            f32x4 yr[4], yi[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                u32x4 bh, bl;
                split4<F16>(v[s][m], bh, bl);
                dft16_cols<F16>(F, bh, bl, yr[m], yi[m]);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int q = 4 * gq + i;
                float2 c[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) c[m] = c_mul(make_float2(yr[m][i], yi[m][i]), tw1[16 * q + 4 * w + m]);
                u32x4 h, l;
                split4<F16>(c, h, l);
                xh[(4 * q + w) * 16 + cn] = h;
                xl[(4 * q + w) * 16 + cn] = l;
            }
            lds_barrier();
            float2 o[2][4][4];   // outputs to store: [half][butterfly a][i]
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int q1 = 4 * w + a;
                f32x4 zr, zi;
                dft16_cols<F16>(F, xh[(4 * q1 + gq) * 16 + cn], xl[(4 * q1 + gq) * 16 + cn], zr, zi);
                if constexpr (R2 == 1) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) o[0][a][i] = sc(make_float2(zr[i], zi[i]), us);
                } else if (s == 0) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) y0[a][i] = make_float2(zr[i], zi[i]);
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float2 t = c_mul(make_float2(zr[i], zi[i]), tw2[q1 + 16 * (4 * gq + i)]);
                        o[0][a][i] = sc(c_add(y0[a][i], t), us);
                        o[1][a][i] = sc(c_sub(y0[a][i], t), us);
                    }
                }
            }
            if (R2 == 1 || s == 1) {
#pragma unroll
                for (int h = 0; h < R2; ++h) {
                    if constexpr (ROWADJ) {
                        // lane group g' gets outputs 4k + i of butterfly 4w + g' in block k
                        xpose4(o[h]);
#pragma unroll
                        for (int k = 0; k < 4; ++k)
#pragma unroll
                            for (int i = 0; i < 4; ++i)
                                bst<NTMP>(rs, vq, 8u * uint32_t(g.fft.off(uint32_t(16 * (4 * k + i) + 256 * h))),
                                          o[h][k][i]);
                    } else {
#pragma unroll
                        for (int a = 0; a < 4; ++a) {
                            const uint32_t oq1 = 8u * uint32_t(g.fft.off(uint32_t(4 * w + a))) + (h ? off256 : 0u);
#pragma unroll
                            for (int i = 0; i < 4; ++i) bst<NTMP>(rs, lb + offq[i], oq1, o[h][a][i]);
                        }
                    }
                }
            }
            if (s + 1 < R2) lds_barrier();
        }
    } else {
        constexpr int LP = 16 / R2;   // lines per phase
        constexpr int LW = LP / 4;    // lines per wave and phase
        const uint32_t ocn = 8u * uint32_t(g.fft.off(uint32_t(cn)));
        uint32_t vo[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) vo[e] = ocn + offr[e];
        // line bases (wave-uniform) and validity of the wave's lines
        uint32_t lbase[R2][LW];
        bool lval[R2][LW];
#pragma unroll
        for (int ph = 0; ph < R2; ++ph)
#pragma unroll
            for (int la = 0; la < LW; ++la) {
                bool ok;
                const int64_t b = line_base(g, false, 16, 1, block, ph * LP + w * LW + la, &ok, false);
                lbase[ph][la] = 8u * uint32_t(ok ? b : 0);
                lval[ph][la] = ok;
            }
        float2 v[R2][LW][R2][4];
#pragma unroll
        for (int ph = 0; ph < R2; ++ph)
#pragma unroll
            for (int la = 0; la < LW; ++la)
#pragma unroll
                for (int jh = 0; jh < R2; ++jh) {
                    const uint32_t so = lbase[ph][la] + 8u * uint32_t(g.fft.off(uint32_t(16 * jh)));
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        v[ph][la][jh][e] = lval[ph][la] ? bld<NTMP>(rs, vo[e], so) : make_float2(0.f, 0.f);
                }
        // fp16: one power-of-two scale per line, from its largest magnitude
        // over the wave (a wave holds whole lines)
        float ul[R2][LW];
#pragma unroll
        for (int ph = 0; ph < R2; ++ph)
#pragma unroll
            for (int la = 0; la < LW; ++la) {
                ul[ph][la] = 1.f;
                if constexpr (F16) {
                    uint32_t mx = 0;
#pragma unroll
                    for (int jh = 0; jh < R2; ++jh)
#pragma unroll
                        for (int e = 0; e < 4; ++e) mx = cmax(mx, v[ph][la][jh][e]);
#pragma unroll
                    for (int x = 1; x < 64; x <<= 1) mx = max(mx, uint32_t(__shfl_xor(int(mx), x)));
                    const float s0 = f16_scale(__uint_as_float(mx));
                    ul[ph][la] = (16.f / s0) * fs;
#pragma unroll
                    for (int jh = 0; jh < R2; ++jh)
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            v[ph][la][jh][e].x *= s0;
                            v[ph][la][jh][e].y *= s0;
                        }
                }
            }
        lds_barrier();   // twiddles in LDS
        u32x4* const wh = xh + w * 256;   // this wave's exchange region
        u32x4* const wl = xl + w * 256;
#pragma unroll
        for (int ph = 0; ph < R2; ++ph) {
            // stage 0: rows = butterflies j = 16 jh + row of line la
#pragma unroll
            for (int la = 0; la < LW; ++la)
#pragma unroll
                for (int jh = 0; jh < R2; ++jh) {
                    u32x4 ah, al;
                    split4<F16>(v[ph][la][jh], ah, al);
                    f32x4 yr, yi;
                    dft16_rows<F16>(F, ah, al, yr, yi);
                    // lane holds output q = cn of butterflies j = 16 jh + 4 gq + i
                    float2 c[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int r = (16 * jh + 4 * gq + i) / R2;   // K slot in the stage-1 butterfly
                        c[i] = c_mul(make_float2(yr[i], yi[i]), tw1[16 * r + cn]);
                    }
                    if constexpr (R2 == 1) {
                        u32x4 h, l;
                        split4<F16>(c, h, l);
                        wh[(la * 4 + gq) * 16 + cn] = h;
                        wl[(la * 4 + gq) * 16 + cn] = l;
                    } else {
#pragma unroll
                        for (int s = 0; s < 2; ++s) {   // outputs i = s, s + 2: K slots 8 jh + 2 gq + {0, 1}
                            uint32_t h0, l0, h1, l1;
                            split_c<F16>(c[s], h0, l0);
                            split_c<F16>(c[s + 2], h1, l1);
                            const int slot = ((la * 2 + s) * 4 + 2 * jh + (gq >> 1)) * 16 + cn;
                            u32x2* const ph2 = reinterpret_cast<u32x2*>(wh + slot) + (gq & 1);
                            u32x2* const pl2 = reinterpret_cast<u32x2*>(wl + slot) + (gq & 1);
                            *ph2 = u32x2{h0, h1};
                            *pl2 = u32x2{l0, l1};
                        }
                    }
                }
            // the wave reads back only what it wrote: LDS accesses of one
            // wavefront complete in order
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int la = 0; la < LW; ++la) {
                if (!lval[ph][la]) continue;
                const uint32_t so = lbase[ph][la];
                f32x4 zr[R2], zi[R2];
#pragma unroll
                for (int s = 0; s < R2; ++s) {
                    const int slot = ((la * R2 + s) * 4 + gq) * 16 + cn;
                    dft16_cols<F16>(F, wh[slot], wl[slot], zr[s], zi[s]);
                }
                // lane holds outputs q' = 4 gq + i of butterflies j1 = cn + 16 s
                if constexpr (R2 == 1) {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        bst<NTMP>(rs, ocn + offq[i], so, sc(make_float2(zr[0][i], zi[0][i]), ul[ph][la]));
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float2 y0 = make_float2(zr[0][i], zi[0][i]);
                        const float2 t = c_mul(make_float2(zr[1][i], zi[1][i]), tw2[cn + 16 * (4 * gq + i)]);
                        bst<NTMP>(rs, ocn + offq[i], so, sc(c_add(y0, t), ul[ph][la]));
                        bst<NTMP>(rs, ocn + offq[i], so + off256, sc(c_sub(y0, t), ul[ph][la]));
                    }
                }
            }
            if (ph + 1 < R2) {   // the next phase overwrites this wave's region
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
    }
}

// Contiguous lines of 1024 = 16 * 16 * 4 and 2048 = 16 * 16 * 8 points,
// split-BF16: the two radix-16 stages as in mf2_kernel's contiguous form, then
// a radix-R2 stage (R2 = N / 256) on VALU.  A workgroup takes 8192 / N lines
// and each wave 8 / R2 whole lines, exchanged through its own 16 KiB of LDS.
// Stage-0 output q of butterfly j is K slot j / R2 of stage-1 butterfly
// q + 16 (j % R2); stage-1 output q' of butterflies j2 % 16 + 16 s (s < R2)
// are the R2 inputs of radix-R2 butterfly j2 = j1 % 16 + 16 q', twiddled by
// W_N^(s j2), written to j2 + 256 q2.
template <int N, int DIR, int FLAGS, bool F16 = false>
__global__ __launch_bounds__(256) void mf2_long_kernel(float2* __restrict__ data, FftGeom g,
                                                       const float2* __restrict__ tw, uint32_t extent) {
    static_assert(N == 1024 || N == 2048, "long lengths");
    constexpr int R2 = N / 256;
    constexpr int B0 = N / 16;
    constexpr int LW = 8 / R2;    // lines per wave
    constexpr int TL = 4 * LW;    // lines per workgroup
    constexpr bool NTMP = (FLAGS & KF_NT) != 0;
    __shared__ u32x4 xh[2048];
    __shared__ u32x4 xl[2048];
    __shared__ float2 tw1[256];

    const int tid = int(threadIdx.x);
    const int lane = tid & 63, cn = lane & 15, gq = lane >> 4;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t block = (FLAGS & KF_XCD) ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(data, 0, int(extent), 0x00020000);
    if constexpr (F16) {   // fp16: stage-1 operands divided by 16 through the twiddles
        const float2 t = tw[16 + tid];
        tw1[tid] = make_float2(t.x * 0.0625f, t.y * 0.0625f);
    } else {
        tw1[tid] = tw[16 + tid];
    }

    F16Op F;
    {
        float2 fr[4], fi[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float2 wv = tw[(cn * (4 * gq + e)) & 15];
            fr[e] = make_float2(wv.x, -wv.y);
            fi[e] = make_float2(wv.y, wv.x);
        }
        split4<F16>(fr, F.rh, F.rl);
        split4<F16>(fi, F.ih, F.il);
    }
    const uint32_t ocn = 8u * uint32_t(g.fft.off(uint32_t(cn)));
    uint32_t vo[4], offq[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        vo[e] = ocn + 8u * uint32_t(g.fft.off(uint32_t(B0 * (4 * gq + e))));
        offq[e] = 8u * uint32_t(g.fft.off(uint32_t(16 * (4 * gq + e))));
    }
    const float fs = g.scale_mode == 0 ? 1.f : g.scale_mode == 2 ? float(1.0 / g.scale) : float(g.scale);

    uint32_t lbase[LW];
    bool lval[LW];
#pragma unroll
    for (int la = 0; la < LW; ++la) {
        bool ok;
        const int64_t b = line_base(g, false, TL, 1, block, w * LW + la, &ok, false);
        lbase[la] = 8u * uint32_t(ok ? b : 0);
        lval[la] = ok;
    }
    float2 v[LW][R2][4];
#pragma unroll
    for (int la = 0; la < LW; ++la)
#pragma unroll
        for (int jh = 0; jh < R2; ++jh) {
            const uint32_t so = lbase[la] + 8u * uint32_t(g.fft.off(uint32_t(16 * jh)));
#pragma unroll
            for (int e = 0; e < 4; ++e) v[la][jh][e] = lval[la] ? bld<NTMP>(rs, vo[e], so) : make_float2(0.f, 0.f);
        }
    // fp16: one power-of-two scale per line (a wave holds whole lines); the
    // stores multiply by ul = fs * 16 / scale
    float ul[LW];
#pragma unroll
    for (int la = 0; la < LW; ++la) {
        ul[la] = 1.f;
        if constexpr (F16) {
            uint32_t mx = 0;
#pragma unroll
            for (int jh = 0; jh < R2; ++jh)
#pragma unroll
                for (int e = 0; e < 4; ++e) mx = cmax(mx, v[la][jh][e]);
#pragma unroll
            for (int x = 1; x < 64; x <<= 1) mx = max(mx, uint32_t(__shfl_xor(int(mx), x)));
            const float s0 = f16_scale(__uint_as_float(mx));
            ul[la] = (16.f / s0) * fs;
#pragma unroll
            for (int jh = 0; jh < R2; ++jh)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    v[la][jh][e].x *= s0;
                    v[la][jh][e].y *= s0;
                }
        }
    }
    lds_barrier();   // twiddles in LDS
    uint32_t* const wh = reinterpret_cast<uint32_t*>(xh + w * 512);
    uint32_t* const wl = reinterpret_cast<uint32_t*>(xl + w * 512);
#pragma unroll
    for (int la = 0; la < LW; ++la) {
        // stage 0: rows = butterflies j = 16 jh + row; the lane holds output
        // q = cn of butterflies 16 jh + 4 gq + i
#pragma unroll
        for (int jh = 0; jh < R2; ++jh) {
            u32x4 ah, al;
            split4<F16>(v[la][jh], ah, al);
            f32x4 yr, yi;
            dft16_rows<F16>(F, ah, al, yr, yi);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int j = 16 * jh + 4 * gq + i;
                const int s = j % R2, r = j / R2;
                uint32_t h, l;
                split_c<F16>(c_mul(make_float2(yr[i], yi[i]), tw1[16 * r + cn]), h, l);
                const int slot = ((la * R2 + s) * 4 + (r >> 2)) * 16 + cn;
                wh[4 * slot + (r & 3)] = h;
                wl[4 * slot + (r & 3)] = l;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // stage 1: butterflies j1 = cn + 16 s; the lane gets outputs q' = 4 gq + i
        float2 z[R2][4];
#pragma unroll
        for (int s = 0; s < R2; ++s) {
            const int slot = ((la * R2 + s) * 4 + gq) * 16 + cn;
            f32x4 zr, zi;
            dft16_cols<F16>(F, xh[w * 512 + slot], xl[w * 512 + slot], zr, zi);
#pragma unroll
            for (int i = 0; i < 4; ++i) z[s][i] = make_float2(zr[i], zi[i]);
        }
        // radix-R2 stage and stores
        if (lval[la]) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int j2 = cn + 16 * (4 * gq + i);
                float2 x[R2];
                x[0] = z[0][i];
#pragma unroll
                for (int s = 1; s < R2; ++s) x[s] = c_mul(z[s][i], tw[16 + 256 + j2 * R2 + s]);
                dft<R2, DIR>(x);
#pragma unroll
                for (int q2 = 0; q2 < R2; ++q2) {
                    const float f = F16 ? ul[la] : fs;   // fp16: fs * 16 / (the line's scale)
                    bst<NTMP>(rs, ocn + offq[i], lbase[la] + 8u * uint32_t(g.fft.off(uint32_t(256 * q2))),
                              make_float2(x[q2].x * f, x[q2].y * f));
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Registry and launch.
// ---------------------------------------------------------------------------
namespace {

struct MfEntry {
    int N, strided, dir, fp16, nw, o32;   // nw: wavefronts per workgroup; o32: 32-bit byte offsets
    const void* fn;
};

template <typename H, int N, bool S>
void mf_add(std::vector<MfEntry>& v) {
    constexpr int T = mf_T(N, S);
    constexpr int f16 = std::is_same<H, _Float16>::value ? 1 : 0;
    if constexpr (N == 256 || N == 512) {
        v.push_back(MfEntry{N, S, -1, f16, 8, 0, (const void*)&mfma_fft_direct<H, N, S, -1, 8, false>});
        v.push_back(MfEntry{N, S, 1, f16, 8, 0, (const void*)&mfma_fft_direct<H, N, S, 1, 8, false>});
        v.push_back(MfEntry{N, S, -1, f16, 8, 1, (const void*)&mfma_fft_direct<H, N, S, -1, 8, true>});
        v.push_back(MfEntry{N, S, 1, f16, 8, 1, (const void*)&mfma_fft_direct<H, N, S, 1, 8, true>});
    } else {
        v.push_back(MfEntry{N, S, -1, f16, 4, 0, (const void*)&mfma_fft_kernel<H, N, T, S, -1>});
        v.push_back(MfEntry{N, S, 1, f16, 4, 0, (const void*)&mfma_fft_kernel<H, N, T, S, 1>});
    }
}

template <typename H, int N>
void mf_add_len(std::vector<MfEntry>& v) {
    mf_add<H, N, false>(v);
    mf_add<H, N, true>(v);
}

const std::vector<MfEntry>& mf_registry() {
    static const std::vector<MfEntry> reg = [] {
        std::vector<MfEntry> v;
        mf_add_len<__bf16, 16>(v);
        mf_add_len<__bf16, 32>(v);
        mf_add_len<__bf16, 64>(v);
        mf_add_len<__bf16, 128>(v);
        mf_add_len<__bf16, 256>(v);
        mf_add_len<__bf16, 512>(v);
        mf_add_len<_Float16, 16>(v);
        mf_add_len<_Float16, 32>(v);
        mf_add_len<_Float16, 64>(v);
        mf_add_len<_Float16, 128>(v);
        mf_add_len<_Float16, 256>(v);
        mf_add_len<_Float16, 512>(v);
        return v;
    }();
    return reg;
}

// The interleaved-K split-BF16 kernel (mf2_kernel): variants by tile order
// and cache policy, each also launched with extra dynamic LDS that limits a
// CU to 3 (kMfOcc3) or 2 (kMfOcc2) workgroups.
constexpr int kMfOcc3 = 1 << 9;
constexpr int kMfOcc2 = 1 << 10;

struct Mf2Entry {
    int N, strided, flags, f16;
    const void* fn;
};

template <int N, bool S, int F, bool F16 = false>
void mf2_add(std::vector<Mf2Entry>& v) {
    v.push_back(Mf2Entry{N, S, F, F16 ? 1 : 0, (const void*)&mf2_kernel<N, S, F, F16>});
}

// Split-FP16 variants: the split-BF16 defaults and their near neighbours.
template <int N>
void mf2_add_len_f16(std::vector<Mf2Entry>& v) {
    mf2_add<N, false, KF_XCD | KF_NT, true>(v);
    mf2_add<N, false, KF_NT, true>(v);
    mf2_add<N, true, KF_XCD | KF_NT, true>(v);
    mf2_add<N, true, KF_NT, true>(v);
    mf2_add<N, true, KF_XCD, true>(v);
    mf2_add<N, true, KF_A2MAJ | KF_NT, true>(v);
    mf2_add<N, true, kMfRowAdj | KF_A2MAJ, true>(v);
    mf2_add<N, true, kMfRowAdj | KF_A2MAJ | KF_NT, true>(v);
    mf2_add<N, true, kMfRowAdj | KF_XCD | KF_NT, true>(v);
}

template <int N>
void mf2_add_len(std::vector<Mf2Entry>& v) {
    mf2_add<N, false, KF_XCD | KF_NT>(v);
    mf2_add<N, false, KF_NT>(v);
    mf2_add<N, false, KF_XCD>(v);
    mf2_add<N, true, KF_XCD | KF_NT>(v);
    mf2_add<N, true, KF_NT>(v);
    mf2_add<N, true, KF_A2MAJ | KF_NT>(v);
    mf2_add<N, true, KF_XCD | KF_A2MAJ | KF_NT>(v);
    mf2_add<N, true, KF_XCD>(v);
    mf2_add<N, true, KF_A2MAJ>(v);
    mf2_add<N, true, KF_XCD | KF_A2MAJ>(v);
    mf2_add<N, true, kMfRowAdj | KF_XCD | KF_NT>(v);
    mf2_add<N, true, kMfRowAdj | KF_XCD | KF_A2MAJ | KF_NT>(v);
    mf2_add<N, true, kMfRowAdj | KF_A2MAJ>(v);
    mf2_add<N, true, kMfRowAdj | KF_A2MAJ | KF_NT>(v);
    mf2_add<N, true, kMfRowAdj | KF_NT>(v);
    mf2_add<N, true, kMfRowAdj>(v);
    mf2_add<N, true, kMfRowAdj | KF_XCD>(v);
}

// Long contiguous lines: one entry per direction (the radix-R2 stage's sign).
struct Mf2LongEntry {
    int N, dir, flags, f16;
    const void* fn;
};

template <int N, int F, bool F16 = false>
void mf2_long_add(std::vector<Mf2LongEntry>& v) {
    v.push_back(Mf2LongEntry{N, -1, F, F16 ? 1 : 0, (const void*)&mf2_long_kernel<N, -1, F, F16>});
    v.push_back(Mf2LongEntry{N, 1, F, F16 ? 1 : 0, (const void*)&mf2_long_kernel<N, 1, F, F16>});
}

const std::vector<Mf2Entry>& mf2_registry() {
    static const std::vector<Mf2Entry> reg = [] {
        std::vector<Mf2Entry> v;
        mf2_add_len<256>(v);
        mf2_add_len<512>(v);
        mf2_add_len_f16<256>(v);
        mf2_add_len_f16<512>(v);
        return v;
    }();
    return reg;
}

const std::vector<Mf2LongEntry>& mf2_long_registry() {
    static const std::vector<Mf2LongEntry> reg = [] {
        std::vector<Mf2LongEntry> v;
        mf2_long_add<1024, KF_XCD | KF_NT>(v);
        mf2_long_add<1024, KF_NT>(v);
        mf2_long_add<2048, KF_XCD | KF_NT>(v);
        mf2_long_add<2048, KF_NT>(v);
        mf2_long_add<1024, KF_XCD | KF_NT, true>(v);
        mf2_long_add<1024, KF_NT, true>(v);
        mf2_long_add<2048, KF_XCD | KF_NT, true>(v);
        mf2_long_add<2048, KF_NT, true>(v);
        return v;
    }();
    return reg;
}

// flags of the built variants of length N (any direction); f16: split-FP16
void mf2_flags(int N, int strided, int f16, std::vector<int>* out) {
    if (N >= 1024) {
        if (strided) return;
        for (const Mf2LongEntry& e : mf2_long_registry())
            if (e.N == N && e.dir == -1 && e.f16 == f16) out->push_back(e.flags);
        return;
    }
    for (const Mf2Entry& e : mf2_registry())
        if (e.N == N && e.strided == (strided != 0) && e.f16 == f16) out->push_back(e.flags);
}

const void* mf2_find(int N, int strided, int flags, int dir, int f16) {
    if (N >= 1024) {
        for (const Mf2LongEntry& e : mf2_long_registry())
            if (e.N == N && !strided && e.flags == flags && e.dir == (dir < 0 ? -1 : 1) && e.f16 == f16) return e.fn;
        return nullptr;
    }
    for (const Mf2Entry& e : mf2_registry())
        if (e.N == N && e.strided == (strided != 0) && e.flags == flags && e.f16 == f16) return e.fn;
    return nullptr;
}

// Lines per workgroup of the interleaved-K kernels.
int mf2_tile_lines(int N) { return N <= 512 ? 16 : 8192 / N; }

struct MfKey {
    int dev, N, dir;
    bool operator<(const MfKey& o) const {
        if (dev != o.dev) return dev < o.dev;
        if (N != o.N) return N < o.N;
        return dir < o.dir;
    }
};
std::mutex g_mf_mu;
std::map<MfKey, void*> g_mf_tw;

}  // namespace


// Twiddles of the MFMA schedule (FP32, rounded once from long double).
const void* mfma_twiddle_table(int N, int dir) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    const MfKey key{dev, N, dir < 0 ? -1 : 1};
    std::lock_guard<std::mutex> lk(g_mf_mu);
    auto it = g_mf_tw.find(key);
    if (it != g_mf_tw.end()) return it->second;
    std::vector<float2> h(size_t(mf_twsize(N)));
    for (int t = 0; t < 16; ++t) {
        long double r, i;
        twiddle_ld(t, 16, key.dir, &r, &i);
        h[size_t(t)] = make_float2(float(r), float(i));
    }
    for (int p = 1; p < mf_nstages(N); ++p) {
        const int R = mf_radix(N, p), Ns = mf_ns(N, p), off = mf_twoff(N, p);
        for (int k = 0; k < Ns; ++k)
            for (int r = 0; r < R; ++r) {
                long double re, im;
                twiddle_ld(int64_t(r) * k, int64_t(Ns) * R, key.dir, &re, &im);
                h[size_t(off + k * R + r)] = make_float2(float(re), float(im));
            }
    }
    void* d = nullptr;
    if (hipMalloc(&d, h.size() * sizeof(float2)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    g_mf_tw[key] = d;
    return d;
}

// The split-K-halves kernels: N in 16..512 and a tile width the axis admits.
static bool legacy_fits(const AxisPass& p, int N) {
    if (!mf_length_ok(N)) return false;
    const int T = mf_T(N, p.strided != 0);
    FftGeom g;
    uint32_t grid = 0;
    return resolve_with(p.fft, p.a1, p.a2, p.strided, N, T, N, &g, &grid, p.a2_start, p.a2_count) == A2AFFT_OK;
}

// One past the largest byte offset a pass touches over its whole axes (all
// strides are positive).
static uint64_t pass_extent(const AxisPass& p, int N) {
    return 8ull * uint64_t(p.a1.off(p.a1.n - 1) + p.a2.off(p.a2.n - 1) + p.fft.off(uint32_t(N - 1))) + 8;
}

// The interleaved-K kernels run split-BF16 and split-FP16 passes of 256 and
// 512 points (16-line tiles) and contiguous lines of 1024 and 2048 points (8
// and 4 lines), whose byte offsets fit 32 bits.
static bool mf2_fits(const AxisPass& p, int N, a2afft_precision prec) {
    if (prec != A2AFFT_SPLIT_BF16 && prec != A2AFFT_SPLIT_FP16) return false;
    if (N != 256 && N != 512 && N != 1024 && N != 2048) return false;
    if (N >= 1024 && p.strided) return false;
    if (pass_extent(p, N) >= (1ull << 32)) return false;
    FftGeom g;
    uint32_t grid = 0;
    return resolve_with(p.fft, p.a1, p.a2, p.strided, N, mf2_tile_lines(N), N, &g, &grid) == A2AFFT_OK;
}

// Geometry check for a split-precision pass.
int mfma_pass_check(const AxisPass& p, int N, a2afft_precision prec) {
    if (mf2_fits(p, N, prec) || legacy_fits(p, N)) return A2AFFT_OK;
    return set_error(A2AFFT_E_UNSUPPORTED,
                     "split precision: no kernel for a pass of length %d (16..512 in tiles of the axis, or "
                     "contiguous 1024 and 2048)", N);
}

// Kernel choices of a split pass: T = tile lines; E = 0 for the interleaved-K
// kernels with variant flags, E = N for the kernels with split K halves.
int mfma_candidates(const AxisPass& p, int N, a2afft_precision prec, KSel* out, int max) {
    int n = 0;
    if (mf2_fits(p, N, prec)) {
        std::vector<int> fl;
        mf2_flags(N, p.strided, prec == A2AFFT_SPLIT_FP16 ? 1 : 0, &fl);
        for (int f : fl)
            for (int occ : {0, kMfOcc3, kMfOcc2})
                if (n < max && (N <= 512 || occ == 0)) out[n++] = KSel{mf2_tile_lines(N), 0, f | occ};
    }
    if (n < max && legacy_fits(p, N)) out[n++] = KSel{mf_T(N, p.strided != 0), N, 0};
    return n;
}

// Defaults of the interleaved-K kernel (split-BF16; split-FP16 takes the
// same), measured on 512^3 complex64 over dims [1,4,4] (ms per pass, whole
// passes unless noted):
//   contiguous (z): XCD order + non-temporal 0.376-0.380 (hardware order 0.408);
//   far-strided (x, rows 128 KiB apart): row-adjacent lanes + z tiles
//     outermost 0.448-0.452 (XCD + non-temporal without row-adjacent lanes
//     0.502-0.528, with them 0.547);
//   near-strided (y): XCD + non-temporal 0.379-0.394 whole; in 128-column
//     windows alternating with x, plain non-temporal order gives the fastest
//     forward (1.104-1.110 ms against 1.126-1.131 with XCD order).
// `windowed`: the pass runs in windows of z columns (the plan's chunking).
KSel mfma_default_sel(const AxisPass& p, int N, a2afft_precision prec, bool windowed) {
    if (mf2_fits(p, N, prec)) {
        if (!p.strided) return KSel{mf2_tile_lines(N), 0, KF_XCD | KF_NT};
        if (p.fft.s_elem >= 4096) return KSel{16, 0, kMfRowAdj | KF_A2MAJ};
        return KSel{16, 0, windowed ? KF_NT : KF_XCD | KF_NT};
    }
    return KSel{mf_T(N, p.strided != 0), N, 0};
}

int launch_mfma_pass(const AxisPass& p, void* data, int N, int dir, a2afft_precision prec, hipStream_t s,
                     const KSel* sel) {
    if (!data) return set_error(A2AFFT_E_INVALID_ARG, "fft: null data");
    const KSel ks = sel ? *sel : mfma_default_sel(p, N, prec, false);
    const void* tw = mfma_twiddle_table(N, dir);
    if (!tw) return set_error(A2AFFT_E_ALLOC, "split precision: twiddle table N=%d", N);
    float2* dp = (float2*)data;
    if (ks.E == 0) {   // interleaved-K kernel
        if (!mf2_fits(p, N, prec)) return set_error(A2AFFT_E_UNSUPPORTED, "split precision: no interleaved-K kernel");
        const void* fn = mf2_find(N, p.strided, ks.flags & ~(kMfOcc3 | kMfOcc2), dir, prec == A2AFFT_SPLIT_FP16 ? 1 : 0);
        if (!fn) return set_error(A2AFFT_E_UNSUPPORTED, "split precision: no variant flags=%d", ks.flags);
        // occupancy variants: extra dynamic LDS leaves room for 3 or 2 workgroups per CU
        const size_t dyn = (ks.flags & kMfOcc2) ? size_t(28 << 10) : (ks.flags & kMfOcc3) ? size_t(16 << 10) : 0;
        FftGeom g;
        uint32_t grid = 0;
        int st = resolve_with(p.fft, p.a1, p.a2, p.strided, N, mf2_tile_lines(N), N, &g, &grid, p.a2_start,
                              p.a2_count);
        if (st < 0) return set_error(A2AFFT_E_UNSUPPORTED, "split precision: geometry (N=%d)", N);
        g.scale_mode = p.scale_mode;
        g.scale = p.scale;
        if (grid == 0) return A2AFFT_OK;
        uint32_t extent = uint32_t(pass_extent(p, N));
        void* args[] = {&dp, &g, (void*)&tw, &extent};
        st = hip_check(int(hipLaunchKernel(fn, dim3(grid), dim3(256), args, dyn, s)), "mfma fft launch");
        if (st < 0) return st;
        return hip_check(int(hipGetLastError()), "mfma fft launch");
    }
    if (!mf_length_ok(N)) return set_error(A2AFFT_E_UNSUPPORTED, "split precision: length %d (16..512)", N);
    const int T = mf_T(N, p.strided != 0);
    FftGeom g;
    uint32_t grid = 0;
    int st = resolve_with(p.fft, p.a1, p.a2, p.strided, N, T, N, &g, &grid, p.a2_start, p.a2_count);
    if (st < 0) return set_error(A2AFFT_E_UNSUPPORTED, "split precision: geometry (N=%d, T=%d)", N, T);
    g.scale_mode = p.scale_mode;
    g.scale = p.scale;
    if (grid == 0) return A2AFFT_OK;
    const int f16 = prec == A2AFFT_SPLIT_FP16 ? 1 : 0;
    const int want_nw = 8;
    const int want_o32 = pass_extent(p, N) < (1ull << 32) ? 1 : 0;
    const void* fn = nullptr;
    int nw = 4;
    for (const MfEntry& e : mf_registry())
        if (e.N == N && e.strided == (p.strided != 0) && e.dir == (dir < 0 ? -1 : 1) && e.fp16 == f16 &&
            (!fn || (e.nw == want_nw && (e.o32 == want_o32 || N < 256)))) {
            fn = e.fn;
            nw = e.nw;
        }
    if (!fn) return set_error(A2AFFT_E_UNSUPPORTED, "split precision: no kernel for N=%d", N);
    void* args[] = {&dp, &g, (void*)&tw};
    st = hip_check(int(hipLaunchKernel(fn, dim3(grid), dim3(64 * nw), args, 0, s)), "mfma fft launch");
    if (st < 0) return st;
    return hip_check(int(hipGetLastError()), "mfma fft launch");
}

}  // namespace a2a
