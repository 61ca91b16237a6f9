// Stockham FFTs for 2^a 3^b 5^c lengths with the butterflies mapped to threads
// pass by pass (KF_MIXL).
//
// A workgroup of 256 threads owns a tile of T lines (contiguous lines, or T
// adjacent columns of a strided axis).  Pass p of radix R has T*N/R
// butterflies; butterfly bb = tid + k*256 (k < ceil(T*N/(R*256))) is line
// bb / (N/R), index j = bb % (N/R) of a contiguous tile, or column bb % T,
// index bb / T of a strided one, so consecutive lanes take consecutive
// butterflies of a line or the same butterfly of neighbouring columns.  A
// pass reads its butterflies' inputs x[j + r*N/R] (pass 0 from global
// memory), applies the Stockham twiddles W_{Ns R}^{r (j mod Ns)}, runs the
// R-point DFT and writes the outputs to (j / Ns)*Ns*R + j mod Ns + q*Ns (the
// last pass straight to global memory, j + q*N/R).  The tile lives in LDS
// between passes, in place: a pass reads all its inputs, waits for the
// workgroup, then writes.  Because no register slot is tied to an element,
// any radix sequence runs: 480 = 16*10*3 in 3 passes, where a kernel that
// keeps element n = t + e*TPL in register e needs every radix to divide E and
// takes 480 as 2*2*2*2*2*3*5.  The arithmetic of a butterfly does not depend
// on the tile or the thread that runs it, so every tile and every layout of a
// line computes the same bits.
//
// LDS layout: a contiguous line is a row of N values plus, for the layout a
// radix-4/8/16 pass writes, one spare slot per R values (its writes step by
// R); a strided tile is N rows of T+1 values.
#include "fft.h"
#include "stockham.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

namespace a2a {

namespace {

constexpr int kMixlThreads = 256;
constexpr int kMixlMaxLds = 131072;   // bytes of LDS a workgroup may take (gfx950: up to 160 KiB)

template <bool STRIDED>
__device__ __forceinline__ int mixl_slot(int line, int d, const MixlPlan& lp, int sh) {
    if constexpr (STRIDED) return d * (lp.T + 1) + line;
    const int row = lp.N + (sh ? lp.N >> sh : 0);
    return line * row + d + (sh ? d >> sh : 0);
}

template <int R, int EM, bool STRIDED, int DIR, bool NTMP, typename V>
__device__ __forceinline__ void mixl_pass(V (&v)[EM], V* __restrict__ lds, V* __restrict__ data, const FftGeom& g,
                                          const V* __restrict__ tw, const MixlPlan& lp, int p, int tid,
                                          uint32_t block) {
    using real = decltype(v[0].x);
    constexpr int KM = EM / R;
    static_assert(KM >= 1, "register budget below the radix");
    const int N = lp.N, T = lp.T, B = N / R, TB = T * B;
    const bool first = p == 0, last = p + 1 == lp.npass;
    const int Ns = lp.ns[p];
    const int shin = first ? 0 : lp.pad[p - 1], shout = lp.pad[p];
    const int rs = lp.rds[p], ws = lp.wrs[p];
    int jj[KM], ll[KM];
    bool ok[KM];
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        const int bb = tid + k * kMixlThreads;
        ok[k] = bb < TB;
        const int b = ok[k] ? bb : 0;
        int line, j;
        if constexpr (STRIDED) {
            line = int(lp.dT.mod(uint32_t(b)));
            j = int(lp.dT.div(uint32_t(b)));
        } else {
            line = int(lp.dB[p].div(uint32_t(b)));
            j = b - line * B;
        }
        jj[k] = j;
        ll[k] = line;
        if (first) {
            bool valid;
            const int64_t base = line_base(g, STRIDED, T, 1, block, line, &valid, false);
            const bool rd = ok[k] && valid;
            // Strided tiles compute every address: with the constant steps
            // they took 8-26 more VGPRs, lost a wave per SIMD and ran their
            // passes 5-30 % slower.
            if (!STRIDED && lp.gin) {
                const V* const src = data + base + g.fft.off(uint32_t(j));
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (rd) v[k * R + r] = ld_global<NTMP>(src + r * lp.gin);
                    else { v[k * R + r].x = 0; v[k * R + r].y = 0; }
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (rd) v[k * R + r] = ld_global<NTMP>(data + base + g.fft.off(uint32_t(j + r * B)));
                    else { v[k * R + r].x = 0; v[k * R + r].y = 0; }
                }
            }
        } else {
            if constexpr (STRIDED) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[k * R + r] = lds[mixl_slot<STRIDED>(line, j + r * B, lp, shin)];
            } else {
                const V* const src = lds + mixl_slot<STRIDED>(line, j, lp, shin);
#pragma unroll
                for (int r = 0; r < R; ++r) v[k * R + r] = src[r * rs];
            }
        }
    }
    if (!first) __syncthreads();   // every read of the previous layout precedes this pass's writes
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        if (!first) {
            const int kk = int(lp.dns[p].mod(uint32_t(jj[k])));
            if (STRIDED && sizeof(real) == 4 && lp.twl) {
                const V* const w = lds + lp.twl + lp.toff[p] + kk;
#pragma unroll
                for (int r = 1; r < R; ++r) v[k * R + r] = c_mul(v[k * R + r], w[r * Ns]);
            } else {
                const V* const w = tw + lp.toff[p] + kk * R;
#pragma unroll
                for (int r = 1; r < R; ++r) v[k * R + r] = c_mul(v[k * R + r], w[r]);
            }
        }
        dft<R, DIR>(&v[k * R]);
    }
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        if (!ok[k]) continue;
        const int jq = int(lp.dns[p].div(uint32_t(jj[k])));
        const int d0 = jq * Ns * R + (jj[k] - jq * Ns);
        if (last) {
            bool valid;
            const int64_t base = line_base(g, STRIDED, T, 1, block, ll[k], &valid, false);
            if (!valid) continue;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                V o = v[k * R + q];
                if (g.scale_mode == 1) {
                    o.x *= real(g.scale);
                    o.y *= real(g.scale);
                } else if (g.scale_mode == 2) {
                    o.x = o.x / real(g.scale);
                    o.y = o.y / real(g.scale);
                }
                if (!STRIDED && lp.gout) st_global<NTMP>(data + base + g.fft.off(uint32_t(d0)) + q * lp.gout, o);
                else st_global<NTMP>(data + base + g.fft.off(uint32_t(d0 + q * Ns)), o);
            }
        } else {
            if constexpr (STRIDED) {
#pragma unroll
                for (int q = 0; q < R; ++q) lds[mixl_slot<STRIDED>(ll[k], d0 + q * Ns, lp, shout)] = v[k * R + q];
            } else {
                V* const dst = lds + mixl_slot<STRIDED>(ll[k], d0, lp, shout);
#pragma unroll
                for (int q = 0; q < R; ++q) dst[q * ws] = v[k * R + q];
            }
        }
    }
    if (!last) __syncthreads();   // the layout is complete before the next pass reads it
}

template <typename real, int EM, bool STRIDED, int DIR, bool XCD, bool NTMP>
__global__ __launch_bounds__(kMixlThreads) void mixl_kernel(typename CT<real>::V* __restrict__ data, FftGeom g,
                                                            const typename CT<real>::V* __restrict__ tw,
                                                            MixlPlan lp) {
    using V = typename CT<real>::V;
    extern __shared__ __align__(16) unsigned char mixl_smem[];
    V* const lds = reinterpret_cast<V*>(mixl_smem);
    const uint32_t block = XCD ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    if (STRIDED && sizeof(real) == 4 && lp.twl) {
        // the twiddles of passes 1.. into LDS, transposed to [r][k] so that
        // neighbouring butterflies read neighbouring slots; the first pass
        // needs none, and its closing barrier orders these writes before the
        // second pass reads them
        for (int p = 1; p < lp.npass; ++p) {
            const int R = lp.radix[p], Ns = lp.ns[p];
            for (int d = int(threadIdx.x); d < R * Ns; d += kMixlThreads) {
                const int r = int(lp.dns[p].div(uint32_t(d))), kk = d - r * Ns;
                lds[lp.twl + lp.toff[p] + d] = tw[lp.toff[p] + kk * R + r];
            }
        }
    }
    V v[EM];
    for (int p = 0; p < lp.npass; ++p) {
        // the thread index as the compiler must re-read it every pass: left
        // visible, every radix's addresses are hoisted out of the loop
        int t = int(threadIdx.x);
        asm volatile("" : "+v"(t));
        switch (lp.radix[p]) {
        case 2: mixl_pass<2, EM, STRIDED, DIR, NTMP>(v, lds, data, g, tw, lp, p, t, block); break;
        case 3: mixl_pass<3, EM, STRIDED, DIR, NTMP>(v, lds, data, g, tw, lp, p, t, block); break;
        case 4: mixl_pass<4, EM, STRIDED, DIR, NTMP>(v, lds, data, g, tw, lp, p, t, block); break;
        case 5: mixl_pass<5, EM, STRIDED, DIR, NTMP>(v, lds, data, g, tw, lp, p, t, block); break;
        case 6: mixl_pass<6, EM, STRIDED, DIR, NTMP>(v, lds, data, g, tw, lp, p, t, block); break;
        case 10: mixl_pass<10, EM, STRIDED, DIR, NTMP>(v, lds, data, g, tw, lp, p, t, block); break;
        case 15:   // never planned on the complex64 strided 16-register kernel
            if constexpr (!(STRIDED && EM == 16 && sizeof(real) == 4))
                mixl_pass<15, EM, STRIDED, DIR, NTMP>(v, lds, data, g, tw, lp, p, t, block);
            break;
        case 8: mixl_pass<8, EM, STRIDED, DIR, NTMP>(v, lds, data, g, tw, lp, p, t, block); break;
        default: mixl_pass<16, EM, STRIDED, DIR, NTMP>(v, lds, data, g, tw, lp, p, t, block); break;
        }
    }
}

// Built variants: register budget EM (complex values a thread) and flags
// (KF_XCD, KF_NT) per layout and direction.
struct MixlEntry {
    int dbl, EM, strided, flags, dir;
    const void* fn;
};

template <typename real, int EM, bool S, int F>
void mixl_add(std::vector<MixlEntry>& v) {
    constexpr bool X = (F & KF_XCD) != 0, NTM = (F & KF_NT) != 0;
    constexpr int dbl = sizeof(real) == 8;
    v.push_back(MixlEntry{dbl, EM, S, F, -1, (const void*)&mixl_kernel<real, EM, S, -1, X, NTM>});
    v.push_back(MixlEntry{dbl, EM, S, F, 1, (const void*)&mixl_kernel<real, EM, S, 1, X, NTM>});
}

template <typename real, int EM>
void mixl_add_em(std::vector<MixlEntry>& v) {
    mixl_add<real, EM, false, KF_XCD>(v);
    mixl_add<real, EM, false, KF_XCD | KF_NT>(v);
    mixl_add<real, EM, true, KF_XCD>(v);
    mixl_add<real, EM, true, KF_XCD | KF_NT>(v);
    mixl_add<real, EM, true, 0>(v);
}

const std::vector<MixlEntry>& mixl_registry() {
    static const std::vector<MixlEntry> reg = [] {
        std::vector<MixlEntry> v;
        mixl_add_em<float, 16>(v);
        mixl_add_em<float, 24>(v);
        mixl_add_em<float, 32>(v);
        mixl_add_em<float, 48>(v);
        mixl_add_em<double, 16>(v);
        mixl_add_em<double, 24>(v);
        return v;
    }();
    return reg;
}

constexpr int kMixlEmF[] = {16, 24, 32, 48};
constexpr int kMixlEmD[] = {16, 24};

int ilog2i(int v) {
    int s = 0;
    while ((1 << s) < v) ++s;
    return s;
}

// The smallest built register budget that holds every pass's butterflies.
int em_for(int N, int T, const int* radix, int npass, bool dbl, bool strided) {
    int need = 0;
    for (int i = 0; i < npass; ++i) {
        const int R = radix[i];
        const int K = (T * (N / R) + kMixlThreads - 1) / kMixlThreads;
        need = std::max(need, K * R);
        // the complex64 strided 16-register kernel is built without radix 15,
        // which raised it from 72 to 81 VGPRs (100x100x4096 x and y passes 7-10 %
        // slower): radix-15 schedules take the 24-register kernel there
        if (R == 15 && strided && !dbl) need = std::max(need, 17);
    }
    if (!dbl) {
        for (int e : kMixlEmF)
            if (e >= need) return e;
    } else {
        for (int e : kMixlEmD)
            if (e >= need) return e;
    }
    return 0;
}

}  // namespace

// Schedule: the power-of-two factor as 16s, 8s and 4s (8 x 4 for 32, 8 x 8 x
// 8 for 512), then 6s, 10s and 15s where they save passes, then the 3s, then
// the 5s.  It depends on N only, so every pass and layout of a length computes
// the same bits.  The default tile, from a
// sweep of every tile at 49 lengths from 72 to 4000 on 128-256 MiB grids:
// contiguous, the most lines that hold at most 4096 complex64 or 2560
// complex128 values (occupancy: the complex128 kernels take 109 VGPRs) and
// need at most 16 data registers a thread (one radix-16 butterfly, two of
// radix 8, ...); strided, the most columns of 16, 8, 4, 2 within 4096
// complex64 values at 24 registers or 3072 complex128 values at 16, or, for
// complex64 rows at least 256 KiB apart, of 8, 4, 2 within 8192 values at 32
// registers (the x passes of 640x640x256, 960x960x128 and 1600x1600x32:
// 0.543 -> 0.456, 0.648 -> 0.529 and 0.689 -> 0.517 ms on 64- and 32-byte
// rows; nearer rows lost, 720^3 y pass 1.86 -> 2.57 ms), and two columns
// rather than one wherever they fit LDS at 32 (complex64) or 24 (complex128)
// registers: 16- and 32-byte row segments read up to 1.6x faster than 8- and
// 16-byte ones.  One line otherwise.  kMixlStridedTile overrides the strided
// rule for lengths where a sweep of every tile on 0.6-13 GB grids (N^3 and
// N x N x 16..64, where the pass is bound by HBM rather than served from the
// 256 MB Infinity Cache) found another tile faster on both near and far rows
// (radix-15 lengths measured on their radix-15 schedules).
struct MixlTile { int N; bool dbl; int T; };
static constexpr MixlTile kMixlStridedTile[] = {
    {600, false, 8},    // 600^3 y pass 1.06 -> 0.80 ms
    {640, false, 8},    // 640^3 y pass 1.25 -> 1.09 ms
    {720, false, 4},    // 720^3 x pass 2.49 -> 2.32 ms, 720x720x64 x 0.24 -> 0.15 ms
    {960, false, 8},    // 960^3 y pass 4.40 -> 3.91 ms
    {1440, false, 4},   // 1440x1440x32 y pass 0.436 -> 0.421 ms
    {1536, false, 4},   // 1536x1536x32 y pass 0.47 -> 0.44 ms
    {1600, false, 4},   // 1600x1600x32 y pass 0.54 -> 0.49 ms
    {1800, false, 4},   // 1800x1800x32 y pass 0.87 -> 0.55 ms
    {1920, false, 4},   // 1920x1920x32 y pass 0.97 -> 0.67 ms
    {1080, false, 2},   // 1080x1080x64 x pass 0.58 -> 0.53 ms
};

bool mixl_plan(const AxisPass& ps, int N, bool dbl, MixlPlan* out, int force_T) {
    if (N < 2 || N > kDftMaxN || is_pow2(N)) return false;
    int m = N, a = 0, b3 = 0, c5 = 0;
    while (m % 2 == 0) { m /= 2; ++a; }
    while (m % 3 == 0) { m /= 3; ++b3; }
    while (m % 5 == 0) { m /= 5; ++c5; }
    if (m != 1) return false;
    // one strided column (N rows of two values) must fit a workgroup's LDS
    // (128 KiB here), so that the length takes this kernel in every layout or
    // in none
    const int esize = dbl ? 16 : 8;
    if (2 * N * esize > kMixlMaxLds) return false;
    MixlPlan q{};
    q.N = N;
    int np = 0;
    static const int kPow2[13][3] = {{0, 0, 0}, {2, 0, 0}, {4, 0, 0}, {8, 0, 0}, {16, 0, 0}, {8, 4, 0}, {8, 8, 0},
                                     {16, 8, 0}, {16, 16, 0}, {8, 8, 8}, {16, 8, 8}, {16, 16, 8}, {16, 16, 16}};
    if (a > 12) return false;
    // 6 = 2*3, 10 = 2*5 and 15 = 3*5 as single prime-factor passes where
    // pairing takes the length to fewer passes (480 = 16*10*3, 360 = 6*6*10,
    // 960 = 16*6*10, 1000 = 10*10*10, 720 = 16*3*15, 900 = 4*15*15): the
    // fewest passes, then the fewest 15s, then the fewest composite radices.
    auto npow2 = [&](int e) { return (kPow2[e][0] != 0) + (kPow2[e][1] != 0) + (kPow2[e][2] != 0); };
    int n6 = 0, n10 = 0, n15 = 0, best = npow2(a) + b3 + c5;
    for (int s6 = 0; s6 <= std::min(a, b3); ++s6)
        for (int s10 = 0; s10 <= std::min(a - s6, c5); ++s10)
            for (int s15 = 0; s15 <= std::min(b3 - s6, c5 - s10); ++s15) {
                const int passes = npow2(a - s6 - s10) + b3 + c5 - s15;
                if (passes < best || (passes == best && (s15 < n15 || (s15 == n15 && s6 + s10 < n6 + n10)))) {
                    best = passes;
                    n6 = s6;
                    n10 = s10;
                    n15 = s15;
                }
            }
    const int a2 = a - n6 - n10;
    for (int i = 0; i < 3; ++i)
        if (kPow2[a2][i]) q.radix[np++] = kPow2[a2][i];
    for (int i = 0; i < n6; ++i) q.radix[np++] = 6;
    for (int i = 0; i < n10; ++i) q.radix[np++] = 10;
    for (int i = 0; i < n15; ++i) q.radix[np++] = 15;
    for (int i = 0; i < b3 - n6 - n15; ++i) q.radix[np++] = 3;
    for (int i = 0; i < c5 - n10 - n15; ++i) q.radix[np++] = 5;
    if (np > kMixlMaxPasses || np < 1) return false;
    q.npass = np;
    q.NT = kMixlThreads;
    // LDS row of a contiguous line: N plus the largest padding of a pass
    int maxpad = 0;
    for (int i = 0; i < np; ++i) {
        const int R = q.radix[i];
        q.pad[i] = (!ps.strided && (R == 4 || R == 8 || R == 16)) ? ilog2i(R) : 0;
        if (q.pad[i]) maxpad = std::max(maxpad, N >> q.pad[i]);
    }
    const int row = N + maxpad;
    auto fits = [&](int T, int emlim) {
        if (T < 1) return false;
        const int lds = ps.strided ? N * (T + 1) * esize : T * row * esize;
        if (lds > kMixlMaxLds) return false;
        const int em = em_for(N, T, q.radix, np, dbl, ps.strided);
        if (em == 0 || em > emlim) return false;
        if (ps.strided && T > 1 && !(ps.a2.s_elem == 1 && ps.a2.l % uint32_t(T) == 0 && ps.a2.n % uint32_t(T) == 0))
            return false;
        return true;
    };
    int T = 0;
    if (force_T > 0) {
        if (!fits(force_T, 1 << 30)) return false;
        T = force_T;
    } else if (!ps.strided) {
        const int cap = dbl ? 2560 : 4096;
        for (int t = 64; t >= 1 && T == 0; --t)
            if (t * N <= cap && fits(t, 16)) T = t;
    } else {
        const int cap = dbl ? 3072 : 4096, emlim = dbl ? 16 : 24;
        const bool far = ps.fft.s_elem * esize >= (256 << 10);   // rows at least 256 KiB apart
        for (const MixlTile& mt : kMixlStridedTile)
            if (mt.N == N && mt.dbl == dbl && fits(mt.T, dbl ? 24 : 32)) T = mt.T;
        // complex128: four columns up to 4000 values at 24 registers (750-1000
        // points; 800^3 x pass 6.72 -> 5.05 ms, 1000^3 13.0 -> 9.0 ms, 960^3
        // 12.2 -> 8.3 ms; 1080 points lose, 0.43 -> 0.60 ms)
        for (int t = 16; t >= 1 && T == 0; t /= 2)
            if ((t * N <= cap && fits(t, emlim)) || (!dbl && far && t <= 8 && t * N <= 8192 && fits(t, 32)) ||
                (dbl && t == 4 && t * N <= 4000 && fits(t, 24)))
                T = t;
        if (T == 1 && fits(2, dbl ? 24 : 32)) T = 2;
    }
    if (T == 0 && fits(1, 1 << 30)) T = 1;
    if (T == 0) return false;
    q.T = T;
    q.EM = em_for(N, T, q.radix, np, dbl, ps.strided);
    q.dT.set(uint32_t(T));
    int ns = 1, off = 0;
    for (int i = 0; i < np; ++i) {
        q.ns[i] = ns;
        q.dns[i].set(uint32_t(ns));
        q.dB[i].set(uint32_t(N / q.radix[i]));
        q.toff[i] = i > 0 ? off : 0;
        if (i > 0) off += ns * q.radix[i];
        ns *= q.radix[i];
    }
    q.twsize = off;
    // Strided complex64 tiles stage the twiddles in LDS when they take at
    // most 16 KiB and there is room beside the tile (long lines read the
    // global table:
    // its size grows like N): 480^3 x and y passes 0.528 -> 0.468 ms, half
    // the vector-memory reads.  Contiguous tiles (the LDS path cost them a
    // wave per SIMD: z pass 0.487 -> 0.501 ms) and complex128 (121-125 ->
    // 173-185 VGPRs, 240^3 0.257 -> 0.351 ms) read the global table.
    if (!dbl && ps.strided) {
        const int tile = mixl_lds_bytes(q, dbl, ps.strided) / esize;
        q.twl = (np > 1 && lds_room_for_twiddles(tile * esize, off * esize) && (tile + off) * esize <= kMixlMaxLds)
                    ? tile
                    : 0;
    }
    // LDS steps.  Padded rows put slot d at d + (d >> sh).  A pass reads
    // d = j + r B, with B = N / R a multiple of the previous pass's 2^sh (the
    // power-of-two radices come first, largest first, so B keeps every factor
    // of 2 the previous radix took), and writes d = d0 + q Ns, with Ns = 1 or
    // a multiple of its own 2^sh: both steps are constant.
    // (Checked for the contiguous padding in every layout, so that a length
    // takes this kernel in all layouts or none.)
    const auto cpad = [&](int i) { return (q.radix[i] == 4 || q.radix[i] == 8 || q.radix[i] == 16) ? ilog2i(q.radix[i]) : 0; };
    const auto lin = [](int step, int sh) { return sh == 0 || step % (1 << sh) == 0; };
    for (int i = 0; i < np; ++i) {
        const int B = N / q.radix[i];
        if (!lin(B, i > 0 ? cpad(i - 1) : 0) || !(q.ns[i] == 1 || lin(q.ns[i], cpad(i)))) return false;
        const int shin = i > 0 ? q.pad[i - 1] : 0, shout = q.pad[i];
        if (ps.strided) {
            q.rds[i] = B * (T + 1);
            q.wrs[i] = q.ns[i] * (T + 1);
        } else {
            q.rds[i] = B + (shin ? B >> shin : 0);
            q.wrs[i] = q.ns[i] + (shout ? q.ns[i] >> shout : 0);
        }
    }
    // Global steps: off(c + m) = off(c) + off(m) when the block extent l is
    // at least N (one segment) or divides m.
    const uint32_t l = ps.fft.l;
    const int B0 = N / q.radix[0], NsL = q.ns[np - 1];
    q.gin = (l >= uint32_t(N) || uint32_t(B0) % l == 0) ? ps.fft.off(uint32_t(B0)) : 0;
    q.gout = (l >= uint32_t(N) || uint32_t(NsL) % l == 0) ? ps.fft.off(uint32_t(NsL)) : 0;
    *out = q;
    return true;
}

const void* mixl_kernel_for(int EM, bool dbl, int strided, int flags, int dir) {
    for (const MixlEntry& e : mixl_registry())
        if (e.dbl == int(dbl) && e.EM == EM && e.strided == (strided != 0) && e.flags == (flags & (KF_XCD | KF_NT)) &&
            e.dir == (dir < 0 ? -1 : 1))
            return e.fn;
    return nullptr;
}

int mixl_lds_bytes(const MixlPlan& lp, bool dbl, int strided) {
    const int esize = dbl ? 16 : 8;
    if (lp.twl) return (lp.twl + lp.twsize) * esize;   // the tile, then the twiddles
    if (strided) return lp.N * (lp.T + 1) * esize;
    int row = lp.N;
    for (int i = 0; i < lp.npass; ++i)
        if (lp.pad[i]) row = std::max(row, lp.N + (lp.N >> lp.pad[i]));
    return lp.T * row * esize;
}

// The same table layout as the runtime-schedule kernel's (toff[p] + k R + r).
const void* mixl_twiddles(const MixlPlan& lp, int dir, a2afft_dtype dt) {
    MixedPlan mp{};
    mp.N = lp.N;
    mp.npass = lp.npass;
    mp.twsize = lp.twsize;
    for (int i = 0; i < lp.npass; ++i) {
        mp.radix[i] = lp.radix[i];
        mp.ns[i] = lp.ns[i];
        mp.toff[i] = lp.toff[i];
    }
    return mixed_twiddles(mp, dir, dt);
}

}  // namespace a2a
