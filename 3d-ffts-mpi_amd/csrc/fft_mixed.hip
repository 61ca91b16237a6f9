// Stockham FFTs for the 2^a 3^b 5^c lengths the fixed-schedule kernels do not
// cover (above the one-thread-per-line limit of fft_len_ok, up to 4096): the
// radix schedule travels in the kernel arguments and only the elements per
// thread E is a template parameter, so nine E values cover every such length.
//
// Schedule: the power-of-two factor as radix-4 passes (one radix 2 for an odd
// exponent), or as radix-2 passes where those allow a smaller E and the
// radix-4 E would need more than 48 data registers, then the
// 3s, then the 5s.  E is the smallest built value (at most 60 complex64 or 30
// complex128 values) that every radix divides and that divides N; a thread holds
// the elements n = t + e*TPL (TPL = N/E) of one line, exactly as in the
// fixed kernels, and each pass's body is unrolled for its radix behind a
// switch, so the registers are indexed with constants only.
#include "fft.h"
#include "stockham.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

namespace a2a {

namespace {

constexpr int kMixE[] = {9, 10, 12, 15, 18, 20, 25, 30, 60};
constexpr int kMixThreads = 256;
// Strided tiles of long lines: with 256 threads a 768-point line (TPL 64)
// leaves room for 4 lines per tile, i.e. 32-byte row segments.  The wide
// kernels take up to 1024 threads and 128 KiB of LDS so such tiles reach 16
// lines; they hold at most 48 data registers per thread (E <= 24 complex64,
// E <= 12 complex128) to stay within the 128 VGPRs that 16 waves allow.
constexpr int kMixWideThreads = 1024;
constexpr int kMixWideLds = 131072;

int gcd_i(int a, int b) { return b ? gcd_i(b, a % b) : a; }

template <bool STRIDED>
__device__ __forceinline__ int mix_slot(int line, int d, int N, int T) {
    return STRIDED ? d * T + line : line * N + d;
}

// Pass p of radix R on the thread's registers, then (unless it is the last
// pass) the exchange through LDS that puts the next pass's inputs in place.
template <int R, int E, bool STRIDED, int DIR, typename V>
__device__ __forceinline__ void mix_pass(V (&v)[E], V* lds, const V* __restrict__ tw, const MixedPlan& mp, int p,
                                         int t, int line, bool active) {
    if constexpr (E % R == 0) {
        constexpr int NB = E / R;
        const int TPL = mp.TPL, Ns = mp.ns[p], toff = mp.toff[p];
        const FastDiv dns = mp.dns[p];   // j / Ns as a multiply-shift, not a division sequence
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            V x[R];
#pragma unroll
            for (int r = 0; r < R; ++r) x[r] = v[b + r * NB];
            if (p > 0) {
                const int j = t + b * TPL;
                const int k = int(dns.mod(uint32_t(j)));
                if (STRIDED && sizeof(x[0].x) == 4 && mp.twl) {
#pragma unroll
                    for (int r = 1; r < R; ++r) x[r] = c_mul(x[r], lds[mp.twl + toff + r * Ns + k]);
                } else {
#pragma unroll
                    for (int r = 1; r < R; ++r) x[r] = c_mul(x[r], tw[toff + k * R + r]);
                }
            }
            dft<R, DIR>(x);
#pragma unroll
            for (int q = 0; q < R; ++q) v[b + q * NB] = x[q];
        }
        if (p + 1 < mp.npass) {
            if (p > 0) __syncthreads();   // the previous exchange's reads are done
            if (active) {
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const int j = t + b * TPL;
                    const int jq = int(dns.div(uint32_t(j)));
                    const int d0 = jq * (Ns * R) + (j - jq * Ns);
#pragma unroll
                    for (int q = 0; q < R; ++q) lds[mix_slot<STRIDED>(line, d0 + q * Ns, mp.N, mp.T)] = v[b + q * NB];
                }
            }
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = lds[mix_slot<STRIDED>(line, t + e * TPL, mp.N, mp.T)];
        }
    }
}

template <typename real, int E, bool STRIDED, int DIR, int NTB>
__global__ __launch_bounds__(NTB) void mixed_kernel(typename CT<real>::V* __restrict__ data, FftGeom g,
                                                            const typename CT<real>::V* __restrict__ tw, MixedPlan mp) {
    using V = typename CT<real>::V;
    extern __shared__ __align__(16) unsigned char mix_smem[];
    V* const lds = reinterpret_cast<V*>(mix_smem);
    const int TPL = mp.TPL, T = mp.T;
    const int tid = int(threadIdx.x);
    const bool active = tid < T * TPL;
    const int tl = active ? tid : 0;
    const int t = STRIDED ? tl / T : tl % TPL;
    const int line = STRIDED ? tl % T : tl / TPL;
    bool valid;
    const int64_t base = line_base(g, STRIDED, T, TPL, blockIdx.x, tl, &valid, false);
    valid = valid && active;
    if (STRIDED && sizeof(real) == 4 && mp.twl) {
        // twiddles of passes 1.. into LDS as [r][k]; the first exchange's
        // barrier orders these writes before pass 1 reads them
        for (int p = 1; p < mp.npass; ++p) {
            const int R = mp.radix[p], Ns = mp.ns[p];
            for (int d = tid; d < R * Ns; d += int(blockDim.x)) {   // wide tiles launch fewer than NTB threads
                const int r = int(mp.dns[p].div(uint32_t(d))), k = d - r * Ns;
                lds[mp.twl + mp.toff[p] + d] = tw[mp.toff[p] + k * R + r];
            }
        }
    }
    V v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        if (valid) v[e] = data[base + elem_off(g, e * TPL)];
        else { v[e].x = 0; v[e].y = 0; }
    }
    for (int p = 0; p < mp.npass; ++p) {
        // t and line as the compiler must re-read them every pass: left
        // visible, the LDS addresses of every radix's exchange are hoisted
        // out of the loop and held in registers across all four bodies.
        int tp = t, lp = line;
        asm volatile("" : "+v"(tp), "+v"(lp));
        switch (mp.radix[p]) {
        case 2: mix_pass<2, E, STRIDED, DIR>(v, lds, tw, mp, p, tp, lp, active); break;
        case 3: mix_pass<3, E, STRIDED, DIR>(v, lds, tw, mp, p, tp, lp, active); break;
        case 4: mix_pass<4, E, STRIDED, DIR>(v, lds, tw, mp, p, tp, lp, active); break;
        default: mix_pass<5, E, STRIDED, DIR>(v, lds, tw, mp, p, tp, lp, active); break;
        }
    }
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
    for (int e = 0; e < E; ++e) data[base + elem_off(g, e * TPL)] = v[e];
}

template <typename real, int E>
const void* mixed_fn(bool strided, int dir) {
    if (strided)
        return dir < 0 ? (const void*)&mixed_kernel<real, E, true, -1, kMixThreads>
                       : (const void*)&mixed_kernel<real, E, true, 1, kMixThreads>;
    return dir < 0 ? (const void*)&mixed_kernel<real, E, false, -1, kMixThreads>
                   : (const void*)&mixed_kernel<real, E, false, 1, kMixThreads>;
}

template <typename real, int E>
const void* mixed_wide_fn(int dir) {
    return dir < 0 ? (const void*)&mixed_kernel<real, E, true, -1, kMixWideThreads>
                   : (const void*)&mixed_kernel<real, E, true, 1, kMixWideThreads>;
}

bool wide_ok(int E, bool dbl) { return E * (dbl ? 4 : 2) <= 48; }

template <typename real>
const void* mixed_wide_fn_e(int E, int dir) {
    if (sizeof(real) == 8 && E > 12) return nullptr;   // beyond wide_ok's register budget
    switch (E) {
    case 9: return mixed_wide_fn<real, 9>(dir);
    case 10: return mixed_wide_fn<real, 10>(dir);
    case 12: return mixed_wide_fn<real, 12>(dir);
    default: break;
    }
    if constexpr (sizeof(real) == 4) {
        switch (E) {
        case 15: return mixed_wide_fn<real, 15>(dir);
        case 18: return mixed_wide_fn<real, 18>(dir);
        case 20: return mixed_wide_fn<real, 20>(dir);
        default: break;
        }
    }
    return nullptr;
}

template <typename real>
const void* mixed_fn_e(int E, bool strided, int dir) {
    switch (E) {
    case 9: return mixed_fn<real, 9>(strided, dir);
    case 10: return mixed_fn<real, 10>(strided, dir);
    case 12: return mixed_fn<real, 12>(strided, dir);
    case 15: return mixed_fn<real, 15>(strided, dir);
    case 18: return mixed_fn<real, 18>(strided, dir);
    case 20: return mixed_fn<real, 20>(strided, dir);
    case 25: return mixed_fn<real, 25>(strided, dir);
    case 30: return mixed_fn<real, 30>(strided, dir);
    default: return nullptr;
    }
}

}  // namespace

// The schedule and E depend on (N, precision) only, so every pass and every
// layout of a line computes the same bits; the pass decides the tile (T) and
// whether the shape fits at all (additive offsets: the axis block extent and
// TPL divide one another).
// The radix schedule and register count of a length (layout-independent).
static bool mixed_schedule(int N, bool dbl, MixedPlan* out) {
    if (N < 2 || N > kDftMaxN || is_pow2(N)) return false;
    int m = N, a = 0, b3 = 0, c5 = 0;
    while (m % 2 == 0) { m /= 2; ++a; }
    while (m % 3 == 0) { m /= 3; ++b3; }
    while (m % 5 == 0) { m /= 5; ++c5; }
    if (m != 1) return false;
    const int emax = dbl ? 30 : 60;
    MixedPlan q{};
    bool found = false;
    for (int twos = 0; twos < 2; ++twos) {   // radix-4 passes first, then radix-2 passes
        MixedPlan c{};
        c.N = N;
        int np = 0;
        if (!twos) {
            for (int i = 0; i < a / 2; ++i) c.radix[np++] = 4;
            if (a % 2) c.radix[np++] = 2;
        } else {
            if (a < 2) continue;   // the same schedule as above
            for (int i = 0; i < a; ++i) c.radix[np++] = 2;
        }
        for (int i = 0; i < b3; ++i) c.radix[np++] = 3;
        for (int i = 0; i < c5; ++i) c.radix[np++] = 5;
        if (np > kMixMaxPasses) continue;
        c.npass = np;
        int G = 1;
        for (int i = 0; i < np; ++i) G = G / gcd_i(G, c.radix[i]) * c.radix[i];
        for (int e : kMixE)
            if (e <= emax && e % G == 0 && N % e == 0) {
                c.E = e;
                break;
            }
        if (c.E == 0 || N / c.E > kMixThreads) continue;
        // The radix-4 schedule (fewer passes, fewer LDS exchanges) unless its
        // data needs more than 48 registers and the radix-2 one needs fewer
        // (640-point complex64 lines: E=20 in 5 passes, not E=10 in 8).
        if (!found || (c.E < q.E && q.E * (dbl ? 4 : 2) > 48)) q = c;
        found = true;
    }
    if (found) *out = q;
    return found;
}

int mixed_plan_regs(int N, bool dbl) {
    MixedPlan q{};
    return mixed_schedule(N, dbl, &q) ? q.E : 0;
}

bool mixed_plan(const AxisPass& ps, int N, bool dbl, MixedPlan* mp, int force_T) {
    const int esize = dbl ? 16 : 8;
    MixedPlan q{};
    if (!mixed_schedule(N, dbl, &q)) return false;
    q.TPL = N / q.E;
    if (!(ps.fft.l % uint32_t(q.TPL) == 0 || uint32_t(q.TPL) % ps.fft.l == 0)) return false;
    const int lds_lines = 65536 / (N * esize);
    if (lds_lines < 1) return false;
    if (force_T > 0) {   // a tile the caller asks for (autotune, tuning cache): accept it if it fits
        const int tt = force_T;
        const bool narrow = tt * q.TPL <= kMixThreads && tt <= lds_lines;
        if (!ps.strided) {
            if (!narrow) return false;
        } else {
            const bool wide = wide_ok(q.E, dbl) && tt * q.TPL <= kMixWideThreads &&
                              tt <= kMixWideLds / (N * esize);
            if (!narrow && !wide) return false;
            if (tt > 1 && !(ps.a2.s_elem == 1 && ps.a2.l % uint32_t(tt) == 0 && ps.a2.n % uint32_t(tt) == 0))
                return false;
        }
        q.T = tt;
    } else if (!ps.strided) {
        q.T = std::min(kMixThreads / q.TPL, lds_lines);
    } else {
        // As many columns as fill the 256 threads: lines of few threads (TPL
        // 4 for 120 points) would otherwise leave most of a 16-column
        // workgroup idle.
        q.T = 1;
        for (int tt : {64, 32, 16, 8, 4, 2})
            if (tt * q.TPL <= kMixThreads && tt <= lds_lines && ps.a2.s_elem == 1 && ps.a2.l % uint32_t(tt) == 0 &&
                ps.a2.n % uint32_t(tt) == 0) {
                q.T = tt;
                break;
            }
        // Wide tiles only where the row segments are shorter than 128 bytes:
        // beyond that they cost occupancy (one 96 KiB workgroup per CU) and
        // gain nothing (384-point complex128 lines: T=16 was slower than 8).
        const int seg_lines = 128 / esize;
        if (q.T < seg_lines && wide_ok(q.E, dbl)) {
            const int wide_lines = kMixWideLds / (N * esize);
            for (int tt : {16, 8, 4, 2}) {
                if (tt > seg_lines) continue;
                if (tt <= q.T) break;
                if (tt * q.TPL <= kMixWideThreads && tt <= wide_lines && ps.a2.s_elem == 1 &&
                    ps.a2.l % uint32_t(tt) == 0 && ps.a2.n % uint32_t(tt) == 0) {
                    q.T = tt;
                    break;
                }
            }
        }
    }
    q.NT = mixed_is_wide(N, q.T, q.E) ? (q.T * q.TPL + 63) / 64 * 64 : kMixThreads;
    int ns = 1, off = 0;
    for (int i = 0; i < q.npass; ++i) {
        q.ns[i] = ns;
        q.dns[i].set(uint32_t(ns));
        q.toff[i] = i > 0 ? off : 0;
        if (i > 0) off += ns * q.radix[i];
        ns *= q.radix[i];
    }
    q.twsize = off;
    // Strided complex64 tiles stage the twiddles in LDS beside the exchange
    // tile where there is room (768x768x256 x pass 0.783 -> 0.753 ms).
    q.twl = (!dbl && ps.strided && q.npass > 1 && lds_room_for_twiddles(q.T * N * esize, off * esize) &&
             (q.T * N + off) * esize <= kMixWideLds)
                ? q.T * N
                : 0;
    *mp = q;
    return true;
}

bool mixed_is_wide(int N, int T, int E) { return E > 0 && T * (N / E) > kMixThreads; }

const void* mixed_kernel_for(int N, int T, int E, bool dbl, int strided, int dir) {
    if (mixed_is_wide(N, T, E)) {
        if (!strided || !wide_ok(E, dbl) || T * (N / E) > kMixWideThreads) return nullptr;
        return dbl ? mixed_wide_fn_e<double>(E, dir) : mixed_wide_fn_e<float>(E, dir);
    }
    if (dbl) return E == 60 ? nullptr : mixed_fn_e<double>(E, strided != 0, dir);
    if (E == 60) return mixed_fn<float, 60>(strided != 0, dir);
    return mixed_fn_e<float>(E, strided != 0, dir);
}

int mixed_lds_bytes(const MixedPlan& mp, bool dbl) {
    return (mp.twl ? mp.twl + mp.twsize : mp.T * mp.N) * (dbl ? 16 : 8);
}

// exp(dir 2 pi i r k / (Ns R)) at toff[p] + k R + r for passes p >= 1,
// long double rounded once; cached per (device, N, dir, dtype).
const void* mixed_twiddles(const MixedPlan& mp, int dir, a2afft_dtype dt) {
    static std::mutex mu;
    static std::map<std::vector<int>, void*> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::vector<int> key{dev, mp.N, dir < 0 ? -1 : 1, int(dt), mp.npass};
    for (int i = 0; i < mp.npass; ++i) key.push_back(mp.radix[i]);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    const size_t n = size_t(mp.twsize > 0 ? mp.twsize : 1);
    std::vector<long double> re(n, 1.0L), im(n, 0.0L);
    for (int p = 1; p < mp.npass; ++p) {
        const int R = mp.radix[p], Ns = mp.ns[p];
        for (int k = 0; k < Ns; ++k)
            for (int r = 0; r < R; ++r)
                twiddle_ld(int64_t(r) * k, int64_t(Ns) * R, dir, &re[size_t(mp.toff[p] + k * R + r)],
                           &im[size_t(mp.toff[p] + k * R + r)]);
    }
    void* d = nullptr;
    if (dt == A2AFFT_C128) {
        std::vector<double2> h(n);
        for (size_t i = 0; i < n; ++i) { h[i].x = double(re[i]); h[i].y = double(im[i]); }
        if (hipMalloc(&d, n * sizeof(double2)) != hipSuccess) return nullptr;
        if (hipMemcpy(d, h.data(), n * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    } else {
        std::vector<float2> h(n);
        for (size_t i = 0; i < n; ++i) { h[i].x = float(re[i]); h[i].y = float(im[i]); }
        if (hipMalloc(&d, n * sizeof(float2)) != hipSuccess) return nullptr;
        if (hipMemcpy(d, h.data(), n * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    }
    cache[key] = d;
    return d;
}

}  // namespace a2a
