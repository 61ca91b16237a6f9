// Batched FFTs of contiguous lines longer than one kernel's 4096 points: the
// four-step decomposition N = N1 * N2 (both <= 4096), the lines' counterpart
// of cuFFT's large-length plans behind forward_1d_fft / inverse_1d_fft
// (a2affts.cu:7-76).  With n = N2*n1 + n2 and k = k1 + N1*k2,
//   X[k1 + N1 k2] = sum_n2 W_N2^(n2 k2) W_N^(n2 k1) sum_n1 W_N1^(n1 k1) x[N2 n1 + n2]
// 1. length-N1 FFTs down the columns of each line seen as [N1][N2], in place
//    (the strided Stockham pass, T adjacent columns per workgroup);
// 2. multiply by W_N^(n2 k1) and transpose to [N2][N1] into a workspace;
// 3. length-N2 FFTs down the columns of the workspace lines, which leaves
//    X[k1 + N1 k2] at k2*N1 + k1, natural order;
// 4. copy the workspace back.
// Lines are processed in chunks of at most 256 MiB; the workspace comes from
// the stream-ordered allocator on the caller's stream (allocated and freed in
// stream order, so calls on different streams do not share it).  When N1 =
// N2 the transpose runs in place over tile pairs: three passes, no
// workspace.
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "fft.h"
#include "remap.h"
#include "stockham.h"

namespace a2a {

namespace {

constexpr int64_t kLongChunkBytes = int64_t(256) << 20;

// W_N^m = hi[m / B] * lo[m % B]: two tables of about sqrt(N) entries, each
// entry from long double rounded once to double.
struct LongTw {
    uint32_t B = 0, nhi = 0;
    double2* hi = nullptr;
    double2* lo = nullptr;
};

const LongTw* long_twiddles(int64_t N, int dir) {
    static std::mutex mu;
    static std::map<std::vector<int64_t>, LongTw> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    const std::vector<int64_t> key{dev, N, dir < 0 ? -1 : 1};
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return &it->second;
    LongTw t;
    t.B = 1;
    while (int64_t(t.B) * t.B < N) ++t.B;
    t.nhi = uint32_t((N + t.B - 1) / t.B);
    std::vector<double2> lo(t.B), hi(t.nhi);
    long double re, im;
    for (uint32_t i = 0; i < t.B; ++i) {
        twiddle_ld(int64_t(i), N, dir, &re, &im);
        lo[i].x = double(re);
        lo[i].y = double(im);
    }
    for (uint32_t j = 0; j < t.nhi; ++j) {
        twiddle_ld(int64_t(j) * t.B, N, dir, &re, &im);
        hi[j].x = double(re);
        hi[j].y = double(im);
    }
    if (hipMalloc(&t.lo, lo.size() * sizeof(double2)) != hipSuccess) return nullptr;
    if (hipMalloc(&t.hi, hi.size() * sizeof(double2)) != hipSuccess) {
        (void)hipFree(t.lo);
        return nullptr;
    }
    if (hipMemcpy(t.lo, lo.data(), lo.size() * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t.hi, hi.data(), hi.size() * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(t.lo);
        (void)hipFree(t.hi);
        return nullptr;
    }
    return &(cache[key] = t);
}

// dst[line][n2][k1] = src[line][k1][n2] * W_N^(n2 k1), through a 32 x 33 LDS
// tile; 256 threads as 32 x 8, one tile per workgroup (long_tile).
template <typename R>
__global__ __launch_bounds__(256) void twiddle_transpose_kernel(const HIP_vector_type<R, 2>* __restrict__ src,
                                                                HIP_vector_type<R, 2>* __restrict__ dst,
                                                                uint32_t N1, uint32_t N2, uint32_t tn2s,
                                                                uint32_t tk1s, const double2* __restrict__ hi,
                                                                const double2* __restrict__ lo, uint32_t B) {
    using V = HIP_vector_type<R, 2>;
    __shared__ V tile[32][33];
    uint32_t tn2, tk1, line;
    long_tile(blockIdx.x, tn2s, tk1s, &tn2, &tk1, &line);
    const uint32_t tx = threadIdx.x & 31u, ty = threadIdx.x >> 5;
    const int64_t base = int64_t(line) * N1 * N2;
#pragma unroll
    for (uint32_t r = 0; r < 32; r += 8) {
        const uint32_t k1 = tk1 * 32u + ty + r, n2 = tn2 * 32u + tx;
        if (k1 < N1 && n2 < N2) {
            const V v = src[base + int64_t(k1) * N2 + n2];
            const uint32_t m = n2 * k1;   // < N1 * N2 <= 2^24
            const double2 h = hi[m / B], l = lo[m % B];
            const double wr = h.x * l.x - h.y * l.y, wi = h.x * l.y + h.y * l.x;
            const double vr = double(v.x), vi = double(v.y);
            V o;
            o.x = R(vr * wr - vi * wi);
            o.y = R(vr * wi + vi * wr);
            tile[ty + r][tx] = o;
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < 32; r += 8) {
        const uint32_t n2 = tn2 * 32u + ty + r, k1 = tk1 * 32u + tx;
        if (k1 < N1 && n2 < N2) dst[base + int64_t(n2) * N1 + k1] = tile[tx][ty + r];
    }
}

// The square case in place: line [M][M], workgroup = tile pair (a, c),
// a <= c (long_pair).  Both tiles are read, multiplied by W_N^(k1 n2) (the
// same for (k1, n2) and (n2, k1)) and staged in LDS before either is written,
// and every element belongs to exactly one pair, so no two workgroups touch
// the same element.
template <typename R>
__global__ __launch_bounds__(256) void twiddle_transpose_square_kernel(HIP_vector_type<R, 2>* __restrict__ data,
                                                                       uint32_t M, uint32_t tiles,
                                                                       const double2* __restrict__ hi,
                                                                       const double2* __restrict__ lo, uint32_t B) {
    using V = HIP_vector_type<R, 2>;
    __shared__ V t0[32][33];
    __shared__ V t1[32][33];
    uint32_t a, c, line;
    long_pair(blockIdx.x, tiles, &a, &c, &line);
    const uint32_t tx = threadIdx.x & 31u, ty = threadIdx.x >> 5;
    V* const L = data + int64_t(line) * M * M;
    auto load = [&](uint32_t row0, uint32_t col0, V (&t)[32][33]) {
#pragma unroll
        for (uint32_t r = 0; r < 32; r += 8) {
            const uint32_t k1 = row0 + ty + r, n2 = col0 + tx;
            if (k1 < M && n2 < M) {
                const V v = L[int64_t(k1) * M + n2];
                const uint32_t m = n2 * k1;
                const double2 h = hi[m / B], l = lo[m % B];
                const double wr = h.x * l.x - h.y * l.y, wi = h.x * l.y + h.y * l.x;
                const double vr = double(v.x), vi = double(v.y);
                V o;
                o.x = R(vr * wr - vi * wi);
                o.y = R(vr * wi + vi * wr);
                t[ty + r][tx] = o;
            }
        }
    };
    // t[i][j] holds element (row0 + i, col0 + j); it goes to (col0 + j, row0 + i)
    auto store = [&](uint32_t row0, uint32_t col0, V (&t)[32][33]) {
#pragma unroll
        for (uint32_t r = 0; r < 32; r += 8) {
            const uint32_t n2 = col0 + ty + r, k1 = row0 + tx;
            if (k1 < M && n2 < M) L[int64_t(n2) * M + k1] = t[tx][ty + r];
        }
    };
    load(a * 32u, c * 32u, t0);
    if (a != c) load(c * 32u, a * 32u, t1);
    __syncthreads();
    store(a * 32u, c * 32u, t0);
    if (a != c) store(c * 32u, a * 32u, t1);
}

// Bluestein's chirp-z form for lengths with no two-factor split:
//   X[k] = c[k] sum_n (x[n] c[n]) conj(c[k - n]),  c[m] = exp(dir pi i m^2 / N),
// a cyclic convolution of length M (a power of two >= 2N - 1) done with two
// M-point FFTs.  chirp: c[0..N) (m^2 reduced mod 2N exactly, long double,
// rounded once); bhat: the forward M-point FFT of b, b[m] = conj(c[|m|]) for
// |m| < N (indices mod M), 0 elsewhere, computed in FP64 on the device.
struct BluePlan {
    int M = 0;
    double2* chirp = nullptr;
    double2* bhat = nullptr;
};

template <typename R>
__global__ __launch_bounds__(256) void blue_pre_kernel(const HIP_vector_type<R, 2>* __restrict__ x,
                                                       HIP_vector_type<R, 2>* __restrict__ a, int N, int M,
                                                       uint64_t total, const double2* __restrict__ c) {
    const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
    if (i >= total) return;
    const uint64_t line = i / uint64_t(M);
    const int m = int(i - line * uint64_t(M));
    HIP_vector_type<R, 2> o;
    o.x = R(0);
    o.y = R(0);
    if (m < N) {
        const HIP_vector_type<R, 2> v = x[line * uint64_t(N) + uint64_t(m)];
        const double2 w = c[m];
        const double vr = double(v.x), vi = double(v.y);
        o.x = R(vr * w.x - vi * w.y);
        o.y = R(vr * w.y + vi * w.x);
    }
    a[i] = o;
}

template <typename R>
__global__ __launch_bounds__(256) void blue_mul_kernel(HIP_vector_type<R, 2>* __restrict__ a, int M, uint64_t total,
                                                       const double2* __restrict__ bhat, double inv_m) {
    const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
    if (i >= total) return;
    const double2 b = bhat[i % uint64_t(M)];
    const HIP_vector_type<R, 2> v = a[i];
    const double vr = double(v.x), vi = double(v.y);
    HIP_vector_type<R, 2> o;
    o.x = R((vr * b.x - vi * b.y) * inv_m);
    o.y = R((vr * b.y + vi * b.x) * inv_m);
    a[i] = o;
}

template <typename R>
__global__ __launch_bounds__(256) void blue_post_kernel(const HIP_vector_type<R, 2>* __restrict__ a,
                                                        HIP_vector_type<R, 2>* __restrict__ x, int N, int M,
                                                        uint64_t total, const double2* __restrict__ c) {
    const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
    if (i >= total) return;
    const uint64_t line = i / uint64_t(N);
    const int k = int(i - line * uint64_t(N));
    const HIP_vector_type<R, 2> v = a[line * uint64_t(M) + uint64_t(k)];
    const double2 w = c[k];
    const double vr = double(v.x), vi = double(v.y);
    HIP_vector_type<R, 2> o;
    o.x = R(vr * w.x - vi * w.y);
    o.y = R(vr * w.y + vi * w.x);
    x[i] = o;
}

// The cached plan for (device, N, dir); built once, synchronously on `s`.
const BluePlan* blue_plan(int N, int dir, hipStream_t s) {
    static std::mutex mu;
    static std::map<std::vector<int64_t>, BluePlan> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    const std::vector<int64_t> key{dev, N, dir < 0 ? -1 : 1};
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return &it->second;
    BluePlan bp;
    bp.M = 1;
    while (bp.M < 2 * N - 1) bp.M *= 2;
    std::vector<double2> c(static_cast<size_t>(N)), b(static_cast<size_t>(bp.M));
    long double re, im;
    for (int m = 0; m < N; ++m) {
        const int64_t sq = int64_t(m) * m % (2 * int64_t(N));
        twiddle_ld(sq, 2 * int64_t(N), dir, &re, &im);   // exp(dir 2 pi i sq / 2N)
        c[size_t(m)].x = double(re);
        c[size_t(m)].y = double(im);
    }
    for (int m = 0; m < bp.M; ++m) b[size_t(m)] = make_double2(0.0, 0.0);
    for (int m = 0; m < N; ++m) {
        const double2 cc = make_double2(c[size_t(m)].x, -c[size_t(m)].y);
        b[size_t(m)] = cc;
        if (m > 0) b[size_t(bp.M - m)] = cc;
    }
    if (hipMalloc(&bp.chirp, c.size() * sizeof(double2)) != hipSuccess) return nullptr;
    if (hipMalloc(&bp.bhat, b.size() * sizeof(double2)) != hipSuccess) {
        (void)hipFree(bp.chirp);
        return nullptr;
    }
    bool ok = hipMemcpy(bp.chirp, c.data(), c.size() * sizeof(double2), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(bp.bhat, b.data(), b.size() * sizeof(double2), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && fft_lines(bp.bhat, bp.M, 1, A2AFFT_C128, -1, s) >= 0 && hipStreamSynchronize(s) == hipSuccess;
    if (!ok) {
        (void)hipFree(bp.chirp);
        (void)hipFree(bp.bhat);
        return nullptr;
    }
    return &(cache[key] = bp);
}

}  // namespace

int launch_chirp_z(void* data, int N, int64_t lines, a2afft_dtype dt, int dir, hipStream_t s) {
    const BluePlan* bp = blue_plan(N, dir, s);
    if (!bp) return set_error(A2AFFT_E_ALLOC, "1d fft: Bluestein tables for length %d", N);
    const int M = bp->M;
    const size_t esize = dt == A2AFFT_C128 ? 16 : 8;
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(lines, kLongChunkBytes / (int64_t(M) * esize)));
    void* work = nullptr;
    A2A_TRY(hip_check(int(hipMallocAsync(&work, size_t(chunk) * M * esize, s)), "1d fft: workspace"));
    int st = A2AFFT_OK;
    for (int64_t l0 = 0; l0 < lines && st >= 0; l0 += chunk) {
        const int64_t cl = std::min<int64_t>(chunk, lines - l0);
        char* D = static_cast<char*>(data) + size_t(l0) * N * esize;
        const uint64_t tm = uint64_t(cl) * M, tn = uint64_t(cl) * N;
        const dim3 gm(uint32_t((tm + 255) / 256)), gn(uint32_t((tn + 255) / 256));
        if (dt == A2AFFT_C128)
            hipLaunchKernelGGL(blue_pre_kernel<double>, gm, dim3(256), 0, s, reinterpret_cast<const double2*>(D),
                               static_cast<double2*>(work), N, M, tm, bp->chirp);
        else
            hipLaunchKernelGGL(blue_pre_kernel<float>, gm, dim3(256), 0, s, reinterpret_cast<const float2*>(D),
                               static_cast<float2*>(work), N, M, tm, bp->chirp);
        st = hip_check(int(hipGetLastError()), "1d fft: Bluestein pre");
        if (st >= 0) st = fft_lines(work, M, cl, dt, -1, s);
        if (st < 0) break;
        if (dt == A2AFFT_C128)
            hipLaunchKernelGGL(blue_mul_kernel<double>, gm, dim3(256), 0, s, static_cast<double2*>(work), M, tm,
                               bp->bhat, 1.0 / double(M));
        else
            hipLaunchKernelGGL(blue_mul_kernel<float>, gm, dim3(256), 0, s, static_cast<float2*>(work), M, tm,
                               bp->bhat, 1.0 / double(M));
        st = hip_check(int(hipGetLastError()), "1d fft: Bluestein product");
        if (st >= 0) st = fft_lines(work, M, cl, dt, +1, s);
        if (st < 0) break;
        if (dt == A2AFFT_C128)
            hipLaunchKernelGGL(blue_post_kernel<double>, gn, dim3(256), 0, s, static_cast<const double2*>(work),
                               reinterpret_cast<double2*>(D), N, M, tn, bp->chirp);
        else
            hipLaunchKernelGGL(blue_post_kernel<float>, gn, dim3(256), 0, s, static_cast<const float2*>(work),
                               reinterpret_cast<float2*>(D), N, M, tn, bp->chirp);
        st = hip_check(int(hipGetLastError()), "1d fft: Bluestein post");
    }
    const int fs = hip_check(int(hipFreeAsync(work, s)), "1d fft: workspace free");
    return st < 0 ? st : fs;
}

int launch_long_fft(void* data, int N, int64_t lines, a2afft_dtype dt, int dir, hipStream_t s) {
    int N1 = 0, N2 = 0;
    if (!long_fft_split(N, &N1, &N2)) {
        if (N <= kBluesteinMaxN) return launch_chirp_z(data, N, lines, dt, dir, s);
        return set_error(A2AFFT_E_UNSUPPORTED,
                         "1d fft: length %d is neither a product of two lengths <= %d nor at most %d", N, kDftMaxN,
                         kBluesteinMaxN);
    }
    // A factor with a prime above 5 runs its column pass as a chirp-z pass
    // (8194 = 241 * 34: both factors).  A direct DFT factor (7) would cost
    // O(7) per element, which the four-step absorbs.
    const size_t esize = dt == A2AFFT_C128 ? 16 : 8;
    const LongTw* tw = long_twiddles(N, dir);
    if (!tw) return set_error(A2AFFT_E_ALLOC, "1d fft: twiddles for length %d", N);
    if (N1 == N2) {   // square: three passes in place, no workspace
        const uint32_t M = uint32_t(N1), tiles = (M + 31) / 32, pairs = tiles * (tiles + 1) / 2;
        const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(lines, ((int64_t(1) << 31) - 1) / pairs));
        for (int64_t l0 = 0; l0 < lines; l0 += chunk) {
            const uint32_t cl = uint32_t(std::min<int64_t>(chunk, lines - l0));
            char* D = static_cast<char*>(data) + size_t(l0) * N * esize;
            const AxisPass p{make_axis(M, M, 0, M), make_axis(cl, cl, 0, N), make_axis(M, M, 0, 1), 1, 0, 1.0};
            A2A_TRY(launch_fft_pass(p, D, dt, N1, dir, s, nullptr));
            const uint32_t blocks = uint32_t(pairs) * cl;
            if (dt == A2AFFT_C128)
                hipLaunchKernelGGL(twiddle_transpose_square_kernel<double>, dim3(blocks), dim3(256), 0, s,
                                   reinterpret_cast<double2*>(D), M, tiles, tw->hi, tw->lo, tw->B);
            else
                hipLaunchKernelGGL(twiddle_transpose_square_kernel<float>, dim3(blocks), dim3(256), 0, s,
                                   reinterpret_cast<float2*>(D), M, tiles, tw->hi, tw->lo, tw->B);
            A2A_TRY(hip_check(int(hipGetLastError()), "1d fft: transpose launch"));
            A2A_TRY(launch_fft_pass(p, D, dt, N2, dir, s, nullptr));
        }
        return A2AFFT_OK;
    }
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(lines, kLongChunkBytes / (int64_t(N) * esize)));
    void* work = nullptr;
    A2A_TRY(hip_check(int(hipMallocAsync(&work, size_t(chunk) * N * esize, s)), "1d fft: workspace"));
    const uint32_t tn2s = uint32_t((N2 + 31) / 32), tk1s = uint32_t((N1 + 31) / 32);
    int st = A2AFFT_OK;
    for (int64_t l0 = 0; l0 < lines && st >= 0; l0 += chunk) {
        const uint32_t cl = uint32_t(std::min<int64_t>(chunk, lines - l0));
        char* D = static_cast<char*>(data) + size_t(l0) * N * esize;
        // 1. columns of [N1][N2], in place
        const AxisPass p1{make_axis(uint32_t(N1), uint32_t(N1), 0, N2), make_axis(cl, cl, 0, N),
                          make_axis(uint32_t(N2), uint32_t(N2), 0, 1), 1, 0, 1.0};
        st = launch_fft_pass(p1, D, dt, N1, dir, s, nullptr);
        if (st < 0) break;
        // 2. twiddle and transpose into the workspace
        const uint64_t blocks = uint64_t(tn2s) * tk1s * cl;
        if (blocks >= (uint64_t(1) << 31)) {
            st = set_error(A2AFFT_E_SIZE, "1d fft: too many transpose tiles");
            break;
        }
        if (dt == A2AFFT_C128)
            hipLaunchKernelGGL(twiddle_transpose_kernel<double>, dim3(uint32_t(blocks)), dim3(256), 0, s,
                               reinterpret_cast<const double2*>(D), static_cast<double2*>(work), uint32_t(N1),
                               uint32_t(N2), tn2s, tk1s, tw->hi, tw->lo, tw->B);
        else
            hipLaunchKernelGGL(twiddle_transpose_kernel<float>, dim3(uint32_t(blocks)), dim3(256), 0, s,
                               reinterpret_cast<const float2*>(D), static_cast<float2*>(work), uint32_t(N1),
                               uint32_t(N2), tn2s, tk1s, tw->hi, tw->lo, tw->B);
        st = hip_check(int(hipGetLastError()), "1d fft: transpose launch");
        if (st < 0) break;
        // 3. columns of [N2][N1] in the workspace: natural order
        const AxisPass p3{make_axis(uint32_t(N2), uint32_t(N2), 0, N1), make_axis(cl, cl, 0, N),
                          make_axis(uint32_t(N1), uint32_t(N1), 0, 1), 1, 0, 1.0};
        st = launch_fft_pass(p3, work, dt, N2, dir, s, nullptr);
        if (st < 0) break;
        // 4. back in place
        st = launch_copy(D, work, int64_t(cl) * N, dt, s);
    }
    const int fs = hip_check(int(hipFreeAsync(work, s)), "1d fft: workspace free");
    return st < 0 ? st : fs;
}


// ---------------------------------------------------------------------------
// Chirp-z passes (KF_CZ) over any axis pass geometry, N <= kDftMaxN:
// launch_chirp_z's convolution for plan passes and batched lines, in one
// kernel, O(N log N) where the direct DFT costs O(N) per element.  Every line
// takes the same arithmetic in every layout, so fused and staged plans agree
// bit for bit.
// ---------------------------------------------------------------------------
namespace {

// v times the chirp value w in FP64, rounded once; then the pass's scale in
// the data's precision, as the Stockham passes and the staged path's scale
// kernel apply it.
template <typename R>
__device__ inline HIP_vector_type<R, 2> cz_mul(HIP_vector_type<R, 2> v, double2 w, int scale_mode, double scale) {
    const double vr = double(v.x), vi = double(v.y);
    HIP_vector_type<R, 2> o;
    o.x = R(vr * w.x - vi * w.y);
    o.y = R(vr * w.y + vi * w.x);
    if (scale_mode == 1) {
        o.x *= R(scale);
        o.y *= R(scale);
    } else if (scale_mode == 2) {
        o.x = o.x / R(scale);
        o.y = o.y / R(scale);
    }
    return o;
}

// The whole convolution of T = 8192 / M lines per workgroup (M <= 8192):
// each line times the chirp into registers (zeros past N), the forward
// M-point Stockham FFT, the product with the filter spectrum / M (FP64,
// rounded once, as blue_mul_kernel), the inverse FFT, then times the chirp and
// the pass's scale back into place.  One read and one write of each line, no
// workspace.  In place: a workgroup reads all its lines before it writes any,
// and owns them alone.
template <typename real, int M, bool STRIDED>
__global__ __launch_bounds__(kCzFusedThreads) void cz_fused_kernel(
    typename CT<real>::V* __restrict__ data, FftGeom g, uint32_t units, int N,
    const typename CT<real>::V* __restrict__ twf, const typename CT<real>::V* __restrict__ twi,
    const double2* __restrict__ chirp, const double2* __restrict__ bhat, double inv_m) {
    using V = typename CT<real>::V;
    constexpr int E = kCzFusedE, T = cz_fused_lines(M), TPL = M / E;
    static_assert(T * TPL == kCzFusedThreads, "512 threads a workgroup");
    __shared__ V lds[lds_slots<M, T, STRIDED, 0>()];
    int t, line;
    cz_fused_thread(STRIDED, M, int(threadIdx.x), &t, &line);
    const uint32_t u = blockIdx.x * uint32_t(T) + uint32_t(line);
    const bool valid = u < units;
    const int64_t base = valid ? cz_line_base(g, u) : 0;
    V v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = t + e * TPL;
        if (valid && n < N) {
            v[e] = cz_mul<real>(data[base + g.fft.off(uint32_t(n))], chirp[n], 0, 1.0);
        } else {
            v[e].x = real(0);
            v[e].y = real(0);
        }
    }
    stockham_passes<real, M, E, T, STRIDED, -1, 0, 0>(v, lds, twf, t, line, true);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double2 b = bhat[t + e * TPL];
        const double vr = double(v[e].x), vi = double(v[e].y);
        v[e].x = real((vr * b.x - vi * b.y) * inv_m);
        v[e].y = real((vr * b.y + vi * b.x) * inv_m);
    }
    __syncthreads();   // the forward transform's last LDS reads before the inverse's first writes
    stockham_passes<real, M, E, T, STRIDED, 1, 0, 0>(v, lds, twi, t, line, true);
    if (!valid) return;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = t + e * TPL;
        if (k < N) data[base + g.fft.off(uint32_t(k))] = cz_mul<real>(v[e], chirp[k], g.scale_mode, g.scale);
    }
}

template <typename real, bool S>
const void* cz_fused_for(int M) {
    switch (M) {
    case 32: return (const void*)&cz_fused_kernel<real, 32, S>;
    case 64: return (const void*)&cz_fused_kernel<real, 64, S>;
    case 128: return (const void*)&cz_fused_kernel<real, 128, S>;
    case 256: return (const void*)&cz_fused_kernel<real, 256, S>;
    case 512: return (const void*)&cz_fused_kernel<real, 512, S>;
    case 1024: return (const void*)&cz_fused_kernel<real, 1024, S>;
    case 2048: return (const void*)&cz_fused_kernel<real, 2048, S>;
    case 4096: return (const void*)&cz_fused_kernel<real, 4096, S>;
    case 8192: return (const void*)&cz_fused_kernel<real, 8192, S>;
    default: return nullptr;
    }
}

}  // namespace

bool chirp_pass_ok(int N) { return N >= 9 && N <= kDftMaxN; }

const void* chirp_pass_kernel(int N, bool dbl, int strided) {
    if (!chirp_pass_ok(N)) return nullptr;
    const int M = cz_length(N);
    if (dbl) return strided ? cz_fused_for<double, true>(M) : cz_fused_for<double, false>(M);
    return strided ? cz_fused_for<float, true>(M) : cz_fused_for<float, false>(M);
}

const void* chirp_pass_tables(int N, int dir, a2afft_dtype dt) {
    if (!chirp_pass_ok(N)) return nullptr;
    const BluePlan* f = blue_plan(N, -1, nullptr);
    const BluePlan* b = blue_plan(N, 1, nullptr);
    const int M = cz_length(N);
    if (!f || !b || !twiddle_table(M, -1, dt) || !twiddle_table(M, 1, dt)) return nullptr;
    return dir < 0 ? f->chirp : b->chirp;
}

int launch_chirp_pass(const FftGeom& g, uint32_t units, int strided, void* data, a2afft_dtype dt, int N, int dir,
                      hipStream_t s) {
    const void* fn = chirp_pass_kernel(N, dt == A2AFFT_C128, strided);
    if (!fn) return set_error(A2AFFT_E_UNSUPPORTED, "chirp-z pass: length %d", N);
    if (units == 0) return A2AFFT_OK;
    const BluePlan* bp = blue_plan(N, dir, s);
    const int M = cz_length(N);
    const void* twf = twiddle_table(M, -1, dt);
    const void* twi = twiddle_table(M, 1, dt);
    if (!bp || !twf || !twi) return set_error(A2AFFT_E_ALLOC, "chirp-z pass: tables for length %d", N);
    const uint32_t T = uint32_t(cz_fused_lines(M)), grid = (units + T - 1) / T;
    const double2* ch = bp->chirp;
    const double2* bh = bp->bhat;
    double inv_m = 1.0 / double(M);
    int n = N;
    uint32_t nu = units;
    void* args[] = {&data, (void*)&g, &nu, &n, (void*)&twf, (void*)&twi, (void*)&ch, (void*)&bh, &inv_m};
    A2A_TRY(hip_check(int(hipLaunchKernel(fn, dim3(grid), dim3(kCzFusedThreads), args, 0, s)), "chirp-z pass launch"));
    return hip_check(int(hipGetLastError()), "chirp-z pass launch");
}

}  // namespace a2a
