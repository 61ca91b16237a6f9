// Stockham FFT dispatch: kernel instantiations, launch shapes, twiddle tables
// and the forward_1d_fft / inverse_1d_fft entry points that replace the cuFFT
// calls of src/a2afft/dfft/kernels/a2affts.cu:7-76.
#include "fft.h"
#include "stockham.h"

#include <math.h>
#include <map>
#include <mutex>
#include <vector>

namespace a2a {

bool fft_length_supported(int N) { return N >= 2 && N <= 4096 && (N & (N - 1)) == 0; }

// ---------------------------------------------------------------------------
// Launch shapes.  Contiguous lines: E elements per thread (at least the
// largest radix), 256 threads per workgroup.  Strided tiles: T adjacent
// columns (128 B of complex64 / complex128 per row) and E chosen so a
// workgroup has at most 256 threads; LDS stays at or below 64 KiB.
// ---------------------------------------------------------------------------
template <typename real> constexpr int cont_E(int N) { return N <= 16 ? N : (N <= 64 ? 8 : 16); }
template <typename real> constexpr int cont_T(int N) {
    return (256 * cont_E<real>(N)) / N > 0 ? (256 * cont_E<real>(N)) / N : 1;
}
template <typename real> constexpr int str_T(int N) {
    return sizeof(real) == 4 ? (N <= 512 ? 16 : 8192 / N) : (N <= 512 ? 8 : 4096 / N);
}
template <typename real> constexpr int str_E(int N) {
    return (N * str_T<real>(N)) / 256 > rmax_of(N) ? (N * str_T<real>(N)) / 256 : rmax_of(N);
}

struct KInfo {
    const void* fn;
    int T, E, NT, lds;
};

template <typename real, int N, int E, int T, bool STRIDED, int DIR>
static KInfo kinfo() {
    KInfo k;
    k.fn = (const void*)&stockham_kernel<real, N, E, T, STRIDED, DIR>;
    k.T = T;
    k.E = E;
    k.NT = T * (N / E);
    k.lds = npass_of(N) > 1 ? N * T * int(2 * sizeof(real)) : 0;
    return k;
}

// variant: 0 contiguous, 1 strided (wide tile), 2 strided single column
template <typename real, int N, int DIR>
static KInfo kinfo_variant(int variant) {
    if (variant == 0) return kinfo<real, N, cont_E<real>(N), cont_T<real>(N), false, DIR>();
    if (variant == 1) return kinfo<real, N, str_E<real>(N), str_T<real>(N), true, DIR>();
    return kinfo<real, N, rmax_of(N), 1, true, DIR>();
}

template <typename real, int DIR>
static KInfo kinfo_n(int N, int variant) {
    switch (N) {
    case 2: return kinfo_variant<real, 2, DIR>(variant);
    case 4: return kinfo_variant<real, 4, DIR>(variant);
    case 8: return kinfo_variant<real, 8, DIR>(variant);
    case 16: return kinfo_variant<real, 16, DIR>(variant);
    case 32: return kinfo_variant<real, 32, DIR>(variant);
    case 64: return kinfo_variant<real, 64, DIR>(variant);
    case 128: return kinfo_variant<real, 128, DIR>(variant);
    case 256: return kinfo_variant<real, 256, DIR>(variant);
    case 512: return kinfo_variant<real, 512, DIR>(variant);
    case 1024: return kinfo_variant<real, 1024, DIR>(variant);
    case 2048: return kinfo_variant<real, 2048, DIR>(variant);
    case 4096: return kinfo_variant<real, 4096, DIR>(variant);
    default: return KInfo{nullptr, 0, 0, 0, 0};
    }
}

static KInfo kinfo_for(a2afft_dtype dt, int N, int dir, int variant) {
    if (dt == A2AFFT_C128) return dir < 0 ? kinfo_n<double, -1>(N, variant) : kinfo_n<double, 1>(N, variant);
    return dir < 0 ? kinfo_n<float, -1>(N, variant) : kinfo_n<float, 1>(N, variant);
}

static bool divides_either(uint32_t a, uint32_t b) { return a && b && (a % b == 0 || b % a == 0); }

// Resolve the variant and final geometry of a pass.
static int resolve(const AxisPass& p, a2afft_dtype dt, int N, KInfo* ki, FftGeom* g, int* grid, int* fallback) {
    if (!fft_length_supported(N)) return set_error(A2AFFT_E_UNSUPPORTED, "FFT length %d not supported", N);
    if (p.fft.n != uint32_t(N)) return set_error(A2AFFT_E_INVALID_ARG, "pass: axis extent %u != N %d", p.fft.n, N);
    int variant = p.strided ? 1 : 0;
    *fallback = 0;
    if (p.strided) {
        KInfo wide = kinfo_for(dt, N, -1, 1);
        if (p.a2.s_elem != 1 || p.a2.l % wide.T || p.a2.n % wide.T) {
            variant = 2;
            *fallback = 1;
        }
    }
    *ki = kinfo_for(dt, N, -1, variant);
    const int TPL = N / ki->E;
    if (!divides_either(p.fft.l, uint32_t(TPL)))
        return set_error(A2AFFT_E_UNSUPPORTED, "pass: block extent %u and %d threads per line do not divide", p.fft.l,
                         TPL);
    g->fft = p.fft;
    g->a1 = p.a1;
    g->a2 = p.a2;
    g->scale_mode = p.scale_mode;
    g->scale = p.scale;
    g->tstride = p.a2.s_elem;
    uint64_t units;
    if (variant == 0) {
        g->n2 = p.a2.n;
        units = uint64_t(p.a1.n) * p.a2.n;
        *grid = int((units + ki->T - 1) / ki->T);
        if ((units + ki->T - 1) / ki->T >= (uint64_t(1) << 31)) return set_error(A2AFFT_E_SIZE, "pass: grid too large");
    } else {
        g->n2 = p.a2.n / ki->T;
        units = uint64_t(p.a1.n) * g->n2;
        if (units >= (uint64_t(1) << 31)) return set_error(A2AFFT_E_SIZE, "pass: grid too large");
        *grid = int(units);
    }
    if (units >= (uint64_t(1) << 32)) return set_error(A2AFFT_E_SIZE, "pass: too many lines");
    g->nunits = uint32_t(units);
    g->dn2.set(g->n2 ? g->n2 : 1);
    return A2AFFT_OK;
}

int fft_pass_shape(const AxisPass& p, a2afft_dtype dt, int N, FftShape* out) {
    KInfo ki;
    FftGeom g;
    int grid, fb;
    A2A_TRY(resolve(p, dt, N, &ki, &g, &grid, &fb));
    out->T = ki.T;
    out->E = ki.E;
    out->NT = ki.NT;
    out->grid = grid;
    out->lds_bytes = ki.lds;
    out->fallback = fb;
    return A2AFFT_OK;
}

int launch_fft_pass(const AxisPass& p, void* data, a2afft_dtype dt, int N, int dir, hipStream_t s) {
    if (!data) return set_error(A2AFFT_E_INVALID_ARG, "fft: null data");
    KInfo ki;
    FftGeom g;
    int grid, fb;
    A2A_TRY(resolve(p, dt, N, &ki, &g, &grid, &fb));
    if (grid == 0) return A2AFFT_OK;
    const void* tw = twiddle_table(N, dir, dt);
    if (!tw) return set_error(A2AFFT_E_ALLOC, "fft: twiddle table for N=%d", N);
    KInfo k = kinfo_for(dt, N, dir, p.strided ? (fb ? 2 : 1) : 0);
    void* args[] = {&data, &g, (void*)&tw};
    int st = hip_check(int(hipLaunchKernel(k.fn, dim3(grid), dim3(k.NT), args, 0, s)), "stockham launch");
    if (st < 0) return st;
    return hip_check(int(hipGetLastError()), "stockham launch");
}

// ---------------------------------------------------------------------------
// Twiddles: computed in long double with an exact quadrant reduction, rounded
// once to the table's precision.
// ---------------------------------------------------------------------------
void twiddle_ld(int64_t m, int64_t M, int dir, long double* re, long double* im) {
    m %= M;
    if (m < 0) m += M;
    const int64_t k = (4 * m + M / 2) / M;          // nearest quadrant
    const int64_t num = 4 * m - k * M;              // angle = 2*pi*num/(4M), |angle| <= pi/4
    const long double pi = 3.14159265358979323846264338327950288L;
    long double c = 1.0L, sn = 0.0L;
    if (num != 0) {
        const long double a = (long double)num * pi / (2.0L * (long double)M);
        c = cosl(a);
        sn = sinl(a);
    }
    long double cr, sr;
    switch (int(k & 3)) {
    case 0: cr = c; sr = sn; break;
    case 1: cr = -sn; sr = c; break;
    case 2: cr = -c; sr = -sn; break;
    default: cr = sn; sr = -c; break;
    }
    *re = cr + 0.0L;
    *im = (dir < 0 ? -sr : sr) + 0.0L;
}

static void build_table(int N, int dir, std::vector<long double>& re, std::vector<long double>& im) {
    const int NP = npass_of(N);
    const int size = twsize_of(N);
    re.assign(size > 0 ? size : 1, 1.0L);
    im.assign(size > 0 ? size : 1, 0.0L);
    for (int p = 1; p < NP; ++p) {
        const int R = radix_of(N, p), Ns = ns_of(N, p), off = twoff_of(N, p);
        for (int k = 0; k < Ns; ++k)
            for (int r = 0; r < R; ++r)
                twiddle_ld(int64_t(r) * k, int64_t(Ns) * R, dir, &re[off + k * R + r], &im[off + k * R + r]);
    }
}

namespace {
struct TwKey {
    int dev, N, dir, dt;
    bool operator<(const TwKey& o) const {
        if (dev != o.dev) return dev < o.dev;
        if (N != o.N) return N < o.N;
        if (dir != o.dir) return dir < o.dir;
        return dt < o.dt;
    }
};
std::mutex g_tw_mu;
std::map<TwKey, void*> g_tw;
}  // namespace

const void* twiddle_table(int N, int dir, a2afft_dtype dt) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    TwKey key{dev, N, dir < 0 ? -1 : 1, int(dt)};
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto it = g_tw.find(key);
    if (it != g_tw.end()) return it->second;
    std::vector<long double> re, im;
    build_table(N, dir, re, im);
    const size_t n = re.size();
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
    g_tw[key] = d;
    return d;
}

}  // namespace a2a

using namespace a2a;

// Batched in-place DFTs over n_elems/Ng contiguous pencils (a2affts.cu:7-76).
static int fft_1d(void* data, int Ng, int64_t n_elems, a2afft_dtype dtype, int dir, void* stream) {
    if (!data || Ng <= 0 || n_elems <= 0 || n_elems % Ng)
        return set_error(A2AFFT_E_INVALID_ARG, "1d fft: bad arguments");
    const int64_t lines = n_elems / Ng;
    if (lines >= (int64_t(1) << 32)) return set_error(A2AFFT_E_SIZE, "1d fft: too many lines");
    AxisPass p;
    p.fft = make_axis(uint32_t(Ng), uint32_t(Ng), 0, 1);
    p.a1 = make_axis(uint32_t(lines), uint32_t(lines), 0, Ng);
    p.a2 = make_axis(1, 1, 0, 0);
    p.strided = 0;
    p.scale_mode = 0;
    p.scale = 1.0;
    return launch_fft_pass(p, data, dtype, Ng, dir, (hipStream_t)stream);
}

extern "C" int a2afft_forward_1d_fft(void* data, int Ng, int64_t n_elems, a2afft_dtype dtype, void* stream) {
    return finish_call(stream, fft_1d(data, Ng, n_elems, dtype, -1, stream));
}

extern "C" int a2afft_inverse_1d_fft(void* data, int Ng, int64_t n_elems, a2afft_dtype dtype, void* stream) {
    return finish_call(stream, fft_1d(data, Ng, n_elems, dtype, +1, stream));
}
