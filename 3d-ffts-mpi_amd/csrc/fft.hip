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

bool fft_length_supported(int N) { return fft_len_ok(N); }

template <typename real, int N, int DIR>
static const void* kernel_variant(int variant) {
    constexpr bool dbl = sizeof(real) == 8;
    if (variant == 0)
        return (const void*)&stockham_kernel<real, N, cont_E(N), cont_T(N), false, DIR>;
    if (variant == 1)
        return (const void*)&stockham_kernel<real, N, str_E(N, dbl), str_T(N, dbl), true, DIR>;
    return (const void*)&stockham_kernel<real, N, rmax_of(N), 1, true, DIR>;
}

template <typename real, int DIR>
static const void* kernel_n(int N, int variant) {
    switch (N) {
    case 2: return kernel_variant<real, 2, DIR>(variant);
    case 4: return kernel_variant<real, 4, DIR>(variant);
    case 8: return kernel_variant<real, 8, DIR>(variant);
    case 16: return kernel_variant<real, 16, DIR>(variant);
    case 32: return kernel_variant<real, 32, DIR>(variant);
    case 64: return kernel_variant<real, 64, DIR>(variant);
    case 128: return kernel_variant<real, 128, DIR>(variant);
    case 256: return kernel_variant<real, 256, DIR>(variant);
    case 512: return kernel_variant<real, 512, DIR>(variant);
    case 1024: return kernel_variant<real, 1024, DIR>(variant);
    case 2048: return kernel_variant<real, 2048, DIR>(variant);
    case 4096: return kernel_variant<real, 4096, DIR>(variant);
    default: return nullptr;
    }
}

static const void* kernel_for(a2afft_dtype dt, int N, int dir, int variant) {
    if (dt == A2AFFT_C128) return dir < 0 ? kernel_n<double, -1>(N, variant) : kernel_n<double, 1>(N, variant);
    return dir < 0 ? kernel_n<float, -1>(N, variant) : kernel_n<float, 1>(N, variant);
}

static int resolve(const AxisPass& p, a2afft_dtype dt, int N, Shape* sh, FftGeom* g, uint32_t* grid) {
    int st = resolve_pass(p.fft, p.a1, p.a2, p.strided, N, dt == A2AFFT_C128, sh, g, grid);
    if (st == A2AFFT_E_UNSUPPORTED)
        return set_error(st, "fft pass: length %d / block extent %u not supported", N, p.fft.l);
    if (st < 0) return set_error(st, "fft pass: bad geometry (N=%d)", N);
    g->scale_mode = p.scale_mode;
    g->scale = p.scale;
    return A2AFFT_OK;
}

int fft_pass_shape(const AxisPass& p, a2afft_dtype dt, int N, FftShape* out) {
    Shape sh;
    FftGeom g;
    uint32_t grid;
    A2A_TRY(resolve(p, dt, N, &sh, &g, &grid));
    out->T = sh.T;
    out->E = sh.E;
    out->NT = sh.NT;
    out->grid = int(grid);
    out->lds_bytes = sh.lds;
    out->fallback = sh.variant == 2;
    return A2AFFT_OK;
}

int launch_fft_pass(const AxisPass& p, void* data, a2afft_dtype dt, int N, int dir, hipStream_t s) {
    if (!data) return set_error(A2AFFT_E_INVALID_ARG, "fft: null data");
    Shape sh;
    FftGeom g;
    uint32_t grid;
    A2A_TRY(resolve(p, dt, N, &sh, &g, &grid));
    if (grid == 0) return A2AFFT_OK;
    const void* tw = twiddle_table(N, dir, dt);
    if (!tw) return set_error(A2AFFT_E_ALLOC, "fft: twiddle table for N=%d", N);
    const void* fn = kernel_for(dt, N, dir, sh.variant);
    void* args[] = {&data, &g, (void*)&tw};
    int st = hip_check(int(hipLaunchKernel(fn, dim3(grid), dim3(sh.NT), args, 0, s)), "stockham launch");
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
    const PassMaps pm = batched_pass_maps(Ng, uint32_t(lines));
    const AxisPass p{pm.fft, pm.a1, pm.a2, 0, 0, 1.0};
    return launch_fft_pass(p, data, dtype, Ng, dir, (hipStream_t)stream);
}

extern "C" int a2afft_forward_1d_fft(void* data, int Ng, int64_t n_elems, a2afft_dtype dtype, void* stream) {
    return finish_call(stream, fft_1d(data, Ng, n_elems, dtype, -1, stream));
}

extern "C" int a2afft_inverse_1d_fft(void* data, int Ng, int64_t n_elems, a2afft_dtype dtype, void* stream) {
    return finish_call(stream, fft_1d(data, Ng, n_elems, dtype, +1, stream));
}
