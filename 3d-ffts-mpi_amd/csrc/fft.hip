// Stockham FFT dispatch: kernel instantiations, launch shapes, twiddle tables
// and the forward_1d_fft / inverse_1d_fft entry points that replace the cuFFT
// calls of src/a2afft/dfft/kernels/a2affts.cu:7-76.
#include "fft.h"
#include "fft_shape.h"
#include "stockham.h"

#include <math.h>
#include <map>
#include <mutex>
#include <vector>

namespace a2a {

int register_kernels_f32(KEntry* out);
int register_kernels_f64(KEntry* out);

// Every length 2..kDftMaxN runs: the Stockham kernels where their schedules
// and builds cover it, the direct DFT otherwise.
bool fft_length_supported(int N) { return N >= 2 && N <= kDftMaxN; }

static bool is_dft(const KSel& s) { return (s.flags & KF_DFT) != 0; }
static bool is_mix(const KSel& s) { return (s.flags & KF_MIX) != 0; }
static bool is_mixl(const KSel& s) { return (s.flags & KF_MIXL) != 0; }
static bool is_cz(const KSel& s) { return (s.flags & KF_CZ) != 0; }
static KSel dft_sel(int N) { return KSel{1, N, KF_DFT}; }
static KSel cz_sel(int N) { return KSel{1, N, KF_CZ}; }

// Kernel choice for a length no fixed-schedule kernel covers: for 2^a 3^b 5^c
// lengths, the runtime-schedule Stockham when it holds at most 24 values a
// thread and the length is not a multiple of 20, else the per-pass-mapped
// Stockham, else the runtime schedule with more; then the chirp-z pass (9
// points and up), else the direct DFT.  The runtime schedule needs E
// divisible by every radix: lengths with both 3 and 5 (120, 360, 480, 960)
// run at E = 30, 1.1-4.3x slower per 3D transform (one contiguous and two
// strided passes) than the per-pass-mapped kernel, which also pairs 2s with
// 5s into radix-10 passes: multiples of 20 (100, 400, 800, 1000, 1600) run
// 1.06x (complex64) and 1.14x (complex128) faster on it, geometric means over
// 12 lengths.  2^a 3^b lengths and odd 3^b 5^c ones (243, 375, 729, 2187)
// stay on the runtime schedule, up to 1.7x faster there (49 lengths, 72 to
// 4000, both precisions, both layouts).  Where the same sweep's cost of a 3D
// transform favours the other family by more than 7 %, kFamilySwap overrides
// the rule.  Per GiB of complex64 lines the
// direct DFT costs O(N): 7.7 ms at N = 11 (one workgroup per line leaves most
// of its threads idle), 3.3 ms at 97, 13 ms at 509, 124 ms at 4093; the
// chirp-z pass 0.9-2.1 ms at every length from 11 to 4093.  complex128 agrees,
// except contiguous lines of 37-53 points, where the DFT is up to 1.3x faster.
// The choice depends on (N, precision) only, never on the layout, so that
// fused and staged plans compute every line alike.
//
// kFamilySwap: (N, precision) pairs whose family is the other one: the sweep
// (one contiguous pass plus two strided passes of 128-256 MiB, best tile of
// each family) favoured it by more than 7 %, and so did the strided passes of
// a grid (ms per x + y pass, this family against the other's best tile).
// Long 2^a 3 lines run their strided passes faster per-pass-mapped, which
// outweighs their contiguous passes; 4000 (complex64) and 1280, 2000
// (complex128) run theirs faster on the runtime schedule.  768 (complex64)
// and 288 (complex128) passed the sweep but not the grids (768x768x256 x pass
// 0.84 per-pass-mapped against 0.75 ms), so they keep the rule.
struct FamilySwap { int N; bool dbl; };
static constexpr FamilySwap kFamilySwap[] = {
    {1536, false},   // 1536x1536x64: 1.94 against 2.17 ms
    {3072, false},   // 3072x3072x16: 3.02 against 3.20 ms
    {1536, true},    // 1536x1536x32: 1.85 against 2.12 ms
    {4000, false},   // 4000x4000x8: 2.49 against 3.27 ms
    {1280, true},    // 1280x1280x64: 2.42 against 2.57 ms
    {2000, true},    // 2000x2000x16: 1.73 against 2.25 ms
};

static bool family_swapped(int N, bool dbl) {
    for (const FamilySwap& f : kFamilySwap)
        if (f.N == N && f.dbl == dbl) return true;
    return false;
}

static KSel fallback_sel(const AxisPass& p, a2afft_dtype dt, int N) {
    const bool dbl = dt == A2AFFT_C128;
    MixedPlan mp;
    const int e = mixed_plan_regs(N, dbl);
    const bool runtime = e > 0 && e <= 24 && (N % 20 != 0) != family_swapped(N, dbl);
    if (runtime && mixed_plan(p, N, dbl, &mp)) return KSel{mp.T, mp.E, KF_MIX};
    MixlPlan lp;
    if (mixl_plan(p, N, dbl, &lp)) return KSel{lp.T, lp.EM, KF_MIXL | KF_XCD};
    if (mixed_plan(p, N, dbl, &mp)) return KSel{mp.T, mp.E, KF_MIX};
    if (chirp_pass_ok(N)) return cz_sel(N);
    return dft_sel(N);
}

// Direct DFT for the lengths no Stockham kernel covers (prime factors other
// than 2, 3 and 5, mixed lengths above the one-thread-per-line limit, mixed
// lengths not built in a precision).  A workgroup reads one line into LDS;
// thread t forms outputs k = t, t + 256, ... as sum over n = 0..N-1, in that
// order, of x[n] * w[(n k) mod N] with explicit fma, so every layout of the
// same line gives the same bits, and writes them in place (all reads of the
// line precede the barrier).  w[m] = exp(DIR 2 pi i m / N), rounded once from
// long double, is staged in LDS when line and table fit in 64 KiB.
template <typename real, int DIR>
__global__ __launch_bounds__(256) void dft_kernel(typename CT<real>::V* __restrict__ data, FftGeom g,
                                                  const typename CT<real>::V* __restrict__ w, int N) {
    using V = typename CT<real>::V;
    extern __shared__ __align__(16) unsigned char dft_smem[];
    V* const x = reinterpret_cast<V*>(dft_smem);
    V* const wl = x + N;
    const bool stage_w = 2 * N * int(sizeof(V)) <= 65536;
    const int tid = int(threadIdx.x);
    const uint32_t ln = blockIdx.x;
    const uint32_t u1 = g.dn2.div(ln);
    const uint32_t u2 = ln - u1 * g.n2;
    const int64_t base = g.a1.off(u1) + g.a2.off(g.a2_start + u2);
    for (int n = tid; n < N; n += 256) {
        x[n] = data[base + g.fft.off(uint32_t(n))];
        if (stage_w) wl[n] = w[n];
    }
    __syncthreads();
    auto run = [&](const V* tw) {
        for (int k = tid; k < N; k += 256) {
            real sr = 0, si = 0;
            int m = 0;
            for (int n = 0; n < N; ++n) {
                const V a = x[n], t = tw[m];
                sr = fma(a.x, t.x, sr);
                sr = fma(-a.y, t.y, sr);
                si = fma(a.x, t.y, si);
                si = fma(a.y, t.x, si);
                m += k;
                if (m >= N) m -= N;
            }
            if (g.scale_mode == 1) {
                sr *= real(g.scale);
                si *= real(g.scale);
            } else if (g.scale_mode == 2) {
                sr = sr / real(g.scale);
                si = si / real(g.scale);
            }
            V o;
            o.x = sr;
            o.y = si;
            data[base + g.fft.off(uint32_t(k))] = o;
        }
    };
    if (stage_w) run(wl);
    else run(w);
}

static const void* dft_kernel_for(bool dbl, int dir) {
    if (dbl) return dir < 0 ? (const void*)&dft_kernel<double, -1> : (const void*)&dft_kernel<double, 1>;
    return dir < 0 ? (const void*)&dft_kernel<float, -1> : (const void*)&dft_kernel<float, 1>;
}

namespace {
constexpr int kMaxKernels = 4096;
KEntry g_kernels[kMaxKernels];
int g_nkernels = 0;
std::once_flag g_reg_once;

void init_registry() {
    std::call_once(g_reg_once, [] {
        int n = register_kernels_f32(g_kernels);
        n += register_kernels_f64(g_kernels + n);
        g_nkernels = n;
    });
}
}  // namespace

const void* find_kernel(int N, bool dbl, int strided, const KSel& s, int dir) {
    if (is_mix(s)) return s.flags == KF_MIX ? mixed_kernel_for(N, s.T, s.E, dbl, strided, dir) : nullptr;
    if (is_mixl(s)) return (s.flags & ~(KF_XCD | KF_NT)) == KF_MIXL ? mixl_kernel_for(s.E, dbl, strided, s.flags, dir)
                                                                     : nullptr;
    if (is_dft(s)) return (s.T == 1 && s.E == N && s.flags == KF_DFT && fft_length_supported(N)) ? dft_kernel_for(dbl, dir)
                                                                                                    : nullptr;
    if (is_cz(s)) return (s.T == 1 && s.E == N && s.flags == KF_CZ) ? chirp_pass_kernel(N, dbl, strided) : nullptr;
    init_registry();
    for (int i = 0; i < g_nkernels; ++i) {
        const KEntry& k = g_kernels[i];
        if (k.N == N && k.dbl == int(dbl) && k.strided == strided && k.T == s.T && k.E == s.E &&
            k.flags == s.flags && k.dir == (dir < 0 ? -1 : 1))
            return k.fn;
    }
    return nullptr;
}

int fft_candidates(const AxisPass& p, a2afft_dtype dt, int N, KSel* out, int max) {
    init_registry();
    const bool dbl = dt == A2AFFT_C128;
    int n = 0;
    for (int i = 0; i < g_nkernels && n < max; ++i) {
        const KEntry& k = g_kernels[i];
        if (k.N != N || k.dbl != int(dbl) || k.strided != p.strided || k.dir != -1) continue;
        if (!shape_fits(p.fft, p.a2, p.strided, N, k.T, k.E)) continue;
        out[n++] = KSel{k.T, k.E, k.flags};
    }
    if (n == 0 && fft_length_supported(N)) {
        MixlPlan lp;
        if (mixl_plan(p, N, dbl, &lp)) {
            // the per-pass-mapped kernel: every tile it accepts on this pass, each
            // in the built flag variants
            for (int tt : {1, 2, 4, 8, 12, 16, 24, 32, 48, 64}) {
                if (!mixl_plan(p, N, dbl, &lp, tt)) continue;
                for (int fl : {int(KF_XCD), KF_XCD | KF_NT, 0})
                    if (n < max && mixl_kernel_for(lp.EM, dbl, p.strided, fl, -1))
                        out[n++] = KSel{tt, lp.EM, KF_MIXL | fl};
            }
        }
        MixedPlan mp;
        if (mixed_plan(p, N, dbl, &mp) && n < max) {
            const KSel f{mp.T, mp.E, KF_MIX};
            out[n++] = f;
            // every other tile the runtime-schedule kernel accepts on this pass
            for (int tt : {1, 2, 4, 8, 16, 32, 64})
                if (tt != f.T && n < max && mixed_plan(p, N, dbl, &mp, tt)) out[n++] = KSel{tt, mp.E, KF_MIX};
        }
        if (chirp_pass_ok(N) && n < max) out[n++] = cz_sel(N);
        if (n < max) out[n++] = dft_sel(N);
    }
    return n;
}

// Measured preferences (MI355X autotune: 512 on 512^3 over dims [1,4,4],
// complex128 256 on 256^3 over dims [1,2,4], the other lengths on single
// grids).  XCD-aware order and non-temporal access
// help the 512 z and y passes; the far-strided x pass prefers z tiles
// outermost; padded LDS helps the long contiguous lines (2048 over 256x256
// lines: 0.384 ms padded, XCD-ordered and non-temporal, 0.474 ms plain).
// Strided 2048-point columns run 16 columns wide (128 B rows) with
// half-width exchanges: 0.50-0.52 ms per 1 GiB pass against 0.69 ms for the
// generic 4-column tiles (2048x256x256 x pass, 256x2048x256 y pass).  The
// first entry that is built and fits the geometry wins; otherwise the generic
// default with XCD order.
// `far`: 1 = the transform axis steps >= 64 KiB between elements (the x
// axis of a block), 2 = >= 1 MiB (the x axis of a whole 512^3 grid; also
// matched by 1), 0 = nearer, -1 = either.  `seg`: 1 = a line runs across
// rank blocks (the block extent along the axis is shorter than the line), 0 =
// a line sits in one block, -1 = either.  Contiguous 512-point lines in 4
// segments (the z pass of 512^3 over dims [1,4,4]) run fastest two lines to
// a 128-thread workgroup: 0.366 ms against 0.380 ms for 16-line tiles
// (pattern copies 6.05 against 5.80 TB/s); whole lines (512^3 in one block)
// do not gain (0.406 against 0.393 ms); whole lines run fastest one to a
// 64-thread workgroup (0.373 against 0.382 ms on a 512^3 grid).  On a whole
// 512^3 grid the x pass (2 MiB between elements) runs 32 columns wide (256 B
// rows, full exchanges, hardware order): 1.195 ms a forward against 1.328 ms
// on the block layout's choice.  Complex128 256-point lines in 4
// segments (256^3 over dims [1,2,4]) likewise prefer 4-line tiles: 0.097
// against 0.102 ms; complex128 512-point lines keep 8-line tiles (two-line
// tiles 0.81 against 0.78 ms).
struct Pref {
    int N, dbl, strided, far, seg;
    KSel s;
};
static const Pref kPrefs[] = {
    {512, 0, 0, -1, 1, {2, 8, KF_XCD | KF_NT}},
    {512, 0, 0, -1, 0, {1, 8, 0}},
    {512, 0, 0, -1, -1, {16, 16, KF_XCD | KF_PAD | KF_NT}},
    {512, 0, 1, 2, -1, {32, 32, KF_PAD}},
    {512, 0, 1, 1, -1, {16, 32, KF_A2MAJ | KF_HALFX}},
    {512, 0, 1, -1, -1, {16, 32, KF_XCD | KF_PAD | KF_NT | KF_HALFX}},
    {512, 1, 0, -1, -1, {8, 16, KF_XCD | KF_PAD | KF_HALFX}},
    {512, 1, 1, 1, -1, {8, 16, KF_A2MAJ | KF_HALFX}},
    {512, 1, 1, -1, -1, {8, 16, KF_XCD | KF_NT}},
    {256, 0, 1, 1, -1, {32, 32, KF_XCD | KF_PAD}},
    {256, 0, 1, -1, -1, {16, 32, KF_XCD}},
    {256, 1, 0, -1, 1, {4, 16, KF_XCD | KF_PAD | KF_HALFX}},
    {256, 1, 0, -1, -1, {16, 16, KF_XCD | KF_PAD | KF_HALFX}},
    {256, 1, 1, 1, -1, {8, 16, KF_XCD | KF_A2MAJ | KF_HALFX}},
    {256, 1, 1, -1, -1, {8, 16, KF_XCD | KF_PAD | KF_HALFX}},
    // strided 1024-point columns on 128 B rows (1024^3: complex64 x pass
    // 6.68 -> 4.81 ms, y 3.45 -> 3.34; complex128 x 12.9 -> 9.3, y 6.85 -> 6.17)
    {1024, 0, 1, 1, -1, {16, 32, KF_XCD | KF_NT | KF_HALFX}},
    {1024, 0, 1, 0, -1, {16, 32, KF_XCD | KF_PAD}},
    {1024, 1, 1, 1, -1, {8, 16, KF_XCD | KF_NT | KF_HALFX}},
    {1024, 1, 1, 0, -1, {8, 16, KF_XCD}},
    {2048, 0, 0, -1, -1, {2, 16, KF_XCD | KF_PAD | KF_NT}},
    // complex128 strided 2048-point columns on 128 B rows (2048^2 x 1024: x pass
    // 95.5 -> 37.8 ms, y 39.7 -> 26.4 ms against 32 B rows)
    {2048, 1, 1, 1, -1, {8, 32, KF_XCD | KF_PAD | KF_NT | KF_HALFX}},
    {2048, 1, 1, 0, -1, {8, 32, KF_XCD | KF_HALFX}},
    {2048, 0, 1, 2, -1, {16, 64, KF_XCD | KF_PAD | KF_NT | KF_HALFX}},   // 2048^3 grid x pass: 46.7 -> 44.8 ms
    {2048, 0, 1, 1, -1, {16, 64, KF_XCD | KF_A2MAJ | KF_HALFX}},
    {2048, 0, 1, -1, -1, {16, 64, KF_XCD | KF_HALFX}},
    {4096, 0, 0, -1, -1, {1, 16, KF_PAD}},
};

KSel fft_default_sel(const AxisPass& p, a2afft_dtype dt, int N) {
    if (!fft_len_ok(N)) return fallback_sel(p, dt, N);
    const bool dbl = dt == A2AFFT_C128;
    const int64_t step = p.fft.s_elem * (dbl ? 16 : 8);
    const int far = step >= (1 << 20) ? 2 : step >= 65536 ? 1 : 0;
    const int seg = p.fft.l < p.fft.n ? 1 : 0;
    for (const Pref& pr : kPrefs) {
        if (pr.N != N || pr.dbl != int(dbl) || pr.strided != p.strided) continue;
        if (pr.far >= 0 && (pr.far == 1 ? far < 1 : pr.far != far)) continue;
        if (pr.seg >= 0 && pr.seg != seg) continue;
        if (shape_fits(p.fft, p.a2, p.strided, N, pr.s.T, pr.s.E) && find_kernel(N, dbl, p.strided, pr.s, -1))
            return pr.s;
    }
    KSel s = default_sel(N, dbl, p.strided, p.fft, p.a2);
    if (npass_of(N) > 1 && s.T > 1 && is_pow2(N)) s.flags |= KF_XCD;   // mixed lengths: plain variants only
    if (!shape_fits(p.fft, p.a2, p.strided, N, s.T, s.E) || !find_kernel(N, dbl, p.strided, s, -1))
        return fallback_sel(p, dt, N);   // no fixed-schedule kernel for this length, precision or block extent
    return s;
}

// Geometry of a direct-DFT or chirp-z pass: lines enumerated (a1, a2) with a2
// fastest over the launch's a2 window (one line per direct-DFT workgroup).
static int resolve_dft(const AxisPass& p, int N, FftGeom* g, uint32_t* grid) {
    if (p.fft.n != uint32_t(N)) return set_error(A2AFFT_E_INVALID_ARG, "fft pass: bad geometry (N=%d)", N);
    // the contiguous enumeration with one line per unit, whatever the layout
    const int st = resolve_geom(p.fft, p.a1, p.a2, 0, 1, g, grid, p.a2_start, p.a2_count);
    if (st == A2AFFT_E_SIZE) return set_error(st, "fft pass: too many lines");
    if (st < 0) return set_error(st, "fft pass: bad window");
    g->scale_mode = p.scale_mode;
    g->scale = p.scale;
    return A2AFFT_OK;
}

static int resolve(const AxisPass& p, a2afft_dtype dt, int N, const KSel& s, FftGeom* g, uint32_t* grid,
                   MixedPlan* mp_out = nullptr, MixlPlan* lp_out = nullptr) {
    if (is_mixl(s)) {
        MixlPlan lp;
        if (!mixl_plan(p, N, dt == A2AFFT_C128, &lp, s.T) || lp.T != s.T || lp.EM != s.E)
            return set_error(A2AFFT_E_UNSUPPORTED, "fft pass: length %d (T=%d, E=%d) not supported", N, s.T, s.E);
        if (p.fft.n != uint32_t(N)) return set_error(A2AFFT_E_INVALID_ARG, "fft pass: bad geometry (N=%d)", N);
        const int st = resolve_geom(p.fft, p.a1, p.a2, p.strided, lp.T, g, grid, p.a2_start, p.a2_count);
        if (st < 0) return set_error(st, "fft pass: bad geometry (N=%d)", N);
        g->scale_mode = p.scale_mode;
        g->scale = p.scale;
        if (lp_out) *lp_out = lp;
        return A2AFFT_OK;
    }
    if (is_mix(s)) {
        MixedPlan mp;
        if (!mixed_plan(p, N, dt == A2AFFT_C128, &mp, s.T) || mp.T != s.T || mp.E != s.E)
            return set_error(A2AFFT_E_UNSUPPORTED, "fft pass: length %d (T=%d, E=%d) not supported", N, s.T, s.E);
        if (p.fft.n != uint32_t(N)) return set_error(A2AFFT_E_INVALID_ARG, "fft pass: bad geometry (N=%d)", N);
        const int st = resolve_geom(p.fft, p.a1, p.a2, p.strided, mp.T, g, grid, p.a2_start, p.a2_count);
        if (st < 0) return set_error(st, "fft pass: bad geometry (N=%d)", N);
        g->scale_mode = p.scale_mode;
        g->scale = p.scale;
        if (mp_out) *mp_out = mp;
        return A2AFFT_OK;
    }
    if (is_dft(s) || is_cz(s)) {
        if (s.T != 1 || s.E != N || !(is_cz(s) ? chirp_pass_ok(N) : fft_length_supported(N)))
            return set_error(A2AFFT_E_UNSUPPORTED, "fft pass: length %d not supported", N);
        return resolve_dft(p, N, g, grid);
    }
    int st = resolve_with(p.fft, p.a1, p.a2, p.strided, N, s.T, s.E, g, grid, p.a2_start, p.a2_count);
    if (st == A2AFFT_E_UNSUPPORTED)
        return set_error(st, "fft pass: length %d (block extent %u, T=%d, E=%d) not supported", N, p.fft.l, s.T, s.E);
    if (st < 0) return set_error(st, "fft pass: bad geometry (N=%d)", N);
    g->scale_mode = p.scale_mode;
    g->scale = p.scale;
    return A2AFFT_OK;
}

int fft_pass_shape(const AxisPass& p, a2afft_dtype dt, int N, const KSel* sel, FftShape* out) {
    const KSel s = sel ? *sel : fft_default_sel(p, dt, N);
    FftGeom g;
    uint32_t grid;
    MixedPlan mp{};
    A2A_TRY(resolve(p, dt, N, s, &g, &grid, &mp));
    out->T = s.T;
    out->E = s.E;
    // the launched thread count: the runtime-schedule kernel rounds its
    // workgroup to its own width (mp.NT)
    out->NT = (is_dft(s) || is_cz(s) || is_mixl(s)) ? 256 : is_mix(s) ? mp.NT : s.T * (N / s.E);
    out->grid = int(grid);
    out->flags = s.flags;
    return A2AFFT_OK;
}

int launch_fft_pass(const AxisPass& p, void* data, a2afft_dtype dt, int N, int dir, hipStream_t s, const KSel* sel) {
    if (!data) return set_error(A2AFFT_E_INVALID_ARG, "fft: null data");
    const KSel ks = sel ? *sel : fft_default_sel(p, dt, N);
    FftGeom g;
    uint32_t grid;
    MixedPlan mp{};
    MixlPlan lp{};
    A2A_TRY(resolve(p, dt, N, ks, &g, &grid, &mp, &lp));
    if (grid == 0) return A2AFFT_OK;
    const void* fn = find_kernel(N, dt == A2AFFT_C128, p.strided, ks, dir);
    if (!fn)
        return set_error(A2AFFT_E_UNSUPPORTED, "fft: no kernel built for N=%d T=%d E=%d flags=%d", N, ks.T, ks.E,
                         ks.flags);
    const void* tw = pass_twiddles(p, N, dir, dt, ks);
    if (!tw) return set_error(A2AFFT_E_ALLOC, "fft: twiddle table for N=%d", N);
    int st;
    if (is_mixl(ks)) {
        void* args[] = {&data, &g, (void*)&tw, &lp};
        st = hip_check(int(hipLaunchKernel(fn, dim3(grid), dim3(256), args,
                                           size_t(mixl_lds_bytes(lp, dt == A2AFFT_C128, p.strided)), s)),
                       "mixed-radix launch");
    } else if (is_mix(ks)) {
        void* args[] = {&data, &g, (void*)&tw, &mp};
        st = hip_check(int(hipLaunchKernel(fn, dim3(grid), dim3(mp.NT), args,
                                           size_t(mixed_lds_bytes(mp, dt == A2AFFT_C128)), s)),
                       "mixed-radix launch");
    } else if (is_cz(ks)) {
        st = launch_chirp_pass(g, grid, p.strided, data, dt, N, dir, s);
    } else if (is_dft(ks)) {
        const size_t line = size_t(N) * size_t(esize_of(dt));
        const size_t lds = 2 * line <= 65536 ? 2 * line : line;
        int n = N;
        void* args[] = {&data, &g, (void*)&tw, &n};
        st = hip_check(int(hipLaunchKernel(fn, dim3(grid), dim3(256), args, lds, s)), "dft launch");
    } else {
        void* args[] = {&data, &g, (void*)&tw};
        st = hip_check(int(hipLaunchKernel(fn, dim3(grid), dim3(ks.T * (N / ks.E)), args, 0, s)), "stockham launch");
    }
    if (st < 0) return st;
    return hip_check(int(hipGetLastError()), "fft launch");
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

// Stockham tables (kind 1) or the N roots of unity of the direct DFT (kind 2).
static const void* cached_table(int N, int dir, a2afft_dtype dt, int kind) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    TwKey key{dev, N, (dir < 0 ? -1 : 1) * kind, int(dt)};
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto it = g_tw.find(key);
    if (it != g_tw.end()) return it->second;
    std::vector<long double> re, im;
    if (kind == 1) {
        build_table(N, dir, re, im);
    } else {
        re.resize(size_t(N));
        im.resize(size_t(N));
        for (int m = 0; m < N; ++m) twiddle_ld(m, N, dir, &re[size_t(m)], &im[size_t(m)]);
    }
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

const void* twiddle_table(int N, int dir, a2afft_dtype dt) { return cached_table(N, dir, dt, 1); }

const void* pass_twiddles(const AxisPass& p, int N, int dir, a2afft_dtype dt, const KSel& s) {
    if (is_mix(s)) {
        MixedPlan mp;
        if (!mixed_plan(p, N, dt == A2AFFT_C128, &mp)) return nullptr;
        return mixed_twiddles(mp, dir, dt);
    }
    if (is_mixl(s)) {
        MixlPlan lp;
        if (!mixl_plan(p, N, dt == A2AFFT_C128, &lp)) return nullptr;
        return mixl_twiddles(lp, dir, dt);
    }
    if (is_cz(s)) return chirp_pass_tables(N, dir, dt);
    return cached_table(N, dir, dt, is_dft(s) ? 2 : 1);
}

}  // namespace a2a

using namespace a2a;

namespace a2a {
int fft_lines(void* data, int N, int64_t lines, a2afft_dtype dtype, int dir, hipStream_t s) {
    if (lines >= (int64_t(1) << 32)) return set_error(A2AFFT_E_SIZE, "1d fft: too many lines");
    if (N > kDftMaxN) return launch_long_fft(data, N, lines, dtype, dir, s);
    const PassMaps pm = batched_pass_maps(N, uint32_t(lines));
    const AxisPass p{pm.fft, pm.a1, pm.a2, 0, 0, 1.0};
    return launch_fft_pass(p, data, dtype, N, dir, s, nullptr);
}
}  // namespace a2a

// Batched in-place DFTs over n_elems/Ng contiguous pencils (a2affts.cu:7-76).
static int fft_1d(void* data, int Ng, int64_t n_elems, a2afft_dtype dtype, int dir, void* stream) {
    if (!data || Ng <= 0 || n_elems <= 0 || n_elems % Ng)
        return set_error(A2AFFT_E_INVALID_ARG, "1d fft: bad arguments");
    return fft_lines(data, Ng, n_elems / Ng, dtype, dir, (hipStream_t)stream);
}

extern "C" int a2afft_forward_1d_fft(void* data, int Ng, int64_t n_elems, a2afft_dtype dtype, void* stream) {
    return finish_call(stream, fft_1d(data, Ng, n_elems, dtype, -1, stream));
}

extern "C" int a2afft_inverse_1d_fft(void* data, int Ng, int64_t n_elems, a2afft_dtype dtype, void* stream) {
    return finish_call(stream, fft_1d(data, Ng, n_elems, dtype, +1, stream));
}
