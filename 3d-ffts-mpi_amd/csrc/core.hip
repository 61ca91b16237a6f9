// Library core: status handling, the virtual-rank distribution that replaces
// a2aDistribution (src/a2afft/distribution/a2adistribution.{h,cpp}), and the
// plan / execute layer that replaces a2adfft (src/a2afft/dfft/a2adfft.{h,cpp}).
#include "common.h"
#include "fft.h"
#include "fft_shape.h"
#include "remap.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>
#include <string>
#include <vector>
#include <unistd.h>

namespace a2a {

static thread_local char g_err[512] = "";
static int g_sync = 0;

int set_error(int status, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return status;
}

int hip_check(int hip_status, const char* what) {
    if (hip_status == hipSuccess) return A2AFFT_OK;
    return set_error(A2AFFT_E_HIP, "%s: %s", what, hipGetErrorString(hipError_t(hip_status)));
}

bool sync_mode() { return g_sync != 0; }

int finish_call(void* stream, int status) {
    if (status < 0 || !g_sync) return status;
    (void)stream;
    int st = hip_check(int(hipDeviceSynchronize()), "device synchronize");
    return st < 0 ? st : status;
}

}  // namespace a2a

using namespace a2a;

// ===========================================================================
// Library
// ===========================================================================

extern "C" const char* a2afft_status_string(int s) {
    switch (s) {
    case A2AFFT_OK: return "ok";
    case A2AFFT_W_DIVERGES_FROM_REFERENCE: return "warning: dims outside the reference y-return domain";
    case A2AFFT_E_INVALID_ARG: return "invalid argument";
    case A2AFFT_E_DIVISIBILITY: return "Ng not divisible by dims";
    case A2AFFT_E_RANKS: return "world_size not compatible with dims";
    case A2AFFT_E_CHUNK: return "all-to-all chunk does not divide";
    case A2AFFT_E_HIP: return "HIP runtime error";
    case A2AFFT_E_ALLOC: return "allocation failed";
    case A2AFFT_E_UNSUPPORTED: return "unsupported length or precision";
    case A2AFFT_E_NO_SCRATCH: return "make_plans not called (no scratch)";
    case A2AFFT_E_SIZE: return "size out of range";
    default: return "unknown status";
    }
}

extern "C" const char* a2afft_last_error(void) { return g_err; }
extern "C" void a2afft_set_sync(int on) { g_sync = on ? 1 : 0; }
extern "C" int a2afft_version(void) { return 100; }

// ===========================================================================
// Distribution (a2aDistribution)
// ===========================================================================

struct a2afft_dist_s {
    int Ng;
    int P;
    int dims[3];
    int lgs[3];
    int64_t nlocal;
    int blockSize;
    a2afft_dtype dtype;
};

// MPI_Dims_create(nnodes, 3) with all dims free: as balanced as possible,
// non-increasing (MPI standard).  Prime factors, largest first, each go to
// the currently smallest dimension.  Gives 4->[2,2,1], 6->[3,2,1],
// 8->[2,2,2], 12->[3,2,2], 16->[4,2,2], 64->[4,4,4].
extern "C" int a2afft_dims_create(int nnodes, int dims[3]) {
    if (nnodes <= 0 || !dims) return set_error(A2AFFT_E_INVALID_ARG, "dims_create: bad arguments");
    int f[64], nf = 0, p = nnodes;
    for (int q = 2; (long)q * q <= p; ++q)
        while (p % q == 0) { f[nf++] = q; p /= q; }
    if (p > 1) f[nf++] = p;
    int d[3] = {1, 1, 1};
    for (int i = nf - 1; i >= 0; --i) {
        int m = 0;
        for (int a = 1; a < 3; ++a)
            if (d[a] < d[m]) m = a;
        d[m] *= f[i];
    }
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b)
            if (d[b] > d[a]) { int t = d[a]; d[a] = d[b]; d[b] = t; }
    dims[0] = d[0]; dims[1] = d[1]; dims[2] = d[2];
    return A2AFFT_OK;
}

extern "C" int a2afft_dist_create(a2afft_dist* out, int Ng, int world_size, const int* dims, int blockSize,
                                  a2afft_dtype dtype) {
    if (!out) return set_error(A2AFFT_E_INVALID_ARG, "dist_create: null out");
    *out = nullptr;
    if (Ng <= 0 || world_size <= 0 || blockSize <= 0 || blockSize > 1024 ||
        (dtype != A2AFFT_C128 && dtype != A2AFFT_C64))
        return set_error(A2AFFT_E_INVALID_ARG, "dist_create: bad arguments");
    int d[3];
    if (dims) {
        for (int a = 0; a < 3; ++a) d[a] = dims[a];
        if (d[0] <= 0 || d[1] <= 0 || d[2] <= 0) return set_error(A2AFFT_E_INVALID_ARG, "dist_create: dims <= 0");
        if (int64_t(d[0]) * d[1] * d[2] != world_size)
            return set_error(A2AFFT_E_RANKS, "dist_create: dims product != world_size");
    } else {
        a2afft_dims_create(world_size, d);
    }
    // a2adistribution.cpp:61-63
    for (int a = 0; a < 3; ++a)
        if (Ng % d[a]) return set_error(A2AFFT_E_DIVISIBILITY, "Ng %% dims[%d] != 0", a);
    int l[3] = {Ng / d[0], Ng / d[1], Ng / d[2]};
    const int64_t nlocal = int64_t(l[0]) * l[1] * l[2];
    // a2adistribution.cpp:74-80
    for (int a = 0; a < 3; ++a)
        if (world_size % d[a]) return set_error(A2AFFT_E_RANKS, "world_size %% dims[%d] != 0", a);
    if ((int64_t(l[0]) * l[1]) % world_size) return set_error(A2AFFT_E_CHUNK, "(lx*ly) %% world_size != 0");
    if ((int64_t(l[0]) * l[2]) % world_size) return set_error(A2AFFT_E_CHUNK, "(lx*lz) %% world_size != 0");
    if ((int64_t(l[1]) * l[2]) % world_size) return set_error(A2AFFT_E_CHUNK, "(ly*lz) %% world_size != 0");
    if (nlocal >= (int64_t(1) << 32)) return set_error(A2AFFT_E_SIZE, "nlocal >= 2^32");
    a2afft_dist_s* s = new (std::nothrow) a2afft_dist_s;
    if (!s) return set_error(A2AFFT_E_ALLOC, "dist_create");
    s->Ng = Ng;
    s->P = world_size;
    for (int a = 0; a < 3; ++a) { s->dims[a] = d[a]; s->lgs[a] = l[a]; }
    s->nlocal = nlocal;
    s->blockSize = blockSize;
    s->dtype = dtype;
    *out = s;
    return A2AFFT_OK;
}

static bool in_reference_domain(const a2afft_dist_s* d) { return d->dims[1] == 2 || d->dims[0] == 1; }

extern "C" int a2afft_dist_get_info(a2afft_dist d, a2afft_dist_info* info) {
    if (!d || !info) return set_error(A2AFFT_E_INVALID_ARG, "dist_get_info: null");
    info->Ng = d->Ng;
    info->nlocal = d->nlocal;
    info->world_size = d->P;
    for (int a = 0; a < 3; ++a) {
        info->local_grid_size[a] = d->lgs[a];
        info->dims[a] = d->dims[a];
    }
    info->blockSize = d->blockSize;
    info->dtype = d->dtype;
    info->reference_domain = in_reference_domain(d) ? 1 : 0;
    return A2AFFT_OK;
}

// rank2coords (a2adistribution.cpp:14-20), get_local_grid_start (:36-42)
extern "C" int a2afft_dist_rank_coords(a2afft_dist d, int rank, int coords[3], int start[3]) {
    if (!d || rank < 0 || rank >= d->P) return set_error(A2AFFT_E_INVALID_ARG, "rank_coords: bad rank");
    int c[3];
    c[0] = rank / (d->dims[1] * d->dims[2]);
    c[1] = (rank - c[0] * (d->dims[1] * d->dims[2])) / d->dims[2];
    c[2] = rank - c[0] * (d->dims[1] * d->dims[2]) - c[1] * d->dims[2];
    for (int a = 0; a < 3; ++a) {
        if (coords) coords[a] = c[a];
        if (start) start[a] = d->lgs[a] * c[a];
    }
    return A2AFFT_OK;
}

// a2adistribution.cpp:101-111
extern "C" int a2afft_dist_write_dims(a2afft_dist d, const char* path) {
    if (!d || !path) return set_error(A2AFFT_E_INVALID_ARG, "write_dims: null");
    if (d->nlocal > 0x7fffffff) return set_error(A2AFFT_E_SIZE, "write_dims: nlocal does not fit int32");
    FILE* f = fopen(path, "wb");
    if (!f) return set_error(A2AFFT_E_INVALID_ARG, "write_dims: cannot open %s", path);
    int32_t hdr[9] = {d->P, d->Ng, int32_t(d->nlocal), d->dims[0], d->dims[1], d->dims[2],
                      d->lgs[0], d->lgs[1], d->lgs[2]};
    size_t w = fwrite(hdr, sizeof(int32_t), 9, f);
    fclose(f);
    return w == 9 ? A2AFFT_OK : set_error(A2AFFT_E_INVALID_ARG, "write_dims: short write");
}

extern "C" int a2afft_dist_destroy(a2afft_dist d) {
    delete d;
    return A2AFFT_OK;
}

// ---- stages (a2adistribution.cpp:115-215) ---------------------------------

enum { Z_A2P = 0, Z_P2A = 1, X_A2P = 2, X_P2A = 3, Y_A2P = 4, Y_P2A = 5 };
enum { T_ZX = 0, T_XY = 1, T_YZ = 2 };

static int remap(const a2afft_dist_s* d, int kind, const void* s, void* t, hipStream_t st) {
    return launch_remap(kind, s, t, d->P, d->nlocal, d->lgs, d->dims, d->dtype, d->P, st);
}
static int transpose(const a2afft_dist_s* d, int which, const void* s, void* t, hipStream_t st) {
    return launch_transpose(which, s, t, d->lgs, d->nlocal, d->dtype, d->P, st);
}
static int all2all(const a2afft_dist_s* d, const void* s, void* t, hipStream_t st) {
    return launch_alltoall(s, t, d->P, d->nlocal, d->dtype, st);
}

static int stage(const a2afft_dist_s* d, int which, void* b1, void* b2, hipStream_t s) {
    if (!d || !b1 || !b2 || b1 == b2) return set_error(A2AFFT_E_INVALID_ARG, "stage: bad buffers");
    switch (which) {
    case 0:  // getZPencils :115-129
        A2A_TRY(all2all(d, b1, b2, s));
        return remap(d, Z_A2P, b2, b1, s);
    case 1:  // returnZPencils :131-145
        A2A_TRY(remap(d, Z_P2A, b1, b2, s));
        return all2all(d, b2, b1, s);
    case 2:  // getXPencils :147-163
        A2A_TRY(transpose(d, T_ZX, b1, b2, s));
        A2A_TRY(all2all(d, b2, b1, s));
        return remap(d, X_A2P, b1, b2, s);
    case 3:  // returnXPencils :165-179
        A2A_TRY(remap(d, X_P2A, b2, b1, s));
        return all2all(d, b1, b2, s);
    case 4:  // getYPencils :181-197
        A2A_TRY(transpose(d, T_XY, b2, b1, s));
        A2A_TRY(all2all(d, b1, b2, s));
        return remap(d, Y_A2P, b2, b1, s);
    default: {  // returnYPencils :199-215
        int w = remap(d, Y_P2A, b2, b1, s);
        if (w < 0) return w;
        A2A_TRY(all2all(d, b1, b2, s));
        A2A_TRY(transpose(d, T_YZ, b2, b1, s));
        return w;
    }
    }
}

#define A2A_STAGE_ENTRY(NAME, WHICH)                                                       \
    extern "C" int NAME(a2afft_dist d, void* b1, void* b2, void* stream) {                 \
        return finish_call(stream, stage(d, WHICH, b1, b2, (hipStream_t)stream));          \
    }
A2A_STAGE_ENTRY(a2afft_get_z_pencils, 0)
A2A_STAGE_ENTRY(a2afft_return_z_pencils, 1)
A2A_STAGE_ENTRY(a2afft_get_x_pencils, 2)
A2A_STAGE_ENTRY(a2afft_return_x_pencils, 3)
A2A_STAGE_ENTRY(a2afft_get_y_pencils, 4)
A2A_STAGE_ENTRY(a2afft_return_y_pencils, 5)

// ===========================================================================
// Plan / execute (a2adfft)
// ===========================================================================

// One kernel launch of a fused transform.
struct Launch {
    int pass;
    uint32_t start, count;
};

struct a2afft_plan_s {
    a2afft_dist_s d;    // the distribution (Ng = 0 for a single-grid plan)
    int n[3];           // global extents x, y, z
    a2afft_precision prec;
    a2afft_exec exec;
    void* scratch;
    AxisPass pass[3];   // fused: z, x, y over the block layout
    KSel sel[3];        // kernel variant per fused pass
    uint32_t chunk;     // z columns per x/y chunk (0: whole passes)
    char report[16384]; // last autotune table
    hipStream_t cap_stream = nullptr;                // graph capture stream (graph_launch)
    hipGraphExec_t graph[2] = {nullptr, nullptr};    // forward, inverse
    void* graph_data[2] = {nullptr, nullptr};        // the buffer each graph was captured for
    std::vector<Launch> launches;             // launch list for `launches_chunk`
    uint32_t launches_chunk = ~0u;
};

// Graphs hold the kernel choices and arguments of their capture: any change
// to the plan drops them, and the next graph_launch captures again.
static void drop_plan_graphs(a2afft_plan_s* p) {
    for (int d = 0; d < 2; ++d) {
        if (p->graph[d]) (void)hipGraphExecDestroy(p->graph[d]);
        p->graph[d] = nullptr;
        p->graph_data[d] = nullptr;
    }
}

// The launches of a fused transform: the z pass, then the x and y passes
// either whole or alternating over windows of `chunk` z columns, so a
// window's data is still in the 256 MB Infinity Cache when the y pass reads
// what the x pass wrote.  Every window holds complete x and y lines, so the
// arithmetic per line -- and the result -- is the same for every chunking.
// The list is rebuilt when the chunk width changes, so execution itself
// allocates nothing once a plan has run.
static const std::vector<Launch>& plan_launches(a2afft_plan_s* p) {
    if (p->launches_chunk == p->chunk) return p->launches;
    std::vector<Launch>& out = p->launches;
    out.clear();
    out.push_back(Launch{0, 0, 0});
    if (!p->chunk) {
        out.push_back(Launch{1, 0, 0});
        out.push_back(Launch{2, 0, 0});
    } else {
        const uint32_t nz = uint32_t(p->n[2]);
        for (uint32_t z0 = 0; z0 < nz; z0 += p->chunk) {
            const uint32_t cnt = nz - z0 < p->chunk ? nz - z0 : p->chunk;
            out.push_back(Launch{1, z0, cnt});
            out.push_back(Launch{2, z0, cnt});
        }
    }
    p->launches_chunk = p->chunk;
    return out;
}

static bool chunk_ok(const a2afft_plan_s* p, uint32_t chunk) {
    if (chunk == 0) return true;
    if (p->n[2] % int(chunk)) return false;
    for (int i = 1; i < 3; ++i)
        if (p->sel[i].T > 0 && chunk % uint32_t(p->sel[i].T)) return false;
    return true;
}

static bool is_split(a2afft_precision prec) { return prec == A2AFFT_SPLIT_BF16 || prec == A2AFFT_SPLIT_FP16; }

// Default chunk width.  x and y passes on the 512-point tiles run faster
// alternating over windows of z columns that fit the 256 MB memory-side
// cache, once the grid is larger than 512 MiB, when the x pass runs the
// half-width-exchange tile (and, for complex64, the y pass too):
//   512^3 complex64:  1.163 -> 1.134 ms with 128-column windows,
//   512^3 complex128: 2.40 -> 2.23 ms with 64-column windows.
// complex64 with full-exchange tiles (1.163 -> 1.189 ms) and smaller grids
// measured slower chunked: whole passes otherwise.  Split-precision plans
// window when both strided passes run the interleaved-K kernel; the
// split-K-halves kernels measured slower windowed (1.41-1.48 against
// 1.30-1.33 ms) and run whole passes.
static uint32_t default_chunk(const a2afft_plan_s* p) {
    const bool c64 = p->prec != A2AFFT_FP64;
    for (int i = 1; i < 3; ++i)
        if (p->pass[i].fft.n != 512u || !p->pass[i].strided) return 0;
    if (is_split(p->prec)) {
        // the interleaved-K kernel on both strided passes (split-BF16 512^3
        // complex64: 1.18-1.19 ms whole, 1.104-1.116 ms in 128-column
        // windows); the split-K-halves kernels run whole passes (windows
        // measured slower for them)
        if (p->sel[1].E != 0 || p->sel[2].E != 0) return 0;
    } else if (!(p->sel[1].flags & KF_HALFX) || (c64 && !(p->sel[2].flags & KF_HALFX))) {
        return 0;
    }
    const uint64_t column = uint64_t(p->n[0]) * uint64_t(p->n[1]) * (c64 ? 8u : 16u);   // bytes per z column
    if (column * uint64_t(p->n[2]) <= (uint64_t(512) << 20)) return 0;
    uint32_t w = 0;
    for (uint32_t c = 16; c < uint32_t(p->n[2]) && column * c <= (uint64_t(256) << 20); c *= 2)
        if (chunk_ok(p, c)) w = c;
    return w;
}

// The kernel choices that can run fused pass i of the plan.
static int pass_candidates(const a2afft_plan_s* p, int i, KSel* out, int max) {
    const int N = int(p->pass[i].fft.n);
    if (is_split(p->prec)) return mfma_candidates(p->pass[i], N, p->prec, out, max);
    return fft_candidates(p->pass[i], p->d.dtype, N, out, max);
}

static bool sel_allowed(const a2afft_plan_s* p, int i, const KSel& s) {
    KSel cand[256];
    const int n = pass_candidates(p, i, cand, 256);
    for (int c = 0; c < n; ++c)
        if (cand[c].T == s.T && cand[c].E == s.E && cand[c].flags == s.flags) return true;
    return false;
}

// Builds the twiddle tables selection `s` of pass i reads (for a chirp-z
// pass its chirp, filter spectrum and M-point tables), both directions, so
// that execution allocates nothing.
static int sel_tables(const a2afft_plan_s* p, int i, const KSel& s) {
    const int N = int(p->pass[i].fft.n);
    const bool ok = is_split(p->prec)
                        ? (mfma_twiddle_table(N, -1) && mfma_twiddle_table(N, 1))
                        : (pass_twiddles(p->pass[i], N, -1, p->d.dtype, s) && pass_twiddles(p->pass[i], N, 1, p->d.dtype, s));
    return ok ? A2AFFT_OK : set_error(A2AFFT_E_ALLOC, "plan: twiddle tables");
}

// Threads per workgroup that selection `s` of pass i launches.
static int sel_threads(const a2afft_plan_s* p, int i, const KSel& s) {
    const int N = int(p->pass[i].fft.n);
    if (is_split(p->prec)) return s.E == 0 ? 256 : (N >= 256 ? 512 : 256);
    FftShape sh;
    return fft_pass_shape(p->pass[i], p->d.dtype, N, &s, &sh) == A2AFFT_OK ? sh.NT : 0;
}

// One launch of fused pass `ps` with selection `sel`.
static int launch_pass(const a2afft_plan_s* p, const AxisPass& ps, void* D, int dir, hipStream_t s, const KSel* sel) {
    if (is_split(p->prec)) return launch_mfma_pass(ps, D, int(ps.fft.n), dir, p->prec, s, sel);
    return launch_fft_pass(ps, D, p->d.dtype, int(ps.fft.n), dir, s, sel);
}

static int plan_init(a2afft_plan_s* p) {
    p->scratch = nullptr;
    p->report[0] = 0;
    p->chunk = 0;
    PassMaps pm[3];
    fused_pass_maps(p->n, p->d.dims, p->d.lgs, p->d.nlocal, pm);
    for (int i = 0; i < 3; ++i) p->pass[i] = AxisPass{pm[i].fft, pm[i].a1, pm[i].a2, pm[i].strided, 0, 1.0};
    FftShape sh;
    for (int i = 0; i < 3; ++i) {
        const int N = int(p->pass[i].fft.n);
        if (is_split(p->prec)) {
            A2A_TRY(mfma_pass_check(p->pass[i], N, p->prec));
            p->sel[i] = mfma_default_sel(p->pass[i], N, p->prec, false);
            A2A_TRY(sel_tables(p, i, p->sel[i]));
            continue;
        }
        if (!fft_length_supported(N)) return set_error(A2AFFT_E_UNSUPPORTED, "plan: length %d not supported", N);
        p->sel[i] = fft_default_sel(p->pass[i], p->d.dtype, N);
        A2A_TRY(fft_pass_shape(p->pass[i], p->d.dtype, N, &p->sel[i], &sh));
        if (!find_kernel(N, p->d.dtype == A2AFFT_C128, p->pass[i].strided, p->sel[i], -1))
            return set_error(A2AFFT_E_UNSUPPORTED, "plan: no kernel built for length %d in this precision", N);
        // build the twiddle tables now so execute allocates nothing
        A2A_TRY(sel_tables(p, i, p->sel[i]));
    }
    // the staged path's batched pencil pass (its default selection)
    if (p->exec == A2AFFT_EXEC_STAGED && p->d.Ng > 0) {
        const int64_t total = int64_t(p->d.P) * p->d.nlocal;
        const PassMaps bm = batched_pass_maps(p->d.Ng, uint32_t(total / p->d.Ng));
        const AxisPass bp{bm.fft, bm.a1, bm.a2, 0, 0, 1.0};
        const KSel bs = fft_default_sel(bp, p->d.dtype, p->d.Ng);
        if (!pass_twiddles(bp, p->d.Ng, -1, p->d.dtype, bs) || !pass_twiddles(bp, p->d.Ng, 1, p->d.dtype, bs))
            return set_error(A2AFFT_E_ALLOC, "plan: twiddle tables");
    }
    p->chunk = default_chunk(p);
    if (p->chunk && is_split(p->prec))   // the windowed choice of the strided passes
        for (int i = 1; i < 3; ++i) p->sel[i] = mfma_default_sel(p->pass[i], int(p->pass[i].fft.n), p->prec, true);
    if (p->chunk && p->prec == A2AFFT_FP32) {
        // Windowed, the near-strided 512-point pass runs fastest in hardware
        // tile order: 512^3 complex64 over dims [1,4,4] 1.110-1.128 ms against
        // 1.143-1.168 ms with XCD-renumbered tiles (same bits).
        const KSel xcd{16, 32, KF_XCD | KF_PAD | KF_NT | KF_HALFX}, hw{16, 32, KF_PAD | KF_NT | KF_HALFX};
        for (int i = 1; i < 3; ++i) {
            const KSel& c = p->sel[i];
            if (p->pass[i].fft.n == 512u && p->pass[i].fft.s_elem * 8 < 65536 && c.T == xcd.T && c.E == xcd.E &&
                c.flags == xcd.flags && sel_allowed(p, i, hw)) {
                p->sel[i] = hw;
                A2A_TRY(sel_tables(p, i, hw));
            }
        }
    }
    return A2AFFT_OK;
}

static int check_precision(a2afft_precision prec, a2afft_dtype dt) {
    if (prec == A2AFFT_FP64 && dt != A2AFFT_C128)
        return set_error(A2AFFT_E_INVALID_ARG, "plan_create: FP64 needs complex128 data");
    if (prec != A2AFFT_FP64 && dt != A2AFFT_C64)
        return set_error(A2AFFT_E_INVALID_ARG, "plan_create: FP32 / split precisions need complex64 data");
    if (prec != A2AFFT_FP64 && prec != A2AFFT_FP32 && prec != A2AFFT_SPLIT_BF16 && prec != A2AFFT_SPLIT_FP16)
        return set_error(A2AFFT_E_INVALID_ARG, "plan_create: precision");
    return A2AFFT_OK;
}

// ---------------------------------------------------------------------------
// Tuning cache (no reference counterpart: the reference re-plans on every call,
// a2affts.cu:15).  Autotune results are kept in a text file, one line per plan
// key (device architecture, grid, decomposition, data type, precision and
// strategy):  "<key> | z=T,E,F x=T,E,F y=T,E,F chunk=C".  A plan created
// while a cache file is set takes its entry when every recorded kernel is
// built and fits; anything else is ignored.
// ---------------------------------------------------------------------------
static std::mutex g_cache_mu;
static std::string g_cache_path;

extern "C" int a2afft_set_tuning_cache(const char* path) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_cache_path = path ? path : "";
    return A2AFFT_OK;
}

// The key's version changes whenever a default kernel family or schedule
// changes (v3: 2^a 3^b 5^c lengths whose runtime schedule needs more than 24
// values a thread moved to the per-pass-mapped kernel), so that entries tuned
// within an old family are ignored and re-tuned rather than applied.
static std::string plan_key(const a2afft_plan_s* p) {
    int dev = 0;
    hipDeviceProp_t prop;
    std::string arch = "unknown";
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        arch = prop.gcnArchName;
        const size_t colon = arch.find(':');
        if (colon != std::string::npos) arch.resize(colon);
    }
    char buf[256];
    snprintf(buf, sizeof(buf), "a2afft-tune-v3 arch=%s grid=%dx%dx%d P=%d dims=%dx%dx%d dtype=%d prec=%d exec=%d",
             arch.c_str(), p->n[0], p->n[1], p->n[2], p->d.P, p->d.dims[0], p->d.dims[1], p->d.dims[2],
             int(p->d.dtype), int(p->prec), int(p->exec));
    return buf;
}

static std::vector<std::string> cache_lines(const std::string& path) {
    std::vector<std::string> out;
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return out;
    char line[1024];
    while (fgets(line, sizeof(line), f)) {
        std::string l(line);
        while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
        if (!l.empty()) out.push_back(l);
    }
    fclose(f);
    return out;
}

// Applies the cached entry of this plan, if any and valid; 1 when applied.
static int cache_apply(a2afft_plan_s* p) {
    std::string path;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        path = g_cache_path;
    }
    if (path.empty() || p->exec != A2AFFT_EXEC_FUSED) return 0;
    const std::string key = plan_key(p) + " | ";
    for (const std::string& l : cache_lines(path)) {
        if (l.compare(0, key.size(), key) != 0) continue;
        KSel sel[3];
        unsigned chunk = 0;
        if (sscanf(l.c_str() + key.size(), "z=%d,%d,%d x=%d,%d,%d y=%d,%d,%d chunk=%u", &sel[0].T, &sel[0].E,
                   &sel[0].flags, &sel[1].T, &sel[1].E, &sel[1].flags, &sel[2].T, &sel[2].E, &sel[2].flags,
                   &chunk) != 10)
            return 0;
        const KSel keep[3] = {p->sel[0], p->sel[1], p->sel[2]};
        const uint32_t keep_chunk = p->chunk;
        // validating an entry must not leave an error message behind
        char saved_err[sizeof(g_err)];
        memcpy(saved_err, g_err, sizeof(g_err));
        bool ok = true;
        for (int i = 0; i < 3 && ok; ++i) {
            ok = sel_allowed(p, i, sel[i]);
            if (ok) p->sel[i] = sel[i];
        }
        if (ok) {
            p->chunk = chunk;
            ok = (chunk == 0 || chunk >= 8) && chunk_ok(p, chunk);
        }
        for (int i = 0; i < 3 && ok; ++i) ok = sel_tables(p, i, p->sel[i]) == A2AFFT_OK;
        memcpy(g_err, saved_err, sizeof(g_err));
        if (!ok) {   // leave the defaults
            for (int i = 0; i < 3; ++i) p->sel[i] = keep[i];
            p->chunk = keep_chunk;
            return 0;
        }
        snprintf(p->report, sizeof(p->report), "tuning from cache %s\n", path.c_str());
        return 1;
    }
    return 0;
}

// Records this plan's selection in the cache file (replacing its old line).
static int cache_store(const a2afft_plan_s* p) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    if (g_cache_path.empty()) return A2AFFT_OK;
    const std::string key = plan_key(p) + " | ";
    char val[160];
    snprintf(val, sizeof(val), "z=%d,%d,%d x=%d,%d,%d y=%d,%d,%d chunk=%u", p->sel[0].T, p->sel[0].E,
             p->sel[0].flags, p->sel[1].T, p->sel[1].E, p->sel[1].flags, p->sel[2].T, p->sel[2].E, p->sel[2].flags,
             p->chunk);
    std::vector<std::string> lines = cache_lines(g_cache_path);
    std::vector<std::string> keep;
    for (const std::string& l : lines)
        if (l.compare(0, key.size(), key) != 0) keep.push_back(l);
    keep.push_back(key + val);
    // write a private temporary file, then rename it over the cache
    const std::string tmp = g_cache_path + ".tmp." + std::to_string(long(getpid()));
    FILE* f = fopen(tmp.c_str(), "w");
    if (!f) return set_error(A2AFFT_E_INVALID_ARG, "tuning cache: cannot write %s", tmp.c_str());
    bool ok = true;
    for (const std::string& l : keep) ok = ok && fprintf(f, "%s\n", l.c_str()) > 0;
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), g_cache_path.c_str()) != 0) {
        (void)remove(tmp.c_str());
        return set_error(A2AFFT_E_INVALID_ARG, "tuning cache: cannot update %s", g_cache_path.c_str());
    }
    return A2AFFT_OK;
}

extern "C" int a2afft_plan_create(a2afft_plan* out, a2afft_dist dist, a2afft_precision prec, a2afft_exec exec) {
    if (!out || !dist) return set_error(A2AFFT_E_INVALID_ARG, "plan_create: null");
    *out = nullptr;
    if (exec != A2AFFT_EXEC_FUSED && exec != A2AFFT_EXEC_STAGED)
        return set_error(A2AFFT_E_INVALID_ARG, "plan_create: bad exec");
    A2A_TRY(check_precision(prec, dist->dtype));
    if (exec == A2AFFT_EXEC_STAGED && is_split(prec))
        return set_error(A2AFFT_E_UNSUPPORTED, "plan_create: split precisions run the fused path only");
    if (int64_t(dist->P) * dist->nlocal >= (int64_t(1) << 40)) return set_error(A2AFFT_E_SIZE, "plan: too large");
    a2afft_plan_s* p = new (std::nothrow) a2afft_plan_s;
    if (!p) return set_error(A2AFFT_E_ALLOC, "plan_create");
    p->d = *dist;
    p->n[0] = p->n[1] = p->n[2] = dist->Ng;
    p->prec = prec;
    p->exec = exec;
    int st = plan_init(p);
    if (st < 0) {
        delete p;
        return st;
    }
    (void)cache_apply(p);
    *out = p;
    return A2AFFT_OK;
}

// One nx x ny x nz grid in C order (z fastest): the fused transform without a
// rank decomposition (dims [1,1,1]).
extern "C" int a2afft_grid_plan_create(a2afft_plan* out, int nx, int ny, int nz, a2afft_precision prec) {
    if (!out || nx <= 0 || ny <= 0 || nz <= 0) return set_error(A2AFFT_E_INVALID_ARG, "grid_plan_create: bad args");
    *out = nullptr;
    const a2afft_dtype dt = prec == A2AFFT_FP64 ? A2AFFT_C128 : A2AFFT_C64;
    A2A_TRY(check_precision(prec, dt));
    // Addresses are 64-bit throughout the fused passes (each pass's line or
    // tile count stays below 2^32, checked per launch), so a grid may fill
    // the device: 2048^3 complex64 is 64 GiB.
    const int64_t total = int64_t(nx) * ny * nz;
    if (total >= (int64_t(1) << 40)) return set_error(A2AFFT_E_SIZE, "grid_plan_create: grid >= 2^40 points");
    a2afft_plan_s* p = new (std::nothrow) a2afft_plan_s;
    if (!p) return set_error(A2AFFT_E_ALLOC, "grid_plan_create");
    p->d.Ng = 0;
    p->d.P = 1;
    p->d.dims[0] = p->d.dims[1] = p->d.dims[2] = 1;
    p->d.lgs[0] = nx;
    p->d.lgs[1] = ny;
    p->d.lgs[2] = nz;
    p->d.nlocal = total;
    p->d.blockSize = 64;
    p->d.dtype = dt;
    p->n[0] = nx;
    p->n[1] = ny;
    p->n[2] = nz;
    p->prec = prec;
    p->exec = A2AFFT_EXEC_FUSED;
    int st = plan_init(p);
    if (st < 0) {
        delete p;
        return st;
    }
    (void)cache_apply(p);
    *out = p;
    return A2AFFT_OK;
}

extern "C" int a2afft_make_plans(a2afft_plan p, void* scratch) {
    if (!p) return set_error(A2AFFT_E_INVALID_ARG, "make_plans: null plan");
    p->scratch = scratch;
    drop_plan_graphs(p);
    return A2AFFT_OK;
}

// a2adfft.cpp:26-158 in the reference stage order.
static int exec_staged(a2afft_plan_s* p, void* D, int dir, hipStream_t s) {
    if (!p->scratch) return set_error(A2AFFT_E_NO_SCRATCH, "staged execute needs scratch (make_plans)");
    if (p->d.Ng <= 0) return set_error(A2AFFT_E_INVALID_ARG, "staged execute needs a distribution plan");
    const a2afft_dist_s* d = &p->d;
    void* S = p->scratch;
    const int64_t total = int64_t(d->P) * d->nlocal;
    auto fft = [&](void* buf) -> int {
        const PassMaps pm = batched_pass_maps(d->Ng, uint32_t(total / d->Ng));
        const AxisPass ps{pm.fft, pm.a1, pm.a2, 0, 0, 1.0};
        return launch_fft_pass(ps, buf, d->dtype, d->Ng, dir, s, nullptr);
    };
    int warn = 0;
    A2A_TRY(stage(d, 0, D, S, s));
    A2A_TRY(fft(D));
    A2A_TRY(stage(d, 1, D, S, s));
    A2A_TRY(stage(d, 2, D, S, s));
    A2A_TRY(fft(S));
    A2A_TRY(stage(d, 3, D, S, s));
    A2A_TRY(stage(d, 4, D, S, s));
    A2A_TRY(fft(D));
    if (dir < 0) {
        A2A_TRY(launch_copy(S, D, total, d->dtype, s));   // a2adfft.cpp:79
    } else {
        const double scale = double(int64_t(d->Ng) * d->Ng * d->Ng);
        A2A_TRY(launch_scale(S, D, total, scale, d->dtype, s));   // a2adfft.cpp:147
    }
    int w = stage(d, 5, D, S, s);
    if (w < 0) return w;
    warn = w;
    return warn;
}

static int run_launch(a2afft_plan_s* p, const Launch& L, void* D, int dir, hipStream_t s) {
    AxisPass ps = p->pass[L.pass];
    ps.a2_start = L.start;
    ps.a2_count = L.count;
    if (L.pass == 2 && dir > 0) {
        ps.scale = double(int64_t(p->n[0]) * p->n[1] * p->n[2]);
        ps.scale_mode = 2;
    }
    return launch_pass(p, ps, D, dir, s, &p->sel[L.pass]);
}

// In-place passes over the block layout in the reference's axis order z, x,
// y, with 1/(nx*ny*nz) folded into the last inverse pass.
static int exec_fused(a2afft_plan_s* p, void* D, int dir, hipStream_t s) {
    for (const Launch& L : plan_launches(p)) A2A_TRY(run_launch(p, L, D, dir, s));
    return A2AFFT_OK;
}

static int execute(a2afft_plan p, void* data, int dir, void* stream) {
    if (!p || !data) return set_error(A2AFFT_E_INVALID_ARG, "execute: null");
    if (p->exec == A2AFFT_EXEC_STAGED) return exec_staged(p, data, dir, (hipStream_t)stream);
    return exec_fused(p, data, dir, (hipStream_t)stream);
}

extern "C" int a2afft_forward(a2afft_plan p, void* data, void* stream) {
    return finish_call(stream, execute(p, data, -1, stream));
}

extern "C" int a2afft_inverse(a2afft_plan p, void* data, void* stream) {
    return finish_call(stream, execute(p, data, +1, stream));
}

extern "C" int a2afft_execute_pass(a2afft_plan p, void* data, int pass, int direction, void* stream) {
    if (!p || !data || pass < 0 || pass > 2 || (direction != -1 && direction != 1))
        return set_error(A2AFFT_E_INVALID_ARG, "execute_pass: bad arguments");
    if (p->exec != A2AFFT_EXEC_FUSED) return set_error(A2AFFT_E_INVALID_ARG, "execute_pass: fused plans only");
    for (const Launch& L : plan_launches(p))
        if (L.pass == pass) {
            int st = run_launch(p, L, data, direction, (hipStream_t)stream);
            if (st < 0) return st;
        }
    return finish_call(stream, A2AFFT_OK);
}

extern "C" int a2afft_plan_num_launches(a2afft_plan p) {
    if (!p) return set_error(A2AFFT_E_INVALID_ARG, "num_launches: null plan");
    if (p->exec != A2AFFT_EXEC_FUSED) return set_error(A2AFFT_E_INVALID_ARG, "num_launches: fused plans only");
    return int(plan_launches(p).size());
}

// Launch `index` of a fused forward / inverse; returns the pass it belongs
// to (0 z, 1 x, 2 y) or a negative status.
extern "C" int a2afft_execute_launch(a2afft_plan p, void* data, int index, int direction, void* stream) {
    if (!p || !data || (direction != -1 && direction != 1))
        return set_error(A2AFFT_E_INVALID_ARG, "execute_launch: bad arguments");
    if (p->exec != A2AFFT_EXEC_FUSED) return set_error(A2AFFT_E_INVALID_ARG, "execute_launch: fused plans only");
    const std::vector<Launch>& L = plan_launches(p);
    if (index < 0 || index >= int(L.size())) return set_error(A2AFFT_E_INVALID_ARG, "execute_launch: index out of range");
    int st = run_launch(p, L[size_t(index)], data, direction, (hipStream_t)stream);
    if (st < 0) return st;
    st = finish_call(stream, A2AFFT_OK);
    return st < 0 ? st : L[size_t(index)].pass;
}

extern "C" int a2afft_plan_set_chunking(a2afft_plan p, int z_columns) {
    if (!p || z_columns < 0) return set_error(A2AFFT_E_INVALID_ARG, "set_chunking: bad arguments");
    if (p->exec != A2AFFT_EXEC_FUSED) return set_error(A2AFFT_E_INVALID_ARG, "set_chunking: fused plans only");
    if (z_columns >= p->n[2]) z_columns = 0;
    if (z_columns > 0 && z_columns < 8) return set_error(A2AFFT_E_INVALID_ARG, "set_chunking: at least 8 columns");
    if (!chunk_ok(p, uint32_t(z_columns)))
        return set_error(A2AFFT_E_INVALID_ARG, "set_chunking: %d columns must divide nz and be a multiple of the "
                         "x and y tile widths", z_columns);
    p->chunk = uint32_t(z_columns);
    drop_plan_graphs(p);
    return A2AFFT_OK;
}

extern "C" int a2afft_plan_set_kernel(a2afft_plan p, int pass, int T, int E, int flags) {
    if (!p || pass < 0 || pass > 2) return set_error(A2AFFT_E_INVALID_ARG, "set_kernel: bad arguments");
    const KSel s{T, E, flags};
    if (!sel_allowed(p, pass, s))
        return set_error(A2AFFT_E_UNSUPPORTED, "set_kernel: no built variant T=%d E=%d flags=%d fits pass %d", T, E,
                         flags, pass);
    const KSel old = p->sel[pass];
    p->sel[pass] = s;
    if (!chunk_ok(p, p->chunk)) {
        p->sel[pass] = old;
        return set_error(A2AFFT_E_INVALID_ARG, "set_kernel: T=%d does not divide the chunk width %u", T, p->chunk);
    }
    const int st = sel_tables(p, pass, s);
    if (st < 0) {
        p->sel[pass] = old;
        return st;
    }
    drop_plan_graphs(p);
    return A2AFFT_OK;
}

// The kernel family of a selection: variants within one family compute every
// line with the same arithmetic (the fixed-schedule Stockham shapes, the
// runtime-schedule tiles, the per-pass-mapped tiles, the chirp-z pass, the direct DFT; the interleaved-K
// or the split-K-halves MFMA kernels).
static int sel_family(const a2afft_plan_s* p, const KSel& s) {
    if (is_split(p->prec)) return s.E == 0 ? 1 : 2;
    return s.flags & (KF_MIX | KF_MIXL | KF_CZ | KF_DFT);
}

// Times every built variant that fits each fused pass (forward direction,
// `iters` launches each, in place on `work`, whose contents are destroyed)
// within the family of the pass's selection, keeps the fastest per pass, then
// the fastest chunk width, and keeps the result only if whole forwards beat
// the plan's choices before the call.  Every variant computes identical bits,
// so tuning never changes a result (set_kernel crosses families).
extern "C" int a2afft_plan_autotune(a2afft_plan p, void* work, int iters, void* stream) {
    if (!p || !work || iters <= 0) return set_error(A2AFFT_E_INVALID_ARG, "autotune: bad arguments");
    if (p->exec != A2AFFT_EXEC_FUSED) return set_error(A2AFFT_E_INVALID_ARG, "autotune: fused plans only");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    A2A_TRY(hip_check(int(hipEventCreate(&e0)), "autotune event"));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return set_error(A2AFFT_E_HIP, "autotune event");
    }
    int len = 0;
    const char* names[3] = {"z", "x", "y"};
    int st = A2AFFT_OK;
    const KSel start_sel[3] = {p->sel[0], p->sel[1], p->sel[2]};
    const uint32_t start_chunk = p->chunk;
    // mean ms of `reps` whole forwards with the plan's current choices
    auto time_forward = [&](int reps, float* ms) -> int {
        A2A_TRY(exec_fused(p, work, -1, s));
        A2A_TRY(hip_check(int(hipEventRecord(e0, s)), "autotune record"));
        for (int it = 0; it < reps; ++it) A2A_TRY(exec_fused(p, work, -1, s));
        A2A_TRY(hip_check(int(hipEventRecord(e1, s)), "autotune record"));
        A2A_TRY(hip_check(int(hipEventSynchronize(e1)), "autotune sync"));
        A2A_TRY(hip_check(int(hipEventElapsedTime(ms, e0, e1)), "autotune elapsed"));
        *ms /= float(reps);
        return A2AFFT_OK;
    };
    for (int i = 0; i < 3 && st >= 0; ++i) {
        KSel cand[256];
        const int n = pass_candidates(p, i, cand, 256);
        float best = 1e30f;
        for (int c = 0; c < n && st >= 0; ++c) {
            if (sel_family(p, cand[c]) != sel_family(p, start_sel[i])) continue;
            st = launch_pass(p, p->pass[i], work, -1, s, &cand[c]);   // warm
            if (st < 0) break;
            st = hip_check(int(hipEventRecord(e0, s)), "autotune record");
            for (int it = 0; it < iters && st >= 0; ++it) st = launch_pass(p, p->pass[i], work, -1, s, &cand[c]);
            if (st < 0) break;
            st = hip_check(int(hipEventRecord(e1, s)), "autotune record");
            if (st >= 0) st = hip_check(int(hipEventSynchronize(e1)), "autotune sync");
            float ms = 0.f;
            if (st >= 0) st = hip_check(int(hipEventElapsedTime(&ms, e0, e1)), "autotune elapsed");
            if (st < 0) break;
            ms /= float(iters);
            if (ms < best && (i == 0 || p->chunk == 0 || p->chunk % uint32_t(cand[c].T) == 0)) {
                best = ms;
                p->sel[i] = cand[c];
            }
            if (len < int(sizeof(p->report)) - 96)
                len += snprintf(p->report + len, sizeof(p->report) - len,
                                "pass %s T=%d E=%d NT=%d flags=%d %.4f ms\n", names[i], cand[c].T, cand[c].E,
                                sel_threads(p, i, cand[c]), cand[c].flags, ms);
        }
    }
    // chunk widths: time whole forwards
    if (st >= 0) {
        uint32_t cands[16];
        int nc = 0;
        cands[nc++] = 0;
        for (uint32_t w = 16; w < uint32_t(p->n[2]) && nc < 16; w *= 2)
            if (chunk_ok(p, w)) cands[nc++] = w;
        const uint32_t keep = p->chunk;
        float best = 1e30f;
        uint32_t best_w = keep;
        for (int k = 0; k < nc && st >= 0; ++k) {
            p->chunk = cands[k];
            st = exec_fused(p, work, -1, s);
            if (st < 0) break;
            st = hip_check(int(hipEventRecord(e0, s)), "autotune record");
            for (int it = 0; it < iters && st >= 0; ++it) st = exec_fused(p, work, -1, s);
            if (st < 0) break;
            st = hip_check(int(hipEventRecord(e1, s)), "autotune record");
            if (st >= 0) st = hip_check(int(hipEventSynchronize(e1)), "autotune sync");
            float ms = 0.f;
            if (st >= 0) st = hip_check(int(hipEventElapsedTime(&ms, e0, e1)), "autotune elapsed");
            if (st < 0) break;
            ms /= float(iters);
            if (ms < best) { best = ms; best_w = cands[k]; }
            if (len < int(sizeof(p->report)) - 96)
                len += snprintf(p->report + len, sizeof(p->report) - len, "forward chunk=%u %.4f ms\n", cands[k], ms);
        }
        p->chunk = best_w;
    }
    // Passes timed one by one can mislead where launches, not bytes, dominate
    // (small grids): keep the tuned choice only if whole forwards confirm it.
    if (st >= 0) {
        float tuned = 0.f, start = 0.f;
        const KSel tuned_sel[3] = {p->sel[0], p->sel[1], p->sel[2]};
        const uint32_t tuned_chunk = p->chunk;
        st = time_forward(2 * iters, &tuned);
        for (int i = 0; i < 3 && st >= 0; ++i) p->sel[i] = start_sel[i];
        p->chunk = start_chunk;
        if (st >= 0) st = time_forward(2 * iters, &start);
        if (st >= 0 && tuned < start) {
            for (int i = 0; i < 3; ++i) p->sel[i] = tuned_sel[i];
            p->chunk = tuned_chunk;
        }
        if (st >= 0 && len < int(sizeof(p->report)) - 96)
            len += snprintf(p->report + len, sizeof(p->report) - len, "forward tuned %.4f ms, before %.4f ms: kept %s\n",
                            tuned, start, tuned < start ? "tuned" : "before");
    }
    for (int i = 0; i < 3 && st >= 0; ++i) st = sel_tables(p, i, p->sel[i]);   // both directions
    if (st < 0) {   // leave the plan as it was
        for (int i = 0; i < 3; ++i) p->sel[i] = start_sel[i];
        p->chunk = start_chunk;
    }
    drop_plan_graphs(p);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st < 0) return st;
    A2A_TRY(cache_store(p));
    return A2AFFT_OK;
}

// Text description of the chosen kernel per pass and the last autotune table.
extern "C" int a2afft_plan_describe(a2afft_plan p, char* buf, int len) {
    if (!p || !buf || len <= 0) return set_error(A2AFFT_E_INVALID_ARG, "describe: bad arguments");
    const char* names[3] = {"z", "x", "y"};
    int n = 0;
    for (int i = 0; i < 3 && n < len; ++i) {
        if (is_split(p->prec)) {
            n += snprintf(buf + n, len - n, "chosen %s split-%s MFMA radix-16 %s T=%d E=%d NT=%d flags=%d\n", names[i],
                          p->prec == A2AFFT_SPLIT_BF16 ? "bf16" : "fp16",
                          p->sel[i].E == 0 ? "interleaved-K" : "split-K-halves", p->sel[i].T, p->sel[i].E,
                          sel_threads(p, i, p->sel[i]), p->sel[i].flags);
            continue;
        }
        FftShape sh;
        if (fft_pass_shape(p->pass[i], p->d.dtype, int(p->pass[i].fft.n), &p->sel[i], &sh) < 0) continue;
        n += snprintf(buf + n, len - n, "chosen %s T=%d E=%d NT=%d flags=%d grid=%d\n", names[i], sh.T, sh.E,
                      sh.NT, sh.flags, sh.grid);
    }
    if (n < len) n += snprintf(buf + n, len - n, "chunk %u z columns\n", p->chunk);
    if (n < len) snprintf(buf + n, len - n, "%s", p->report);
    return A2AFFT_OK;
}

// Pattern copy of one fused pass: mean ms of `iters` launches of a copy
// kernel with the pass kernel's addresses, tile order, cache policy and LDS
// footprint, in place and leaving `data` unchanged.
extern "C" int a2afft_plan_pattern_copy(a2afft_plan p, void* data, int pass, int iters, float* ms, void* stream) {
    if (!p || !data || !ms || pass < 0 || pass > 2 || iters <= 0)
        return set_error(A2AFFT_E_INVALID_ARG, "pattern copy: bad arguments");
    if (p->exec != A2AFFT_EXEC_FUSED || is_split(p->prec))
        return set_error(A2AFFT_E_UNSUPPORTED, "pattern copy: fused FP32/FP64 plans only");
    hipStream_t s = (hipStream_t)stream;
    const AxisPass& ap = p->pass[pass];
    const int N = int(ap.fft.n);
    A2A_TRY(launch_pattern_copy(ap, data, p->d.dtype, N, p->sel[pass], s));   // warm-up
    hipEvent_t e0, e1;
    A2A_TRY(hip_check(int(hipEventCreate(&e0)), "pattern copy event"));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return set_error(A2AFFT_E_HIP, "pattern copy event");
    }
    int st = hip_check(int(hipEventRecord(e0, s)), "pattern copy record");
    for (int it = 0; it < iters && st >= 0; ++it) st = launch_pattern_copy(ap, data, p->d.dtype, N, p->sel[pass], s);
    if (st >= 0) st = hip_check(int(hipEventRecord(e1, s)), "pattern copy record");
    if (st >= 0) st = hip_check(int(hipEventSynchronize(e1)), "pattern copy sync");
    if (st >= 0) st = hip_check(int(hipEventElapsedTime(ms, e0, e1)), "pattern copy elapsed");
    if (st >= 0) *ms /= float(iters);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

// Captures forward or inverse for `data` into a hipGraph on first use (on an
// internal stream) and replays it on `stream`; a call for another buffer, or
// after the plan's kernels or chunking change, captures again.
extern "C" int a2afft_graph_launch(a2afft_plan p, void* data, int direction, void* stream) {
    if (!p || !data || (direction != -1 && direction != 1))
        return set_error(A2AFFT_E_INVALID_ARG, "graph_launch: bad arguments");
    const int d = direction < 0 ? 0 : 1;
    int warn = 0;
    if (!p->graph[d] || p->graph_data[d] != data) {
        if (!p->cap_stream) A2A_TRY(hip_check(int(hipStreamCreate(&p->cap_stream)), "graph stream"));
        if (p->graph[d]) (void)hipGraphExecDestroy(p->graph[d]);
        p->graph[d] = nullptr;
        hipGraph_t g = nullptr;
        A2A_TRY(hip_check(int(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal)), "capture"));
        const int st = execute(p, data, direction, p->cap_stream);
        const int st2 = hip_check(int(hipStreamEndCapture(p->cap_stream, &g)), "end capture");
        if (st < 0 || st2 < 0) {
            if (g) (void)hipGraphDestroy(g);
            return st < 0 ? st : st2;
        }
        warn = st;
        const int st3 = hip_check(int(hipGraphInstantiate(&p->graph[d], g, nullptr, nullptr, 0)), "graph instantiate");
        (void)hipGraphDestroy(g);
        if (st3 < 0) return st3;
        p->graph_data[d] = data;
    }
    const int st = finish_call(stream, hip_check(int(hipGraphLaunch(p->graph[d], (hipStream_t)stream)), "graph launch"));
    return st < 0 ? st : warn;
}

extern "C" int a2afft_plan_destroy(a2afft_plan p) {
    if (p) {
        drop_plan_graphs(p);
        if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    }
    delete p;
    return A2AFFT_OK;
}

// ===========================================================================
// FP64 error reduction
// ===========================================================================

template <typename V>
__global__ __launch_bounds__(256) void err_partials(const V* __restrict__ y, const double2* __restrict__ ref,
                                                    uint64_t n, double* __restrict__ part) {
    __shared__ double s_d[256], s_r[256], s_m[256];
    double sd = 0.0, sr = 0.0, mx = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256) {
        const double2 r = ref[i];
        const double dx = double(y[i].x) - r.x, dy = double(y[i].y) - r.y;
        const double d2 = dx * dx + dy * dy;
        sd += d2;
        sr += r.x * r.x + r.y * r.y;
        mx = fmax(mx, sqrt(d2));
    }
    s_d[threadIdx.x] = sd;
    s_r[threadIdx.x] = sr;
    s_m[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
            s_d[threadIdx.x] += s_d[threadIdx.x + w];
            s_r[threadIdx.x] += s_r[threadIdx.x + w];
            s_m[threadIdx.x] = fmax(s_m[threadIdx.x], s_m[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = s_d[0];
        part[3 * blockIdx.x + 1] = s_r[0];
        part[3 * blockIdx.x + 2] = s_m[0];
    }
}

extern "C" int a2afft_error_metrics(const void* y, a2afft_dtype dtype, const void* ref, int64_t n, double* rel_l2,
                                    double* max_abs, void* stream) {
    if (!y || !ref || n <= 0 || !rel_l2 || !max_abs) return set_error(A2AFFT_E_INVALID_ARG, "error_metrics: bad args");
    const int blocks = 1024;
    double* part = nullptr;
    A2A_TRY(hip_check(int(hipMalloc(&part, sizeof(double) * 3 * blocks)), "error_metrics alloc"));
    hipStream_t s = (hipStream_t)stream;
    if (dtype == A2AFFT_C128)
        hipLaunchKernelGGL(err_partials<double2>, dim3(blocks), dim3(256), 0, s, (const double2*)y,
                           (const double2*)ref, uint64_t(n), part);
    else
        hipLaunchKernelGGL(err_partials<float2>, dim3(blocks), dim3(256), 0, s, (const float2*)y,
                           (const double2*)ref, uint64_t(n), part);
    int st = hip_check(int(hipGetLastError()), "error_metrics launch");
    double h[3 * 1024];
    if (st == A2AFFT_OK) st = hip_check(int(hipStreamSynchronize(s)), "error_metrics sync");
    if (st == A2AFFT_OK)
        st = hip_check(int(hipMemcpy(h, part, sizeof(h), hipMemcpyDeviceToHost)), "error_metrics copy");
    (void)hipFree(part);
    if (st < 0) return st;
    long double sd = 0.0L, sr = 0.0L;
    double mx = 0.0;
    for (int b = 0; b < blocks; ++b) {
        sd += h[3 * b];
        sr += h[3 * b + 1];
        mx = h[3 * b + 2] > mx ? h[3 * b + 2] : mx;
    }
    *rel_l2 = sr > 0 ? double(sqrtl(sd / sr)) : double(sqrtl(sd));
    *max_abs = mx;
    return A2AFFT_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free error reduction.  a2afft_error_metrics allocates its
// per-workgroup partial sums on every call; this variant keeps them in a
// file-scope device array instead, for callers that cannot allocate (code
// under stream capture, serving loops).  Each GPU holds its own copy of the
// array, found once per device with hipGetSymbolAddress; every call on a
// device shares it, so calls that overlap on one device race.  Every
// workgroup writes its three partials on every call, so nothing needs
// clearing between calls.
// ---------------------------------------------------------------------------
constexpr int kErrBlocks = 1024;
constexpr int64_t kErrNoAllocMaxElems = int64_t(1) << 36;   // 1 TiB of complex128: beyond any one GPU
static __device__ __align__(256) double scratch[3 * kErrBlocks];

static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;   // per device: the address of `scratch`

static int scratch_address(double** out) {
    int dev = 0;
    A2A_TRY(hip_check(int(hipGetDevice(&dev)), "error_metrics_noalloc: device"));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (size_t(dev) >= g_scratch_ptr.size()) g_scratch_ptr.resize(size_t(dev) + 1, nullptr);
    if (!g_scratch_ptr[size_t(dev)]) {
        void* p = nullptr;
        A2A_TRY(hip_check(int(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))), "error_metrics_noalloc: scratch"));
        if (!p) return set_error(A2AFFT_E_HIP, "error_metrics_noalloc: scratch has no device address");
        g_scratch_ptr[size_t(dev)] = p;
    }
    *out = static_cast<double*>(g_scratch_ptr[size_t(dev)]);
    return A2AFFT_OK;
}

extern "C" int a2afft_error_metrics_noalloc(const void* y, a2afft_dtype dtype, const void* ref, int64_t n,
                                            double* rel_l2, double* max_abs, void* stream) {
    if (!y || !ref || n <= 0 || !rel_l2 || !max_abs || (dtype != A2AFFT_C64 && dtype != A2AFFT_C128))
        return set_error(A2AFFT_E_INVALID_ARG, "error_metrics_noalloc: bad args");
    if (n > kErrNoAllocMaxElems)
        return set_error(A2AFFT_E_SIZE, "error_metrics_noalloc: %lld elements exceed the maximum of 2^36",
                         (long long)n);
    double* part = nullptr;
    A2A_TRY(scratch_address(&part));
    hipStream_t s = (hipStream_t)stream;
    if (dtype == A2AFFT_C128)
        hipLaunchKernelGGL(err_partials<double2>, dim3(kErrBlocks), dim3(256), 0, s, (const double2*)y,
                           (const double2*)ref, uint64_t(n), part);
    else
        hipLaunchKernelGGL(err_partials<float2>, dim3(kErrBlocks), dim3(256), 0, s, (const float2*)y,
                           (const double2*)ref, uint64_t(n), part);
    A2A_TRY(hip_check(int(hipGetLastError()), "error_metrics_noalloc launch"));
    double h[3 * kErrBlocks];
    A2A_TRY(hip_check(int(hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, s)), "error_metrics_noalloc copy"));
    A2A_TRY(hip_check(int(hipStreamSynchronize(s)), "error_metrics_noalloc sync"));
    long double sd = 0.0L, sr = 0.0L;
    double mx = 0.0;
    for (int b = 0; b < kErrBlocks; ++b) {
        sd += h[3 * b];
        sr += h[3 * b + 1];
        mx = h[3 * b + 2] > mx ? h[3 * b + 2] : mx;
    }
    *rel_l2 = sr > 0 ? double(sqrtl(sd / sr)) : double(sqrtl(sd));
    *max_abs = mx;
    return A2AFFT_OK;
}
