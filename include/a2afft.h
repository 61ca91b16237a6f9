/*
 * a2afft -- MI355X-native pencil 3D FFT (C ABI).
 *
 * Drop-in for the a2afft hot path of humz2k/3D-FFTs-MPI (src/a2afft/).  The
 * reference's P MPI ranks become P *virtual ranks* whose blocks sit side by
 * side in one HBM allocation, so every buffer handed to this library is
 * [P][nlocal] interleaved complex (re, im) of the distribution's dtype, and
 * each MPI_Alltoall becomes a device-local permutation.
 *
 * Conventions
 *   - Device pointers are plain `void*`; the element type is the dtype enum.
 *   - `stream` is a hipStream_t passed as `void*` (NULL = the null stream).
 *   - Every entry point returns an int status (A2AFFT_OK == 0).  Negative
 *     values are errors; positive values are warnings and the call has run.
 *   - Calls are stream-ordered and return without synchronising, unless
 *     a2afft_set_sync(1) was called, in which case every call ends with a
 *     device synchronise as the reference's launchers do
 *     (e.g. src/a2afft/distribution/kernels/a2apencils.cu:34).
 *   - Handles are not thread-safe; distinct plans may run on distinct streams.
 */
#ifndef A2AFFT_H
#define A2AFFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define A2AFFT_OK                          0
/* y-return ran with dims outside the domain where the reference's formula
 * (a2apencils.cu:197) is a bijection; the true inverse was applied instead. */
#define A2AFFT_W_DIVERGES_FROM_REFERENCE   1
#define A2AFFT_E_INVALID_ARG              -1
#define A2AFFT_E_DIVISIBILITY             -2  /* Ng % dims[i] != 0           (a2adistribution.cpp:61-63) */
#define A2AFFT_E_RANKS                    -3  /* P % dims[i] != 0 / dims != P (a2adistribution.cpp:74-76) */
#define A2AFFT_E_CHUNK                    -4  /* (l_a*l_b) % P != 0           (a2adistribution.cpp:78-80) */
#define A2AFFT_E_HIP                      -5
#define A2AFFT_E_ALLOC                    -6
#define A2AFFT_E_UNSUPPORTED              -7  /* length / precision not built */
#define A2AFFT_E_NO_SCRATCH               -8  /* execute before make_plans    (a2adfft.cpp:18-24) */
#define A2AFFT_E_SIZE                     -9  /* index range exceeded */

typedef enum { A2AFFT_C128 = 0, A2AFFT_C64 = 1 } a2afft_dtype;

typedef enum {
    A2AFFT_FP64 = 0,        /* c128 data, FP64 Stockham (the reference's precision) */
    A2AFFT_FP32 = 1,        /* c64 data, FP32 Stockham */
    A2AFFT_SPLIT_BF16 = 2,  /* c64 data, radix-16 stages on MFMA, split bf16 hi+lo */
    A2AFFT_SPLIT_FP16 = 3   /* c64 data, radix-16 stages on MFMA, split fp16 hi+lo */
} a2afft_precision;

typedef enum {
    A2AFFT_EXEC_FUSED = 0,  /* 3 in-place axis passes over the block layout */
    A2AFFT_EXEC_STAGED = 1  /* the reference stage order: 18 launches (a2adfft.cpp:26-158) */
} a2afft_exec;

typedef struct a2afft_dist_s* a2afft_dist;
typedef struct a2afft_plan_s* a2afft_plan;

/* Public fields of a2aDistribution (a2adistribution.h:8-17). */
typedef struct {
    int Ng;
    int64_t nlocal;
    int world_size;
    int local_grid_size[3];
    int dims[3];
    int blockSize;
    int dtype;               /* a2afft_dtype */
    int reference_domain;    /* 1 when every remap equals the reference bit for bit */
} a2afft_dist_info;

/* ---- library ------------------------------------------------------------ */
const char* a2afft_status_string(int status);
/* Detail of the last error raised on the calling thread ("" if none). */
const char* a2afft_last_error(void);
/* 1: synchronise the device after every call (the reference's behaviour). */
void a2afft_set_sync(int on);
int a2afft_version(void);

/* ---- decomposition: a2aDistribution ------------------------------------- */
/* MPI_Dims_create(nnodes, 3, dims) with all dims free
 * (replaces the call at a2adistribution.cpp:59). */
int a2afft_dims_create(int nnodes, int dims[3]);

/* a2aDistribution::a2aDistribution(MPI_Comm, Ng, blockSize)
 * (a2adistribution.cpp:44-113).  The communicator becomes `world_size`
 * virtual ranks; dims == NULL applies the MPI_Dims_create rule, otherwise
 * dims[0]*dims[1]*dims[2] must equal world_size.  Checks the reference's
 * asserted preconditions and returns their codes instead of aborting. */
int a2afft_dist_create(a2afft_dist* out, int Ng, int world_size, const int* dims,
                       int blockSize, a2afft_dtype dtype);
int a2afft_dist_get_info(a2afft_dist dist, a2afft_dist_info* info);
/* rank2coords + get_local_grid_start (a2adistribution.cpp:14-42). */
int a2afft_dist_rank_coords(a2afft_dist dist, int rank, int coords[3],
                            int local_coordinates_start[3]);
/* The `dims` file written under saveDims (a2adistribution.cpp:101-111):
 * int32 world_size, Ng, nlocal, dims[3], local_grid_size[3]. */
int a2afft_dist_write_dims(a2afft_dist dist, const char* path);
/* a2aDistribution::finalize (a2adistribution.cpp:217-223). */
int a2afft_dist_destroy(a2afft_dist dist);

/* Stage methods (a2adistribution.cpp:115-215) over all P virtual ranks.
 * buff1/buff2 play d_Buff1/d_Buff2; each is [P][nlocal] complex. */
int a2afft_get_z_pencils(a2afft_dist dist, void* buff1, void* buff2, void* stream);
int a2afft_return_z_pencils(a2afft_dist dist, void* buff1, void* buff2, void* stream);
int a2afft_get_x_pencils(a2afft_dist dist, void* buff1, void* buff2, void* stream);
int a2afft_return_x_pencils(a2afft_dist dist, void* buff1, void* buff2, void* stream);
int a2afft_get_y_pencils(a2afft_dist dist, void* buff1, void* buff2, void* stream);
int a2afft_return_y_pencils(a2afft_dist dist, void* buff1, void* buff2, void* stream);

/* ---- launcher level (distribution/kernels/kernels.h:3-17) --------------- */
/* Each applies the reference kernel's index map to `nbuffers` consecutive
 * rank buffers of `nlocal` elements (nbuffers = world_size for all virtual
 * ranks, 1 for one rank's share).  blockSize is accepted for signature
 * parity and only bounds-checked: the launch shape is the library's. */
int a2afft_launch_d_z_a2a_to_z_pencils(const void* source, void* dest, int blockSize,
        int world_size, int64_t nlocal, const int local_grid_size[3], const int dims[3],
        a2afft_dtype dtype, int nbuffers, void* stream);   /* a2apencils.cu:23-35 */
int a2afft_launch_d_z_pencils_to_z_a2a(const void* source, void* dest, int blockSize,
        int world_size, int64_t nlocal, const int local_grid_size[3], const int dims[3],
        a2afft_dtype dtype, int nbuffers, void* stream);   /* a2apencils.cu:58-71 */
int a2afft_launch_d_x_a2a_to_x_pencils(const void* source, void* dest, int blockSize,
        int world_size, int64_t nlocal, const int local_grid_size[3], const int dims[3],
        a2afft_dtype dtype, int nbuffers, void* stream);   /* a2apencils.cu:94-108 */
int a2afft_launch_d_x_pencils_to_x_a2a(const void* source, void* dest, int blockSize,
        int world_size, int64_t nlocal, const int local_grid_size[3], const int dims[3],
        a2afft_dtype dtype, int nbuffers, void* stream);   /* a2apencils.cu:132-146 */
int a2afft_launch_d_y_a2a_to_y_pencils(const void* source, void* dest, int blockSize,
        int world_size, int64_t nlocal, const int local_grid_size[3], const int dims[3],
        a2afft_dtype dtype, int nbuffers, void* stream);   /* a2apencils.cu:170-186 */
/* Returns A2AFFT_W_DIVERGES_FROM_REFERENCE when dims are outside the
 * reference formula's valid domain (a2apencils.cu:197); the true inverse runs. */
int a2afft_launch_d_y_pencils_to_y_a2a(const void* source, void* dest, int blockSize,
        int world_size, int64_t nlocal, const int local_grid_size[3], const int dims[3],
        a2afft_dtype dtype, int nbuffers, void* stream);   /* a2apencils.cu:210-226 */
int a2afft_launch_d_fast_z_to_x(const void* source, void* dest, const int local_grid_size[3],
        int blockSize, int64_t nlocal, a2afft_dtype dtype, int nbuffers,
        void* stream);                                     /* a2achangefastaxis.cu:24-32 */
int a2afft_launch_d_fast_x_to_y(const void* source, void* dest, const int local_grid_size[3],
        int blockSize, int64_t nlocal, a2afft_dtype dtype, int nbuffers,
        void* stream);                                     /* a2achangefastaxis.cu:55-63 */
int a2afft_launch_d_fast_y_to_z(const void* source, void* dest, const int local_grid_size[3],
        int blockSize, int64_t nlocal, a2afft_dtype dtype, int nbuffers,
        void* stream);                                     /* a2achangefastaxis.cu:86-94 */

/* MPI_Alltoall(send, nlocal/P, TYPE_COMPLEX, recv, ...) over the P virtual
 * ranks of [P][nlocal] buffers: recv[s][q*ns+i] = send[q][s*ns+i]
 * (a2adistribution.cpp:117-125 and the five other call sites). */
int a2afft_alltoall(const void* send, void* recv, int world_size, int64_t nlocal,
                    a2afft_dtype dtype, void* stream);

/* ---- FFT launchers (dfft/kernels/kernels.h:5-11) ------------------------ */
/* forward_1d_fft / inverse_1d_fft (a2affts.cu:7-76): n_elems/Ng in-place
 * unnormalised DFTs of contiguous length-Ng pencils (replaces cuFFT Z2Z/C2C).
 * Lengths: every Ng in 2..4096 (cuFFT's domain up to that length).  Every
 * 2^a*3^b*5^c length runs as a Stockham FFT (fixed schedules for powers of
 * two and short mixed lengths, the radix schedule in the kernel arguments for
 * the rest); lengths with other prime factors as a direct O(Ng^2) DFT per
 * line, or, once Ng^2 >= 108 M (M the power of two >= 2Ng - 1, where it is
 * measured faster), as the chirp-z convolution described below.  Longer
 * lines, up to 4096^2, run as four steps over Ng = N1 * N2
 * (both factors <= 4096): N1-point FFTs in place, a twiddle-and-transpose
 * into a workspace, N2-point FFTs there, a copy back.  The workspace (at most
 * 256 MiB, lines are processed in chunks) comes from the stream-ordered
 * allocator on `stream` and is released in stream order; square lengths
 * (N1 = N2) transpose in place and take none.  Lines above 4096 with no such
 * split (a prime factor above 4096), up to 2^23, run as Bluestein's chirp-z
 * convolution over two power-of-two FFTs of length M >= 2Ng - 1, with a
 * workspace from the same allocator; their chirp and filter tables are built
 * on the first call for (Ng, direction), which synchronizes `stream` once.
 * Longer lines with no split return A2AFFT_E_UNSUPPORTED. */
int a2afft_forward_1d_fft(void* data, int Ng, int64_t n_elems, a2afft_dtype dtype, void* stream);
int a2afft_inverse_1d_fft(void* data, int Ng, int64_t n_elems, a2afft_dtype dtype, void* stream);
/* launch_scale_fft (a2affts.cu:78-92,108-116): out = data / Ng^3 (64-bit Ng^3). */
int a2afft_launch_scale_fft(void* out, const void* data, int Ng, int64_t n_elems,
                            int blockSize, a2afft_dtype dtype, void* stream);
/* launch_fast_copy_fft (a2affts.cu:94-106,118-126): out = data. */
int a2afft_launch_fast_copy_fft(void* out, const void* data, int64_t n_elems,
                                int blockSize, a2afft_dtype dtype, void* stream);

/* ---- plan / execute: a2adfft -------------------------------------------- */
/* a2adfft::a2adfft(dist) (a2adfft.cpp:7-16) plus the precision and
 * execution strategy.  Builds twiddle tables and kernel choices. */
int a2afft_plan_create(a2afft_plan* out, a2afft_dist dist, a2afft_precision precision,
                       a2afft_exec exec);
/* A fused plan over one nx x ny x nz grid in C order (z fastest), without a
 * rank decomposition; no reference counterpart (the reference is cubic, one
 * grid per distribution, a2adistribution.h:9).  Each extent a transform
 * length of forward_1d_fft; complex128 for FP64, complex64 otherwise.  make_plans is not
 * needed; forward / inverse / autotune / execute_* apply as for other fused
 * plans, and the inverse scales by 1/(nx*ny*nz).  Up to 2^40 points (addresses
 * are 64-bit; 2048^3 complex64 takes 64 GiB); A2AFFT_E_SIZE beyond, or when
 * one pass would need 2^32 lines or tiles. */
int a2afft_grid_plan_create(a2afft_plan* out, int nx, int ny, int nz, a2afft_precision precision);
/* a2adfft::make_plans(d_scratch) (a2adfft.cpp:18-24).  scratch is a
 * [P][nlocal] device buffer; the staged strategy needs it, the fused one
 * accepts NULL. */
int a2afft_make_plans(a2afft_plan plan, void* d_scratch);
/* a2adfft::forward / inverse (a2adfft.cpp:26-158): in place on d_data in the
 * block layout; inverse applies the 1/Ng^3 of scale_fft. */
int a2afft_forward(a2afft_plan plan, void* d_data, void* stream);
int a2afft_inverse(a2afft_plan plan, void* d_data, void* stream);
/* forward (direction -1) or inverse (+1) as a hipGraph replay (no reference
 * counterpart: launch overhead matters for small grids).  The first call for
 * a buffer and direction captures the plan's launches on an internal stream;
 * later calls replay it on `stream`.  A call for another buffer, or after
 * make_plans / set_kernel / set_chunking / autotune, captures again.  Same
 * bits as forward / inverse. */
int a2afft_graph_launch(a2afft_plan plan, void* d_data, int direction, void* stream);
int a2afft_plan_destroy(a2afft_plan plan);

/* ---- batched grids (no reference counterpart) ---------------------------- */
/* Many independent C-order grids of uneven shapes, complex64, forward or
 * inverse in ONE persistent launch (the reference transforms one cubic grid
 * per distribution, a2adistribution.h:9).  shapes: ngrids x {nx, ny, nz},
 * powers of two 16..512.  The grids sit back to back in one buffer; offsets
 * (ngrids + 1 entries, in complex elements) says where.  execute is
 * stream-ordered; graph_launch captures the same work into a hipGraph on first
 * use for a given buffer and direction and replays it.  status synchronises
 * and reports whether every dependency wait finished within its bound. */
typedef struct a2afft_batch_s* a2afft_batch;
int a2afft_batch_create(a2afft_batch* out, int ngrids, const int* shapes, a2afft_precision precision);
/* The same with a virtual-rank decomposition per grid (dims: ngrids x {d0, d1,
 * d2}, NULL = all {1,1,1}): grid g is stored as d0*d1*d2 rank blocks back to
 * back, each (nx/d0) x (ny/d1) x (nz/d2) in C order, rank-major -- the layout
 * of a distribution's [P][nlocal] buffer (a2adistribution.cpp:14-42, 59-69)
 * -- and its passes address the blocks directly, as a fused plan over that
 * distribution does (no separate all-to-all or pack). */
int a2afft_batch_create_ranks(a2afft_batch* out, int ngrids, const int* shapes, const int* dims,
                              a2afft_precision precision);
int a2afft_batch_offsets(a2afft_batch batch, int64_t* offsets);
int a2afft_batch_execute(a2afft_batch batch, void* d_data, int direction, void* stream);
int a2afft_batch_graph_launch(a2afft_batch batch, void* d_data, int direction, void* stream);
int a2afft_batch_status(a2afft_batch batch, void* stream);
/* Schedule of the persistent launch: workgroups (0 = every resident slot on
 * every CU); queue layout (1 = each pass's units split into one contiguous
 * range per XCD, 0 = one global queue, the default); units handed out per
 * take of the queue (1..64).  Results are identical for every schedule. */
int a2afft_batch_set_schedule(a2afft_batch batch, int workgroups, int xcd_queues, int chunk);
/* Order of the work: grids are packed, in order, into groups of at most
 * group_points points (0 = one group of all grids, the default: each pass of
 * every grid before the next pass; a grid larger than group_points forms a
 * group of its own).  A group's next pass waits only for the same group's
 * previous pass, so a group that fits the device's memory-side cache re-reads
 * its data from there.  skewed = 1 interleaves the (group, pass) phases along
 * diagonals, so that another group's work runs while a pass completes.  At
 * most 128 groups.  Results are identical for every order. */
int a2afft_batch_set_grouping(a2afft_batch batch, int64_t group_points, int skewed);
/* Times `iters` forward executions per candidate on `stream` and keeps the
 * fastest schedule, order and tile shapes: workgroup counts x queue layouts x
 * takes of 1, 2, 4 units (pass-major), then skewed groups of about one and
 * two of the largest grids, then per (length, strided) class the tile shapes
 * (elements per thread, LDS padding).  d_data (the batch buffer) is
 * overwritten.  Synchronises. */
int a2afft_batch_autotune(a2afft_batch batch, void* d_data, int iters, void* stream);
/* Tile shape of one (length N, strided 0/1) class of passes: T lines or
 * columns per unit, E elements per thread, LDS padding (the kernel holds a
 * primary shape per class, the same unpadded, and half the tile); passes it
 * does not fit keep the primary.  T = 0 restores the primary.  Results are
 * identical for every shape. */
int a2afft_batch_set_shape(a2afft_batch batch, int N, int strided, int T, int E, int pad);
/* First line: the current schedule as key=value pairs; second line: the tile
 * shape of every (length, strided) class; then the last autotune's timings. */
int a2afft_batch_describe(a2afft_batch batch, char* buf, int len);
int a2afft_batch_info(a2afft_batch batch, int64_t* total_units, int* workgroups, int* max_workgroups);
int a2afft_batch_destroy(a2afft_batch batch);

/* ---- measurement helpers (no reference counterpart) --------------------- */
/* rel_l2 = ||y - ref||_2 / ||ref||_2 and max_abs = max |y - ref| over n
 * complex values, reduced on the device in FP64.  y has dtype; ref is
 * complex128 on the device.  Blocks until the result is on the host. */
int a2afft_error_metrics(const void* y, a2afft_dtype dtype, const void* ref, int64_t n,
                         double* rel_l2, double* max_abs, void* stream);

/* a2afft_error_metrics without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc).  The partial sums live in a
 * fixed 24 KiB device array of the library (one per GPU, held until the
 * process ends) instead of a per-call allocation; the result is the same,
 * bit for bit.  n <= 2^36, else A2AFFT_E_SIZE.  Calls that overlap on one
 * device race on that array.  It synchronizes `stream`, so it cannot be
 * recorded into a graph. */
int a2afft_error_metrics_noalloc(const void* y, a2afft_dtype dtype, const void* ref, int64_t n,
                                 double* rel_l2, double* max_abs, void* stream);

/* One axis pass of a fused plan, exactly as a2afft_forward (direction -1) or
 * a2afft_inverse (+1) runs it: pass 0 = z, 1 = x, 2 = y (the reference's axis
 * order, a2adfft.cpp:31-85); the inverse's 1/Ng^3 rides on pass 2.  For
 * per-kernel timing; the three passes in order equal one transform. */
int a2afft_execute_pass(a2afft_plan plan, void* d_data, int pass, int direction, void* stream);

/* Kernel variants (no reference counterpart: cuFFT chose its own kernels).
 * Every variant of a pass computes identical bits; they differ in tile width
 * T, elements per thread E and flags: 1 XCD-aware workgroup order, 2 padded
 * LDS, 4 non-temporal loads/stores, 8 z-tiles-outermost unit order, 32 LDS
 * exchanges one component at a time (half the LDS per workgroup).  Kernel
 * families, which compute different bits from one another: the fixed-schedule
 * Stockham shapes (no family flag), 128 runtime radix schedule, 4096
 * butterflies mapped per pass (E = data registers a thread), 2048 chirp-z,
 * 64 direct DFT; autotune stays within the family of each pass's selection.
 * autotune times each built variant that fits each fused pass, `iters`
 * launches in place on `work` (a [P][nlocal] buffer whose contents are
 * destroyed), keeps the fastest per pass and then the fastest chunk width,
 * and keeps the result only if whole forwards beat the choices it started
 * from; set_kernel picks one by hand. */
int a2afft_plan_autotune(a2afft_plan plan, void* work, int iters, void* stream);
/* Process-wide tuning cache file (NULL: none); no reference counterpart (the
 * reference re-plans on every call, a2affts.cu:15).  a2afft_plan_autotune
 * records its plan's kernels and chunk width under a key of device
 * architecture, grid, decomposition, data type, precision and strategy;
 * fused plans created later with the same key start from that choice when
 * every recorded kernel is built and fits.  The file is text, one line per
 * key, rewritten through a temporary file and a rename. */
int a2afft_set_tuning_cache(const char* path);
int a2afft_plan_set_kernel(a2afft_plan plan, int pass, int T, int E, int flags);
/* Chunked execution: after the z pass, the x and y passes alternate over
 * windows of `z_columns` z columns (0 = whole passes), so each window is
 * re-read from the Infinity Cache.  Results are identical for every width;
 * the width must divide Ng and be a multiple of both strided tile widths.
 * A new plan starts at 0, or, for grids above 512 MiB whose x and y passes
 * are 512-point lines on half-width-exchange tiles (for complex128 the x pass
 * alone), at the widest window of at most 256 MiB.  autotune also times
 * whole forwards over the valid widths. */
int a2afft_plan_set_chunking(a2afft_plan plan, int z_columns);
/* Launches in one fused forward / inverse, and launch `index` alone (returns
 * its pass: 0 z, 1 x, 2 y), for per-kernel timing. */
int a2afft_plan_num_launches(a2afft_plan plan);
int a2afft_execute_launch(a2afft_plan plan, void* d_data, int index, int direction, void* stream);
/* Writes the chosen variant per pass and the last autotune timings as text. */
int a2afft_plan_describe(a2afft_plan plan, char* buf, int len);
/* Pattern copy of fused pass `pass` (0 = z, 1 = x, 2 = y): the mean
 * milliseconds over `iters` launches of a copy kernel that reads and writes
 * every element of the pass in place with the pass kernel's addresses, tile
 * order, cache policy and LDS footprint, but no arithmetic; `data` is left
 * unchanged.  Fused FP32/FP64 plans whose pass runs a fixed-schedule Stockham
 * kernel; A2AFFT_E_UNSUPPORTED otherwise.  No counterpart in the reference: a
 * measurement aid that sets each pass's time beside the copy time of its own
 * access pattern. */
int a2afft_plan_pattern_copy(a2afft_plan plan, void* d_data, int pass, int iters, float* ms, void* stream);

/* Seeded synthetic input (no reference counterpart: the reference's test
 * fills with an unseeded rand(), tests/test.cpp:12-17).  Philox4x32-10 with
 * key (seed lo, seed hi) and counter (g lo, g hi, kind, 0) for global index
 * g = g0 + i; kind 0: re, im uniform in [-1, 1); kind 1: standard normal by
 * Box-Muller.  Writes n complex values of dtype. */
int a2afft_fill_random(void* out, a2afft_dtype dtype, int64_t n, uint64_t seed, int kind, int64_t g0,
                       void* stream);

#ifdef __cplusplus
}
#endif
#endif /* A2AFFT_H */
