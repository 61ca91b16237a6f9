/*
 * a2afft oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C (C99 + OpenMP) restatement of the reference a2afft dataflow
 * (humz2k/3D-FFTs-MPI, src/a2afft/), used as the checker for the HIP product
 * library.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.  The product path never calls into this library.
 *
 * Every function cites the reference file:line it restates.  How each one is
 * pinned (golden vectors from the compiled reference C twins, closed forms,
 * brute force, textbook identities) is listed in tests/test_oracle_*.py.
 *
 * Conventions
 *   - P virtual ranks; each owns `nlocal` complex values; a "buffer" is
 *     [P][nlocal] complex128 (interleaved re,im doubles) unless a function
 *     says it moves raw elements of `esize` bytes.
 *   - All indices are int64 (the reference uses int32; values agree wherever
 *     the reference does not overflow).
 */
#ifndef A2AFFT_ORACLE_H
#define A2AFFT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int Ng;                  /* a2adistribution.h:9 */
    int P;                   /* world_size, a2adistribution.h:11 */
    int dims[3];             /* a2adistribution.h:14 */
    int lgs[3];              /* local_grid_size, a2adistribution.h:13 */
    int64_t nlocal;          /* a2adistribution.h:10 */
} orc_decomp;

/* Remap kinds: the 9 index kernels of src/a2afft/distribution/kernels/. */
enum {
    ORC_Z_A2A_TO_Z_PENCILS = 0,   /* a2apencils.cu:4-21 */
    ORC_Z_PENCILS_TO_Z_A2A = 1,   /* a2apencils.cu:37-56 */
    ORC_X_A2A_TO_X_PENCILS = 2,   /* a2apencils.cu:73-92 */
    ORC_X_PENCILS_TO_X_A2A = 3,   /* a2apencils.cu:110-130 */
    ORC_Y_A2A_TO_Y_PENCILS = 4,   /* a2apencils.cu:148-168 */
    ORC_Y_PENCILS_TO_Y_A2A = 5,   /* a2apencils.cu:188-208 */
    ORC_FAST_Z_TO_X = 6,          /* a2achangefastaxis.cu:3-22 */
    ORC_FAST_X_TO_Y = 7,          /* a2achangefastaxis.cu:34-53 */
    ORC_FAST_Y_TO_Z = 8           /* a2achangefastaxis.cu:65-84 */
};

/* Stage kinds: a2adistribution.cpp:115-215 */
enum {
    ORC_GET_Z = 0, ORC_RETURN_Z = 1, ORC_GET_X = 2,
    ORC_RETURN_X = 3, ORC_GET_Y = 4, ORC_RETURN_Y = 5
};

/* --- decomposition (a2adistribution.cpp:14-80) --------------------------- */
void orc_dims_create(int P, int dims[3]);
void orc_rank2coords(int rank, const int dims[3], int out[3]);
/* Returns 0 on success, or a negative code naming the failed precondition
 * (-1 Ng%d, -2 P%d, -3 (lx*ly)%P, -4 (lx*lz)%P, -5 (ly*lz)%P, -6 bad args). */
int orc_decompose(int Ng, int P, const int* dims_or_null, orc_decomp* d);
/* 1 when dims lie inside the domain where the reference's y-return formula
 * (a2apencils.cu:197) is the exact inverse of y_a2a_to_y_pencils. */
int orc_y_return_in_reference_domain(const orc_decomp* d);

/* --- remaps (index form, one rank) ---------------------------------------- */
/* For the six pack/unpack kernels (scatters: dest[new]=src[idx]) returns the
 * destination of source index idx.  For the three transposes (gathers:
 * dest[idx]=src[f(idx)]) returns the SOURCE index of destination idx.
 * ref_formula: for ORC_Y_PENCILS_TO_Y_A2A only, 1 = the reference's rank
 * expression exactly as written (a2apencils.cu:197, may leave [0,nlocal)),
 * 0 = the true inverse of ORC_Y_A2A_TO_Y_PENCILS. */
int64_t orc_remap_index(int kind, int64_t idx, const orc_decomp* d, int ref_formula);
/* 1 if `kind` is a gather (transpose), 0 if a scatter. */
int orc_remap_is_gather(int kind);
/* Applies a remap to every virtual rank: src,dst are [P][nlocal] elements of
 * esize bytes.  Uses the true inverse for y-return. */
void orc_remap_apply(int kind, const orc_decomp* d, const void* src, void* dst, int esize);
/* MPI_Alltoall over P virtual ranks with equal chunks (a2adistribution.cpp:117-125):
 * recv[s][q*ns+i] = send[q][s*ns+i], ns = nlocal/P. */
void orc_alltoall(const orc_decomp* d, const void* send, void* recv, int esize);
/* Stage compositions (a2adistribution.cpp:115-215) on complex buffers of esize bytes. */
void orc_stage(int stage, const orc_decomp* d, void* b1, void* b2, int esize);

/* --- 1D FFT (the published algorithm behind dfft/kernels/a2affts.cu:7-76) -- */
/* In-place batched complex128 DFT of `nlines` contiguous lines of length N:
 * X[k] = sum_n x[n] exp(sign*2*pi*i*n*k/N), unnormalised.  Mixed-radix
 * Stockham over the factors of N with long-double twiddles; a length with a
 * prime factor above 64 takes orc_dft_naive line by line. */
int orc_fft_batched(double* data, int N, int64_t nlines, int sign);
/* O(N^2) long-double DFT of one line (tests pin orc_fft_batched with it). */
void orc_dft_naive(const double* in, double* out, int N, int sign);
/* Twiddle exp(sign*2*pi*i*m/M) in long double, rounded once to double. */
void orc_twiddle(int64_t m, int64_t M, int sign, double* re, double* im);

/* --- 3D transform in the reference stage order (a2adfft.cpp:26-158) ------- */
/* D, S: [P][nlocal] complex128. Result in D (block layout), S clobbered. */
int orc_forward(const orc_decomp* d, double* D, double* S);
/* Inverse with the 1/Ng^3 scaling of a2affts.cu:78-92 (64-bit Ng^3). */
int orc_inverse(const orc_decomp* d, double* D, double* S);

/* --- layout helpers (tests/stitch.py:28-47, helpers/helpers.c:37-55) ------ */
void orc_blocks_to_global(const orc_decomp* d, const void* blocks, void* global, int esize);
void orc_global_to_blocks(const orc_decomp* d, const void* global, void* blocks, int esize);

/* --- counter-based input generator (Philox4x32-10, Salmon et al. SC'11) --- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Global C-order grid of n complex128 values, re/im i.i.d. U[-1,1) keyed by
 * (seed, global index g). */
void orc_fill_uniform(uint64_t seed, int64_t n, double* out);
/* Same, standard normal by Box-Muller. */
void orc_fill_normal(uint64_t seed, int64_t n, double* out);

/* Thread count used by the OpenMP loops (0 = runtime default). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
