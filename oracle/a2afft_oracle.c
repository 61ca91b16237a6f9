/*
 * a2afft oracle -- TEST INFRASTRUCTURE ONLY (see a2afft_oracle.h).
 *
 * Plain C restatement of the reference a2afft dataflow.  Reference paths are
 * relative to the humz2k/3D-FFTs-MPI source tree.  Nothing in the product
 * library links or loads this file; it is the checker.
 */
#include "a2afft_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 0;

void orc_set_threads(int n) {
    g_threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Decomposition                                                            */
/* ------------------------------------------------------------------------ */

/* MPI_Dims_create(P, 3, dims) with all dims free (a2adistribution.cpp:59).
 * MPI-standard rule: dims as close to each other as possible, non-increasing.
 * Restated as: prime factors of P taken largest first, each multiplied into
 * the currently smallest dimension, then sorted non-increasing.  Pinned by
 * the table 4->[2,2,1], 6->[3,2,1], 8->[2,2,2], 12->[3,2,2], 16->[4,2,2],
 * 64->[4,4,4] in the tests. */
void orc_dims_create(int P, int dims[3]) {
    int f[64], nf = 0, p = P;
    dims[0] = dims[1] = dims[2] = 1;
    if (P <= 0) return;
    for (int q = 2; (long)q * q <= p; ++q)
        while (p % q == 0) { f[nf++] = q; p /= q; }
    if (p > 1) f[nf++] = p;
    for (int i = nf - 1; i >= 0; --i) {          /* factors ascend; walk down */
        int m = 0;
        for (int a = 1; a < 3; ++a)
            if (dims[a] < dims[m]) m = a;
        dims[m] *= f[i];
    }
    for (int a = 0; a < 3; ++a)                   /* sort non-increasing */
        for (int b = a + 1; b < 3; ++b)
            if (dims[b] > dims[a]) { int t = dims[a]; dims[a] = dims[b]; dims[b] = t; }
}

/* a2adistribution.cpp:14-20 */
void orc_rank2coords(int rank, const int dims[3], int out[3]) {
    out[0] = rank / (dims[1] * dims[2]);
    out[1] = (rank - out[0] * (dims[1] * dims[2])) / dims[2];
    out[2] = rank - out[0] * (dims[1] * dims[2]) - out[1] * dims[2];
}

/* a2adistribution.cpp:44-80: dims, local grid (topology2localgrid :25-31),
 * nlocal (:69) and the asserted preconditions (:61-63, :74-80). */
int orc_decompose(int Ng, int P, const int* dims_or_null, orc_decomp* d) {
    if (Ng <= 0 || P <= 0 || !d) return -6;
    d->Ng = Ng;
    d->P = P;
    if (dims_or_null) {
        for (int a = 0; a < 3; ++a) d->dims[a] = dims_or_null[a];
        if (d->dims[0] <= 0 || d->dims[1] <= 0 || d->dims[2] <= 0) return -6;
        if ((long)d->dims[0] * d->dims[1] * d->dims[2] != P) return -6;
    } else {
        orc_dims_create(P, d->dims);
    }
    for (int a = 0; a < 3; ++a)
        if (Ng % d->dims[a]) return -1;
    for (int a = 0; a < 3; ++a) d->lgs[a] = Ng / d->dims[a];
    d->nlocal = (int64_t)d->lgs[0] * d->lgs[1] * d->lgs[2];
    for (int a = 0; a < 3; ++a)
        if (P % d->dims[a]) return -2;
    if (((int64_t)d->lgs[0] * d->lgs[1]) % P) return -3;
    if (((int64_t)d->lgs[0] * d->lgs[2]) % P) return -4;
    if (((int64_t)d->lgs[1] * d->lgs[2]) % P) return -5;
    return 0;
}

/* The reference's y-return rank expression (a2apencils.cu:197) equals the
 * true inverse exactly when dims[1]==2 or dims[0]==1. */
int orc_y_return_in_reference_domain(const orc_decomp* d) {
    return d->dims[1] == 2 || d->dims[0] == 1;
}

/* ------------------------------------------------------------------------ */
/* Remaps                                                                   */
/* ------------------------------------------------------------------------ */

int orc_remap_is_gather(int kind) {
    return kind == ORC_FAST_Z_TO_X || kind == ORC_FAST_X_TO_Y || kind == ORC_FAST_Y_TO_Z;
}

int64_t orc_remap_index(int kind, int64_t idx, const orc_decomp* d, int ref_formula) {
    const int64_t P = d->P, nlocal = d->nlocal;
    const int64_t lx = d->lgs[0], ly = d->lgs[1], lz = d->lgs[2];
    const int64_t cpr = nlocal / P;                 /* n_cells_per_rank, a2apencils.cu:27 */
    switch (kind) {
    case ORC_Z_A2A_TO_Z_PENCILS: {                  /* a2apencils.cu:4-35 */
        const int64_t L = lz, mpr = cpr / L, st = d->dims[2];
        int64_t rank = idx / cpr;
        int64_t mp = (idx / L) % mpr;
        int64_t mpo = (rank % st) * L;
        int64_t stack = idx / (st * cpr);
        int64_t off = stack * st * cpr;
        return off + mpo + mp * st * L + idx % L;
    }
    case ORC_Z_PENCILS_TO_Z_A2A: {                  /* a2apencils.cu:37-71 */
        const int64_t L = lz, mpr = cpr / L, st = d->dims[2];
        int64_t stack = idx / (st * cpr);
        int64_t mp = (idx / L) % st;
        int64_t rank = stack * st + mp;
        int64_t pid = (idx / (L * st)) % mpr;
        return rank * cpr + pid * L + idx % L;
    }
    case ORC_X_A2A_TO_X_PENCILS: {                  /* a2apencils.cu:73-108 */
        const int64_t L = lx, st = d->dims[0], cps = nlocal / st;
        int64_t mp = (idx / L) % (cps / L);
        int64_t stack = idx / cps;
        return stack * L + mp * st * L + idx % L;
    }
    case ORC_X_PENCILS_TO_X_A2A: {                  /* a2apencils.cu:110-146 */
        const int64_t L = lx, mpr = cpr / L, st = d->dims[0], cps = nlocal / st;
        int64_t mp = idx / L;
        int64_t stack = mp % st;
        int64_t rank = idx / (cpr * st) + stack * (cps / cpr);
        int64_t pid = (idx / (st * L)) % mpr;
        return rank * cpr + pid * L + idx % L;
    }
    case ORC_Y_A2A_TO_Y_PENCILS: {                  /* a2apencils.cu:148-186 */
        const int64_t L = ly, mpr = cpr / L, st = d->dims[1];
        const int64_t rpt = (P / st) / d->dims[0];  /* n_ranks_per_tower, :179-180 */
        int64_t mp = (idx / L) % (rpt * mpr);
        int64_t stack = idx / (rpt * st * cpr);
        int64_t lst = (idx / (rpt * cpr)) % st;
        return mp * st * L + lst * L + stack * rpt * st * cpr + idx % L;
    }
    case ORC_Y_PENCILS_TO_Y_A2A: {                  /* a2apencils.cu:188-226 */
        const int64_t L = ly, mpr = cpr / L, st = d->dims[1];
        const int64_t rpt = (P / st) / d->dims[0];
        int64_t mp = ((idx / L) % st) * rpt;
        int64_t stack = (idx / cpr) / st;
        int64_t rank;
        if (ref_formula)
            rank = mp + stack + (stack / rpt) * rpt;          /* :197 as written */
        else
            rank = mp + stack % rpt + (stack / rpt) * rpt * st; /* true inverse */
        int64_t lst = (idx / (L * st)) % mpr;
        return rank * cpr + lst * L + idx % L;
    }
    case ORC_FAST_Z_TO_X: {                         /* a2achangefastaxis.cu:3-22 */
        int64_t i = idx / (lx * ly);
        int64_t j = (idx - i * lx * ly) / lx;
        int64_t k = idx - i * lx * ly - j * lx;
        return k * ly * lz + j * lz + i;
    }
    case ORC_FAST_X_TO_Y: {                         /* a2achangefastaxis.cu:34-53 */
        int64_t i = idx / (lz * ly);
        int64_t j = (idx - i * lz * ly) / ly;
        int64_t k = idx - i * lz * ly - j * ly;
        return j * lx * ly + k * lx + i;
    }
    case ORC_FAST_Y_TO_Z: {                         /* a2achangefastaxis.cu:65-84 */
        int64_t i = idx / (lz * ly);
        int64_t j = (idx - i * lz * ly) / lz;
        int64_t k = idx - i * lz * ly - j * lz;
        return i * lz * ly + k * ly + j;
    }
    default:
        return -1;
    }
}

void orc_remap_apply(int kind, const orc_decomp* d, const void* src, void* dst, int esize) {
    const int64_t nlocal = d->nlocal;
    const int gather = orc_remap_is_gather(kind);
    for (int r = 0; r < d->P; ++r) {
        const char* s = (const char*)src + (size_t)r * nlocal * esize;
        char* t = (char*)dst + (size_t)r * nlocal * esize;
        #pragma omp parallel for schedule(static)
        for (int64_t idx = 0; idx < nlocal; ++idx) {
            int64_t m = orc_remap_index(kind, idx, d, 0);
            if (gather)
                memcpy(t + idx * esize, s + m * esize, esize);
            else
                memcpy(t + m * esize, s + idx * esize, esize);
        }
    }
}

/* MPI_Alltoall with nsends = nlocal/P per peer (a2adistribution.cpp:117-125;
 * every stage's nsends reduces to nlocal/P, :117,133,149,167,183,201). */
void orc_alltoall(const orc_decomp* d, const void* send, void* recv, int esize) {
    const int64_t P = d->P, nlocal = d->nlocal, ns = nlocal / P;
    const size_t chunk = (size_t)ns * esize;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int64_t s = 0; s < P; ++s)
        for (int64_t q = 0; q < P; ++q)
            memcpy((char*)recv + ((size_t)s * nlocal + q * ns) * esize,
                   (const char*)send + ((size_t)q * nlocal + s * ns) * esize, chunk);
}

/* a2adistribution.cpp:115-215.  b1 = d_Buff1, b2 = d_Buff2. */
void orc_stage(int stage, const orc_decomp* d, void* b1, void* b2, int esize) {
    switch (stage) {
    case ORC_GET_Z:        /* :115-129 */
        orc_alltoall(d, b1, b2, esize);
        orc_remap_apply(ORC_Z_A2A_TO_Z_PENCILS, d, b2, b1, esize);
        break;
    case ORC_RETURN_Z:     /* :131-145 */
        orc_remap_apply(ORC_Z_PENCILS_TO_Z_A2A, d, b1, b2, esize);
        orc_alltoall(d, b2, b1, esize);
        break;
    case ORC_GET_X:        /* :147-163 */
        orc_remap_apply(ORC_FAST_Z_TO_X, d, b1, b2, esize);
        orc_alltoall(d, b2, b1, esize);
        orc_remap_apply(ORC_X_A2A_TO_X_PENCILS, d, b1, b2, esize);
        break;
    case ORC_RETURN_X:     /* :165-179 */
        orc_remap_apply(ORC_X_PENCILS_TO_X_A2A, d, b2, b1, esize);
        orc_alltoall(d, b1, b2, esize);
        break;
    case ORC_GET_Y:        /* :181-197 */
        orc_remap_apply(ORC_FAST_X_TO_Y, d, b2, b1, esize);
        orc_alltoall(d, b1, b2, esize);
        orc_remap_apply(ORC_Y_A2A_TO_Y_PENCILS, d, b2, b1, esize);
        break;
    case ORC_RETURN_Y:     /* :199-215 */
        orc_remap_apply(ORC_Y_PENCILS_TO_Y_A2A, d, b2, b1, esize);
        orc_alltoall(d, b1, b2, esize);
        orc_remap_apply(ORC_FAST_Y_TO_Z, d, b2, b1, esize);
        break;
    default:
        break;
    }
}

/* ------------------------------------------------------------------------ */
/* 1D FFT                                                                   */
/* ------------------------------------------------------------------------ */

/* exp(sign*2*pi*i*m/M): reduce to the nearest quadrant in integers so the
 * angle evaluated in long double lies in [-pi/4, pi/4], then rotate exactly.
 * Quadrant points come out as exact 0/+-1. */
void orc_twiddle(int64_t m, int64_t M, int sign, double* re, double* im) {
    m %= M;
    if (m < 0) m += M;
    /* k = nearest quadrant index: round(4m/M) */
    int64_t k = (4 * m + M / 2) / M;
    int64_t num = 4 * m - k * M;                   /* angle = 2*pi*num/(4M) */
    long double a = (long double)num * 3.14159265358979323846264338327950288L /
                    (2.0L * (long double)M);
    long double c = cosl(a), s = sinl(a);
    if (num == 0) { c = 1.0L; s = 0.0L; }
    long double cr, sr;
    switch ((int)(k & 3)) {
    case 0: cr = c; sr = s; break;
    case 1: cr = -s; sr = c; break;
    case 2: cr = -c; sr = -s; break;
    default: cr = s; sr = -c; break;
    }
    *re = (double)cr;
    *im = (double)(sign < 0 ? -sr : sr);
    if (*im == 0.0) *im = 0.0;                     /* no negative zero */
    if (*re == 0.0) *re = 0.0;
}

void orc_dft_naive(const double* in, double* out, int N, int sign) {
    for (int k = 0; k < N; ++k) {
        long double sr = 0.0L, si = 0.0L;
        for (int n = 0; n < N; ++n) {
            double wr, wi;
            orc_twiddle((int64_t)n * k % N, N, sign, &wr, &wi);
            long double xr = in[2 * n], xi = in[2 * n + 1];
            sr += xr * wr - xi * wi;
            si += xr * wi + xi * wr;
        }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
}

static int factorize(int N, int* rad) {
    int n = 0, m = N;
    while (m % 4 == 0) { rad[n++] = 4; m /= 4; }
    while (m % 2 == 0) { rad[n++] = 2; m /= 2; }
    for (int p = 3; m > 1; p += 2)
        while (m % p == 0) { rad[n++] = p; m /= p; }
    return n;
}

/* Stockham autosort, decimation in time: pass with radix R and sub-length Ns
 * reads a[j + r*N/R], twiddles by w_{Ns*R}^{r*(j mod Ns)}, takes an R-point
 * DFT and writes b[(j/Ns)*Ns*R + (j mod Ns) + q*Ns]. */
static void fft_line(double* a, double* b, int N, const double* tw, const int* rad, int nrad) {
    double vr[64], vi[64];
    int Ns = 1;
    double* src = a;
    double* dst = b;
    for (int p = 0; p < nrad; ++p) {
        const int R = rad[p];
        const int stride = N / R;
        for (int j = 0; j < stride; ++j) {
            const int k = j % Ns;
            for (int r = 0; r < R; ++r) {
                double xr = src[2 * (j + r * stride)], xi = src[2 * (j + r * stride) + 1];
                int64_t m = (int64_t)r * k * (N / (Ns * R));
                double wr = tw[2 * (m % N)], wi = tw[2 * (m % N) + 1];
                vr[r] = xr * wr - xi * wi;
                vi[r] = xr * wi + xi * wr;
            }
            const int base = (j / Ns) * Ns * R + k;
            for (int q = 0; q < R; ++q) {
                double sr = 0.0, si = 0.0;
                for (int r = 0; r < R; ++r) {
                    int64_t m = (int64_t)((r * q) % R) * stride;
                    double wr = tw[2 * m], wi = tw[2 * m + 1];
                    sr += vr[r] * wr - vi[r] * wi;
                    si += vr[r] * wi + vi[r] * wr;
                }
                dst[2 * (base + q * Ns)] = sr;
                dst[2 * (base + q * Ns) + 1] = si;
            }
        }
        Ns *= R;
        double* t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, sizeof(double) * 2 * N);
}

int orc_fft_batched(double* data, int N, int64_t nlines, int sign) {
    if (N <= 0) return -1;
    int rad[64];
    int nrad = factorize(N, rad);
    int naive = 0;
    for (int p = 0; p < nrad; ++p)
        if (rad[p] > 64) naive = 1;               /* a prime factor above the butterfly scratch bound */
    if (naive) {                                  /* the textbook sum, line by line */
        int err = 0;
        #pragma omp parallel
        {
            double* out = (double*)malloc(sizeof(double) * 2 * N);
            if (!out) {
                #pragma omp atomic write
                err = -3;
            } else {
                #pragma omp for schedule(static)
                for (int64_t l = 0; l < nlines; ++l) {
                    orc_dft_naive(data + 2 * l * (int64_t)N, out, N, sign);
                    memcpy(data + 2 * l * (int64_t)N, out, sizeof(double) * 2 * N);
                }
                free(out);
            }
        }
        return err;
    }
    double* tw = (double*)malloc(sizeof(double) * 2 * N);
    if (!tw) return -3;
    for (int m = 0; m < N; ++m) orc_twiddle(m, N, sign, &tw[2 * m], &tw[2 * m + 1]);
    int err = 0;
    #pragma omp parallel
    {
        double* scratch = (double*)malloc(sizeof(double) * 2 * N);
        if (!scratch) {
            #pragma omp atomic write
            err = -3;
        } else {
            #pragma omp for schedule(static)
            for (int64_t l = 0; l < nlines; ++l)
                fft_line(data + 2 * l * (int64_t)N, scratch, N, tw, rad, nrad);
            free(scratch);
        }
    }
    free(tw);
    return err;
}

/* ------------------------------------------------------------------------ */
/* 3D transform, reference stage order                                      */
/* ------------------------------------------------------------------------ */

/* a2adfft.cpp:26-91 with forward_1d_fft (a2affts.cu:7-40: nlocal/Ng lines
 * per rank, in place) and fast_copy_fft (a2affts.cu:94-106). */
int orc_forward(const orc_decomp* d, double* D, double* S) {
    const int64_t total = (int64_t)d->P * d->nlocal, lines = total / d->Ng;
    const int esz = 16;
    int e = 0;
    orc_stage(ORC_GET_Z, d, D, S, esz);
    e |= orc_fft_batched(D, d->Ng, lines, -1);
    orc_stage(ORC_RETURN_Z, d, D, S, esz);
    orc_stage(ORC_GET_X, d, D, S, esz);
    e |= orc_fft_batched(S, d->Ng, lines, -1);
    orc_stage(ORC_RETURN_X, d, D, S, esz);
    orc_stage(ORC_GET_Y, d, D, S, esz);
    e |= orc_fft_batched(D, d->Ng, lines, -1);
    memcpy(S, D, sizeof(double) * 2 * total);      /* launch_fast_copy_fft(S, D) */
    orc_stage(ORC_RETURN_Y, d, D, S, esz);
    return e;
}

/* a2adfft.cpp:93-158; scale_fft (a2affts.cu:78-92) divides by Ng^3,
 * evaluated here in 64-bit so it does not overflow for Ng >= 1291. */
int orc_inverse(const orc_decomp* d, double* D, double* S) {
    const int64_t total = (int64_t)d->P * d->nlocal, lines = total / d->Ng;
    const int esz = 16;
    int e = 0;
    orc_stage(ORC_GET_Z, d, D, S, esz);
    e |= orc_fft_batched(D, d->Ng, lines, +1);
    orc_stage(ORC_RETURN_Z, d, D, S, esz);
    orc_stage(ORC_GET_X, d, D, S, esz);
    e |= orc_fft_batched(S, d->Ng, lines, +1);
    orc_stage(ORC_RETURN_X, d, D, S, esz);
    orc_stage(ORC_GET_Y, d, D, S, esz);
    e |= orc_fft_batched(D, d->Ng, lines, +1);
    const double scale = (double)((int64_t)d->Ng * d->Ng * d->Ng);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 2 * total; ++i) S[i] = D[i] / scale;
    orc_stage(ORC_RETURN_Y, d, D, S, esz);
    return e;
}

/* ------------------------------------------------------------------------ */
/* Layout                                                                   */
/* ------------------------------------------------------------------------ */

/* Block r holds [lx][ly][lz] at global origin coords(r)*lgs
 * (a2adistribution.cpp:36-42, tests/stitch.py:28-47). */
static void blocks_global(const orc_decomp* d, void* blocks, void* global, int esize, int to_global) {
    const int64_t Ng = d->Ng, lx = d->lgs[0], ly = d->lgs[1], lz = d->lgs[2];
    for (int r = 0; r < d->P; ++r) {
        int c[3];
        orc_rank2coords(r, d->dims, c);
        char* b = (char*)blocks + (size_t)r * d->nlocal * esize;
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < lx; ++i)
            for (int64_t j = 0; j < ly; ++j) {
                char* g = (char*)global +
                    ((size_t)((c[0] * lx + i) * Ng + c[1] * ly + j) * Ng + c[2] * lz) * esize;
                char* bl = b + (size_t)((i * ly + j) * lz) * esize;
                if (to_global) memcpy(g, bl, (size_t)lz * esize);
                else memcpy(bl, g, (size_t)lz * esize);
            }
    }
}

void orc_blocks_to_global(const orc_decomp* d, const void* blocks, void* global, int esize) {
    blocks_global(d, (void*)blocks, global, esize, 1);
}

void orc_global_to_blocks(const orc_decomp* d, const void* global, void* blocks, int esize) {
    blocks_global(d, blocks, (void*)global, esize, 0);
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10                                                            */
/* ------------------------------------------------------------------------ */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void philox_at(uint64_t seed, int64_t g, uint32_t stream, uint32_t out[4]) {
    uint32_t ctr[4] = {(uint32_t)g, (uint32_t)((uint64_t)g >> 32), stream, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

/* 53 random bits from two words: ((hi<<32)|lo) >> 11. */
static inline uint64_t bits53(uint32_t hi, uint32_t lo) {
    return (((uint64_t)hi << 32) | lo) >> 11;
}

void orc_fill_uniform(uint64_t seed, int64_t n, double* out) {
    const double s = 1.0 / 4503599627370496.0;   /* 2^-52 */
    #pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < n; ++g) {
        uint32_t w[4];
        philox_at(seed, g, 0u, w);
        out[2 * g] = (double)bits53(w[0], w[1]) * s - 1.0;
        out[2 * g + 1] = (double)bits53(w[2], w[3]) * s - 1.0;
    }
}

void orc_fill_normal(uint64_t seed, int64_t n, double* out) {
    const double s = 1.0 / 9007199254740992.0;   /* 2^-53 */
    const double two_pi = 6.283185307179586476925286766559;
    #pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < n; ++g) {
        uint32_t w[4];
        philox_at(seed, g, 1u, w);
        double u1 = (double)(bits53(w[0], w[1]) + 1) * s;   /* (0,1] */
        double u2 = (double)bits53(w[2], w[3]) * s;         /* [0,1) */
        double r = sqrt(-2.0 * log(u1));
        out[2 * g] = r * cos(two_pi * u2);
        out[2 * g + 1] = r * sin(two_pi * u2);
    }
}
