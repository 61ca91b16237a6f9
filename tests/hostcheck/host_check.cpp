// CPU-side checker of the HIP library's index arithmetic.  Compiled with g++
// against the library's own host+device headers, so it evaluates exactly the
// expressions the kernels evaluate: FastDiv, the six pack/unpack index maps,
// the launch shapes the dispatcher picks and every address an FFT launch
// touches.
#include <stdint.h>

#include "fft_shape.h"
#include "remap_index.h"

using namespace a2a;

extern "C" {

void hc_fastdiv(uint32_t d, const uint32_t* n, uint32_t* q, int64_t count) {
    FastDiv f(d);
    for (int64_t i = 0; i < count; ++i) q[i] = f.div(n[i]);
}

int hc_remap_dest(int kind, int P, int64_t nlocal, const int* lgs, const int* dims, uint32_t* out) {
    RemapParams p = make_params(kind, P, nlocal, lgs, dims);
    for (int64_t i = 0; i < nlocal; ++i) {
        uint32_t idx = uint32_t(i);
        switch (kind) {
        case 0: out[i] = remap_dest<0>(idx, p); break;
        case 1: out[i] = remap_dest<1>(idx, p); break;
        case 2: out[i] = remap_dest<2>(idx, p); break;
        case 3: out[i] = remap_dest<3>(idx, p); break;
        case 4: out[i] = remap_dest<4>(idx, p); break;
        default: out[i] = remap_dest<5>(idx, p); break;
        }
    }
    return 0;
}

// Enumerates (address, n, line) for every element an FFT launch loads and
// stores.  pass 0..2: the fused passes of a (Ng, P, dims) decomposition;
// pass -1: forward_1d_fft over P contiguous pencils of length Ng.
// T = E = 0 picks the default shape; xcd bit 0 applies the XCD renumbering,
// bit 1 the a2-major unit order.
// shape_out: T, E, NT, grid.  Returns the count, or a negative status.
int64_t hc_fft_enumerate(int Ng, int P, const int* dims, int dbl, int pass, int T, int E, int xcd, int64_t* addr,
                         int32_t* n_out, int64_t* line_out, int64_t cap, int64_t* shape_out, uint32_t a2_start,
                         uint32_t a2_count) {
    PassMaps pm;
    if (pass < 0) {
        pm = batched_pass_maps(Ng, uint32_t(P));
    } else {
        int lgs[3] = {Ng / dims[0], Ng / dims[1], Ng / dims[2]};
        int64_t nlocal = int64_t(lgs[0]) * lgs[1] * lgs[2];
        PassMaps all[3];
        const int n3[3] = {Ng, Ng, Ng};
        fused_pass_maps(n3, dims, lgs, nlocal, all);
        pm = all[pass];
    }
    if (T == 0) {
        KSel d = default_sel(Ng, dbl != 0, pm.strided, pm.fft, pm.a2);
        T = d.T;
        E = d.E;
    }
    FftGeom g;
    uint32_t grid;
    int st = resolve_with(pm.fft, pm.a1, pm.a2, pm.strided, Ng, T, E, &g, &grid, a2_start, a2_count);
    if (st < 0) return st;
    const int NT = T * (Ng / E);
    shape_out[0] = T;
    shape_out[1] = E;
    shape_out[2] = NT;
    shape_out[3] = grid;
    const bool strided = pm.strided != 0;
    const int TPL = Ng / E;
    int64_t c = 0;
    for (uint32_t hb = 0; hb < grid; ++hb) {
        const uint32_t b = (xcd & 1) ? xcd_block(hb, grid) : hb;
        for (int tid = 0; tid < NT; ++tid) {
            bool valid;
            const int64_t base = line_base(g, strided, T, TPL, b, tid, &valid, (xcd & 2) != 0);
            if (!valid) continue;
            const int t = strided ? tid / T : tid % TPL;
            const int line = strided ? tid % T : tid / TPL;
            for (int e = 0; e < E; ++e) {
                if (c >= cap) return -100;
                addr[c] = base + elem_off(g, e * TPL);
                n_out[c] = t + e * TPL;
                line_out[c] = int64_t(b) * T + line;
                ++c;
            }
        }
    }
    return c;
}

int hc_shape_fits(int Ng, int P, const int* dims, int pass, int T, int E) {
    PassMaps pm;
    if (pass < 0) {
        pm = batched_pass_maps(Ng, uint32_t(P));
    } else {
        int lgs[3] = {Ng / dims[0], Ng / dims[1], Ng / dims[2]};
        int64_t nlocal = int64_t(lgs[0]) * lgs[1] * lgs[2];
        PassMaps all[3];
        const int n3[3] = {Ng, Ng, Ng};
        fused_pass_maps(n3, dims, lgs, nlocal, all);
        pm = all[pass];
    }
    return shape_fits(pm.fft, pm.a2, pm.strided, Ng, T, E) ? 1 : 0;
}

// Every address of pass `pass` of a plain nx x ny x nz C-order grid with
// shape (T, E); returns the count or a negative status.
int64_t hc_grid_enumerate(int nx, int ny, int nz, int pass, int T, int E, int64_t* addr, int32_t* n_out,
                          int64_t* line_out, int64_t cap) {
    const int n3[3] = {nx, ny, nz};
    const int dims[3] = {1, 1, 1};
    PassMaps all[3];
    fused_pass_maps(n3, dims, n3, int64_t(nx) * ny * nz, all);
    const PassMaps pm = all[pass];
    const int N = int(pm.fft.n);
    FftGeom g;
    uint32_t grid;
    int st = resolve_with(pm.fft, pm.a1, pm.a2, pm.strided, N, T, E, &g, &grid);
    if (st < 0) return st;
    const int NT = T * (N / E), TPL = N / E;
    const bool strided = pm.strided != 0;
    int64_t c = 0;
    for (uint32_t b = 0; b < grid; ++b)
        for (int tid = 0; tid < NT; ++tid) {
            bool valid;
            const int64_t base = line_base(g, strided, T, TPL, b, tid, &valid);
            if (!valid) continue;
            const int t = strided ? tid / T : tid % TPL;
            const int line = strided ? tid % T : tid / TPL;
            for (int e = 0; e < E; ++e) {
                if (c >= cap) return -100;
                addr[c] = base + elem_off(g, e * TPL);
                n_out[c] = t + e * TPL;
                line_out[c] = int64_t(b) * T + line;
                ++c;
            }
        }
    return c;
}

// Grids of 2^32 points and more: for the hardware workgroups [b_begin,
// b_end) of pass `pass` of an nx x ny x nz grid in the virtual-rank block
// layout over `dims` ([1,1,1]: one C-order grid) with shape (T, E) and flags
// (KF_XCD, KF_A2MAJ honoured), every address the kernel computes (line_base
// + elem_off) against the element's index computed directly in 64 bits from
// its global (x, y, z): rank (x / lx, y / ly, z / lz) row-major times the
// block size, plus the C-order offset inside the block.  grid_out: the
// launch's workgroup count.  Returns the number of mismatches or
// out-of-range addresses (0 = all agree), or a negative status.
int64_t hc_grid_check(int nx, int ny, int nz, const int* dims, int pass, int T, int E, int flags, uint32_t b_begin,
                      uint32_t b_end, int64_t* grid_out, int64_t* checked_out) {
    const int n3[3] = {nx, ny, nz};
    const int lgs[3] = {nx / dims[0], ny / dims[1], nz / dims[2]};
    const uint64_t lx = uint64_t(lgs[0]), ly = uint64_t(lgs[1]), lz = uint64_t(lgs[2]);
    PassMaps all[3];
    const int64_t nlocal = int64_t(lx * ly * lz);
    const int64_t total = int64_t(nx) * ny * nz;
    fused_pass_maps(n3, dims, lgs, nlocal, all);
    const PassMaps pm = all[pass];
    const int N = int(pm.fft.n);
    FftGeom g;
    uint32_t grid;
    const int st = resolve_geom(pm.fft, pm.a1, pm.a2, pm.strided, T, &g, &grid, 0, 0);
    if (st < 0) return st;
    *grid_out = grid;
    const int TPL = N / E, NT = T * TPL;
    const bool strided = pm.strided != 0, a2maj = (flags & KF_A2MAJ) != 0;
    int64_t bad = 0, checked = 0;
    for (uint32_t hb = b_begin; hb < b_end && hb < grid; ++hb) {
        const uint32_t b = (flags & KF_XCD) ? xcd_block(hb, grid) : hb;
        for (int tid = 0; tid < NT; ++tid) {
            bool valid;
            const int64_t base = line_base(g, strided, T, TPL, b, tid, &valid, a2maj);
            if (!valid) continue;
            const int t = strided ? tid / T : tid % TPL;
            const int line = strided ? tid % T : tid / TPL;
            // the unit's (a1, a2) coordinates, straight from the unit order
            uint64_t u1, u2;
            if (strided) {
                const uint64_t n2 = uint64_t(pm.a2.n) / uint64_t(T), n1 = pm.a1.n;
                u1 = a2maj ? b % n1 : b / n2;
                u2 = (a2maj ? b / n1 : b % n2) * uint64_t(T) + uint64_t(line);
            } else {
                const uint64_t ln = uint64_t(b) * uint64_t(T) + uint64_t(line);
                u1 = ln / pm.a2.n;
                u2 = ln % pm.a2.n;
            }
            for (int e = 0; e < E; ++e) {
                const uint64_t n = uint64_t(t) + uint64_t(e) * uint64_t(TPL);
                uint64_t x, y, z;
                if (pass == 0) { x = u1; y = u2; z = n; }
                else if (pass == 1) { x = n; y = u1; z = u2; }
                else { x = u1; y = n; z = u2; }
                const uint64_t rank = ((x / lx) * uint64_t(dims[1]) + y / ly) * uint64_t(dims[2]) + z / lz;
                const int64_t want =
                    int64_t(rank * uint64_t(nlocal) + ((x % lx) * ly + (y % ly)) * lz + (z % lz));
                const int64_t got = base + elem_off(g, e * TPL);   // line_base includes off(t)
                if (got != want || got < 0 || got >= total) ++bad;
                ++checked;
            }
        }
    }
    *checked_out = checked;
    return bad;
}

void hc_xcd_blocks(uint32_t G, uint32_t* out) {
    for (uint32_t b = 0; b < G; ++b) out[b] = xcd_block(b, G);
}

// The four-step of lines longer than 4096: its factor split, and for every
// workgroup of a chunk of `lines` lines every (source, destination) element
// pair its twiddle-transpose moves, with the tile slot the pair passes through
// (the writer thread (tx, ty + r) of phase 1 and the reader of phase 2 meet at
// tile[ty + r][tx]).  Returns the count, or a negative status.
int hc_long_split(int64_t N, int* n1, int* n2) { return long_fft_split(N, n1, n2) ? 1 : 0; }

int64_t hc_long_transpose(int N1, int N2, int lines, int64_t* src, int64_t* dst, int64_t cap) {
    const uint32_t tn2s = uint32_t((N2 + 31) / 32), tk1s = uint32_t((N1 + 31) / 32);
    const uint64_t blocks = uint64_t(tn2s) * tk1s * uint32_t(lines);
    const int64_t N = int64_t(N1) * N2;
    int64_t c = 0;
    for (uint64_t b = 0; b < blocks; ++b) {
        uint32_t tn2, tk1, line;
        long_tile(uint32_t(b), tn2s, tk1s, &tn2, &tk1, &line);
        for (uint32_t i = 0; i < 32; ++i)        // tile row = ty + r of the phase-1 writer
            for (uint32_t j = 0; j < 32; ++j) {  // tile column = tx
                const uint32_t k1 = tk1 * 32u + i, n2 = tn2 * 32u + j;
                if (!(k1 < uint32_t(N1) && n2 < uint32_t(N2))) continue;
                if (c >= cap) return -100;
                src[c] = int64_t(line) * N + int64_t(k1) * N2 + n2;
                // phase 2: thread (tx = i, ty + r = j) reads tile[i][j], n2' = tn2*32 + j, k1' = tk1*32 + i
                dst[c] = int64_t(line) * N + int64_t(tn2 * 32u + j) * N1 + (tk1 * 32u + i);
                ++c;
            }
    }
    return c;
}

// The square in-place case: for every workgroup (tile pair) of `lines` lines
// of [M][M], every (source, destination) element pair, and the workgroup that
// moves it.  Returns the count, or a negative status.
int64_t hc_long_square(int M, int lines, int64_t* src, int64_t* dst, int64_t* wg, int64_t cap) {
    const uint32_t tiles = uint32_t((M + 31) / 32), pairs = tiles * (tiles + 1) / 2;
    const uint64_t blocks = uint64_t(pairs) * uint32_t(lines);
    const int64_t N = int64_t(M) * M;
    int64_t c = 0;
    for (uint64_t b = 0; b < blocks; ++b) {
        uint32_t a, cc, line;
        long_pair(uint32_t(b), tiles, &a, &cc, &line);
        for (int k = 0; k < (a == cc ? 1 : 2); ++k) {
            const uint32_t row0 = (k == 0 ? a : cc) * 32u, col0 = (k == 0 ? cc : a) * 32u;
            for (uint32_t i = 0; i < 32; ++i)
                for (uint32_t j = 0; j < 32; ++j) {
                    const uint32_t k1 = row0 + i, n2 = col0 + j;
                    if (!(k1 < uint32_t(M) && n2 < uint32_t(M))) continue;
                    if (c >= cap) return -100;
                    src[c] = int64_t(line) * N + int64_t(k1) * M + n2;
                    dst[c] = int64_t(line) * N + int64_t(n2) * M + k1;
                    wg[c] = int64_t(b);
                    ++c;
                }
        }
    }
    return c;
}

// Every address the four-step's first FFT pass touches: `lines` lines of
// N1 * N2 seen as [N1][N2], N1-point FFTs down the N2 columns, shape (T, E)
// (T = E = 0: the default).  n_out = the element's index along its column,
// col_out = line * N2 + column.  Returns the count, or a negative status.
int64_t hc_long_pass_enumerate(int N1, int N2, int lines, int dbl, int T, int E, int64_t* addr, int32_t* n_out,
                               int64_t* col_out, int64_t cap, int64_t* shape_out) {
    const int64_t N = int64_t(N1) * N2;
    const AxisMap fft = make_axis(uint32_t(N1), uint32_t(N1), 0, N2);
    const AxisMap a1 = make_axis(uint32_t(lines), uint32_t(lines), 0, N);
    const AxisMap a2 = make_axis(uint32_t(N2), uint32_t(N2), 0, 1);
    if (T == 0) {
        const KSel d = default_sel(N1, dbl != 0, 1, fft, a2);
        T = d.T;
        E = d.E;
    }
    FftGeom g;
    uint32_t grid;
    int st;
    if (fft_len_ok(N1)) {
        st = resolve_with(fft, a1, a2, 1, N1, T, E, &g, &grid);
    } else {
        // the runtime-schedule kernel and the direct DFT: the geometry of
        // their chosen (T, E), resolved as the dispatcher does
        if (T <= 0 || E <= 0 || N1 % E) return A2AFFT_E_UNSUPPORTED;
        const uint32_t tpl = uint32_t(N1 / E);
        if (!(fft.l % tpl == 0 || tpl % fft.l == 0) || T * int(tpl) > 1024) return A2AFFT_E_UNSUPPORTED;
        st = resolve_geom(fft, a1, a2, 1, T, &g, &grid, 0, 0);
    }
    if (st < 0) return st;
    const int NT = T * (N1 / E), TPL = N1 / E;
    shape_out[0] = T;
    shape_out[1] = E;
    int64_t c = 0;
    for (uint32_t b = 0; b < grid; ++b)
        for (int tid = 0; tid < NT; ++tid) {
            bool valid;
            const int64_t base = line_base(g, true, T, TPL, b, tid, &valid);
            if (!valid) continue;
            const int t = tid / T;
            for (int e = 0; e < E; ++e) {
                if (c >= cap) return -100;
                addr[c] = base + elem_off(g, e * TPL);
                n_out[c] = t + e * TPL;
                col_out[c] = int64_t(b) * T + tid % T;
                ++c;
            }
        }
    return c;
}

// Every element the chirp-z pass (cz_fused_kernel) loads and stores: pass `pass` of an n3 grid over `dims` virtual ranks,
// window [a2_start, a2_start + a2_count) of its a2 axis.  Per element: its
// grid address, its index n along the line, and the window line u.  Returns
// the count, or a negative status.
int64_t hc_cz_fused_enumerate(const int* n3, const int* dims, int pass, uint32_t a2_start, uint32_t a2_count,
                              int64_t* addr, int32_t* n_out, int64_t* line_out, int64_t cap) {
    int lgs[3];
    for (int i = 0; i < 3; ++i) lgs[i] = n3[i] / dims[i];
    const int64_t nlocal = int64_t(lgs[0]) * lgs[1] * lgs[2];
    PassMaps all[3];
    fused_pass_maps(n3, dims, lgs, nlocal, all);
    const PassMaps& pm = all[pass];
    const int N = int(pm.fft.n), M = cz_length(N);
    if (N < 9 || N > kDftMaxN) return -101;
    FftGeom g;
    uint32_t units;
    const int st = resolve_geom(pm.fft, pm.a1, pm.a2, 0, 1, &g, &units, a2_start, a2_count);
    if (st < 0) return st;
    const uint32_t T = uint32_t(cz_fused_lines(M)), TPL = uint32_t(M / kCzFusedE);
    const uint32_t grid = (units + T - 1) / T;
    int64_t c = 0;
    for (uint32_t b = 0; b < grid; ++b)
        for (int tid = 0; tid < kCzFusedThreads; ++tid) {
            int t, line;
            cz_fused_thread(pm.strided != 0, M, tid, &t, &line);
            const uint32_t u = b * T + uint32_t(line);
            if (u >= units) continue;
            for (int e = 0; e < kCzFusedE; ++e) {
                const int n = t + e * int(TPL);
                if (n >= N) continue;
                if (c >= cap) return -100;
                addr[c] = cz_line_base(g, u) + g.fft.off(uint32_t(n));
                n_out[c] = n;
                line_out[c] = u;
                ++c;
            }
        }
    return c;
}

}  // extern "C"
