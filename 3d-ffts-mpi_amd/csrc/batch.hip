// Batched transform of many independent grids of uneven shapes in ONE
// persistent launch (no reference counterpart: the reference transforms one
// cubic grid per a2aDistribution, a2adistribution.h:9).
//
// Work queue.  Every (pass, grid) pair is split into the units its Stockham
// kernel would use (lines or column tiles).  The grids are packed into groups
// (one group of all grids by default) and the work is cut into phases: one
// phase per (group, pass).  Units are numbered phase by phase, in an order
// the host chooses: group-major (z, x, y of group 0, then of group 1, ...) or
// skewed along diagonals, so that the phase between a pass and the next pass
// of the same group belongs to another group.  A group's next pass then reads
// data its previous pass wrote shortly before, which may still sit in the
// device's memory-side cache (measured: 8 grids of 256^3 take 1.14 ms one grid
// per group, skewed, against 1.29 ms pass-major).  Workgroups take units
// in queue order from atomic counters; a table maps each unit to its
// (phase, grid) item.  The take of the next unit is issued before the current
// unit is computed and examined after it, so the latency of the device-scope
// atomic hides behind the unit's own loads.
//
// Synchronisation is per phase, not per unit.  The 8 XCDs of the device have
// L2 caches that are not coherent with one another, and a device-scope fence
// writes back or invalidates a whole L2, so the kernel uses none: every unit
// load and store is a relaxed device-scope access (sc1), coherent across the
// XCDs by itself.  A workgroup counts the units it finishes and, only when
// the next unit it takes belongs to another phase (or the queues are empty),
// waits for its own stores to complete, adds its count to that phase's total
// and, if the new phase is a later pass of a group, waits until the phase of
// the previous pass of that group is complete.  Every wait ends: every unit of
// phase q is taken before any unit of a later phase, a phase depends only on
// an earlier one, and the workgroup holding a unit of phase q finishes it
// after at most a wait for that earlier phase (it may hold one taken unit of
// a later phase meanwhile, which nobody waits on yet), so by induction over q
// every phase completes.  The wait is still bounded, and a bound that trips
// sets an error flag, after which every other wait gives up within 16 polls,
// instead of hanging.
#include "fft.h"
#include "stockham.h"

#include <algorithm>
#include <cstdio>
#include <new>
#include <string>
#include <vector>

namespace a2a {

namespace {

constexpr int kBatchThreads = 256;
constexpr uint32_t kSpinLimit = 1u << 20;   // polls of ~8K cycles each: seconds
constexpr uint32_t kMaxUnits = 1u << 28;   // the unit -> item table holds 4 bytes per unit
constexpr int kMaxPhases = 384;            // (group, pass) phases: at most 128 groups

// Kernel shapes per (N, strided): the primary (first of its class in this
// table), alternatives the autotuner may choose instead (the same tile
// unpadded, half the tile, wider tiles for short lines), and a single-column
// fallback for z extents below the tile width.  Alternatives keep the primary's elements per thread and
// stay within its LDS, so the kernel's registers (103) and LDS (39.9 KB) are
// those of the primaries.  Twice the elements per thread was measured slower
// for every class (the 64-grid batch, 8 x 256^3, 15 x 256x64x512) and made
// the kernel need 164 VGPRs.
struct BShape {
    int N, strided, T, E, pad;
};
constexpr BShape kShapes[] = {
    // primary
    {16, 0, 256, 16, 1}, {32, 0, 64, 8, 1}, {64, 0, 32, 8, 1}, {128, 0, 32, 16, 1}, {256, 0, 16, 16, 1},
    {512, 0, 8, 16, 1}, {16, 1, 32, 16, 1}, {32, 1, 32, 8, 1}, {64, 1, 32, 8, 1}, {128, 1, 32, 16, 1},
    {256, 1, 16, 16, 1}, {512, 1, 8, 16, 1},
    // single-column fallbacks
    {16, 1, 1, 16, 1}, {32, 1, 1, 8, 1}, {64, 1, 1, 8, 1}, {128, 1, 1, 16, 1}, {256, 1, 1, 16, 1}, {512, 1, 1, 8, 1},
    // alternatives: unpadded
    {32, 0, 64, 8, 0}, {64, 0, 32, 8, 0}, {128, 0, 32, 16, 0}, {256, 0, 16, 16, 0}, {512, 0, 8, 16, 0},
    {32, 1, 32, 8, 0}, {64, 1, 32, 8, 0}, {128, 1, 32, 16, 0}, {256, 1, 16, 16, 0}, {512, 1, 8, 16, 0},
    // alternatives: half the tile (lines or columns), 128 threads
    {32, 0, 32, 8, 1}, {64, 0, 16, 8, 1}, {128, 0, 16, 16, 1}, {256, 0, 8, 16, 1}, {512, 0, 4, 16, 1},
    {64, 1, 16, 8, 1}, {128, 1, 16, 16, 1}, {256, 1, 8, 16, 1}, {512, 1, 4, 16, 1},
    // alternatives: wider tiles for short lines (more columns or lines per unit)
    {32, 1, 64, 8, 1}, {64, 1, 64, 16, 1}, {32, 1, 128, 16, 1}, {16, 1, 256, 16, 1}, {64, 0, 64, 16, 1},
    {32, 0, 128, 16, 1},
};
constexpr int kNumShapes = int(sizeof(kShapes) / sizeof(kShapes[0]));
constexpr int kFirstAlt = 18;

}  // namespace

struct BatchItem {
    FftGeom g;
    int64_t base;        // element offset of the grid in the batch buffer
    const float2* tw;    // twiddles of this pass's length and direction
    uint32_t units;
    uint32_t first;      // number of this item's first unit in the queue
    int shape;           // index into kShapes
    int phase;           // the (group, pass) phase the item belongs to
};

static_assert(kNumShapes == 43, "run_unit lists every shape");

constexpr int shape_flags(int i) { return (kShapes[i].pad ? KF_PAD : 0) | KF_COH; }

template <int DIR, typename Hook>
__device__ __forceinline__ void run_unit(int shape, float2* data, const FftGeom& g, const float2* tw, uint32_t u,
                                         float2* lds, int tid, Hook&& hook) {
    switch (shape) {
#define A2A_B(I)                                                                                                \
    case I:                                                                                                    \
        stockham_unit<float, kShapes[I].N, kShapes[I].E, kShapes[I].T, kShapes[I].strided != 0, DIR,          \
                      shape_flags(I)>(data, g, tw, u, lds, tid, hook);                                         \
        break;
        A2A_B(0) A2A_B(1) A2A_B(2) A2A_B(3) A2A_B(4) A2A_B(5) A2A_B(6) A2A_B(7) A2A_B(8)
        A2A_B(9) A2A_B(10) A2A_B(11) A2A_B(12) A2A_B(13) A2A_B(14) A2A_B(15) A2A_B(16) A2A_B(17)
        A2A_B(18) A2A_B(19) A2A_B(20) A2A_B(21) A2A_B(22) A2A_B(23) A2A_B(24) A2A_B(25) A2A_B(26)
        A2A_B(27) A2A_B(28) A2A_B(29) A2A_B(30) A2A_B(31) A2A_B(32) A2A_B(33) A2A_B(34) A2A_B(35)
        A2A_B(36) A2A_B(37) A2A_B(38) A2A_B(39) A2A_B(40) A2A_B(41) A2A_B(42)
#undef A2A_B
    default:   // the host only stores valid shape indices
        __builtin_unreachable();
    }
}

template <int I>
constexpr int batch_lds_max() {
    if constexpr (I == kNumShapes) {
        return 1;
    } else {
        constexpr int here = lds_slots<kShapes[I].N, kShapes[I].T, kShapes[I].strided != 0, kShapes[I].pad>();
        constexpr int rest = batch_lds_max<I + 1>();
        return here > rest ? here : rest;
    }
}

// Unit queues.  Phase q owns units [start[q], start[q+1]); with xcd_queues
// they are split into 8 contiguous sub-ranges, one per XCD (the hardware
// hands workgroup b to XCD b % 8), taken through 8 counters; a workgroup
// whose own sub-range is empty takes from the others.  One take hands out
// `chunk` consecutive units.  A workgroup moves to phase q+1 only after every
// counter of phase q is exhausted, so all units of a phase are taken before
// any unit of a later one.
//
// Control words, each in its own 256-byte slot so that takes on different
// counters do not serialise on one cache line: slot 0 error flag, slot 1 + q
// units of phase q finished and released, slot 1 + kMaxPhases + 8 q + x take
// counter of queue x of phase q (queue 0 only without xcd_queues).
constexpr int kCtrlErr = 0, kCtrlDone = 1, kCtrlTake = 1 + kMaxPhases, kCtrlSlots = 1 + 9 * kMaxPhases;
constexpr int kCtrlStride = 64;
constexpr uint32_t kNoUnit = 0xffffffffu;

__host__ __device__ constexpr int ctrl_word(int slot) { return slot * kCtrlStride; }

struct BatchQueues {
    int nphases;
    int xcd_queues;
    uint32_t chunk;
};

// Where thread 0 of a workgroup takes its next unit: the rest [next, end) of
// its current chunk, else phase `scan`, the i-th queue counted from its own.
// Exhausted queues stay exhausted (counters only grow), so the cursor only
// moves forward.
struct Cursor {
    int scan, i;
    uint32_t next, end;
};

// The phase table in LDS.  Thread 0 must not read it from global memory
// inside the unit loop: the wait for such a load would also wait for its own
// outstanding stores of the previous unit, while an LDS read waits for LDS
// only.
struct PhaseTable {
    const uint32_t* start;   // nphases + 1 unit numbers
    const int* dep;          // phase that must be complete first, or -1
    int n;
    uint32_t chunk;
};

__device__ __forceinline__ void queue_range(const PhaseTable& pt, int p, int x, int nq, uint32_t& lo, uint32_t& hi) {
    const uint32_t a = pt.start[p];
    const uint32_t len = pt.start[p + 1] - a;
    lo = a + uint32_t((uint64_t(len) * uint32_t(x)) / uint32_t(nq));
    hi = a + uint32_t((uint64_t(len) * uint32_t(x + 1)) / uint32_t(nq));
}

// Issues the take from the cursor's queue and returns the raw ticket without
// examining it, so the atomic's latency overlaps the unit being computed.  No
// take while the current chunk has units left.
__device__ __forceinline__ uint32_t claim_issue(const Cursor& c, const PhaseTable& pt, int myq, int nq,
                                                uint32_t* ctrl) {
    if (c.next < c.end || c.scan >= pt.n) return kNoUnit;
    const int x = (myq + c.i) % nq;
    uint32_t lo, hi;
    queue_range(pt, c.scan, x, nq, lo, hi);
    return lo < hi ? atomicAdd(&ctrl[ctrl_word(kCtrlTake + c.scan * 8 + x)], 1u) : kNoUnit;
}

// Turns a ticket into a unit number, moving on through the queues (with
// further takes) while the ticket falls past the end of its queue; kNoUnit
// when every queue of every phase is exhausted.
__device__ __forceinline__ uint32_t claim_finish(Cursor& c, uint32_t ticket, const PhaseTable& pt, int myq, int nq,
                                                 uint32_t* ctrl) {
    if (c.next < c.end) return c.next++;
    for (;;) {
        if (c.scan >= pt.n) return kNoUnit;
        const int x = (myq + c.i) % nq;
        uint32_t lo, hi;
        queue_range(pt, c.scan, x, nq, lo, hi);
        if (lo < hi && ticket < (hi - lo + pt.chunk - 1) / pt.chunk) {
            const uint32_t first = lo + ticket * pt.chunk;
            c.next = first + 1;
            c.end = min(first + pt.chunk, hi);
            return first;
        }
        if (++c.i == nq) {
            c.i = 0;
            ++c.scan;
        }
        ticket = claim_issue(c, pt, myq, nq, ctrl);
    }
}

// Waits until `need` units of phase q are finished and released, giving up
// (and raising the error flag) after kSpinLimit polls, or within 16 polls once
// another wait has given up.
__device__ __forceinline__ void wait_phase(uint32_t* ctrl, int q, uint32_t need) {
    uint32_t spins = 0;
    while (__hip_atomic_load(&ctrl[ctrl_word(kCtrlDone + q)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
        __builtin_amdgcn_s_sleep(127);   // long polls measured a little faster: less contention on the word
        ++spins;
        if ((spins & 15u) == 0 &&
            __hip_atomic_load(&ctrl[ctrl_word(kCtrlErr)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
            return;
        if (spins > kSpinLimit) {
            __hip_atomic_store(&ctrl[ctrl_word(kCtrlErr)], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
    }
}

template <int DIR>
// At least 3 wavefronts per SIMD (at most 168 VGPRs); the kernel needs 103, so LDS
// (39.9 KB per workgroup) is what limits a CU to four workgroups.
__global__ __launch_bounds__(kBatchThreads, 3) void batch_kernel(float2* __restrict__ data,
                                                              const BatchItem* __restrict__ items,
                                                              const uint32_t* __restrict__ unit_item,
                                                              const uint32_t* __restrict__ phases, BatchQueues qs,
                                                              uint32_t* __restrict__ ctrl) {
    __shared__ float2 lds[batch_lds_max<0>()];
    __shared__ uint32_t s_start[kMaxPhases + 1];
    __shared__ int s_dep[kMaxPhases];
    __shared__ uint32_t s_unit[2];   // alternate per unit: one barrier per unit suffices
    const int tid = int(threadIdx.x);
    const int nph = __builtin_amdgcn_readfirstlane(qs.nphases);
    for (int i = tid; i <= nph; i += kBatchThreads) s_start[i] = phases[i];
    for (int i = tid; i < nph; i += kBatchThreads) s_dep[i] = int(phases[nph + 1 + i]);
    __syncthreads();
    const int nq = qs.xcd_queues ? 8 : 1;
    const int myq = qs.xcd_queues ? int(blockIdx.x & 7u) : 0;
    const PhaseTable pt{s_start, s_dep, nph, uint32_t(__builtin_amdgcn_readfirstlane(int(qs.chunk)))};
    Cursor c{0, 0, 0u, 0u};
    if (tid == 0) s_unit[0] = claim_finish(c, claim_issue(c, pt, myq, nq, ctrl), pt, myq, nq, ctrl);
    __syncthreads();
    uint32_t unit = uint32_t(__builtin_amdgcn_readfirstlane(int(s_unit[0])));
    int cur = 0;          // phase of the units this workgroup is finishing
    uint32_t count = 0;   // units of phase `cur` it has finished
    for (int it_no = 1;; ++it_no) {
        const int k = unit == kNoUnit ? -1 : int(unit_item[unit]);
        const int q = k < 0 ? nph : __builtin_amdgcn_readfirstlane(items[k].phase);   // nph: queues empty
        if (q != cur) {
            // Release: the unit data is written with device-coherent stores, so
            // waiting for this wavefront's stores to complete is enough (no L2
            // write-back); the barrier then covers every wavefront.
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (tid == 0) {
                if (count) atomicAdd(&ctrl[ctrl_word(kCtrlDone + cur)], count);
                const int dq = q < nph ? pt.dep[q] : -1;
                if (dq >= 0) wait_phase(ctrl, dq, pt.start[dq + 1] - pt.start[dq]);
            }
            __syncthreads();   // acquire: the unit loads that follow are device-coherent too
            cur = q;
            count = 0;
        }
        if (q == nph) break;
        uint32_t ticket;
        if (tid == 0) ticket = claim_issue(c, pt, myq, nq, ctrl);
        const BatchItem& it = items[k];
        // The thread index as the compiler must re-read it every unit: left
        // visible, the per-thread LDS addresses of every shape are hoisted out
        // of the loop and kept live across all 18 shapes (168 VGPRs with
        // spills, whose reloads wait for every outstanding global access)
        // instead of 103 VGPRs recomputed per unit.
        int utid = tid;
        asm volatile("" : "+v"(utid));
        // The ticket is examined once the unit's loads are consumed and before
        // its stores are issued, so thread 0 never waits for its own stores.
        run_unit<DIR>(it.shape, data + it.base, it.g, it.tw, unit - it.first, lds, utid, [&]() {
            if (tid == 0) s_unit[it_no & 1] = claim_finish(c, ticket, pt, myq, nq, ctrl);
        });
        ++count;
        __syncthreads();   // the unit is done with LDS; s_unit is written
        unit = uint32_t(__builtin_amdgcn_readfirstlane(int(s_unit[it_no & 1])));
    }
}

// Clears the first word of every control slot (the only word used).
__global__ void batch_reset_kernel(uint32_t* ctrl, int nslots) {
    for (int i = int(threadIdx.x); i < nslots; i += int(blockDim.x)) ctrl[ctrl_word(i)] = 0u;
}

}  // namespace a2a

using namespace a2a;

struct a2afft_batch_s {
    int G = 0;
    std::vector<int> shapes;          // 3 per grid
    std::vector<int64_t> offsets;     // G + 1
    std::vector<AxisPass> passes;     // pass p of grid g at p * G + g
    int choice[2][10];                // [strided][log2 N]: chosen shape of the class, -1 = the primary
    std::vector<BatchItem> proto[2];  // forward, inverse: item of (pass p, grid g) at p * G + g
    BatchItem* d_items[2] = {nullptr, nullptr};
    uint32_t* d_unit_item = nullptr;  // item of every unit
    uint32_t* d_phases = nullptr;     // phase starts (nphases + 1), then dependencies (nphases)
    uint32_t* d_ctrl = nullptr;
    std::vector<uint32_t> pstart;     // host copy of the phase starts
    uint32_t total = 0;
    BatchQueues qs{0, 0, 1};
    int64_t group_points = 0;
    int skewed = 0;
    int grid = 0;                     // persistent workgroups
    int max_grid = 0;
    hipStream_t cap_stream = nullptr;
    struct Graph {
        void* data;
        int dir;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;        // captured transforms, most recent last
    std::string report;               // the last autotune's timings
};

// Captured graphs kept per batch: one per (buffer, direction), the least
// recently used dropped beyond this count.
constexpr size_t kMaxBatchGraphs = 32;

static void drop_graphs(a2afft_batch_s* b) {
    for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.exec);
    b->graphs.clear();
}

static void batch_free(a2afft_batch_s* b) {
    if (!b) return;
    drop_graphs(b);
    for (int d = 0; d < 2; ++d)
        if (b->d_items[d]) (void)hipFree(b->d_items[d]);
    if (b->d_unit_item) (void)hipFree(b->d_unit_item);
    if (b->d_phases) (void)hipFree(b->d_phases);
    if (b->d_ctrl) (void)hipFree(b->d_ctrl);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    delete b;
}

static int log2i(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

static bool shape_ok(int i, const AxisPass& p, int N) {
    const BShape& s = kShapes[i];
    return s.N == N && int(p.fft.n) == N && s.strided == p.strided && shape_fits(p.fft, p.a2, p.strided, N, s.T, s.E);
}

// The class's chosen shape if it fits the pass, else the primary, else the
// single-column fallback.
static int pick_shape(const a2afft_batch_s* b, const AxisPass& p, int N) {
    const int c = b->choice[p.strided][log2i(N)];
    if (c >= 0 && shape_ok(c, p, N)) return c;
    for (int i = 0; i < kFirstAlt; ++i)
        if (shape_ok(i, p, N)) return i;
    return -1;
}

// Items of every (pass, grid) for the current shape choices.
static int build_protos(a2afft_batch_s* b) {
    const int G = b->G;
    for (int p = 0; p < 3; ++p)
        for (int g = 0; g < G; ++g) {
            const int k = p * G + g;
            const AxisPass& ps = b->passes[size_t(k)];
            const int N = int(ps.fft.n);
            const int si = pick_shape(b, ps, N);
            if (si < 0) return set_error(A2AFFT_E_UNSUPPORTED, "batch: no kernel shape for grid %d pass %d", g, p);
            const int64_t total = b->offsets[size_t(g) + 1] - b->offsets[size_t(g)];
            for (int d = 0; d < 2; ++d) {
                BatchItem& it = b->proto[d][size_t(k)];
                uint32_t units = 0;
                const int st = resolve_with(ps.fft, ps.a1, ps.a2, ps.strided, N, kShapes[si].T, kShapes[si].E, &it.g,
                                            &units);
                if (st < 0) return set_error(st, "batch: geometry of grid %d pass %d", g, p);
                it.g.scale_mode = (d == 1 && p == 2) ? 2 : 0;
                it.g.scale = (d == 1 && p == 2) ? double(total) : 1.0;
                it.base = b->offsets[size_t(g)];
                it.units = units;
                it.first = 0;
                it.shape = si;
                it.phase = 0;
                it.tw = (const float2*)twiddle_table(N, d == 0 ? -1 : 1, A2AFFT_C64);
                if (!it.tw) return set_error(A2AFFT_E_ALLOC, "batch: twiddle table N=%d", N);
            }
        }
    return A2AFFT_OK;
}

template <typename T>
static int upload(T** dst, const std::vector<T>& v) {
    *dst = nullptr;
    A2A_TRY(hip_check(int(hipMalloc(dst, sizeof(T) * v.size())), "batch alloc"));
    return hip_check(int(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice)), "batch copy");
}

// Builds the queue: groups of consecutive grids of at most group_points
// points (0 = one group; a larger grid forms a group of its own), one phase
// per (group, pass), phases group-major or skewed along the diagonals
// group + pass (the later pass first), units numbered phase by phase.  The
// device tables are replaced only when every new one is in place.
// First grid of every group, then G.
static std::vector<int> group_starts(const a2afft_batch_s* b, int64_t group_points) {
    std::vector<int> first{0};
    if (group_points > 0) {
        int64_t acc = 0;
        for (int g = 0; g < b->G; ++g) {
            const int64_t n = b->offsets[size_t(g) + 1] - b->offsets[size_t(g)];
            if (g > first.back() && acc + n > group_points) {
                first.push_back(g);
                acc = 0;
            }
            acc += n;
        }
    }
    first.push_back(b->G);
    return first;
}

static int batch_tables(a2afft_batch_s* b, int64_t group_points, int skewed) {
    const int G = b->G;
    const std::vector<int> first = group_starts(b, group_points);
    const int W = int(first.size()) - 1;
    if (3 * W > kMaxPhases)
        return set_error(A2AFFT_E_SIZE, "batch: %d groups of at most %lld points (at most %d groups)", W,
                         (long long)group_points, kMaxPhases / 3);
    std::vector<int> ph_group, ph_pass;
    for (int d = 0; d < (skewed ? W + 2 : W); ++d)
        for (int i = 0; i < 3; ++i) {
            const int p = skewed ? 2 - i : i;
            const int w = skewed ? d - p : d;
            if (w < 0 || w >= W) continue;
            ph_group.push_back(w);
            ph_pass.push_back(p);
        }
    const int nph = int(ph_group.size());
    std::vector<int> index(size_t(3 * W), -1);
    for (int q = 0; q < nph; ++q) index[size_t(3 * ph_group[q] + ph_pass[q])] = q;
    std::vector<BatchItem> items[2];
    std::vector<uint32_t> phases(size_t(2 * nph + 1)), unit_item;
    uint64_t total = 0;
    for (int q = 0; q < nph; ++q) {
        const int w = ph_group[q], p = ph_pass[q];
        phases[size_t(q)] = uint32_t(total);
        phases[size_t(nph + 1 + q)] = p > 0 ? uint32_t(index[size_t(3 * w + p - 1)]) : 0xffffffffu;
        for (int g = first[size_t(w)]; g < first[size_t(w) + 1]; ++g) {
            const uint32_t units = b->proto[0][size_t(p * G + g)].units;
            if (total + units > kMaxUnits) return set_error(A2AFFT_E_SIZE, "batch: more than %u work units", kMaxUnits);
            for (int d = 0; d < 2; ++d) {
                BatchItem it = b->proto[d][size_t(p * G + g)];
                it.first = uint32_t(total);
                it.phase = q;
                items[d].push_back(it);
            }
            unit_item.insert(unit_item.end(), units, uint32_t(items[0].size() - 1));
            total += units;
        }
    }
    phases[size_t(nph)] = uint32_t(total);
    BatchItem* d_items[2] = {nullptr, nullptr};
    uint32_t *d_unit_item = nullptr, *d_phases = nullptr;
    int st = upload(&d_items[0], items[0]);
    if (st == A2AFFT_OK) st = upload(&d_items[1], items[1]);
    if (st == A2AFFT_OK) st = upload(&d_unit_item, unit_item);
    if (st == A2AFFT_OK) st = upload(&d_phases, phases);
    if (st < 0) {
        for (void* q : {(void*)d_items[0], (void*)d_items[1], (void*)d_unit_item, (void*)d_phases})
            if (q) (void)hipFree(q);
        return st;
    }
    drop_graphs(b);   // they hold the old tables
    for (int d = 0; d < 2; ++d) {
        if (b->d_items[d]) (void)hipFree(b->d_items[d]);
        b->d_items[d] = d_items[d];
    }
    if (b->d_unit_item) (void)hipFree(b->d_unit_item);
    if (b->d_phases) (void)hipFree(b->d_phases);
    b->d_unit_item = d_unit_item;
    b->d_phases = d_phases;
    b->pstart.assign(phases.begin(), phases.begin() + nph + 1);
    b->total = uint32_t(total);
    b->qs.nphases = nph;
    b->group_points = group_points;
    b->skewed = skewed;
    return A2AFFT_OK;
}

extern "C" int a2afft_batch_create_ranks(a2afft_batch* out, int ngrids, const int* shapes, const int* dims,
                                         a2afft_precision prec) {
    if (!out || ngrids <= 0 || !shapes) return set_error(A2AFFT_E_INVALID_ARG, "batch_create: bad arguments");
    *out = nullptr;
    if (prec != A2AFFT_FP32) return set_error(A2AFFT_E_UNSUPPORTED, "batch_create: FP32 (complex64) only");
    a2afft_batch_s* b = new (std::nothrow) a2afft_batch_s;
    if (!b) return set_error(A2AFFT_E_ALLOC, "batch_create");
    b->G = ngrids;
    b->shapes.assign(shapes, shapes + 3 * ngrids);
    b->offsets.assign(ngrids + 1, 0);
    const int K = 3 * ngrids;
    b->proto[0].resize(K);
    b->proto[1].resize(K);
    b->passes.resize(K);
    for (int s = 0; s < 2; ++s)
        for (int l = 0; l < 10; ++l) b->choice[s][l] = -1;
    for (int g = 0; g < ngrids; ++g) {
        const int* n = shapes + 3 * g;
        const int64_t total = int64_t(n[0]) * n[1] * n[2];
        for (int a = 0; a < 3; ++a)
            if (n[a] < 16 || n[a] > 512 || (n[a] & (n[a] - 1))) {
                batch_free(b);
                return set_error(A2AFFT_E_UNSUPPORTED, "batch: grid %d extent %d (powers of two 16..512)", g, n[a]);
            }
        // the grid's virtual-rank block decomposition: P = d0*d1*d2 blocks of
        // (nx/d0) x (ny/d1) x (nz/d2), rank-major (a2adistribution.cpp:14-42)
        const int d[3] = {dims ? dims[3 * g] : 1, dims ? dims[3 * g + 1] : 1, dims ? dims[3 * g + 2] : 1};
        int lgs[3];
        for (int a = 0; a < 3; ++a) {
            if (d[a] <= 0 || n[a] % d[a]) {
                batch_free(b);
                return set_error(A2AFFT_E_DIVISIBILITY, "batch: grid %d extent %d not divisible by dims %d", g, n[a],
                                 d[a]);
            }
            lgs[a] = n[a] / d[a];
        }
        b->offsets[g + 1] = b->offsets[g] + total;
        PassMaps pm[3];
        fused_pass_maps(n, d, lgs, total / (int64_t(d[0]) * d[1] * d[2]), pm);
        for (int p = 0; p < 3; ++p)
            b->passes[size_t(p * ngrids + g)] = AxisPass{pm[p].fft, pm[p].a1, pm[p].a2, pm[p].strided, 0, 1.0};
    }
    if (const int st0 = build_protos(b); st0 < 0) {
        batch_free(b);
        return st0;
    }
    int st = batch_tables(b, 0, 0);
    if (st == A2AFFT_OK)
        st = hip_check(int(hipMalloc(&b->d_ctrl, sizeof(uint32_t) * size_t(ctrl_word(kCtrlSlots)))), "batch alloc");
    int dev = 0, cus = 0, occ = 0;
    if (st == A2AFFT_OK) st = hip_check(int(hipGetDevice(&dev)), "batch device");
    if (st == A2AFFT_OK)
        st = hip_check(int(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)), "batch CUs");
    if (st == A2AFFT_OK)
        st = hip_check(int(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_kernel<-1>, kBatchThreads, 0)),
                       "batch occupancy");
    if (st == A2AFFT_OK) st = hip_check(int(hipStreamCreate(&b->cap_stream)), "batch stream");
    if (st < 0) {
        batch_free(b);
        return st;
    }
    b->max_grid = cus * (occ > 0 ? occ : 1);
    b->grid = b->max_grid;
    *out = b;
    return A2AFFT_OK;
}

extern "C" int a2afft_batch_create(a2afft_batch* out, int ngrids, const int* shapes, a2afft_precision prec) {
    return a2afft_batch_create_ranks(out, ngrids, shapes, nullptr, prec);
}

extern "C" int a2afft_batch_offsets(a2afft_batch b, int64_t* offsets) {
    if (!b || !offsets) return set_error(A2AFFT_E_INVALID_ARG, "batch_offsets: null");
    for (int g = 0; g <= b->G; ++g) offsets[g] = b->offsets[g];
    return A2AFFT_OK;
}

static int batch_launch(a2afft_batch_s* b, void* data, int dir, hipStream_t s) {
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(1024), 0, s, b->d_ctrl, kCtrlSlots);
    A2A_TRY(hip_check(int(hipGetLastError()), "batch reset launch"));
    if (dir < 0)
        hipLaunchKernelGGL(batch_kernel<-1>, dim3(b->grid), dim3(kBatchThreads), 0, s, (float2*)data, b->d_items[0],
                           b->d_unit_item, b->d_phases, b->qs, b->d_ctrl);
    else
        hipLaunchKernelGGL(batch_kernel<1>, dim3(b->grid), dim3(kBatchThreads), 0, s, (float2*)data, b->d_items[1],
                           b->d_unit_item, b->d_phases, b->qs, b->d_ctrl);
    return hip_check(int(hipGetLastError()), "batch launch");
}

extern "C" int a2afft_batch_execute(a2afft_batch b, void* data, int direction, void* stream) {
    if (!b || !data || (direction != -1 && direction != 1))
        return set_error(A2AFFT_E_INVALID_ARG, "batch_execute: bad arguments");
    return finish_call(stream, batch_launch(b, data, direction, (hipStream_t)stream));
}

// Captures reset + persistent launch for `data` into a hipGraph on first use
// for that buffer and direction, and replays it; up to kMaxBatchGraphs
// captures are kept, so callers can rotate over several buffers.
extern "C" int a2afft_batch_graph_launch(a2afft_batch b, void* data, int direction, void* stream) {
    if (!b || !data || (direction != -1 && direction != 1))
        return set_error(A2AFFT_E_INVALID_ARG, "batch_graph_launch: bad arguments");
    const int d = direction < 0 ? 0 : 1;
    size_t k = 0;
    while (k < b->graphs.size() && !(b->graphs[k].data == data && b->graphs[k].dir == d)) ++k;
    if (k == b->graphs.size()) {
        hipGraph_t g = nullptr;
        A2A_TRY(hip_check(int(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal)), "capture"));
        int st = batch_launch(b, data, direction, b->cap_stream);
        int st2 = hip_check(int(hipStreamEndCapture(b->cap_stream, &g)), "end capture");
        if (st < 0 || st2 < 0) {
            if (g) (void)hipGraphDestroy(g);
            return st < 0 ? st : st2;
        }
        hipGraphExec_t ex = nullptr;
        st = hip_check(int(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0)), "graph instantiate");
        (void)hipGraphDestroy(g);
        if (st < 0) return st;
        if (b->graphs.size() >= kMaxBatchGraphs) {
            (void)hipGraphExecDestroy(b->graphs.front().exec);
            b->graphs.erase(b->graphs.begin());
        }
        b->graphs.push_back(a2afft_batch_s::Graph{data, d, ex});
        k = b->graphs.size() - 1;
    } else if (k + 1 != b->graphs.size()) {   // most recently used last
        const a2afft_batch_s::Graph g = b->graphs[k];
        b->graphs.erase(b->graphs.begin() + long(k));
        b->graphs.push_back(g);
        k = b->graphs.size() - 1;
    }
    return finish_call(stream, hip_check(int(hipGraphLaunch(b->graphs[k].exec, (hipStream_t)stream)), "graph launch"));
}

// 0 when the last execution finished normally, 1 when a dependency wait hit
// its bound.  Synchronises the stream.
extern "C" int a2afft_batch_status(a2afft_batch b, void* stream) {
    if (!b) return set_error(A2AFFT_E_INVALID_ARG, "batch_status: null");
    A2A_TRY(hip_check(int(hipStreamSynchronize((hipStream_t)stream)), "batch sync"));
    const int nph = b->qs.nphases;
    std::vector<uint32_t> h(size_t(ctrl_word(kCtrlDone + nph)));
    A2A_TRY(hip_check(int(hipMemcpy(h.data(), b->d_ctrl, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost)),
                      "batch status copy"));
    if (h[ctrl_word(kCtrlErr)]) return set_error(A2AFFT_E_HIP, "batch: a dependency wait timed out");
    for (int q = 0; q < nph; ++q) {
        const uint32_t need = b->pstart[size_t(q) + 1] - b->pstart[size_t(q)];
        const uint32_t done = h[ctrl_word(kCtrlDone + q)];
        if (done != need) return set_error(A2AFFT_E_HIP, "batch: phase %d finished %u of %u units", q, done, need);
    }
    return A2AFFT_OK;
}

// Persistent workgroups per launch (0 = one per resident slot on every CU),
// the queue layout (1 = one queue range per XCD, 0 = one global queue) and
// the units handed out per take.
extern "C" int a2afft_batch_set_schedule(a2afft_batch b, int workgroups, int xcd_queues, int chunk) {
    if (!b || workgroups < 0 || chunk < 1 || chunk > 64)
        return set_error(A2AFFT_E_INVALID_ARG, "batch_set_schedule: bad arguments");
    b->grid = workgroups == 0 ? b->max_grid : std::min(workgroups, b->max_grid);
    b->qs.xcd_queues = xcd_queues ? 1 : 0;
    b->qs.chunk = uint32_t(chunk);
    drop_graphs(b);
    return A2AFFT_OK;
}

// Groups of consecutive grids of at most group_points points each (0 = all
// grids in one group: every pass of every grid before the next pass), and
// the phase order (1 = skewed along the diagonals).
extern "C" int a2afft_batch_set_grouping(a2afft_batch b, int64_t group_points, int skewed) {
    if (!b || group_points < 0 || (skewed != 0 && skewed != 1))
        return set_error(A2AFFT_E_INVALID_ARG, "batch_set_grouping: bad arguments");
    return batch_tables(b, group_points, skewed);
}

namespace {

struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int create() {
        A2A_TRY(hip_check(int(hipEventCreate(&e0)), "batch autotune event"));
        return hip_check(int(hipEventCreate(&e1)), "batch autotune event");
    }
    ~EventPair() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// Mean ms of `iters` forward executions with the current schedule (after one
// untimed run); fails if a dependency wait hit its bound.
int batch_time(a2afft_batch_s* b, void* data, int iters, hipStream_t s, const EventPair& ev, float* ms) {
    A2A_TRY(batch_launch(b, data, -1, s));
    A2A_TRY(hip_check(int(hipEventRecord(ev.e0, s)), "batch autotune record"));
    for (int it = 0; it < iters; ++it) A2A_TRY(batch_launch(b, data, -1, s));
    A2A_TRY(hip_check(int(hipEventRecord(ev.e1, s)), "batch autotune record"));
    A2A_TRY(hip_check(int(hipEventSynchronize(ev.e1)), "batch autotune sync"));
    A2A_TRY(hip_check(int(hipEventElapsedTime(ms, ev.e0, ev.e1)), "batch autotune elapsed"));
    *ms /= float(iters);
    uint32_t err = 0;
    A2A_TRY(hip_check(int(hipMemcpy(&err, b->d_ctrl + ctrl_word(kCtrlErr), sizeof(err), hipMemcpyDeviceToHost)),
                      "batch autotune status"));
    return err ? set_error(A2AFFT_E_HIP, "batch autotune: a dependency wait timed out") : A2AFFT_OK;
}

struct Schedule {
    int grid, xcd_queues;
    uint32_t chunk;
    int64_t group_points;
    int skewed;
};

}  // namespace

// Times forward executions over schedules, phase orders and tile shapes and
// keeps the fastest: workgroup counts (all resident slots, 3/4, 1/2, 512,
// 256) x queue layouts x takes of 1, 2, 4 units in pass-major order, then
// groups of about one and two of the largest grids, skewed, at the best
// workgroup count, then each (length, strided) class's tile shapes (elements
// per thread, LDS padding) one class at a time.  Every candidate computes the
// same bits; `data` is overwritten.
extern "C" int a2afft_batch_autotune(a2afft_batch b, void* data, int iters, void* stream) {
    if (!b || !data || iters <= 0) return set_error(A2AFFT_E_INVALID_ARG, "batch_autotune: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    EventPair ev;
    A2A_TRY(ev.create());
    const Schedule keep{b->grid, b->qs.xcd_queues, b->qs.chunk, b->group_points, b->skewed};
    Schedule best = keep;
    float best_ms = 1e30f;
    std::string rep;
    char line[160];
    auto apply = [&](const Schedule& c) -> int {
        if (c.group_points != b->group_points || c.skewed != b->skewed)
            A2A_TRY(batch_tables(b, c.group_points, c.skewed));
        b->grid = c.grid;
        b->qs.xcd_queues = c.xcd_queues;
        b->qs.chunk = c.chunk;
        return A2AFFT_OK;
    };
    auto trial = [&](const Schedule& c) -> int {
        A2A_TRY(apply(c));
        float ms = 0.f;
        A2A_TRY(batch_time(b, data, iters, s, ev, &ms));
        snprintf(line, sizeof(line), "workgroups=%d xcd_queues=%d chunk=%u group_points=%lld skewed=%d %.4f ms\n",
                 c.grid, c.xcd_queues, c.chunk, (long long)c.group_points, c.skewed, ms);
        rep += line;
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
        return A2AFFT_OK;
    };
    std::vector<int> ws{b->max_grid, 3 * b->max_grid / 4, b->max_grid / 2, 512, 256};
    std::sort(ws.begin(), ws.end());
    ws.erase(std::unique(ws.begin(), ws.end()), ws.end());
    int st = A2AFFT_OK;
    for (int w : ws) {
        if (w <= 0 || w > b->max_grid) continue;
        for (int xq = 0; xq < 2 && st >= 0; ++xq)
            for (uint32_t chunk : {1u, 2u, 4u})
                if (st >= 0) st = trial(Schedule{w, xq, chunk, 0, 0});
    }
    int64_t nmax = 0;
    for (int g = 0; g < b->G; ++g) nmax = std::max(nmax, b->offsets[size_t(g) + 1] - b->offsets[size_t(g)]);
    const int w = best.grid;
    for (int64_t gp : {nmax, 2 * nmax}) {
        if (st < 0) break;
        if (3 * (int(group_starts(b, gp).size()) - 1) > kMaxPhases) continue;   // too many groups
        st = batch_tables(b, gp, 1);
        for (int xq = 0; xq < 2 && st >= 0; ++xq)
            for (uint32_t chunk : {1u, 2u})
                if (st >= 0) st = trial(Schedule{w, xq, chunk, gp, 1});
    }
    int keep_choice[2][10];
    std::copy(&b->choice[0][0], &b->choice[0][0] + 20, &keep_choice[0][0]);
    int st2 = apply(st >= 0 ? best : keep);
    // Tile shapes, one (N, strided) class at a time at the best schedule:
    // every shape of the class that fits one of its passes, kept if faster.
    for (int sd = 0; sd < 2 && st >= 0 && st2 >= 0; ++sd)
        for (int l = 4; l <= 9 && st >= 0 && st2 >= 0; ++l) {
            const int N = 1 << l;
            std::vector<int> cands;
            for (int i = 0; i < kNumShapes; ++i) {
                if (kShapes[i].N != N || kShapes[i].strided != sd || kShapes[i].T == 1) continue;
                bool fits = false;   // fits at least one pass of the class
                for (const AxisPass& ps : b->passes) fits = fits || shape_ok(i, ps, N);
                if (fits) cands.push_back(i);
            }
            if (cands.size() < 2) continue;
            int best_shape = b->choice[sd][l];
            for (int c : cands) {
                b->choice[sd][l] = c;
                st = build_protos(b);
                if (st >= 0) st = batch_tables(b, b->group_points, b->skewed);
                float ms = 0.f;
                if (st >= 0) st = batch_time(b, data, iters, s, ev, &ms);
                if (st < 0) break;
                snprintf(line, sizeof(line), "shape N=%d strided=%d T=%d E=%d pad=%d %.4f ms\n", N, sd, kShapes[c].T,
                         kShapes[c].E, kShapes[c].pad, ms);
                rep += line;
                if (ms < best_ms) {
                    best_ms = ms;
                    best_shape = c;
                }
            }
            b->choice[sd][l] = best_shape;
        }
    if (st < 0) std::copy(&keep_choice[0][0], &keep_choice[0][0] + 20, &b->choice[0][0]);
    if (st2 >= 0) st2 = build_protos(b);
    if (st2 >= 0) st2 = batch_tables(b, b->group_points, b->skewed);
    if (st < 0 && st2 >= 0) st2 = apply(keep);
    drop_graphs(b);
    if (st >= 0) b->report = rep;
    return st < 0 ? st : st2;
}

// Tile shape of one (length, strided) class: T lines or columns, E elements
// per thread, LDS padding.  The shape must be in the kernel's table; passes
// of the class it does not fit keep the primary (or the single-column
// fallback).  T = 0 restores the primary.
extern "C" int a2afft_batch_set_shape(a2afft_batch b, int N, int strided, int T, int E, int pad) {
    if (!b || N < 16 || N > 512 || (N & (N - 1)) || (strided != 0 && strided != 1))
        return set_error(A2AFFT_E_INVALID_ARG, "batch_set_shape: bad arguments");
    const int l = log2i(N);
    int pick = -1;
    if (T != 0) {
        for (int i = 0; i < kNumShapes && pick < 0; ++i) {
            const BShape& k = kShapes[i];
            if (k.N == N && k.strided == strided && k.T == T && k.E == E && k.pad == (pad ? 1 : 0)) pick = i;
        }
        if (pick < 0)
            return set_error(A2AFFT_E_UNSUPPORTED, "batch_set_shape: no shape N=%d strided=%d T=%d E=%d pad=%d", N,
                             strided, T, E, pad);
    }
    const int keep = b->choice[strided][l];
    b->choice[strided][l] = pick;
    int st = build_protos(b);
    if (st >= 0) st = batch_tables(b, b->group_points, b->skewed);
    if (st < 0) {
        b->choice[strided][l] = keep;
        (void)build_protos(b);
        (void)batch_tables(b, b->group_points, b->skewed);
    }
    drop_graphs(b);
    return st;
}

// The current schedule on the first line ("key=value" pairs), the tile shape
// of every (length, strided) class on the second, then the last autotune's
// timings.
extern "C" int a2afft_batch_describe(a2afft_batch b, char* buf, int len) {
    if (!b || !buf || len <= 0) return set_error(A2AFFT_E_INVALID_ARG, "batch_describe: bad arguments");
    const int n = snprintf(buf, size_t(len),
                           "grids=%d units=%u phases=%d workgroups=%d max_workgroups=%d xcd_queues=%d chunk=%u "
                           "group_points=%lld skewed=%d\n",
                           b->G, b->total, b->qs.nphases, b->grid, b->max_grid, b->qs.xcd_queues, b->qs.chunk,
                           (long long)b->group_points, b->skewed);
    std::string shapes = "shapes:";
    for (int sd = 0; sd < 2; ++sd)
        for (int l = 4; l <= 9; ++l) {
            bool used = false;
            for (const AxisPass& ps : b->passes) used = used || (int(ps.fft.n) == (1 << l) && ps.strided == sd);
            if (!used) continue;
            const int c = b->choice[sd][l];
            char t[64];
            if (c >= 0)
                snprintf(t, sizeof(t), " %d%s=T%dE%d%s", 1 << l, sd ? "s" : "c", kShapes[c].T, kShapes[c].E,
                         kShapes[c].pad ? "p" : "");
            else
                snprintf(t, sizeof(t), " %d%s=primary", 1 << l, sd ? "s" : "c");
            shapes += t;
        }
    if (n >= 0 && n < len)
        snprintf(buf + n, size_t(len - n), "%s\n%s", shapes.c_str(), b->report.c_str());
    return A2AFFT_OK;
}

extern "C" int a2afft_batch_info(a2afft_batch b, int64_t* total_units, int* workgroups, int* max_workgroups) {
    if (!b) return set_error(A2AFFT_E_INVALID_ARG, "batch_info: null");
    if (total_units) *total_units = b->total;
    if (workgroups) *workgroups = b->grid;
    if (max_workgroups) *max_workgroups = b->max_grid;
    return A2AFFT_OK;
}

extern "C" int a2afft_batch_destroy(a2afft_batch b) {
    batch_free(b);
    return A2AFFT_OK;
}
