// Internal definitions shared by the a2afft HIP sources.
//
// The index helpers (FastDiv, AxisMap) are __host__ __device__ so the host
// code that sizes a launch evaluates exactly the arithmetic the kernels do.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define A2A_HD __host__ __device__ __forceinline__
#else
#define A2A_HD inline
#endif

#include "../../include/a2afft.h"

namespace a2a {

// Workgroups of a grid-stride element kernel (256 threads each): a dispatch
// holds fewer than 2^32 work-items, so kernels over more elements loop.
constexpr uint64_t kStrideBlocks = uint64_t(1) << 20;

// Division of a uint32 by a runtime-invariant divisor with one mul_hi, an
// add and a shift (round-up multiplier method, Granlund & Montgomery 1994).
// Exact for every 32-bit dividend.
struct FastDiv {
    uint32_t d;
    uint32_t m;
    uint32_t s;

    FastDiv() : d(1), m(1), s(0) {}
    explicit FastDiv(uint32_t div) { set(div); }

    void set(uint32_t div) {
        d = div ? div : 1;
        s = 0;
        while ((uint64_t(1) << s) < d) ++s;
        m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1);
    }

    A2A_HD uint32_t div(uint32_t n) const {
#if defined(__HIP_DEVICE_COMPILE__)
        uint32_t t = __umulhi(n, m);
#else
        uint32_t t = uint32_t((uint64_t(n) * m) >> 32);
#endif
        return uint32_t((uint64_t(t) + n) >> s);
    }

    A2A_HD uint32_t mod(uint32_t n) const { return n - div(n) * d; }
};

// Offset of coordinate n along one axis of the virtual-rank block layout:
// (n / l) * s_rank + (n % l) * s_elem, where l is the block extent along the
// axis, s_rank the element distance between neighbouring ranks along it and
// s_elem the element stride inside a block.  The global element (x, y, z) of
// a [P][lx][ly][lz] buffer sits at off_x(x) + off_y(y) + off_z(z)
// (rank2coords / block origin, a2adistribution.cpp:14-42).
struct AxisMap {
    uint32_t l;
    uint32_t n;        // global extent along the axis
    int64_t s_rank;
    int64_t s_elem;
    FastDiv dl;

    A2A_HD int64_t off(uint32_t c) const {
        uint32_t q = dl.div(c);
        uint32_t r = c - q * l;
        return int64_t(q) * s_rank + int64_t(r) * s_elem;
    }
};

inline AxisMap make_axis(uint32_t l, uint32_t n, int64_t s_rank, int64_t s_elem) {
    AxisMap a;
    a.l = l;
    a.n = n;
    a.s_rank = s_rank;
    a.s_elem = s_elem;
    a.dl.set(l);
    return a;
}

// Error state (per host thread) and the sync-after-every-call switch.
int set_error(int status, const char* fmt, ...);
int hip_check(int hip_status, const char* what);
int finish_call(void* stream, int status);
bool sync_mode();

inline int esize_of(a2afft_dtype t) { return t == A2AFFT_C128 ? 16 : 8; }

}  // namespace a2a

#define A2A_TRY(expr)                     \
    do {                                  \
        int _st = (expr);                 \
        if (_st < 0) return _st;          \
    } while (0)
