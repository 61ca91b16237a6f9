// Index maps of the six pack/unpack kernels (src/a2afft/distribution/
// kernels/a2apencils.cu), shared by the device kernels and by the CPU test
// that compares them with the oracle for every element.
#pragma once

#include "common.h"

namespace a2a {

// ---------------------------------------------------------------------------
// Index maps (one rank, scatter form: returns where source element idx goes).
// ---------------------------------------------------------------------------

enum RemapKind {
    Z_A2A_TO_PENCILS = 0,
    Z_PENCILS_TO_A2A = 1,
    X_A2A_TO_PENCILS = 2,
    X_PENCILS_TO_A2A = 3,
    Y_A2A_TO_PENCILS = 4,
    Y_PENCILS_TO_A2A = 5,
};

struct RemapParams {
    uint32_t L;      // n_cells_mini_pencils (l_axis)
    uint32_t cpr;    // n_cells_per_rank = nlocal / world_size
    uint32_t mpr;    // n_mini_pencils_per_rank = cpr / L
    uint32_t st;     // n_mini_pencils_stacked (dims[axis])
    uint32_t cps;    // n_cells_per_stack (x maps) = nlocal / dims[0]
    uint32_t rpt;    // n_ranks_per_tower (y maps) = (P / dims[1]) / dims[0]
    FastDiv dL, dcpr, dmpr, dst, dst_cpr, dL_st, dcps, dcps_L, dcpr_st, drpt_mpr,
        drpt_st_cpr, drpt_cpr, drpt;
};

template <int KIND>
A2A_HD uint32_t remap_dest(uint32_t idx, const RemapParams& p) {
    const uint32_t run = p.dL.div(idx);
    const uint32_t i = idx - run * p.L;
    if (KIND == Z_A2A_TO_PENCILS) {            // a2apencils.cu:9-15
        uint32_t rank = p.dcpr.div(idx);
        uint32_t mp = p.dmpr.mod(run);
        uint32_t mpo = p.dst.mod(rank) * p.L;
        uint32_t stack = p.dst_cpr.div(idx);
        return stack * p.st * p.cpr + mpo + mp * p.st * p.L + i;
    } else if (KIND == Z_PENCILS_TO_A2A) {     // a2apencils.cu:42-50
        uint32_t stack = p.dst_cpr.div(idx);
        uint32_t mp = p.dst.mod(run);
        uint32_t rank = stack * p.st + mp;
        uint32_t pid = p.dmpr.mod(p.dL_st.div(idx));
        return rank * p.cpr + pid * p.L + i;
    } else if (KIND == X_A2A_TO_PENCILS) {     // a2apencils.cu:78-85
        uint32_t mp = p.dcps_L.mod(run);
        uint32_t stack = p.dcps.div(idx);
        return stack * p.L + mp * p.st * p.L + i;
    } else if (KIND == X_PENCILS_TO_A2A) {     // a2apencils.cu:115-123
        uint32_t stack = p.dst.mod(run);
        uint32_t rank = p.dcpr_st.div(idx) + stack * (p.cps / p.cpr);
        uint32_t pid = p.dmpr.mod(p.dL_st.div(idx));
        return rank * p.cpr + pid * p.L + i;
    } else if (KIND == Y_A2A_TO_PENCILS) {     // a2apencils.cu:153-161
        uint32_t mp = p.drpt_mpr.mod(run);
        uint32_t stack = p.drpt_st_cpr.div(idx);
        uint32_t lst = p.dst.mod(p.drpt_cpr.div(idx));
        return mp * p.st * p.L + lst * p.L + stack * p.rpt * p.st * p.cpr + i;
    } else {                                   // Y_PENCILS_TO_A2A, a2apencils.cu:193-201
        // The reference's rank expression (:197) is
        //   mp + stack + (stack / rpt) * rpt,
        // which inverts y_a2a_to_y_pencils only when dims[1] == 2 or
        // dims[0] == 1.  This is the exact inverse for every dims, and equals
        // the reference's expression inside that domain.
        uint32_t mp = p.dst.mod(run) * p.rpt;
        uint32_t stack = p.dst.div(p.dcpr.div(idx));
        uint32_t rank = mp + p.drpt.mod(stack) + p.drpt.div(stack) * p.rpt * p.st;
        uint32_t lst = p.dmpr.mod(p.dL_st.div(idx));
        return rank * p.cpr + lst * p.L + i;
    }
}

inline RemapParams make_params(int kind, int P, int64_t nlocal, const int lgs[3], const int dims[3]) {
    RemapParams p{};
    int axis = (kind <= 1) ? 2 : (kind <= 3 ? 0 : 1);
    p.L = uint32_t(lgs[axis]);
    p.cpr = uint32_t(nlocal / P);
    p.mpr = p.cpr / p.L;
    p.st = uint32_t(dims[axis]);
    p.cps = uint32_t(nlocal / dims[0]);
    p.rpt = uint32_t((P / dims[1]) / dims[0]);
    if (p.rpt == 0) p.rpt = 1;
    p.dL.set(p.L);
    p.dcpr.set(p.cpr);
    p.dmpr.set(p.mpr ? p.mpr : 1);
    p.dst.set(p.st);
    p.dst_cpr.set(p.st * p.cpr);
    p.dL_st.set(p.L * p.st);
    p.dcps.set(p.cps);
    p.dcps_L.set(p.cps / p.L ? p.cps / p.L : 1);
    p.dcpr_st.set(p.cpr * p.st);
    p.drpt_mpr.set(p.rpt * p.mpr ? p.rpt * p.mpr : 1);
    p.drpt_st_cpr.set(p.rpt * p.st * p.cpr);
    p.drpt_cpr.set(p.rpt * p.cpr);
    p.drpt.set(p.rpt);
    return p;
}

}  // namespace a2a
