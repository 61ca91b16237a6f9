// complex64 Stockham kernel variants.
#include "stockham.h"

namespace a2a {

int register_kernels_f32(KEntry* out) {
    int n = 0;
    add_defaults<float, 2>(out, &n);
    add_defaults<float, 4>(out, &n);
    add_defaults<float, 8>(out, &n);
    add_defaults<float, 16>(out, &n);
    add_defaults<float, 32>(out, &n);
    add_defaults<float, 64>(out, &n);
    add_defaults<float, 128>(out, &n);
    add_defaults<float, 256>(out, &n);
    add_defaults<float, 512>(out, &n);
    add_extra_flags<float, 512, 32, 16, true>(out, &n);  // the strided default shape
    add_variant<float, 512, 32, 16, true, KF_A2MAJ | KF_PAD>(out, &n);
    add_variant<float, 512, 32, 16, true, KF_A2MAJ | KF_NT>(out, &n);
    add_variant<float, 512, 32, 16, true, KF_A2MAJ | KF_PAD | KF_NT>(out, &n);
    add_variant<float, 512, 32, 16, true, KF_XCD | KF_A2MAJ | KF_PAD | KF_NT>(out, &n);
    add_defaults<float, 1024>(out, &n);
    add_defaults<float, 2048>(out, &n);
    add_defaults<float, 4096>(out, &n);
    add_mixed_all<float, 3, 5, 6, 9, 10, 12, 15, 18, 20, 24, 25, 27, 30, 36, 40, 45, 48, 50, 54, 60>(out, &n);
    // tuning candidates for the benchmark lengths
    add_shape<float, 512, 16, 16, true>(out, &n);    // 512 threads, 64 KiB LDS
    add_shape<float, 512, 16, 8, true>(out, &n);     // 256 threads, 32 KiB LDS, 64 B rows
    add_shape<float, 512, 8, 8, true>(out, &n);      // 512 threads, 32 KiB LDS
    add_shape<float, 512, 32, 32, true>(out, &n);    // 512 threads, 128 KiB LDS, 256 B rows
    add_extra_flags<float, 512, 32, 32, true>(out, &n);
    add_shape<float, 512, 8, 4, true>(out, &n);      // 256 threads, 16 KiB LDS, 32 B rows
    add_shape<float, 512, 8, 4, false>(out, &n);     // 256 threads, 16 KiB LDS
    add_shape<float, 512, 8, 2, false>(out, &n);     // 128 threads, 8 KiB LDS
    add_extra_flags<float, 512, 8, 2, false>(out, &n);
    add_shape<float, 512, 8, 1, false>(out, &n);     // 64 threads, one line
    add_variant<float, 512, 8, 1, false, KF_XCD | KF_NT>(out, &n);
    add_variant<float, 512, 8, 1, false, KF_XCD | KF_PAD | KF_NT>(out, &n);
    add_shape<float, 512, 16, 16, false>(out, &n);   // 512 threads, 64 KiB LDS
    add_extra_flags<float, 512, 16, 16, false>(out, &n);
    add_shape<float, 512, 32, 16, false>(out, &n);   // 256 threads, 64 KiB LDS
    add_shape<float, 64, 16, 64, true>(out, &n);    // 256 threads, 512 B rows
    add_shape<float, 32, 16, 128, true>(out, &n);   // 256 threads, 1 KiB rows
    add_shape<float, 128, 16, 32, true>(out, &n);   // 256 threads, 256 B rows
    add_shape<float, 256, 16, 8, true>(out, &n);
    add_shape<float, 256, 16, 32, true>(out, &n);
    add_shape<float, 256, 32, 32, true>(out, &n);
    add_shape<float, 2048, 16, 4, false>(out, &n);   // 512 threads, 64 KiB LDS
    add_shape<float, 2048, 32, 4, false>(out, &n);   // 256 threads, 64 KiB LDS
    add_extra_flags<float, 2048, 32, 4, false>(out, &n);
    add_extra_flags<float, 2048, 16, 2, false>(out, &n);   // the contiguous default shape
    add_shape<float, 2048, 16, 1, false>(out, &n);   // 128 threads, 16 KiB LDS
    add_shape<float, 2048, 32, 8, true>(out, &n);    // 512 threads, 128 KiB LDS, 64 B rows
    // far-strided 2048-point columns with half-width exchanges: 64 B rows in
    // 64 KiB of LDS (16-column tiles spill at 512 or 1024 threads)
    add_halfx<float, 2048, 32, 8, true>(out, &n);    // 512 threads
    add_halfx<float, 2048, 64, 8, true>(out, &n);    // 256 threads, 256 VGPRs
    add_halfx<float, 2048, 64, 16, true>(out, &n);   // 512 threads, 128 B rows, 128 KiB LDS
    add_variant<float, 2048, 32, 8, true, KF_A2MAJ>(out, &n);
    add_variant<float, 2048, 32, 8, true, KF_XCD | KF_A2MAJ>(out, &n);
    add_variant<float, 2048, 32, 8, true, KF_XCD | KF_NT>(out, &n);
    // half-width exchanges for the 512 and 2048 default shapes
    add_halfx<float, 512, 32, 16, true>(out, &n);
    // hardware tile order (no XCD renumbering) with half-width exchanges
    add_variant<float, 512, 32, 16, true, KF_HALFX>(out, &n);
    add_variant<float, 512, 32, 16, true, KF_NT | KF_HALFX>(out, &n);
    add_variant<float, 512, 32, 16, true, KF_PAD | KF_NT | KF_HALFX>(out, &n);
    add_variant<float, 512, 32, 16, true, KF_PAD | KF_HALFX>(out, &n);
    add_variant<float, 512, 32, 16, true, KF_A2MAJ | KF_NT | KF_HALFX>(out, &n);
    add_variant<float, 512, 32, 16, true, KF_A2MAJ | KF_PAD | KF_HALFX>(out, &n);
    add_variant<float, 512, 32, 16, true, KF_A2MAJ | KF_PAD | KF_NT | KF_HALFX>(out, &n);
    add_halfx<float, 512, 16, 16, false>(out, &n);
    add_halfx<float, 2048, 16, 2, false>(out, &n);
    // strided 1024-point columns in 16-column tiles (128 B rows; the default
    // shape reads 64 B rows)
    add_shape<float, 1024, 32, 16, true>(out, &n);   // 512 threads, 128 KiB LDS
    add_halfx<float, 1024, 32, 16, true>(out, &n);   // 512 threads, 64 KiB LDS
    add_variant<float, 1024, 32, 16, true, KF_HALFX>(out, &n);
    add_variant<float, 1024, 32, 16, true, KF_NT | KF_HALFX>(out, &n);
    add_variant<float, 1024, 32, 16, true, KF_XCD | KF_NT | KF_HALFX>(out, &n);
    add_halfx<float, 1024, 64, 16, true>(out, &n);   // 256 threads
    add_halfx<float, 1024, 64, 32, true>(out, &n);   // 512 threads, 256 B rows, 128 KiB LDS
    return n;
}

}  // namespace a2a
