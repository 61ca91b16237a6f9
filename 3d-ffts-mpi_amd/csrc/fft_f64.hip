// complex128 Stockham kernel variants.
#include "stockham.h"

namespace a2a {

int register_kernels_f64(KEntry* out) {
    int n = 0;
    add_defaults<double, 2>(out, &n);
    add_defaults<double, 4>(out, &n);
    add_defaults<double, 8>(out, &n);
    add_defaults<double, 16>(out, &n);
    add_defaults<double, 32>(out, &n);
    add_defaults<double, 64>(out, &n);
    add_defaults<double, 128>(out, &n);
    add_defaults<double, 256>(out, &n);
    add_defaults<double, 512>(out, &n);
    add_extra_flags<double, 512, 16, 8, true>(out, &n);   // the strided default shape
    add_defaults<double, 1024>(out, &n);
    add_defaults<double, 2048>(out, &n);
    add_defaults<double, 4096>(out, &n);
    add_mixed_all<double, 3, 5, 6, 9, 10, 12, 15, 18, 20, 24, 25, 27, 30>(out, &n);   // a line fits in registers
    // tuning candidates for the benchmark lengths
    add_extra_flags<double, 256, 16, 8, true>(out, &n);    // the strided default shape
    add_extra_flags<double, 256, 16, 16, false>(out, &n);  // the contiguous default shape
    add_shape<double, 256, 16, 16, true>(out, &n);   // 256 threads, 64 KiB LDS
    add_shape<double, 256, 16, 4, true>(out, &n);    // 64 threads, 16 KiB LDS
    add_shape<double, 256, 32, 16, true>(out, &n);   // 128 threads
    add_shape<double, 256, 16, 8, false>(out, &n);   // 128 threads, 32 KiB LDS
    add_shape<double, 256, 32, 16, false>(out, &n);  // 128 threads, 64 KiB LDS
    add_shape<double, 512, 8, 8, true>(out, &n);     // 512 threads
    add_shape<double, 512, 8, 4, true>(out, &n);     // 256 threads, 32 KiB LDS
    add_shape<double, 512, 8, 4, false>(out, &n);
    add_shape<double, 512, 16, 4, false>(out, &n);
    // few-line contiguous tiles (lines that cross rank blocks)
    add_shape<double, 512, 8, 2, false>(out, &n);    // 128 threads
    add_extra_flags<double, 512, 8, 2, false>(out, &n);
    add_halfx<double, 512, 8, 2, false>(out, &n);
    add_shape<double, 256, 16, 4, false>(out, &n);   // 64 threads
    add_extra_flags<double, 256, 16, 4, false>(out, &n);
    add_halfx<double, 256, 16, 4, false>(out, &n);
    // half-width exchanges: the 152-VGPR FP64 kernels fit 3 waves a SIMD, a
    // full complex128 exchange leaves room for 2
    add_halfx<double, 256, 16, 8, false>(out, &n);
    add_halfx<double, 256, 16, 16, false>(out, &n);
    add_halfx<double, 256, 16, 8, true>(out, &n);
    add_halfx<double, 256, 16, 4, true>(out, &n);
    add_halfx<double, 512, 16, 8, false>(out, &n);
    add_halfx<double, 512, 16, 8, true>(out, &n);
    // strided 1024-point columns in 8-column tiles (128 B rows; the default
    // shape reads 64 B rows)
    add_shape<double, 1024, 16, 8, true>(out, &n);   // 512 threads, 128 KiB LDS
    add_halfx<double, 1024, 16, 8, true>(out, &n);   // 512 threads, 64 KiB LDS
    add_variant<double, 1024, 16, 8, true, KF_XCD | KF_NT | KF_HALFX>(out, &n);
    add_halfx<double, 1024, 32, 8, true>(out, &n);   // 256 threads
    // strided 2048-point columns on 64-128 B rows (the default shape reads 32 B)
    add_halfx<double, 2048, 32, 8, true>(out, &n);   // 512 threads, 128 KiB LDS
    add_halfx<double, 2048, 16, 4, true>(out, &n);   // 512 threads, 64 KiB LDS
    add_halfx<double, 2048, 32, 4, true>(out, &n);   // 256 threads, 64 KiB LDS
    return n;
}

}  // namespace a2a
