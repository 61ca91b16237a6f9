// Host-side interface of the remap kernels (remap.hip).
#pragma once

#include "common.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

namespace a2a {

// kind: 0 z_a2a_to_z_pencils, 1 z_pencils_to_z_a2a, 2 x_a2a_to_x_pencils,
//       3 x_pencils_to_x_a2a, 4 y_a2a_to_y_pencils, 5 y_pencils_to_y_a2a.
int launch_remap(int kind, const void* src, void* dst, int world_size, int64_t nlocal, const int lgs[3],
                 const int dims[3], a2afft_dtype dtype, int nbuffers, hipStream_t s);
// which: 0 fast_z_to_x, 1 fast_x_to_y, 2 fast_y_to_z.
int launch_transpose(int which, const void* src, void* dst, const int lgs[3], int64_t nlocal, a2afft_dtype dtype,
                     int nbuffers, hipStream_t s);
int launch_alltoall(const void* send, void* recv, int world_size, int64_t nlocal, a2afft_dtype dtype, hipStream_t s);
int launch_copy(void* out, const void* in, int64_t n, a2afft_dtype dtype, hipStream_t s);
int launch_scale(void* out, const void* in, int64_t n, double scale, a2afft_dtype dtype, hipStream_t s);

}  // namespace a2a
