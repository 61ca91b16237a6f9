// C++ mirror of the reference's public classes over the a2afft C ABI.
//
// Source-compatible with the reference's use of a2aDistribution and a2adfft
// (src/a2afft/a2a_3D_FFTs.h:1, src/a2afft/distribution/a2adistribution.h:5-42,
// src/a2afft/dfft/a2adfft.h:9-30): same class names, public fields, method
// names and `fftPrecision**` buffer arguments (only [0] is used, as in the
// reference).  The MPI communicator becomes a virtual-rank count: one object
// describes all `world_size` ranks, whose blocks sit side by side in one
// device allocation of world_size * nlocal complex values.  Preconditions
// the reference asserts throw std::runtime_error instead.
#pragma once

#include <stdexcept>
#include <string>

#include "a2afft.h"

#ifndef fftPrecision
#define fftPrecision double
#endif

namespace a2afft_cpp {

inline void check(int status, const char* where) {
    if (status < 0)
        throw std::runtime_error(std::string(where) + ": " + a2afft_status_string(status) + " (" +
                                 a2afft_last_error() + ")");
}

template <typename T> struct dtype_of;
template <> struct dtype_of<double> { static constexpr a2afft_dtype value = A2AFFT_C128; };
template <> struct dtype_of<float> { static constexpr a2afft_dtype value = A2AFFT_C64; };

}  // namespace a2afft_cpp

class a2aDistribution {
public:
    int ndims = 3;
    int Ng = 0;
    long long nlocal = 0;
    int world_size = 0;
    int world_rank = 0;   // the rank whose coords / start are in the fields below
    int local_grid_size[3] = {0, 0, 0};
    int dims[3] = {0, 0, 0};
    int coords[3] = {0, 0, 0};
    int local_coordinates_start[3] = {0, 0, 0};
    int blockSize = 64;
    a2afft_dist handle = nullptr;
    void* stream = nullptr;   // hipStream_t; null = the null stream

    // a2aDistribution(MPI_Comm, Ng, blockSize) (a2adistribution.cpp:44-113)
    // with the communicator replaced by `world_size` virtual ranks; `dims`
    // null applies the MPI_Dims_create rule.
    a2aDistribution(int input_world_size, int input_Ng, int input_blockSize, const int* input_dims = nullptr) {
        a2afft_cpp::check(a2afft_dist_create(&handle, input_Ng, input_world_size, input_dims, input_blockSize,
                                             a2afft_cpp::dtype_of<fftPrecision>::value),
                          "a2aDistribution");
        a2afft_dist_info info;
        a2afft_cpp::check(a2afft_dist_get_info(handle, &info), "a2aDistribution");
        Ng = info.Ng;
        nlocal = info.nlocal;
        world_size = info.world_size;
        blockSize = info.blockSize;
        for (int a = 0; a < 3; ++a) {
            local_grid_size[a] = info.local_grid_size[a];
            dims[a] = info.dims[a];
        }
        set_rank(0);
    }

    a2aDistribution(const a2aDistribution&) = delete;
    a2aDistribution& operator=(const a2aDistribution&) = delete;
    ~a2aDistribution() { finalize(); }

    // Point world_rank / coords / local_coordinates_start at virtual rank r.
    void set_rank(int r) {
        a2afft_cpp::check(a2afft_dist_rank_coords(handle, r, coords, local_coordinates_start), "set_rank");
        world_rank = r;
    }

    void write_dims(const char* path) const { a2afft_cpp::check(a2afft_dist_write_dims(handle, path), "dims"); }

    void getZPencils(fftPrecision** d_Buff1, fftPrecision** d_Buff2) {
        a2afft_cpp::check(a2afft_get_z_pencils(handle, d_Buff1[0], d_Buff2[0], stream), "getZPencils");
    }
    void returnZPencils(fftPrecision** d_Buff1, fftPrecision** d_Buff2) {
        a2afft_cpp::check(a2afft_return_z_pencils(handle, d_Buff1[0], d_Buff2[0], stream), "returnZPencils");
    }
    void getXPencils(fftPrecision** d_Buff1, fftPrecision** d_Buff2) {
        a2afft_cpp::check(a2afft_get_x_pencils(handle, d_Buff1[0], d_Buff2[0], stream), "getXPencils");
    }
    void returnXPencils(fftPrecision** d_Buff1, fftPrecision** d_Buff2) {
        a2afft_cpp::check(a2afft_return_x_pencils(handle, d_Buff1[0], d_Buff2[0], stream), "returnXPencils");
    }
    void getYPencils(fftPrecision** d_Buff1, fftPrecision** d_Buff2) {
        a2afft_cpp::check(a2afft_get_y_pencils(handle, d_Buff1[0], d_Buff2[0], stream), "getYPencils");
    }
    void returnYPencils(fftPrecision** d_Buff1, fftPrecision** d_Buff2) {
        a2afft_cpp::check(a2afft_return_y_pencils(handle, d_Buff1[0], d_Buff2[0], stream), "returnYPencils");
    }

    // a2aDistribution::finalize (a2adistribution.cpp:217-223)
    void finalize() {
        if (handle) a2afft_dist_destroy(handle);
        handle = nullptr;
    }
};

class a2adfft {
public:
    int Ng;
    long long nlocal;
    int world_size;
    int blockSize;
    a2aDistribution& distribution;
    fftPrecision** d_scratch = nullptr;
    int plans_made = 0;
    a2afft_plan plan = nullptr;

    // a2adfft(a2aDistribution&) (a2adfft.cpp:7-16); the execution strategy
    // defaults to the fused three-pass transform.
    explicit a2adfft(a2aDistribution& dist, a2afft_exec exec = A2AFFT_EXEC_FUSED)
        : Ng(dist.Ng), nlocal(dist.nlocal), world_size(dist.world_size), blockSize(dist.blockSize),
          distribution(dist) {
        const a2afft_precision prec = a2afft_cpp::dtype_of<fftPrecision>::value == A2AFFT_C128 ? A2AFFT_FP64
                                                                                                : A2AFFT_FP32;
        a2afft_cpp::check(a2afft_plan_create(&plan, dist.handle, prec, exec), "a2adfft");
    }

    a2adfft(const a2adfft&) = delete;
    a2adfft& operator=(const a2adfft&) = delete;
    ~a2adfft() {
        if (plan) a2afft_plan_destroy(plan);
    }

    // a2adfft::make_plans (a2adfft.cpp:18-24)
    void make_plans(fftPrecision** input_d_scratch) {
        d_scratch = input_d_scratch;
        a2afft_cpp::check(a2afft_make_plans(plan, input_d_scratch ? input_d_scratch[0] : nullptr), "make_plans");
        plans_made = 1;
    }

    // a2adfft::forward / inverse (a2adfft.cpp:26-158)
    void forward(fftPrecision** d_data) {
        a2afft_cpp::check(a2afft_forward(plan, d_data[0], distribution.stream), "forward");
    }
    void inverse(fftPrecision** d_data) {
        a2afft_cpp::check(a2afft_inverse(plan, d_data[0], distribution.stream), "inverse");
    }
};
