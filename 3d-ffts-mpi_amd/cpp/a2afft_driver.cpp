// Command-line driver with the reference test's behaviour (src/a2afft/tests/
// test.cpp:19-143): build the distribution and plan, fill the grid, dump it,
// run forward then inverse, dump after each, and write the `dims` header.
// Output files use the reference's formats so its tooling can read them:
//   dims       int32 world_size, Ng, nlocal, dims[3], local_grid_size[3]
//   proc<r>    3 x nlocal interleaved complex (input, forward, inverse)
// (proc names use the full rank number; the reference's single-character
// name breaks at 10 ranks, test.cpp:69).
//
//   a2afft_driver -Ng 32 -P 8 [-dims 1 2 4] [-blockSize 64] [-seed 1]
//                 [-exec fused|staged] [-out DIR] [-reps 10]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "a2a_3D_FFTs.hpp"

static void hip_or_die(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(2);
    }
}

int main(int argc, char** argv) {
    int Ng = 8, P = 8, blockSize = 64, reps = 0;
    unsigned long long seed = 1;
    int dims[3] = {0, 0, 0};
    bool have_dims = false;
    a2afft_exec exec = A2AFFT_EXEC_FUSED;
    std::string out = ".";
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value after %s\n", a.c_str());
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "-Ng") Ng = std::atoi(next());
        else if (a == "-P") P = std::atoi(next());
        else if (a == "-blockSize") blockSize = std::atoi(next());
        else if (a == "-seed") seed = std::strtoull(next(), nullptr, 10);
        else if (a == "-reps") reps = std::atoi(next());
        else if (a == "-out") out = next();
        else if (a == "-exec") exec = std::strcmp(next(), "staged") == 0 ? A2AFFT_EXEC_STAGED : A2AFFT_EXEC_FUSED;
        else if (a == "-dims") {
            for (int k = 0; k < 3; ++k) dims[k] = std::atoi(next());
            have_dims = true;
        } else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 1;
        }
    }

    try {
        a2aDistribution dist(P, Ng, blockSize, have_dims ? dims : nullptr);
        a2adfft dfft(dist, exec);
        const size_t n = size_t(dist.world_size) * size_t(dist.nlocal);
        std::printf("DIMS [%d,%d,%d] LOCAL_GRID_SIZE [%d,%d,%d] NLOCAL %lld RANKS %d\n", dist.dims[0], dist.dims[1],
                    dist.dims[2], dist.local_grid_size[0], dist.local_grid_size[1], dist.local_grid_size[2],
                    dist.nlocal, dist.world_size);

        fftPrecision** d_data = (fftPrecision**)std::malloc(sizeof(fftPrecision*));
        fftPrecision** d_scratch = (fftPrecision**)std::malloc(sizeof(fftPrecision*));
        hip_or_die(hipMalloc(&d_data[0], n * 2 * sizeof(fftPrecision)), "hipMalloc data");
        hip_or_die(hipMalloc(&d_scratch[0], n * 2 * sizeof(fftPrecision)), "hipMalloc scratch");
        std::vector<fftPrecision> host(n * 2);

        // seeded input, re/im uniform in [-1,1) per buffer position
        a2afft_cpp::check(a2afft_fill_random(d_data[0], a2afft_cpp::dtype_of<fftPrecision>::value, (int64_t)n, seed,
                                             0, 0, nullptr),
                          "fill");
        dist.write_dims((out + "/dims").c_str());
        std::vector<FILE*> files(P);
        for (int r = 0; r < P; ++r) {
            files[r] = std::fopen((out + "/proc" + std::to_string(r)).c_str(), "wb");
            if (!files[r]) {
                std::fprintf(stderr, "cannot open %s/proc%d\n", out.c_str(), r);
                return 1;
            }
        }
        auto dump = [&]() {
            hip_or_die(hipMemcpy(host.data(), d_data[0], n * 2 * sizeof(fftPrecision), hipMemcpyDeviceToHost),
                       "copy to host");
            for (int r = 0; r < P; ++r)
                std::fwrite(host.data() + size_t(r) * dist.nlocal * 2, sizeof(fftPrecision), dist.nlocal * 2,
                            files[r]);
        };
        dump();
        dfft.make_plans(d_scratch);
        dfft.forward(d_data);
        dump();
        dfft.inverse(d_data);
        dump();
        for (FILE* f : files) std::fclose(f);

        if (reps > 0) {
            hipEvent_t e0, e1;
            hip_or_die(hipEventCreate(&e0), "event");
            hip_or_die(hipEventCreate(&e1), "event");
            dfft.forward(d_data);
            hip_or_die(hipEventRecord(e0, nullptr), "record");
            for (int k = 0; k < reps; ++k) dfft.forward(d_data);
            hip_or_die(hipEventRecord(e1, nullptr), "record");
            hip_or_die(hipEventSynchronize(e1), "sync");
            float ms = 0.f;
            hip_or_die(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            std::printf("forward %.4f ms\n", ms / reps);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        hip_or_die(hipFree(d_data[0]), "free");
        hip_or_die(hipFree(d_scratch[0]), "free");
        std::free(d_data);
        std::free(d_scratch);
        dist.finalize();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 0;
}
