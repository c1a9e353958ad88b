// Streaming bandwidth of one RK stage's memory traffic without its arithmetic: the
// measurement beside the stage kernels' HBM roofline (what the memory system delivers for the
// same mix of reads and writes on this device, as opposed to its published peak).
//   stage 1: reads 9 planes and writes 9 (U -> U*);
//   stage 2: reads 9 planes of U* and 9 of U and writes the 9 of U in place.
// One element per lane and plane, grid-stride loops over `cells` elements per plane, planes
// contiguous. No counterpart in the reference.
#include "mhd_abi.hpp"

#include <hip/hip_runtime.h>

#include <cstddef>

namespace mhdhip {
namespace {

constexpr int kStreamBlock = 256;
constexpr int kStreamBlocks = 16384;  // best of 1024 .. 16384 for both mixes on MI355X
constexpr int kPlanes = 9;

template <typename R>
__global__ __launch_bounds__(kStreamBlock) void k_stream1(const R* __restrict__ u,
                                                          R* __restrict__ us, long long n)
{
    const long long step = static_cast<long long>(gridDim.x) * kStreamBlock;
    for (long long i = static_cast<long long>(blockIdx.x) * kStreamBlock + threadIdx.x; i < n;
         i += step) {
#pragma unroll
        for (int c = 0; c < kPlanes; ++c) {
            us[c * n + i] = u[c * n + i];
        }
    }
}

template <typename R>
__global__ __launch_bounds__(kStreamBlock) void k_stream2(const R* __restrict__ us,
                                                          R* __restrict__ u, long long n)
{
    const long long step = static_cast<long long>(gridDim.x) * kStreamBlock;
    for (long long i = static_cast<long long>(blockIdx.x) * kStreamBlock + threadIdx.x; i < n;
         i += step) {
        R a[kPlanes], b[kPlanes];
#pragma unroll
        for (int c = 0; c < kPlanes; ++c) {
            a[c] = us[c * n + i];
            b[c] = u[c * n + i];
        }
#pragma unroll
        for (int c = 0; c < kPlanes; ++c) {
            u[c * n + i] = R(0.5) * (a[c] + b[c]);
        }
    }
}

template <typename R>
double stream_gbs(long long cells, int stage, int reps)
{
    const size_t bytes = sizeof(R) * kPlanes * static_cast<size_t>(cells);
    R* a = nullptr;
    R* b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t s = nullptr;
    double gbs = 0.0;
    auto release = [&] {
        if (s) (void)hipStreamSynchronize(s);
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    };
    try {
        check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        check(hipMalloc(reinterpret_cast<void**>(&a), bytes), "hipMalloc(stream a)");
        check(hipMalloc(reinterpret_cast<void**>(&b), bytes), "hipMalloc(stream b)");
        check(hipMemsetAsync(a, 0, bytes, s), "hipMemsetAsync");
        check(hipMemsetAsync(b, 0, bytes, s), "hipMemsetAsync");
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        auto launch = [&] {
            if (stage == 1) {
                hipLaunchKernelGGL((k_stream1<R>), dim3(kStreamBlocks), dim3(kStreamBlock), 0, s,
                                   a, b, cells);
            } else {
                hipLaunchKernelGGL((k_stream2<R>), dim3(kStreamBlocks), dim3(kStreamBlock), 0, s,
                                   a, b, cells);
            }
            check(hipGetLastError(), "stream kernel launch");
        };
        launch();
        launch();
        check(hipEventRecord(e0, s), "hipEventRecord");
        for (int r = 0; r < reps; ++r) {
            launch();
        }
        check(hipEventRecord(e1, s), "hipEventRecord");
        check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        const double moved = (stage == 1 ? 18.0 : 27.0) * sizeof(R) * static_cast<double>(cells);
        gbs = moved / (static_cast<double>(ms) / reps * 1e-3) / 1e9;
    } catch (...) {
        release();
        throw;
    }
    release();
    return gbs;
}

}  // namespace
}  // namespace mhdhip

extern "C" int mhd_hip_stream_bandwidth(int precision, long long cells, int stage, int reps,
                                        double* gbs)
{
    using namespace mhdhip;
    return guarded([&] {
        if (gbs == nullptr) {
            throw InvalidArgument("null argument");
        }
        if (cells < 1 || (stage != 1 && stage != 2) || reps < 1) {
            throw InvalidArgument("cells and reps must be positive and stage 1 or 2");
        }
        if (precision != MHD_HIP_FP64 && precision != MHD_HIP_FP32 &&
            precision != MHD_HIP_FP64_FAST) {
            throw InvalidArgument("unknown precision");
        }
        *gbs = precision == MHD_HIP_FP32 ? stream_gbs<float>(cells, stage, reps)
                                         : stream_gbs<double>(cells, stage, reps);
    });
}
