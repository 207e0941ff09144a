// measure.hip -- a plain device copy of a known byte count, the calibration
// kernel for HBM byte counters (FETCH_SIZE / WRITE_SIZE): the counters of this
// kernel divided by its known bytes give the scale applied to the counters of
// the Monte Carlo kernels.
#include <hip/hip_runtime.h>

#include "ucmc_core.hpp"

namespace ucmc {
namespace {

__global__ __launch_bounds__(256) void ucmc_copy_kernel(const uint4* __restrict__ a, uint4* __restrict__ b,
                                                        std::size_t n)
{
    for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<std::size_t>(gridDim.x) * blockDim.x)
        b[i] = a[i];
}

}  // namespace
}  // namespace ucmc

extern "C" int ucmc_measure_copy(size_t bytes, int repeats, double* gbps)
{
    using namespace ucmc;
    try {
        if (bytes < 16 || repeats < 1 || !gbps) fail(UCMC_E_INVALID_ARGUMENT, "bad argument");
        const std::size_t n = bytes / 16;
        DevBuf a, b;
        a.alloc(n * 16);
        b.alloc(n * 16);
        UCMC_HIP(hipMemset(a.p, 1, n * 16));
        EventPair ev;
        const dim3 grid(8192), block(256);
        hipLaunchKernelGGL(ucmc_copy_kernel, grid, block, 0, nullptr, a.as<uint4>(), b.as<uint4>(), n);
        UCMC_HIP(hipGetLastError());
        UCMC_HIP(hipEventRecord(ev.e0, nullptr));
        for (int r = 0; r < repeats; ++r)
            hipLaunchKernelGGL(ucmc_copy_kernel, grid, block, 0, nullptr, a.as<uint4>(), b.as<uint4>(), n);
        UCMC_HIP(hipEventRecord(ev.e1, nullptr));
        UCMC_HIP(hipEventSynchronize(ev.e1));
        float ms = 0.f;
        UCMC_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
        *gbps = 2.0 * static_cast<double>(n * 16) * repeats / (ms * 1e-3) / 1e9;
        return UCMC_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    }
}
