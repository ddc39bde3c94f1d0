// First moment along x of one species' counts (inhomogeneous_computeMomentsX +
// inhomogeneous_reduceMomentsBlocks, StochasticInhomogeneousDriftDiffusion.cl:5-82): the sums
// of x * n(x, y) and of n(x, y) over the grid, x the cell column.  The reference adds float
// products in a per-work-group tree; here the sums are exact (64-bit integers: the counts
// are integers), so they do not depend on the grid shape or launch.  HBM-bound: 4 bytes
// read per cell.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gpgmp_hip.h"
#include "host.h"

namespace gpgmp {

__global__ void k_moments_zero(unsigned long long *sums) {
    if (threadIdx.x < 2) sums[threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(256) void k_moments_x(const int32_t *__restrict__ counts, int W, int H,
                                                   unsigned long long *__restrict__ sums) {
    long long m = 0, t = 0;
    const int A = W * H;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A; c += gridDim.x * blockDim.x) {
        const long long n = counts[c];
        m += (long long)(c % W) * n;
        t += n;
    }
    for (int o = 32; o > 0; o >>= 1) {
        m += __shfl_down(m, o, 64);
        t += __shfl_down(t, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&sums[0], (unsigned long long)m);
        atomicAdd(&sums[1], (unsigned long long)t);
    }
}

}  // namespace gpgmp

using namespace gpgmp;

extern "C" int gpgmp_moments_x_device(const int32_t *state, int32_t W, int32_t H, int32_t species,
                                      int64_t *sums, void *stream) {
    try {
        GPGMP_REQUIRE(state && sums, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(W >= 1 && H >= 1 && species >= 0, GPGMP_E_INVALID, "bad grid or species");
        GPGMP_REQUIRE((int64_t)W * H < ((int64_t)1 << 31), GPGMP_E_CAPACITY, "grid too large");
        const hipStream_t s = (hipStream_t)stream;
        const int A = W * H;
        unsigned long long *out = reinterpret_cast<unsigned long long *>(sums);
        hipLaunchKernelGGL(k_moments_zero, dim3(1), dim3(64), 0, s, out);
        const int blocks = (A + 255) / 256 < 2048 ? (A + 255) / 256 : 2048;
        hipLaunchKernelGGL(k_moments_x, dim3(blocks), dim3(256), 0, s, state + (size_t)species * A, W, H, out);
        GPGMP_HIP_CHECK(hipGetLastError());
        return GPGMP_OK;
    } catch (const Error &e) {
        gpgmp_set_last_error(e.what());
        return e.code;
    }
}
