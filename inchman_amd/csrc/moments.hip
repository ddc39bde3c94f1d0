// First moment along x of one species' counts (inhomogeneous_computeMomentsX +
// inhomogeneous_reduceMomentsBlocks, StochasticInhomogeneousDriftDiffusion.cl:5-82): the sums
// of x * n(x, y) and of n(x, y) over the grid, x the cell column.  The reference adds float
// products in a per-work-group tree; here the sums are exact (64-bit integers: the counts
// are integers), so they do not depend on the grid shape or launch.  HBM-bound: 4 bytes
// read per cell (16-byte row loads when the width is a multiple of 4).
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

// Rows of W4 = W/4 int4 words (W % 4 == 0, 16-byte aligned counts): blocks stride over
// rows, threads over a row's words, so the column of every count is known without a
// division and each load is 16 bytes.
__global__ __launch_bounds__(256) void k_moments_x4(const int4 *__restrict__ counts, int W4, int H,
                                                    unsigned long long *__restrict__ sums) {
    __shared__ long long part[2][4];
    long long m = 0, t = 0;
    for (int y = blockIdx.x; y < H; y += gridDim.x) {
        const int4 *row = counts + (size_t)y * W4;
#pragma unroll 4
        for (int j = threadIdx.x; j < W4; j += blockDim.x) {
            const int4 v = row[j];
            const long long x = 4ll * j;
            m += x * v.x + (x + 1) * v.y + (x + 2) * v.z + (x + 3) * v.w;
            t += (long long)v.x + v.y + v.z + v.w;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        m += __shfl_down(m, o, 64);
        t += __shfl_down(t, o, 64);
    }
    // one pair of atomics per block (the waves' sums meet in LDS)
    const int wave = (int)(threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0) {
        part[0][wave] = m;
        part[1][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long ms = 0, ts = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            ms += part[0][w];
            ts += part[1][w];
        }
        atomicAdd(&sums[0], (unsigned long long)ms);
        atomicAdd(&sums[1], (unsigned long long)ts);
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
        const int32_t *counts = state + (size_t)species * A;
        if (W % 4 == 0 && ((uintptr_t)counts & 15u) == 0) {
            const int blocks = H < 1024 ? H : 1024;
            hipLaunchKernelGGL(k_moments_x4, dim3(blocks), dim3(256), 0, s, reinterpret_cast<const int4 *>(counts),
                               W / 4, H, out);
        } else {
            const int blocks = (A + 255) / 256 < 2048 ? (A + 255) / 256 : 2048;
            hipLaunchKernelGGL(k_moments_x, dim3(blocks), dim3(256), 0, s, counts, W, H, out);
        }
        GPGMP_HIP_CHECK(hipGetLastError());
        return GPGMP_OK;
    } catch (const Error &e) {
        gpgmp_set_last_error(e.what());
        return e.code;
    }
}
