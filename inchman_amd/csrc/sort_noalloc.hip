// Allocation-free variant of the stable radix sort of (key, value) pairs on device pointers,
// for callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops): the per-tile digit histograms and digit totals live in a file-scope
// device array instead of a per-call allocation.
//
// The array is one per GPU and holds the last call's contents; every histogram word a
// pass reads is written by that pass first, so nothing needs clearing between calls.
// Calls that overlap on one GPU (different streams, or threads) share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "host.h"
#include "particles_launch.h"

using namespace gpgmp;

// keys per call at most; 4096 keys per radix tile, 256 digit counters per tile
static constexpr int kNoAllocMaxKeys = 1 << 24;
static constexpr size_t kNoAllocTiles = (size_t)kNoAllocMaxKeys / 4096;
static constexpr size_t kNoAllocWords = 256 * kNoAllocTiles + 256;  // histograms, then digit totals (4 MiB + 1 KiB)

static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

namespace {

// Device address of `scratch` per GPU.  The host-side name of a __device__ array is not a
// device address, so the address is looked up once per device, with that device current,
// and only the looked-up pointer is handed to kernels.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t *g_scratch_ptr[kMaxDevices] = {};

uint32_t *scratch_on_current_device() {
    int dev = 0;
    GPGMP_HIP_CHECK(hipGetDevice(&dev));
    GPGMP_REQUIRE(dev >= 0 && dev < kMaxDevices, GPGMP_E_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        GPGMP_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        GPGMP_REQUIRE(p != nullptr, GPGMP_E_HIP, "no device address for the sort scratch");
        g_scratch_ptr[dev] = (uint32_t *)p;
    }
    return g_scratch_ptr[dev];
}

template <typename F>
int guarded_na(F &&f) {
    try {
        f();
        return GPGMP_OK;
    } catch (const Error &e) {
        gpgmp_set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        gpgmp_set_last_error(e.what());
        return GPGMP_E_INVALID;
    }
}

}  // namespace

extern "C" int gpgmp_sort_pairs_device_noalloc(uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp,
                                               uint32_t *vals_tmp, int32_t n, int32_t key_bits, void *stream) {
    return guarded_na([&] {
        GPGMP_REQUIRE(n >= 0, GPGMP_E_INVALID, "n must be non-negative");
        GPGMP_REQUIRE(n <= kNoAllocMaxKeys, GPGMP_E_CAPACITY, "the allocation-free sort takes at most 2^24 keys");
        GPGMP_REQUIRE(key_bits >= 1 && key_bits <= 32, GPGMP_E_INVALID, "key_bits must be in [1, 32]");
        uint32_t *ws = scratch_on_current_device();
        if (n == 0) return;
        GPGMP_REQUIRE(keys && vals && keys_tmp && vals_tmp, GPGMP_E_INVALID, "null buffer");
        GPGMP_REQUIRE(radix_tiles(n) <= (int)kNoAllocTiles, GPGMP_E_CAPACITY, "too many radix tiles");
        hipStream_t s = (hipStream_t)stream;
        bool in_tmp = false;
        GPGMP_HIP_CHECK(radix_sort_pairs(keys, vals, keys_tmp, vals_tmp, nullptr, n, key_bits, ws,
                                         ws + 256 * kNoAllocTiles, s, &in_tmp));
        if (in_tmp) {  // an odd number of digit passes ends in the scratch buffers
            GPGMP_HIP_CHECK(hipMemcpyAsync(keys, keys_tmp, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
            GPGMP_HIP_CHECK(hipMemcpyAsync(vals, vals_tmp, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        }
    });
}
