// Wave-wide (64-lane) integer scans and sums over the DPP lane moves of gfx9 (row shifts
// within 16-lane rows, then the row_bcast:15 / row_bcast:31 broadcasts across rows), so
// a scan is six VALU adds with no LDS round trip (the __shfl_* forms go through
// ds_bpermute).  Integer addition is exact, so any order gives the same result.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpgmp {

// DPP controls (GFX9 encoding)
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

// Inclusive scan over the 64 lanes (every lane active).  A lane whose DPP source lies
// outside its row, or whose row is masked off, adds the `old` operand 0.
__device__ __forceinline__ int32_t wave_incl_scan_i32(int32_t v) {
    v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr1, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr2, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr4, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr8, 0xf, 0xf, false);
    // rows 1 and 3 add the last lane of rows 0 and 2; rows 2 and 3 then add lane 31
    v += __builtin_amdgcn_update_dpp(0, v, kDppRowBcast15, 0xa, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, kDppRowBcast31, 0xc, 0xf, false);
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
    return (uint32_t)wave_incl_scan_i32((int32_t)v);
}

// Sum over the 64 lanes, the same value in every lane.
__device__ __forceinline__ int32_t wave_sum_i32(int32_t v) {
    return __builtin_amdgcn_readlane(wave_incl_scan_i32(v), 63);
}

}  // namespace gpgmp
