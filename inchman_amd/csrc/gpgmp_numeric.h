// Numeric contract of the GPGMP HIP path, shared by host and device code.
//
// The reference leaves several results implementation-defined (OpenCL built-in
// log, literal promotion, contraction).  This header fixes them so that integer
// molecule counts are bit-exact between this library and its CPU oracle:
//   * Threefry-2x32-20 (Salmon et al., SC'11), the generator the reference calls as
//     threefry2x32(ctr, key) (Random123/threefry.h:733, Gillespie.cl:69).
//   * u = (float)((double)(int32)w * 2^-32 + 0.5) (Gillespie.cl:72-73 with
//     M_RAN_INVM32 = 2^-32, Solver.cpp:469-471).  (int32)w * 2^-32 + 0.5 equals
//     (w ^ 2^31) / 2^32 exactly, so the float result is the round-to-nearest
//     conversion of the unsigned integer w ^ 2^31, scaled by 2^-32.
//   * log: double-precision 2*atanh((m-1)/(m+1)) series by a fixed sequence of
//     IEEE operations, rounded to float.
//   * Everything is compiled with -ffp-contract=off and IEEE division.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GPGMP_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define GPGMP_HD inline
#endif

namespace gpgmp {

GPGMP_HD uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

// Threefry-2x32, 20 rounds: rotations {13,15,26,6,17,29,16,24}, key parity
// 0x1BD11BDA, key injection s after every fourth round adds (ks[s%3], ks[(s+1)%3]+s).
struct u32x2 {
    uint32_t x, y;
};

#define GPGMP_TF_ROUND(R) \
    x0 += x1;             \
    x1 = rotl32(x1, R);   \
    x1 ^= x0;

GPGMP_HD u32x2 threefry2x32_20(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1) {
    const uint32_t k2 = 0x1BD11BDAu ^ k0 ^ k1;
    uint32_t x0 = c0 + k0;
    uint32_t x1 = c1 + k1;
    GPGMP_TF_ROUND(13) GPGMP_TF_ROUND(15) GPGMP_TF_ROUND(26) GPGMP_TF_ROUND(6)
    x0 += k1; x1 += k2 + 1u;
    GPGMP_TF_ROUND(17) GPGMP_TF_ROUND(29) GPGMP_TF_ROUND(16) GPGMP_TF_ROUND(24)
    x0 += k2; x1 += k0 + 2u;
    GPGMP_TF_ROUND(13) GPGMP_TF_ROUND(15) GPGMP_TF_ROUND(26) GPGMP_TF_ROUND(6)
    x0 += k0; x1 += k1 + 3u;
    GPGMP_TF_ROUND(17) GPGMP_TF_ROUND(29) GPGMP_TF_ROUND(16) GPGMP_TF_ROUND(24)
    x0 += k1; x1 += k2 + 4u;
    GPGMP_TF_ROUND(13) GPGMP_TF_ROUND(15) GPGMP_TF_ROUND(26) GPGMP_TF_ROUND(6)
    x0 += k2; x1 += k0 + 5u;
    return u32x2{x0, x1};
}
#undef GPGMP_TF_ROUND

// Reference counter/key convention: key {k0, 0xdecafbad + seed}, counter {n, 0xf00dcafe}.
constexpr uint32_t kKeyBase = 0xdecafbadu;
constexpr uint32_t kCtrHi = 0xf00dcafeu;

// uniform in [0,1]: the float nearest (w ^ 2^31) * 2^-32
GPGMP_HD float u01(uint32_t w) { return (float)(w ^ 0x80000000u) * 2.3283064365386963e-10f; }

// (double)(int32)w * 2^-32 + 0.5, exact in double
GPGMP_HD double u01_d(uint32_t w) { return (double)(w ^ 0x80000000u) * 2.3283064365386962890625e-10; }

// natural log in double: fixed operation sequence (see file header)
GPGMP_HD double log_d(double y) {
    if (!(y > 0.0)) return (y == 0.0) ? -__builtin_inf() : __builtin_nan("");
    if (y == __builtin_inf()) return y;
    // y = m * 2^e, m in [1,2): exact bit manipulation of a normal double
    uint64_t b;
    __builtin_memcpy(&b, &y, 8);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) {  // subnormal: scale into the normal range first
        y = y * 18014398509481984.0;  // 2^54
        __builtin_memcpy(&b, &y, 8);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    b = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    __builtin_memcpy(&m, &b, 8);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 15.0;
    p = 1.0 / 13.0 + z * p;
    p = 1.0 / 11.0 + z * p;
    p = 1.0 / 9.0 + z * p;
    p = 1.0 / 7.0 + z * p;
    p = 1.0 / 5.0 + z * p;
    p = 1.0 / 3.0 + z * p;
    p = z * p;
    const double two_s = 2.0 * s;
    return (double)e * 0.6931471805599453094 + (two_s + two_s * p);
}

GPGMP_HD float log_f(float y) { return (float)log_d((double)y); }

// log_f for y >= 1 or y = +inf (the SSA's log(1/r1), r1 in [0, 1]) without branches: the
// same operation sequence as log_d on a normal y (no subnormal scaling is needed at
// y >= 1), the reduction m > sqrt(2) as a select, and +inf selected at the end -- the
// same bits as log_f on that domain, with nothing to diverge on.
GPGMP_HD float log_f_ge1(float yf) {
    const double y = (double)yf;
    uint64_t b;
    __builtin_memcpy(&b, &y, 8);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    b = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    __builtin_memcpy(&m, &b, 8);
    const bool big = m > 1.4142135623730951;
    m = big ? m * 0.5 : m;
    e = big ? e + 1 : e;
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 15.0;
    p = 1.0 / 13.0 + z * p;
    p = 1.0 / 11.0 + z * p;
    p = 1.0 / 9.0 + z * p;
    p = 1.0 / 7.0 + z * p;
    p = 1.0 / 5.0 + z * p;
    p = 1.0 / 3.0 + z * p;
    p = z * p;
    const double two_s = 2.0 * s;
    const float r = (float)((double)e * 0.6931471805599453094 + (two_s + two_s * p));
    return (y == __builtin_inf()) ? __builtin_inff() : r;
}

// exp(x) in double: x = k ln2 + r with |r| <= ln2/2 (ln2 split in two for an exact
// k*ln2_hi), Taylor series of exp(r) to r^14 (truncation below 1e-17), scaled by 2^k.
// IEEE add/mul/div only, so CPU and GPU builds give the same bits.  |x| <= 700.
GPGMP_HD double exp_d(double x) {
    const double k = __builtin_rint(x * 1.4426950408889634074);
    const double r = (x - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
    double p = 1.0 / 87178291200.0;
    p = 1.0 / 6227020800.0 + r * p;
    p = 1.0 / 479001600.0 + r * p;
    p = 1.0 / 39916800.0 + r * p;
    p = 1.0 / 3628800.0 + r * p;
    p = 1.0 / 362880.0 + r * p;
    p = 1.0 / 40320.0 + r * p;
    p = 1.0 / 5040.0 + r * p;
    p = 1.0 / 720.0 + r * p;
    p = 1.0 / 120.0 + r * p;
    p = 1.0 / 24.0 + r * p;
    p = 1.0 / 6.0 + r * p;
    p = 0.5 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    return __builtin_ldexp(p, (int)k);
}

// exp(x) - 1 for x >= 0 without cancellation: Taylor series below ln2/2 (to x^16),
// exp_d(x) - 1 above.
GPGMP_HD double expm1_d(double x) {
    if (x >= 0.34657359027997264) return exp_d(x) - 1.0;
    double p = 1.0 / 355687428096000.0;
    p = 1.0 / 20922789888000.0 + x * p;
    p = 1.0 / 1307674368000.0 + x * p;
    p = 1.0 / 87178291200.0 + x * p;
    p = 1.0 / 6227020800.0 + x * p;
    p = 1.0 / 479001600.0 + x * p;
    p = 1.0 / 39916800.0 + x * p;
    p = 1.0 / 3628800.0 + x * p;
    p = 1.0 / 362880.0 + x * p;
    p = 1.0 / 40320.0 + x * p;
    p = 1.0 / 5040.0 + x * p;
    p = 1.0 / 720.0 + x * p;
    p = 1.0 / 120.0 + x * p;
    p = 1.0 / 24.0 + x * p;
    p = 1.0 / 6.0 + x * p;
    p = 0.5 + x * p;
    return x + x * (x * p);
}

// Admissible time step of one cell for a particle-based species
// (computeIndividualTimestepX/Y, StochasticInhomogeneousDriftDiffusion.cl:130-151): the
// formula in double with tanh(a) = E2/(E2+2), E2 = expm1(2a), and sinh(a) = (E + E/(E+1))/2,
// E = expm1(a), from expm1_d (so the CPU and GPU evaluate the same bits), rounded to
// float.  An empty cell (D = r = 0) gives +inf.
GPGMP_HD float ind_cell_dt(float Df, float rf, float hf) {
    const double D = Df, r = rf, h = hf;
    if (r == 0.0) return (float)(h * h / (2.0 * D) * (1.0 - 2.0 / 3.0));
    if (D == 0.0) return (float)(h / __builtin_fabs(r));
    const double eps = h * r / (2.0 * D);
    const double a = __builtin_fabs(eps);
    double th = 1.0, ish2 = 0.0;
    if (!(a > 350.0)) {
        const double e2 = expm1_d(2.0 * a), e1 = expm1_d(a);
        th = e2 / (e2 + 2.0);
        const double sh = (e1 + e1 / (e1 + 1.0)) * 0.5;
        ish2 = 1.0 / (sh * sh);
    }
    if (eps < 0.0) th = -th;
    double temp = 1.0 + (-1.0 / (th * eps) + ish2);
    if (temp <= 0.0) temp = 1.0 / 3.0;
    return (float)(h / r * th * temp);
}

// log(u) in float for a normal u > 0: u = 2^e m with m in (sqrt(1/2), sqrt(2)], f = m - 1
// (exact), s = f / (2 + f), log(m) = 2 atanh(s) = 2 (s + s^3/3 + ... + s^9/9) (|s| < 0.172:
// truncation below 4e-10), Horner in s^2 with fused multiply-adds.  IEEE operations only, so
// the CPU and the GPU give the same bits.
GPGMP_HD float log_pf(float u) {
    uint32_t b;
    __builtin_memcpy(&b, &u, 4);
    int e = (int)((b >> 23) & 0xffu) - 127;
    b = (b & 0x007fffffu) | 0x3f800000u;
    float m;
    __builtin_memcpy(&m, &b, 4);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float p = 1.0f / 9.0f;
    p = __builtin_fmaf(z, p, 1.0f / 7.0f);
    p = __builtin_fmaf(z, p, 1.0f / 5.0f);
    p = __builtin_fmaf(z, p, 1.0f / 3.0f);
    const float two_s = 2.0f * s;
    return __builtin_fmaf((float)e, 0.693147182f, __builtin_fmaf(two_s, z * p, two_s));
}

// sin(pi t2) in float for t2 in [0,2): folded to r in [0,1/2], x = pi r, Taylor series to
// x^13 (truncation below 7e-10) in Horner form with fused multiply-adds.
GPGMP_HD float sinpi_02(float t2) {
    float sign = 1.0f;
    if (t2 >= 1.0f) {
        sign = -1.0f;
        t2 = t2 - 1.0f;
    }
    const float r = (t2 > 0.5f) ? (1.0f - t2) : t2;
    const float x = 3.14159274f * r;
    const float z = x * x;
    float p = 1.0f / 6227020800.0f;
    p = __builtin_fmaf(z, p, -1.0f / 39916800.0f);
    p = __builtin_fmaf(z, p, 1.0f / 362880.0f);
    p = __builtin_fmaf(z, p, -1.0f / 5040.0f);
    p = __builtin_fmaf(z, p, 1.0f / 120.0f);
    p = __builtin_fmaf(z, p, -1.0f / 6.0f);
    return sign * __builtin_fmaf(x, z * p, x);
}

// sin(pi t) in float: t reduced to t - 2 floor(t / 2) in [0,2), then sinpi_02.
GPGMP_HD float sinpi_pf(float t) { return sinpi_02(t - 2.0f * __builtin_floorf(t * 0.5f)); }

// sinpi_pf(t) for 0 <= t < 4, bit for bit, with the reduction as one compare: below 2 the
// floor is 0, and from 2 on t - 2 is exact (t and 2 within a factor of two).
GPGMP_HD float sinpi_pf4(float t) { return sinpi_02((t >= 2.0f) ? t - 2.0f : t); }

// Gaussian pair by the Box-Muller transform of one Threefry call (key {id, seed}, counter
// {step, 0}), in float: u1 = (w.x / 2^9 + 1/2) 2^-23 and u2 likewise from w.y, in (0, 1)
// and exact, g_x = sqrt(-2 ln u1) cos(2 pi u2), g_y = sqrt(-2 ln u1) sin(2 pi u2), with
// cos(2 pi u2) = sin(pi (2 u2 + 1/2)); no rejection loop, so the lanes of a wave never wait
// for each other.  The largest |g| is sqrt(48 ln 2) = 5.77.
GPGMP_HD void gauss2_f(uint32_t id, uint32_t seed, uint32_t step, float *gx, float *gy) {
    const u32x2 w = threefry2x32_20(step, 0u, id, seed);
    const float u1 = ((float)(w.x >> 9) + 0.5f) * 0x1p-23f;
    const float u2 = ((float)(w.y >> 9) + 0.5f) * 0x1p-23f;
    const float r = __builtin_sqrtf(-2.0f * log_pf(u1));
    *gx = r * sinpi_pf4(2.0f * u2 + 0.5f);  // arguments in (0, 2.5) and (0, 2)
    *gy = r * sinpi_pf4(2.0f * u2);
}

}  // namespace gpgmp
