// log_f_ge1 against log_f on the waiting-time domain y = 1/r1, r1 = u01(w): every
// 1021st word w, the words next to the edges of u01, and y = 1, 2^32 and +inf.
#include <cstdio>
#include <cstring>

#include "../../inchman_amd/csrc/gpgmp_numeric.h"

using namespace gpgmp;

static bool same(float a, float b) { return std::memcmp(&a, &b, sizeof a) == 0; }

int main() {
    long n = 0, bad = 0;
    auto check = [&](float y) {
        ++n;
        if (!same(log_f(y), log_f_ge1(y))) {
            if (bad < 5) std::printf("mismatch at y=%a: %a vs %a\n", y, log_f(y), log_f_ge1(y));
            ++bad;
        }
    };
    for (uint64_t w = 0; w <= 0xffffffffull; w += 1021) check(1.f / u01((uint32_t)w));
    for (uint32_t d = 0; d < 4096; ++d) {
        check(1.f / u01(0x80000000u + d));  // r1 near 0 (y near 2^32, and +inf at d = 0)
        check(1.f / u01(0x7fffffffu - d));  // r1 near 1
    }
    check(1.f);
    check(4294967296.f);
    check(__builtin_inff());
    std::printf("%ld values, %ld mismatches\n", n, bad);
    return bad == 0 ? 0 : 1;
}
