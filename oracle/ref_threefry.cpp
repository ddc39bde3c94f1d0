// Exposes the reference's vendored Random123 Threefry-2x32-20
// (exec/src/gpgmp/src/cl_include/Random123/threefry.h:733) through a C symbol so the
// oracle's own restatement can be checked against it.  Compiled from the header
// where it lies under the reference tree; nothing of it is copied here.
#include <Random123/threefry.h>
#include <stdint.h>

extern "C" void ref_threefry2x32_20(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1,
                                    uint32_t *o0, uint32_t *o1) {
    threefry2x32_ctr_t c = {{c0, c1}};
    threefry2x32_key_t k = {{k0, k1}};
    threefry2x32_ctr_t r = threefry2x32(c, k);
    *o0 = r.v[0];
    *o1 = r.v[1];
}
