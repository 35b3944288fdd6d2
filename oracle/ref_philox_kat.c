/*
 * Known-answer generator for Philox4x32 with 8 rounds, compiled against the
 * reference's vendored Random123 header (src/third_party/Random123/philox.h),
 * the generator STEPS wraps in src/steps/rng/r123.hpp:76. Test infrastructure:
 * prints "c0 c1 c2 c3 k0 k1 -> o0 o1 o2 o3" lines for counters/keys read from
 * stdin so that the oracle's own Philox can be checked against it.
 */
#include <stdio.h>
#include <Random123/philox.h>

int main(void) {
    unsigned c0, c1, c2, c3, k0, k1;
    while (scanf("%u %u %u %u %u %u", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        philox4x32_ctr_t ctr = {{c0, c1, c2, c3}};
        philox4x32_key_t key = {{k0, k1}};
        philox4x32_ctr_t r = philox4x32_R(8, ctr, key);
        printf("%u %u %u %u %u %u %u %u %u %u\n", c0, c1, c2, c3, k0, k1, r.v[0], r.v[1], r.v[2],
               r.v[3]);
    }
    return 0;
}
