/* Compares cffc_pow_glibc (cffc_amd/csrc/glibc_pow.h) with this host's libm
 * pow() bit for bit on seeded random inputs.
 *   glibc_pow_check N LO HI Y  -> x ~ U[LO, HI), exponent Y
 *   glibc_pow_check N LO HI    -> x ~ U[LO, HI), exponent ~ U[-3, 3)
 * Prints "samples mismatches" and the first few mismatches. */
#include <stdio.h>
#include <stdlib.h>
#include "../../cffc_amd/csrc/glibc_pow.h"

static uint64_t state = 0x9e3779b97f4a7c15ULL;
static double urand(void) {            /* xorshift64, 53-bit uniform in [0, 1) */
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return (double)(state >> 11) * 0x1p-53;
}

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    const long n = atol(argv[1]);
    const double lo = atof(argv[2]), hi = atof(argv[3]);
    const int fixed = argc > 4;
    const double yf = fixed ? atof(argv[4]) : 0.0;
    long bad = 0;
    for (long j = 0; j < n; j++) {
        const double x = lo + (hi - lo) * urand();
        const double y = fixed ? yf : -3.0 + 6.0 * urand();
        const double a = pow(x, y), b = cffc_pow_glibc(x, y);
        if (cffc_pow_double_to_bits(a) != cffc_pow_double_to_bits(b)) {
            if (bad < 5) printf("x=%a y=%a libm=%a replica=%a\n", x, y, a, b);
            bad++;
        }
    }
    printf("%ld %ld\n", n, bad);
    return 0;
}
