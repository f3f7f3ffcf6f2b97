/* Host restatement of the launch-uniform division of the fused joint kernels (udiv in
 * reconflow_amd/csrc/rf_device.hpp), checked against IEEE FP32 division.
 *
 * usage: udiv_check d [d ...]
 * For each divisor d it divides every FP32 numerator of the given binades (all 2^23 subnormals,
 * then every significand at each listed exponent, both signs, and the special values) and prints
 * "d mismatches checked plain_mismatches nosign_mismatches": the last two count the mismatches
 * of RN32(RN64(n rd)) (no snap to the 2^-176 grid) and of the sequence without the zero-sign
 * fix, the mistakes the check must be able to see. d must be positive. A mismatch is a different bit pattern (NaNs
 * match any NaN). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint32_t bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
static float from_bits(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* the device sequence: s = RN64(n rd + C), q = RN64(s - C) with C = 1.5 2^-124, RN32(q), the
 * numerator's sign bit or-ed in */
static float udiv(float n, double rd, int sign_fix) {
    const double C = 0x1.8p-124;
    const double q = fma((double)n, rd, C) - C;
    return from_bits(bits((float)q) | (sign_fix ? (bits(n) & 0x80000000u) : 0u));
}

static int differs(float ref, float got) { return !(isnan(ref) && isnan(got)) && bits(ref) != bits(got); }

static long mism, checked, plain_mism, nosign_mism;
static void one(float n, float d, double rd) {
    volatile float ref = n / d;
    const float got = udiv(n, rd, 1);
    ++checked;
    if (differs(ref, got)) {
        if (mism < 5) fprintf(stderr, "d=%a n=%a ref=%a got=%a\n", (double)d, (double)n, (double)ref, (double)got);
        ++mism;
    }
    plain_mism += differs(ref, (float)((double)n * rd));
    nosign_mism += differs(ref, udiv(n, rd, 0));
}

int main(int argc, char** argv) {
    static const int exps[] = {-126, -125, -110, -100, -60, -1, 0, 1, 23, 60, 100, 126, 127};
    for (int a = 1; a < argc; ++a) {
        const float d = strtof(argv[a], NULL);
        const double rd = 1.0 / (double)d;
        mism = checked = plain_mism = nosign_mism = 0;
        for (uint32_t m = 0; m < (1u << 23); ++m) { /* zero and the subnormals */
            one(from_bits(m), d, rd);
            one(from_bits(m | 0x80000000u), d, rd);
        }
        for (size_t k = 0; k < sizeof(exps) / sizeof(exps[0]); ++k) {
            const uint32_t e = (uint32_t)(exps[k] + 127) << 23;
            for (uint32_t m = 0; m < (1u << 23); ++m) {
                one(from_bits(e | m), d, rd);
                one(from_bits(e | m | 0x80000000u), d, rd);
            }
        }
        const float sp[] = {INFINITY, -INFINITY, NAN, 3.4028235e38f, -3.4028235e38f, 1.1754944e-38f};
        for (size_t k = 0; k < sizeof(sp) / sizeof(sp[0]); ++k) one(sp[k], d, rd);
        printf("%a %ld %ld %ld %ld\n", (double)d, mism, checked, plain_mism, nosign_mism);
    }
    return 0;
}
