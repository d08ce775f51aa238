/* The simplex projection's quotients by 3, 5 and 7 (potts_math.hpp,
 * div_small): q0 = x * RN(1/D), q = fma(fma(-D, q0, x), RN(1/D), q0) equals
 * the IEEE quotient x / D for every finite x other than -0.  With no argument
 * this checks every mantissa of a spread of exponents (subnormal to huge) and
 * 2^26 random patterns; with "all" it checks every 32-bit pattern.  Prints
 * the number of differing results per divisor. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float q_of(float x, float d)
{
    const float c = 1.0f / d, q0 = x * c;
    return fmaf(fmaf(-d, q0, x), c, q0);
}

static int differs(uint32_t u, float d)
{
    float x;
    memcpy(&x, &u, 4);
    if (!isfinite(x) || u == 0x80000000u) return 0;
    const float a = q_of(x, d), b = x / d;
    return memcmp(&a, &b, 4) != 0;
}

int main(int argc, char **argv)
{
    const int all = argc > 1 && strcmp(argv[1], "all") == 0;
    const float ds[3] = {3.0f, 5.0f, 7.0f};
    for (int k = 0; k < 3; ++k) {
        long bad = 0;
        if (all) {
#pragma omp parallel for reduction(+ : bad) schedule(static)
            for (int64_t u = 0; u <= 0xffffffffll; ++u) bad += differs((uint32_t)u, ds[k]);
        } else {
            const uint32_t exps[] = {0, 1, 24, 126, 127, 128, 253, 254};
            const int ne = (int)(sizeof(exps) / sizeof(exps[0]));
#pragma omp parallel for reduction(+ : bad) schedule(static)
            for (int64_t i = 0; i < (int64_t)ne << 24; ++i) {
                const uint32_t e = exps[i >> 24], m = (uint32_t)i & 0x7fffffu, s = (uint32_t)(i >> 23) & 1u;
                bad += differs((s << 31) | (e << 23) | m, ds[k]);
            }
#pragma omp parallel for reduction(+ : bad) schedule(static)
            for (int64_t i = 0; i < (1 << 24); ++i) {
                uint64_t z = (uint64_t)i * 0x9e3779b97f4a7c15ull + 0x632be59bd9b4e019ull;
                z ^= z >> 29;
                z *= 0xbf58476d1ce4e5b9ull;
                z ^= z >> 32;
                bad += differs((uint32_t)z, ds[k]);
            }
        }
        printf("%g %ld\n", ds[k], bad);
    }
    return 0;
}
