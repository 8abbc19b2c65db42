// Check of hbp::div_recip(x, d, 1 / (double)d) against the IEEE quotient x / d: every float
// x >= +0 for a set of divisors (the message kernel's divisors max(dc1, 1) and 12, powers of
// two, odd integers, values near 1 and a few extremes), then 2^28 random pairs of arbitrary
// bit patterns (signs, subnormals, zeros, infinities, NaNs).  Equal bit for bit except that a
// NaN may come back as any NaN.  Prints the number of mismatches.  Built and run by
// tests/test_div3.py.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "hbp_common.h"

static bool same(float a, float b) {
    if (a != a) return b != b;
    uint32_t x, y;
    memcpy(&x, &a, 4);
    memcpy(&y, &b, 4);
    return x == y;
}

int main() {
    const float divisors[] = {1.0f, 12.0f, 3.0f, 1.5f, 7.0f, 1000.1f, 1.0000001f, 0.99999994f, 2.0f, 65535.0f,
                              13.37f, 3.4028235e38f, 1.1754944e-38f, 1.0e-45f, 255.75f, 0.1f};
    unsigned long long bad = 0;
    for (float d : divisors) {
        const double r = 1.0 / (double)d;
#pragma omp parallel for reduction(+ : bad) schedule(static)
        for (long long u = 0; u <= 0x7fffffffLL; ++u) {
            const uint32_t b = (uint32_t)u;
            float x;
            memcpy(&x, &b, 4);
            if (!same(hbp::div_recip(x, d, r), x / d)) ++bad;
        }
    }
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long long i = 0; i < (1LL << 28); ++i) {
        uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const uint32_t bx = (uint32_t)z, bd = (uint32_t)(z >> 32);
        float x, d;
        memcpy(&x, &bx, 4);
        memcpy(&d, &bd, 4);
        if (!same(hbp::div_recip(x, d, 1.0 / (double)d), x / d)) ++bad;
    }
    printf("%llu\n", bad);
    return bad != 0;
}
