// Exhaustive check of hbp::div3_nonneg against the IEEE quotient x / 3 over every float
// x >= +0 (sign bit clear): finite x must match bit for bit; +inf and NaN may give any NaN
// (the kernel clamps them with fminf).  Built and run by tests/test_div3.py.
#include <cstdio>
#include <cstring>

#include "hbp_common.h"

int main() {
    unsigned long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long long u = 0; u <= 0x7fffffffLL; ++u) {
        const uint32_t b = (uint32_t)u;
        float x;
        memcpy(&x, &b, 4);
        const float ref = x / 3.0f;
        const float got = hbp::div3_nonneg(x);
        if (x != x || x > 3.4028234663852886e38f) {   // NaN or +inf
            if (got == got && !(got > 3.4028234663852886e38f) && got != ref) ++bad;
            continue;
        }
        uint32_t rb, gb;
        memcpy(&rb, &ref, 4);
        memcpy(&gb, &got, 4);
        if (rb != gb) ++bad;
    }
    printf("%llu\n", bad);
    return bad != 0;
}
