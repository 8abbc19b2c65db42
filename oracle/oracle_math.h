// Oracle (test infrastructure only): pinned scalar arithmetic shared by the oracle's stages.
//
// Every function here is a restatement of a documented, deterministic rule built only from
// IEEE-754 operations that are correctly rounded on every target (+ - * / sqrt, rint,
// ldexp, floor, integer ops).  The HIP library carries its own, independently written copy
// of the same rules; the parity tests compare the two and pin each rule against libm.
//
// Build with -ffp-contract=off: none of the expressions below may be fused.
#pragma once
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

// Cube root on (216/24389, 1.0] (the CIE Lab "f(t)" domain; imutil.cu:28-30 uses
// pow(t, 1/3.0) there).  Newton's iteration y <- (2y + x/y^2)/3 from the linear start
// 0.45 + 0.55x, a fixed 8 steps (7 already reach 1 ulp over the whole domain).
inline double pin_cbrt(double x) {
    double y = 0.45 + 0.55 * x;
    for (int i = 0; i < 8; ++i) y = (2.0 * y + x / (y * y)) / 3.0;
    return y;
}

// t^2.4 on the sRGB linearisation domain ((0.04045 + 0.055)/1.055, 1 + 2^-20] (imutil.cu:15-17
// evaluate pow((c + 0.055)/1.055, 2.4) for blurred, non-integer c).  t^2.4 = t^2 * (t^2)^(1/5);
// the fifth root by Newton's step y <- y + (a/y^4 - y)/5 from 0.6 + 0.4a, a fixed 8 steps
// (7 already converge over the whole domain); within 4 ulp of libm's pow(t, 2.4).
inline double pin_pow24(double t) {
    double a = t * t;
    double y = 0.6 + 0.4 * a;
    for (int i = 0; i < 8; ++i) {
        double y2 = y * y;
        y = y + (a / (y2 * y2) - y) / 5.0;
    }
    return a * y;
}

// exp(x) for x <= 0 (hbp_isp.cu:119 evaluates exp(-dist/gamma) in FP64).
// Cody-Waite reduction x = n*ln2 + r with the split ln2 = LN2_HI + LN2_LO (LN2_HI has its
// low 32 bits zero so n*LN2_HI is exact), then the degree-13 Taylor polynomial of e^r on
// |r| <= ln2/2 in Horner form, then scaling by 2^n.
inline double pin_exp(double x) {
    if (x < -745.0) return 0.0;
    const double INV_LN2 = 1.4426950408889634;
    const double LN2_HI = 6.93147180369123816490e-01;
    const double LN2_LO = 1.90821492927058770002e-10;
    double n = std::rint(x * INV_LN2);
    double r = (x - n * LN2_HI) - n * LN2_LO;
    double p = 1.6059043836821613e-10;
    p = 2.08767569878681e-09 + r * p;
    p = 2.505210838544172e-08 + r * p;
    p = 2.755731922398589e-07 + r * p;
    p = 2.7557319223985893e-06 + r * p;
    p = 2.48015873015873e-05 + r * p;
    p = 0.0001984126984126984 + r * p;
    p = 0.001388888888888889 + r * p;
    p = 0.008333333333333333 + r * p;
    p = 0.041666666666666664 + r * p;
    p = 0.16666666666666666 + r * p;
    p = 0.5 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    return std::ldexp(p, (int)n);
}

// __float2int_rd semantics (hbp_isp.cu:217): floor, NaN -> 0, saturate out of range.
inline int f2i_rd(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT_MAX;
    if (x < -2147483648.0f) return INT_MIN;
    return (int)std::floor(x);
}

// Round-to-nearest-even int64 -> float, written with integer ops only.
inline float i64_to_f32(int64_t q) {
    if (q == 0) return 0.0f;
    bool neg = q < 0;
    uint64_t u = neg ? (uint64_t)0 - (uint64_t)q : (uint64_t)q;
    int msb = 63 - __builtin_clzll(u);
    float f;
    if (msb <= 23) {
        f = (float)(uint32_t)u;
    } else {
        int sh = msb - 23;
        uint64_t m = u >> sh;
        uint64_t rem = u & ((1ull << sh) - 1ull);
        uint64_t half = 1ull << (sh - 1);
        if (rem > half || (rem == half && (m & 1ull))) m += 1ull;
        f = std::ldexp((float)(uint32_t)m, sh);
    }
    return neg ? -f : f;
}

// Fixed-point cluster sums (pinned rule for the plane-cost reduction, whose order the
// reference leaves to thrust::reduce_by_key, hbp_isp.cu:1167-1173): each per-pixel value v
// becomes q = rint(v * 2^s) as int64; a cluster's cost is i64_to_f32(sum q) * 2^-s, or NaN
// if any of its values is NaN.  s is chosen so that every partial sum stays below 2^52 (exact
// in int64 and in double, in any order):
//   s = 52 - e(vmax) - e(N),  x < 2^e(x) from frexp,  vmax = |lambda| * (tau or 1) * 1.0001.
inline int fixed_shift(float lambda, float tau, int use_cost_vol, int npix) {
    double vmax = std::fabs((double)lambda) * (use_cost_vol ? std::fabs((double)tau) : 1.0) * 1.0001;
    int ev = 0;
    if (vmax > 0.0) std::frexp(vmax, &ev);
    int en = 0;
    std::frexp((double)(npix > 0 ? npix : 1), &en);
    int s = 52 - ev - en;
    if (s > 120) s = 120;
    if (s < -60) s = -60;
    return s;
}
inline int64_t to_fixed(float v, int s) { return (int64_t)std::rint(std::ldexp(v, s)); }
inline float from_fixed(int64_t q, int nan_flag, int s) {
    if (nan_flag) {
        uint32_t b = 0x7FC00000u;
        float f;
        std::memcpy(&f, &b, 4);
        return f;
    }
    return std::ldexp(i64_to_f32(q), -s);
}

// atomicMinFloat (util.h:141-148) applied as a left fold, NaN taken as the positive quiet
// NaN: non-negative values take a signed-int min, negative values (and NaN) an unsigned max.
inline float amin_fold(float stored, float v) {
    uint32_t vb, sb;
    if (v != v) vb = 0x7FC00000u; else std::memcpy(&vb, &v, 4);
    std::memcpy(&sb, &stored, 4);
    float vv;
    std::memcpy(&vv, &vb, 4);
    uint32_t out;
    if (vv >= 0.0f) out = ((int32_t)vb < (int32_t)sb) ? vb : sb;
    else out = (vb > sb) ? vb : sb;
    float r;
    std::memcpy(&r, &out, 4);
    return r;
}

// BSD LCG (hbp_isp.cu:551-569): s' = (1103515245 s + 12345) mod 2^31 in uint32 arithmetic.
inline uint32_t lcg(uint32_t s) { return (1103515245u * s + 12345u) % 2147483648u; }
inline float lcg_uniform(uint32_t &s, float f) {
    s = lcg(s);
    return (float)s / 2147483648.0f * f;
}

}  // namespace orc
