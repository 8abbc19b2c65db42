// Oracle (test infrastructure only): the refinements after the left-right check, restating
// PostProcess::median_filter (post_process.cu:127-185, weighted median of planes) and the
// CLI's 7-tap separable median (main_stereo_hbp_isp.cu:120-193); and the Middlebury
// bad-pixel counts of Evaluator (Evaluator.h:49-76).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "hbp_oracle.h"
#include "oracle_math.h"

namespace {

// computePatchWeight (post_process.cu:112-119): exp(-(|dB| + |dG| + |dR|) / omega), omega = 10,
// in float with the pinned exp of the project (FP64, rounded to float)
float wm_weight(float absdiff) {
    const float t = -absdiff / 10.0f;
    return (float)orc::pin_exp((double)t);
}

// the sort key of a plane's disparity at the filtered pixel: ascending float order with -0 and
// +0 equal and NaN after +inf
uint32_t z_key(float z) {
    if (z != z) return 0xffffffffu;
    if (z == 0.0f) z = 0.0f;
    uint32_t b;
    std::memcpy(&b, &z, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

}  // namespace

extern "C" {

float orc_wm_weight(float absdiff) { return wm_weight(absdiff); }

// Weighted median of the failed pixels' planes (pinned form of post_process.cu:127-185):
// for a pixel p with fail != 0, the window [x-20, x+20] x [y-20, y+20] clipped to the image,
// its pixels q in row-major order k; weight w(q) = wm_weight(|I(p) - I(q)|_1) on the BGR
// values as floats, quantised to m(q) = rint(w(q) 2^40) (an exact integer; all sums of them
// are exact, where the reference adds doubles); z(q) = plane(q) evaluated at p as
// (a x + b y) + c.  In the order of (z_key(z), k) the result is the plane of the first q whose
// running sum C satisfies 2 C > sum of all m.  Planes are read from a copy made before the
// pass (post_process.cu:137), so the pass does not depend on the order of the pixels.
void orc_weighted_median(const uint8_t *bgr, const uint8_t *fail, float *plane3, int W, int H) {
    const int R = 20;
    std::vector<float> in(plane3, plane3 + (size_t)W * H * 3);
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < H; ++y) {
        std::vector<std::pair<uint64_t, int64_t>> el;   // ((z_key << 32) | k, m)
        for (int x = 0; x < W; ++x) {
            const long p = (long)y * W + x;
            if (!fail[p]) continue;
            const int x0 = std::max(x - R, 0), x1 = std::min(x + R, W - 1);
            const int y0 = std::max(y - R, 0), y1 = std::min(y + R, H - 1);
            const float pb = bgr[3 * p], pg = bgr[3 * p + 1], pr = bgr[3 * p + 2];
            el.clear();
            int64_t S = 0;
            int k = 0;
            for (int yy = y0; yy <= y1; ++yy)
                for (int xx = x0; xx <= x1; ++xx, ++k) {
                    const long q = (long)yy * W + xx;
                    const float d0 = pb - (float)bgr[3 * q], d1 = pg - (float)bgr[3 * q + 1],
                                d2 = pr - (float)bgr[3 * q + 2];
                    const float absdiff = std::fabs(d0) + std::fabs(d1) + std::fabs(d2);
                    const int64_t m = (int64_t)std::rint(std::ldexp((double)wm_weight(absdiff), 40));
                    const float z = in[3 * q] * (float)x + in[3 * q + 1] * (float)y + in[3 * q + 2];
                    el.push_back({((uint64_t)z_key(z) << 32) | (uint32_t)k, m});
                    S += m;
                }
            std::sort(el.begin(), el.end(),
                      [](const std::pair<uint64_t, int64_t> &a, const std::pair<uint64_t, int64_t> &b) {
                          return a.first < b.first;
                      });
            int64_t C = 0;
            const int ww = x1 - x0 + 1;
            for (const auto &e : el) {
                C += e.second;
                if (2 * C > S) {
                    const int kk = (int)(e.first & 0xffffffffu);
                    const long q = (long)(y0 + kk / ww) * W + (x0 + kk % ww);
                    plane3[3 * p] = in[3 * q];
                    plane3[3 * p + 1] = in[3 * q + 1];
                    plane3[3 * p + 2] = in[3 * q + 2];
                    break;
                }
            }
        }
    }
}

// medianFilter (main_stereo_hbp_isp.cu:120-193): for u in [3, W-3) and v in [3, H-3), a
// horizontal then a vertical 7-tap median by insertion sort (vals[i] > temp shifts; NaN never
// shifts), applied where D(v, u) >= 0, else copied; the temporary image starts at 0, so the
// vertical pass reads zeros in its top and bottom 3 rows; the border is left as it is.
static float median7_sorted(const float *w) {
    float vals[7];
    for (int j = 0; j < 7; ++j) {
        const float t = w[j];
        int i = j - 1;
        while (i >= 0 && vals[i] > t) {
            vals[i + 1] = vals[i];
            --i;
        }
        vals[i + 1] = t;
    }
    return vals[3];
}

void orc_median7(float *D, int W, int H) {
    const int r = 3;
    std::vector<float> T((size_t)W * H, 0.0f);
    for (int u = r; u < W - r; ++u)
        for (int v = r; v < H - r; ++v) {
            const float d = D[(long)v * W + u];
            if (d >= 0) {
                float w[7];
                for (int t = 0; t < 7; ++t) w[t] = D[(long)v * W + u - r + t];
                T[(long)v * W + u] = median7_sorted(w);
            } else {
                T[(long)v * W + u] = d;
            }
        }
    for (int u = r; u < W - r; ++u)
        for (int v = r; v < H - r; ++v) {
            if (D[(long)v * W + u] >= 0) {
                float w[7];
                for (int t = 0; t < 7; ++t) w[t] = T[(long)(v - r + t) * W + u];
                D[(long)v * W + u] = median7_sorted(w);
            }
        }
}

// Evaluator.h:49-76 with OpenCV's element rules: validMask = (GT > 0) & (GT != inf),
// occMask = ~nonocc & validMask, errorMap = |disp - GT| <= threshold where cv::compare takes the
// double threshold in the float Mat's depth (a NaN difference is never good).  The nonocc counts
// are over the whole mask, valid or not, as the reference counts them.
void orc_eval_bad_pixels(const float *disp, const float *gt, const uint8_t *nonocc, int W, int H, double thr,
                         long long *counts) {
    const float t = (float)thr;
    for (int k = 0; k < 6; ++k) counts[k] = 0;
    for (long i = 0; i < (long)W * H; ++i) {
        const float g = gt[i];
        const bool valid = g > 0.0f && g != INFINITY;
        const bool no = nonocc[i] != 0;
        const bool occ = !no && valid;
        const bool good = std::fabs(disp[i] - g) <= t;
        counts[0] += valid;
        counts[1] += good && valid;
        counts[2] += no;
        counts[3] += good && no;
        counts[4] += occ;
        counts[5] += good && occ;
    }
}

}  // extern "C"
