// Oracle (test infrastructure only): stage (d), left-right consistency check and fill.
// Restates post_process.cu of the reference (citations per function).
#include <cmath>
#include <cstring>
#include <vector>

#include "hbp_oracle.h"

extern "C" {

// doConsistencyCheck (post_process.cu:4-41): rx = int(x - d*sign + 0.5f) truncated toward 0;
// a pixel fails (255) when rx lies in the image and |d_other(rx) - d| > 1.5.
void orc_lr_consistency(const float *dispL, const float *dispR, int W, int H, uint8_t *failL,
                        uint8_t *failR) {
    const float *disp[2] = {dispL, dispR};
    uint8_t *fail[2] = {failL, failR};
    for (int i = 0; i < 2; ++i) {
        float sign = i ? -1.0f : 1.0f;
        std::memset(fail[i], 0, (size_t)W * H);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                float ds = disp[i][(long)y * W + x];
                int rx = (int)((float)x - ds * sign + 0.5f);
                if (rx >= 0 && rx < W) {
                    float dsr = disp[1 - i][(long)y * W + rx];
                    if ((double)std::fabs(dsr - ds) > 1.5) fail[i][(long)y * W + x] = 255;
                }
            }
    }
}

// cv::dilate(src, dst, cv::Mat()) (post_process.cu:68-69): 3x3 rectangle, anchor at the
// centre, pixels outside the image do not contribute.
void orc_dilate3(const uint8_t *in, int W, int H, uint8_t *out) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            uint8_t m = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int yy = y + dy, xx = x + dx;
                    if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
                    uint8_t v = in[(long)yy * W + xx];
                    if (v > m) m = v;
                }
            out[(long)y * W + x] = m;
        }
}

// post_process.cu:43-46
static inline float get_z(const float *v, int x, int y) { return v[0] * (float)x + v[1] * (float)y + v[2]; }

// Horizontal fill of the left view (post_process.cu:71-109): each failed pixel takes the plane
// of the nearest pixel to its left / right whose fail2 is not 255, the one with the smaller
// GetZ at this pixel (ties and "left missing" go right); nothing found -> unchanged.
// Reads only non-fail2 pixels, which the fill never writes.
void orc_lr_fill(const uint8_t *fail, const uint8_t *fail2, float *plane3, int W, int H) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            if (fail[(long)y * W + x] == 0) continue;
            int xl = x;
            while (xl >= 0 && fail2[(long)y * W + xl] == 255) xl--;
            int xr = x;
            while (xr < W && fail2[(long)y * W + xr] == 255) xr++;
            const float *pl = xl >= 0 ? plane3 + 3 * ((long)y * W + xl) : nullptr;
            const float *pr = xr < W ? plane3 + 3 * ((long)y * W + xr) : nullptr;
            float *dst = plane3 + 3 * ((long)y * W + x);
            const float *src = nullptr;
            if (!pl && !pr) continue;
            if (!pl) src = pr;
            else if (!pr) src = pl;
            else if (get_z(pl, x, y) < get_z(pr, x, y)) src = pl;
            else src = pr;
            dst[0] = src[0];
            dst[1] = src[1];
            dst[2] = src[2];
        }
}

// plane_label_to_disp (post_process.cu:373-390): disp = GetZ(plane, p), no clamping.
void orc_plane_to_disp(const float *plane3, int W, int H, float *disp) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) disp[(long)y * W + x] = get_z(plane3 + 3 * ((long)y * W + x), x, y);
}

// left_right_check (post_process.cu:47-110): check both views, fail = fail > 0, dilate,
// fill the left view's planes.  The right planes are not modified.
void orc_left_right_check(const float *dispL, const float *dispR, float *planeL, float *planeR,
                          int W, int H, uint8_t *failL, uint8_t *failR, uint8_t *fail2L) {
    (void)planeR;
    orc_lr_consistency(dispL, dispR, W, H, failL, failR);
    std::vector<uint8_t> f2((size_t)W * H);
    orc_dilate3(failL, W, H, f2.data());
    if (fail2L) std::memcpy(fail2L, f2.data(), (size_t)W * H);
    orc_lr_fill(failL, f2.data(), planeL, W, H);
}

}  // extern "C"
