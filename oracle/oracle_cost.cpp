// Oracle (test infrastructure only): stage (a), census transform + Hamming cost volume.
// Restates stereo_cost.cu and imutil.cu of the reference (citations per function).
#ifdef _OPENMP
#include <omp.h>
#endif
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hbp_oracle.h"
#include "oracle_math.h"

extern "C" {

// compute_stereo_cost("MC-CNN") (stereo_cost.cu:399-445): the given f32 volume [D][H][W], the
// left view's columns x <= d set to column d (fill_out_of_view_k, 366-395), every entry
// truncated with fminf at tau (truncate_data_k, 230-246), then WTA_D_k (248-266: the first d
// of the minimum under <, starting from FLT_MAX).  right_view: cost_to_right (471-516),
// C_R[d][y][x] = C_L[d][y][min(x+d, W-1)], and its WTA.
int orc_cost_volume_f32(const float *raw, int W, int H, int D, float tau, int right_view, float *vol,
                        float *dwta) {
    if (W <= 0 || H <= 0 || D < 1 || D > W) return -1;
    const long hw = (long)W * H;
    std::vector<float> left((size_t)hw * D);
    for (int d = 0; d < D; ++d)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const long i = (long)d * hw + (long)y * W + x;
                const float v = x <= d ? raw[(long)d * hw + (long)y * W + d] : raw[i];
                left[i] = std::fmin(v, tau);
            }
    for (int d = 0; d < D; ++d)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const int xs = right_view ? std::min(x + d, W - 1) : x;
                vol[(long)d * hw + (long)y * W + x] = left[(long)d * hw + (long)y * W + xs];
            }
    for (long i = 0; i < hw; ++i) {
        float mc = 3.402823466e+38f;
        int mi = 0;
        for (int d = 0; d < D; ++d)
            if (vol[(long)d * hw + i] < mc) {
                mc = vol[(long)d * hw + i];
                mi = d;
            }
        dwta[i] = (float)mi;
    }
    return 0;
}

// the "AD" cost (ad_cost, stereo_cost.cu:9-43, tau = 30 from compute_stereo_cost 423-428): the
// float greys of imutil.cu:116-130; in the columns x < D-1 every disparity costs 0, elsewhere
// C[d][y][x] = fminf(|gL(y, x) - gR(y, x-d)|, tau).  Then cost_to_right for the right view
// (C_R[d][y][x] = C_L[d][y][min(x+d, W-1)], 471-516) and WTA_D_k (248-266).
int orc_cost_volume_ad(const uint8_t *bgrL, const uint8_t *bgrR, int W, int H, int D, int right_view, float *vol,
                       float *dwta) {
    if (W <= 0 || H <= 0 || D < 1 || D > W) return -1;
    const long hw = (long)W * H;
    std::vector<float> gl(hw), gr(hw);
    for (long i = 0; i < hw; ++i) {
        gl[i] = (float)((double)bgrL[3 * i] * 0.114 + (double)bgrL[3 * i + 1] * 0.587 +
                        (double)bgrL[3 * i + 2] * 0.299);
        gr[i] = (float)((double)bgrR[3 * i] * 0.114 + (double)bgrR[3 * i + 1] * 0.587 +
                        (double)bgrR[3 * i + 2] * 0.299);
    }
    const float tau = 30.0f;
    std::vector<float> left((size_t)hw * D);
    for (int d = 0; d < D; ++d)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const long i = (long)y * W + x;
                left[(long)d * hw + i] = x < D - 1 ? 0.0f : std::fmin(std::fabs(gl[i] - gr[i - d]), tau);
            }
    for (int d = 0; d < D; ++d)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const int xs = right_view ? std::min(x + d, W - 1) : x;
                vol[(long)d * hw + (long)y * W + x] = left[(long)d * hw + (long)y * W + xs];
            }
    for (long i = 0; i < hw; ++i) {
        float mc = 3.402823466e+38f;
        int mi = 0;
        for (int d = 0; d < D; ++d)
            if (vol[(long)d * hw + i] < mc) {
                mc = vol[(long)d * hw + i];
                mi = d;
            }
        dwta[i] = (float)mi;
    }
    return 0;
}

// imutil.cu:116-130 (grey = B*0.114 + G*0.587 + R*0.299 in FP64, stored as f32) followed by
// imutil.cu:154-168 (u8 = f < 255 ? trunc(f) : 255).
void orc_grey(const uint8_t *bgr, int W, int H, uint8_t *grey) {
    long n = (long)W * H;
    for (long i = 0; i < n; ++i) {
        double v = (double)bgr[3 * i] * 0.114 + (double)bgr[3 * i + 1] * 0.587 +
                   (double)bgr[3 * i + 2] * 0.299;
        float f = (float)v;
        grey[i] = f < 255.f ? (uint8_t)f : (uint8_t)255;
    }
}

// stereo_cost.cu:45-146, centre-symmetric census on a 9x7 window, 31 bits:
//   bit k*9+m (k<3, m<9): I(y-3+k, x-4+m) >= I(y+3-k, x+4-m)
//   bit 27+m (m<4):       I(y, x-4+m)     >= I(y, x+4-m)
// pixels outside the image read as 0 (stereo_cost.cu:65-67).
void orc_census(const uint8_t *grey, int W, int H, uint32_t *census) {
    auto px = [&](int y, int x) -> int {
        return (y >= 0 && y < H && x >= 0 && x < W) ? grey[(long)y * W + x] : 0;
    };
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            uint32_t c = 0;
            for (int k = 0; k < 3; ++k)
                for (int m = 0; m < 9; ++m)
                    c |= (uint32_t)(px(y - 3 + k, x - 4 + m) >= px(y + 3 - k, x + 4 - m))
                         << (k * 9 + m);
            for (int m = 0; m < 4; ++m)
                c |= (uint32_t)(px(y, x - 4 + m) >= px(y, x + 4 - m)) << (27 + m);
            census[(long)y * W + x] = c;
        }
}

// stereo_cost.cu:148-207 + 209-228: cost(x, d) = popc(CL(x) ^ CR(max(x - d, 0))) (for x < d
// the kernel's shared-memory window still holds CR(0), lines 163/169), then min(cost, tau)
// written disparity-major [D][H][W].
void orc_hamming(const uint32_t *cl, const uint32_t *cr, int W, int H, int D, int tau,
                 uint8_t *vol) {
    long n = (long)W * H;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int d = 0; d < D; ++d)
            for (int x = 0; x < W; ++x) {
                int xr = x - d < 0 ? 0 : x - d;
                int c = __builtin_popcount(cl[(long)y * W + x] ^ cr[(long)y * W + xr]);
                vol[(long)d * n + (long)y * W + x] = (uint8_t)(c < tau ? c : tau);
            }
}

// stereo_cost.cu:248-266: argmin over d, strict '<' from FLT_MAX, so the smallest d wins ties.
void orc_wta(const uint8_t *vol, int W, int H, int D, float *dwta) {
    long n = (long)W * H;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        float best = FLT_MAX;
        int arg = 0;
        for (int d = 0; d < D; ++d) {
            float c = (float)vol[(long)d * n + i];
            if (c < best) {
                best = c;
                arg = d;
            }
        }
        dwta[i] = (float)arg;
    }
}

// stereo_cost.cu:471-486 then 366-395 with is_left=false: in place, per (d, y) row,
// C[x] <- C[x + d] for x + d < W, then columns W-1-d..W-1 <- C[W-1-d].  Together this is
// C_R[d][y][x] = C_L[d][y][min(x + d, W - 1)] (requires D <= W).
void orc_vol_to_right(uint8_t *vol, int W, int H, int D) {
    long n = (long)W * H;
    for (int d = 0; d < D; ++d)
        for (int y = 0; y < H; ++y) {
            uint8_t *row = vol + (long)d * n + (long)y * W;
            for (int x = 0; x + d < W; ++x) row[x] = row[x + d];
            uint8_t c = row[W - 1 - d];
            for (int x = 0; x <= d; ++x) row[W - 1 - x] = c;
        }
}

// compute_stereo_cost("Census") (stereo_cost.cu:421-445, 299-331): tau = 25, then WTA.
// right_view applies cost_to_right (stereo_cost.cu:488-516) before its WTA.
void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_cost_volume(const uint8_t *bgrL, const uint8_t *bgrR, int W, int H, int D,
                    int right_view, uint8_t *vol, float *dwta) {
    if (W < 9 || H < 7 || D < 1 || D > 1024 || D > W) return -1;
    long n = (long)W * H;
    std::vector<uint8_t> gl(n), gr(n);
    std::vector<uint32_t> cl(n), cr(n);
    orc_grey(bgrL, W, H, gl.data());
    orc_grey(bgrR, W, H, gr.data());
    orc_census(gl.data(), W, H, cl.data());
    orc_census(gr.data(), W, H, cr.data());
    orc_hamming(cl.data(), cr.data(), W, H, D, 25, vol);
    if (right_view) orc_vol_to_right(vol, W, H, D);
    orc_wta(vol, W, H, D, dwta);
    return 0;
}

}  // extern "C"
