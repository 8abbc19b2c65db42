// Refinements after the left-right check, on gfx950:
//  * PostProcess::median_filter (post_process.cu:127-185): each failed pixel of the left view
//    takes the weighted median, in the order of the planes' disparities at that pixel, of the
//    planes in its 41x41 window, weights exp(-|I(p) - I(q)|_1 / 10);
//  * medianFilter of the CLI (main_stereo_hbp_isp.cu:120-193): a horizontal then a vertical
//    7-tap median of the disparity image.
//
// Pinned arithmetic (the same in the oracle): the weight is the project's FP64 exp rounded to
// float; weights are quantised to m = rint(w 2^40) so that every sum is exact (the reference
// adds doubles in two different orders); the order is (disparity key, row-major index), with
// -0 == +0 and NaN after +inf (the reference's std::sort leaves ties unspecified); the result
// is the first plane in that order whose running sum C satisfies 2 C > total.
//
// Weighted median: one wavefront per failed pixel.  Lane l holds the window's pixels
// k = l + 64 i (i < 27): its quantised weight (a double, exact) and disparity key.  The
// weighted median is found without sorting, by two bisections: the smallest key K whose
// weight at or below K exceeds half the total, then the smallest index among the pixels with
// key K; every step is a compare-and-add over the lane's 27 pixels and an exact DPP wave sum.
#include "hbp_internal.h"

namespace hbp {

namespace {

constexpr int WM_R = 20;
constexpr int WM_WIN = 2 * WM_R + 1;
constexpr int WM_PER_LANE = (WM_WIN * WM_WIN + 63) / 64;   // 27

__device__ __forceinline__ float wm_weight(float absdiff) {
    const float t = -absdiff / 10.0f;
    return (float)exp_cody_waite((double)t);
}

// ascending float order, -0 == +0, NaN after +inf
__device__ __forceinline__ uint32_t z_key(float z) {
    if (z != z) return 0xffffffffu;
    const uint32_t b = __float_as_uint(z == 0.0f ? 0.0f : z);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    return __hiloint2double(__builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true),
                            __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true));
}
// exact wave sum of integer valued doubles below 2^53 (every lane gets it)
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x124>(v);
    v += dpp_f64<0x128>(v);
    double r = 0.0;
#pragma unroll
    for (int row = 0; row < 4; ++row)
        r += __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 16 * row),
                              __builtin_amdgcn_readlane(__double2loint(v), 16 * row));
    return r;
}
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__global__ void wm_list_kernel(const uint8_t *__restrict__ fail, long n, int *__restrict__ list,
                               int *__restrict__ count) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        if (fail[i]) list[atomicAdd(count, 1)] = (int)i;
}

__global__ __launch_bounds__(256) void wm_kernel(const uint8_t *__restrict__ bgr, const float *__restrict__ pin,
                                                 float *__restrict__ pout, const int *__restrict__ list,
                                                 const int *__restrict__ count, int W, int H) {
    const int wv = (int)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (wv >= *count) return;
    const int lane = threadIdx.x & 63;
    const int p = __builtin_amdgcn_readfirstlane(list[wv]);
    const int y = p / W, x = p - y * W;
    const int x0 = max(x - WM_R, 0), x1 = min(x + WM_R, W - 1);
    const int y0 = max(y - WM_R, 0), y1 = min(y + WM_R, H - 1);
    const int ww = x1 - x0 + 1, n = ww * (y1 - y0 + 1);
    const float pb = bgr[3L * p], pg = bgr[3L * p + 1], pr = bgr[3L * p + 2];
    const float fx = (float)x, fy = (float)y;
    // this lane's pixels k = lane + 64 i: (row, col) advanced by 64 per step
    uint32_t key[WM_PER_LANE];
    double m[WM_PER_LANE];
    int row = lane / ww, col = lane - (lane / ww) * ww;
    const int drow = 64 / ww, dcol = 64 - drow * ww;
    double tot = 0.0;
    uint32_t kmin = 0xffffffffu, kmax = 0u;
#pragma unroll
    for (int i = 0; i < WM_PER_LANE; ++i) {
        const bool in = lane + 64 * i < n;
        const long q = in ? (long)(y0 + row) * W + (x0 + col) : (long)p;
        const float d0 = pb - (float)bgr[3 * q], d1 = pg - (float)bgr[3 * q + 1], d2 = pr - (float)bgr[3 * q + 2];
        const float absdiff = fabsf(d0) + fabsf(d1) + fabsf(d2);
        const double mq = rint(ldexp((double)wm_weight(absdiff), 40));
        const float z = pin[3 * q] * fx + pin[3 * q + 1] * fy + pin[3 * q + 2];
        const uint32_t kq = z_key(z);
        key[i] = kq;
        m[i] = in ? mq : 0.0;
        tot += m[i];
        if (in) {
            kmin = min(kmin, kq);
            kmax = max(kmax, kq);
        }
        col += dcol;
        row += drow;
        if (col >= ww) {
            col -= ww;
            ++row;
        }
    }
    const double S = wave_sum(tot);
    // the smallest key K with 2 * (weight of keys <= K) > S
    uint32_t lo = wave_min(kmin), hi = wave_max(kmax);
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        double f = 0.0;
#pragma unroll
        for (int i = 0; i < WM_PER_LANE; ++i) f += key[i] <= mid ? m[i] : 0.0;
        if (2.0 * wave_sum(f) > S) hi = mid;
        else lo = mid + 1;
    }
    const uint32_t K = lo;
    double b = 0.0;
#pragma unroll
    for (int i = 0; i < WM_PER_LANE; ++i) b += key[i] < K ? m[i] : 0.0;
    const double base = wave_sum(b);
    // the smallest index t among the keys equal to K with 2 * (base + their weight up to t) > S
    int tl = 0, th = n - 1;
    while (tl < th) {
        const int mid = tl + ((th - tl) >> 1);
        double g = 0.0;
#pragma unroll
        for (int i = 0; i < WM_PER_LANE; ++i) g += (key[i] == K && lane + 64 * i <= mid) ? m[i] : 0.0;
        if (2.0 * (base + wave_sum(g)) > S) th = mid;
        else tl = mid + 1;
    }
    const long qs = (long)(y0 + tl / ww) * W + (x0 + tl % ww);
    if (lane < 3) pout[3L * p + lane] = pin[3 * qs + lane];
}

// insertion sort of 7 values exactly as medianFilter writes it (vals[i] > t shifts), middle
// element; unrolled with compile-time indices
__device__ __forceinline__ float median7(const float (&w)[7]) {
    float v[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const float t = w[j];
        bool moving = true;
#pragma unroll
        for (int i = j - 1; i >= 0; --i) {
            const bool shift = moving && v[i] > t;
            const float vi = v[i];
            v[i + 1] = shift ? vi : (moving ? t : v[i + 1]);
            moving = shift;
        }
        v[0] = moving ? t : v[0];
    }
    return v[3];
}

__global__ void med7_h_kernel(const float *__restrict__ D, float *__restrict__ T, int W, int H) {
    const long n = (long)W * H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int v = (int)(i / W), u = (int)(i - (long)v * W);
        float out = 0.0f;
        if (u >= 3 && u < W - 3 && v >= 3 && v < H - 3) {
            const float d = D[i];
            if (d >= 0) {
                float w[7];
#pragma unroll
                for (int t = 0; t < 7; ++t) w[t] = D[i - 3 + t];
                out = median7(w);
            } else {
                out = d;
            }
        }
        T[i] = out;
    }
}

__global__ void med7_v_kernel(float *__restrict__ D, const float *__restrict__ T, int W, int H) {
    const long n = (long)W * H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int v = (int)(i / W), u = (int)(i - (long)v * W);
        if (u >= 3 && u < W - 3 && v >= 3 && v < H - 3 && D[i] >= 0) {
            float w[7];
#pragma unroll
            for (int t = 0; t < 7; ++t) w[t] = T[i + (long)(t - 3) * W];
            D[i] = median7(w);
        }
    }
}

// Evaluator::evaluate (Evaluator.h:46-52, 66-78): valid = GT > 0 and GT != +inf, nonocc as
// given, occ = valid and not nonocc; a pixel is good when |disp - GT| <= threshold, all in
// float (cv::compare converts the double threshold to the float Mat's depth; NaN is never
// good).  counts: valid, good & valid, nonocc, good & nonocc, occ, good & occ.
__global__ void eval_kernel(const float *__restrict__ disp, const float *__restrict__ gt,
                            const uint8_t *__restrict__ nonocc, long n, float thr,
                            unsigned long long *__restrict__ counts) {
    unsigned int c[6] = {0u, 0u, 0u, 0u, 0u, 0u};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float g = gt[i];
        const bool valid = g > 0.0f && g != INFINITY;
        const bool no = nonocc[i] != 0;
        const bool occ = valid && !no;
        const bool good = fabsf(disp[i] - g) <= thr;
        c[0] += valid;
        c[1] += valid && good;
        c[2] += no;
        c[3] += no && good;
        c[4] += occ;
        c[5] += occ && good;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        unsigned int v = c[k];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v += (unsigned int)__shfl_xor((int)v, o);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&counts[k], (unsigned long long)v);
    }
}

int grid_of(const Ctx &c, long n) {
    const long b = (n + 255) / 256, cap = (long)c.num_cu * 16;
    return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace

// planes [H][W][3] in place (left view); fail u8 [H][W], nonzero = failed
void weighted_median(Ctx &c, const uint8_t *d_bgr, const uint8_t *d_fail, float *d_plane3, int W, int H) {
    const long n = (long)W * H;
    DevBuf<float> pin(c.pool, 3 * (size_t)n);
    DevBuf<int> list(c.pool, (size_t)n), count(c.pool, 1);
    HBP_CHECK(hipMemcpyAsync(pin.p, d_plane3, sizeof(float) * 3 * n, hipMemcpyDeviceToDevice, c.stream));
    HBP_CHECK(hipMemsetAsync(count.p, 0, sizeof(int), c.stream));
    hipLaunchKernelGGL(wm_list_kernel, dim3(grid_of(c, n)), dim3(256), 0, c.stream, d_fail, n, list.p, count.p);
    int nfail = 0;
    HBP_CHECK(hipMemcpyAsync(&nfail, count.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    HBP_CHECK(hipStreamSynchronize(c.stream));
    HBP_REQUIRE(nfail >= 0 && nfail <= n, "weighted median: bad failed-pixel count");
    if (nfail > 0) {
        // algorithmic bytes: per failed pixel its window's colours and planes (15 B each) and
        // its plane out; the windows of neighbouring pixels overlap, so most come from L2
        ProfScope ps(c, "weighted_median", (double)nfail * (15.0 * WM_WIN * WM_WIN + 12.0));
        hipLaunchKernelGGL(wm_kernel, dim3((nfail + 3) / 4), dim3(256), 0, c.stream, d_bgr, pin.p, d_plane3, list.p,
                           count.p, W, H);
    }
    HBP_CHECK(hipGetLastError());
}

void eval_bad_pixels(Ctx &c, const float *d_disp, const float *d_gt, const uint8_t *d_nonocc, int W, int H,
                     double thr, long long *counts) {
    const long n = (long)W * H;
    DevBuf<unsigned long long> cnt(c.pool, 6);
    HBP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long) * 6, c.stream));
    hipLaunchKernelGGL(eval_kernel, dim3(grid_of(c, n)), dim3(256), 0, c.stream, d_disp, d_gt, d_nonocc, n, (float)thr,
                       cnt.p);
    HBP_CHECK(hipGetLastError());
    unsigned long long h[6];
    HBP_CHECK(hipMemcpyAsync(h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    HBP_CHECK(hipStreamSynchronize(c.stream));
    for (int k = 0; k < 6; ++k) counts[k] = (long long)h[k];
}

// disparity [H][W] in place
void median7_filter(Ctx &c, float *d_disp, int W, int H) {
    const long n = (long)W * H;
    DevBuf<float> T(c.pool, (size_t)n);
    ProfScope ps(c, "median7", 16.0 * n);
    hipLaunchKernelGGL(med7_h_kernel, dim3(grid_of(c, n)), dim3(256), 0, c.stream, d_disp, T.p, W, H);
    hipLaunchKernelGGL(med7_v_kernel, dim3(grid_of(c, n)), dim3(256), 0, c.stream, d_disp, T.p, W, H);
    HBP_CHECK(hipGetLastError());
}

}  // namespace hbp
