// Stage (a): census transform and Hamming cost volume (stereo_cost.cu of the reference).
//
// Volume layout in HBM: u8 [D][H][pitch], pitch = W rounded up to 64, one disparity plane
// every H*pitch bytes.  Costs are min(Hamming, 25) <= 25, so u8 holds the reference's f32
// values exactly at a quarter of the bytes; the plane-cost gathers of BP read two bytes per
// pixel and plane from this layout.
#include <cstring>

#include "hbp_internal.h"

namespace hbp {

// ----------------------------------------------------------------------------------------
// grey + centre-symmetric census, both images in one launch.
// Reference: img_BGR_to_grey (imutil.cu:116-130), img_float1_to_uchar1 (imutil.cu:154-168),
// CenterSymmetricCensusKernelSM2 (stereo_cost.cu:45-146).
// Tile: 64 x 16 pixels per 256-thread block, each thread 4 pixels in a column; the block
// stages the (64+8) x (16+6) grey window of both images in LDS.
// ----------------------------------------------------------------------------------------
constexpr int CT_W = 64, CT_H = 16, CT_LW = CT_W + 8, CT_LH = CT_H + 6;

__device__ __forceinline__ uint8_t grey_of(const uint8_t *bgr, long i) {
    double v = (double)bgr[3 * i] * 0.114 + (double)bgr[3 * i + 1] * 0.587 + (double)bgr[3 * i + 2] * 0.299;
    float f = (float)v;
    return f < 255.f ? (uint8_t)f : (uint8_t)255;
}

__global__ __launch_bounds__(256) void census_kernel(const uint8_t *__restrict__ bgrL,
                                                     const uint8_t *__restrict__ bgrR, int W, int H,
                                                     uint32_t *__restrict__ cl, uint32_t *__restrict__ cr) {
    __shared__ uint8_t tl[CT_LH][CT_LW];
    __shared__ uint8_t tr[CT_LH][CT_LW];
    const int x0 = blockIdx.x * CT_W, y0 = blockIdx.y * CT_H;
    for (int idx = threadIdx.x; idx < CT_LW * CT_LH; idx += 256) {
        int ty = idx / CT_LW, tx = idx % CT_LW;
        int gy = y0 - 3 + ty, gx = x0 - 4 + tx;
        bool in = gy >= 0 && gy < H && gx >= 0 && gx < W;
        long gi = (long)gy * W + gx;
        tl[ty][tx] = in ? grey_of(bgrL, gi) : (uint8_t)0;
        tr[ty][tx] = in ? grey_of(bgrR, gi) : (uint8_t)0;
    }
    __syncthreads();
    const int tx = threadIdx.x & 63;
    const int tyb = (threadIdx.x >> 6) * 4;
    const int gx = x0 + tx;
    if (gx >= W) return;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int ty = tyb + r;
        int gy = y0 + ty;
        if (gy >= H) break;
        uint32_t a = 0, b = 0;
        // window origin (ty, tx) is image (gy-3, gx-4)
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int m = 0; m < 9; ++m) {
                a |= (uint32_t)(tl[ty + k][tx + m] >= tl[ty + 6 - k][tx + 8 - m]) << (k * 9 + m);
                b |= (uint32_t)(tr[ty + k][tx + m] >= tr[ty + 6 - k][tx + 8 - m]) << (k * 9 + m);
            }
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            a |= (uint32_t)(tl[ty + 3][tx + m] >= tl[ty + 3][tx + 8 - m]) << (27 + m);
            b |= (uint32_t)(tr[ty + 3][tx + m] >= tr[ty + 3][tx + 8 - m]) << (27 + m);
        }
        cl[(long)gy * W + gx] = a;
        cr[(long)gy * W + gx] = b;
    }
}

// ----------------------------------------------------------------------------------------
// Hamming cost volume + WTA.
// Reference: HammingDistanceCostKernel (stereo_cost.cu:148-207): cost(x,d) =
// popc(CL(x) ^ CR(max(x-d,0))); cost_u1_to_f1_transpose (209-228): min(cost, tau), [D][H][W];
// WTA_D_k (248-266): argmin_d, smallest d on ties.  RIGHT = cost_to_right
// (stereo_cost.cu:488-516): C_R[d][y][x] = C_L[d][y][min(x+d, W-1)].
// Block: 128 threads x 8 columns = 1024 columns of one row; the right census (and, for the
// right view, the left census) of the row segment plus a D-wide halo sit in LDS.  For every d a
// wave writes 512 contiguous bytes of the disparity plane.
// ----------------------------------------------------------------------------------------
constexpr int HM_THREADS = 128, HM_XT = 8, HM_SEG = HM_THREADS * HM_XT;

template <bool RIGHT>
__global__ __launch_bounds__(HM_THREADS) void hamming_wta_kernel(const uint32_t *__restrict__ cl,
                                                                 const uint32_t *__restrict__ cr, int W, int H,
                                                                 int D, int pitch, long plane,
                                                                 uint8_t *__restrict__ vol,
                                                                 float *__restrict__ dwta) {
    extern __shared__ uint32_t lds[];
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * HM_SEG;
    const uint32_t *clr = cl + (long)y * W;
    const uint32_t *crr = cr + (long)y * W;
    // right-census window: LDS[i] = CR(clamp(x0 - D + 1 + i, 0, W-1)), i < HM_SEG + D - 1
    const int base = x0 - D + 1;
    const int nwin = HM_SEG + D - 1;
    uint32_t *crw = lds;
    uint32_t *clw = lds + nwin;  // RIGHT only: LDS[i] = CL(clamp(x0 + i, 0, W-1)), i < HM_SEG + D - 1
    for (int i = threadIdx.x; i < nwin; i += HM_THREADS) {
        int xi = base + i;
        xi = xi < 0 ? 0 : (xi > W - 1 ? W - 1 : xi);
        crw[i] = crr[xi];
        if (RIGHT) {
            int xl = x0 + i;
            xl = xl > W - 1 ? W - 1 : xl;
            clw[i] = clr[xl];
        }
    }
    __syncthreads();
    const int xs = x0 + threadIdx.x * HM_XT;
    if (xs >= pitch) return;
    uint32_t c_l[HM_XT];
    int best[HM_XT], arg[HM_XT];
#pragma unroll
    for (int j = 0; j < HM_XT; ++j) {
        int x = xs + j;
        c_l[j] = clr[x < W ? x : W - 1];
        best[j] = 1 << 30;
        arg[j] = 0;
    }
    uint8_t *out = vol + (long)y * pitch + xs;
    for (int d = 0; d < D; ++d) {
        uint32_t packed[2] = {0u, 0u};
#pragma unroll
        for (int j = 0; j < HM_XT; ++j) {
            int x = xs + j;
            uint32_t a, b;
            if (!RIGHT) {
                a = c_l[j];
                b = crw[x - d - base];  // x - d >= base always; clamped to CR(0) below 0
            } else {
                int xl = x + d;
                xl = xl > W - 1 ? W - 1 : xl;
                a = clw[xl - x0 < nwin ? xl - x0 : nwin - 1];
                int xr = xl - d;
                b = crw[xr - base];
            }
            int c = __popc(a ^ b);
            c = c < 25 ? c : 25;
            if (c < best[j]) {
                best[j] = c;
                arg[j] = d;
            }
            packed[j >> 2] |= (uint32_t)c << (8 * (j & 3));
        }
        *(uint2 *)(out + (long)d * plane) = make_uint2(packed[0], packed[1]);
    }
#pragma unroll
    for (int j = 0; j < HM_XT; ++j) {
        int x = xs + j;
        if (x < W) dwta[(long)y * W + x] = (float)arg[j];
    }
}

// ----------------------------------------------------------------------------------------
// a given f32 volume [D][H][W] (the "MC-CNN" cost: read_volume_from_file, stereo_cost.cu:
// 399-419): the left view's out-of-view columns x <= d take column d's value
// (fill_out_of_view_k, 366-395), then every entry is truncated at tau with fminf
// (truncate_data_k, 230-246); one pass from the given volume.
// ----------------------------------------------------------------------------------------
__global__ void f32_volume_kernel(const float *__restrict__ raw, int W, int H, int D, float tau,
                                  float *__restrict__ out) {
    const long hw = (long)W * H, total = hw * D;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long d = i / hw, r = i - d * hw;
        const long y = r / W, x = r - y * W;
        const float v = x <= d ? raw[d * hw + y * W + d] : raw[i];
        out[i] = fminf(v, tau);
    }
}
// the right view (convert_vol_L2R_k + fill_out_of_view_k(false), stereo_cost.cu:471-516):
// C_R[d][y][x] = C_L[d][y][min(x+d, W-1)]
__global__ void f32_to_right_kernel(const float *__restrict__ in, int W, int H, int D, float *__restrict__ out) {
    const long hw = (long)W * H, total = hw * D;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long d = i / hw, r = i - d * hw;
        const long y = r / W, x = r - y * W;
        const long xs = x + d < W - 1 ? x + d : W - 1;
        out[i] = in[d * hw + y * W + xs];
    }
}
// the largest |entry| (as float bits: non-negative floats order as integers; +inf included)
__global__ void f32_absmax_kernel(const float *__restrict__ vol, long total, unsigned int *__restrict__ out) {
    unsigned int m = 0u;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x)
        m = max(m, __float_as_uint(fabsf(vol[i])));
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = max(m, (unsigned int)__shfl_xor((int)m, o));
    if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}
// WTA_D_k (stereo_cost.cu:248-266): the first d of the minimum under <, starting from FLT_MAX
__global__ void f32_wta_kernel(const float *__restrict__ vol, int W, int H, int D, float *__restrict__ dwta) {
    const long hw = (long)W * H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < hw; i += (long)gridDim.x * blockDim.x) {
        float mc = 3.402823466e+38f;
        int mi = 0;
        for (int d = 0; d < D; ++d) {
            const float c = vol[(long)d * hw + i];
            if (c < mc) {
                mc = c;
                mi = d;
            }
        }
        dwta[i] = (float)mi;
    }
}

// ----------------------------------------------------------------------------------------
// the "AD" cost (ad_cost, stereo_cost.cu:9-43): grey in float (img_BGR_to_grey, imutil.cu:
// 116-130, the double sum rounded once), then C[d][y][x] = min(|gL(x) - gR(x-d)|, tau) for
// x >= D-1 and 0 for every d in the columns x < D-1.
// ----------------------------------------------------------------------------------------
__global__ void grey_f32_kernel(const uint8_t *__restrict__ bgrL, const uint8_t *__restrict__ bgrR, long n,
                                float *__restrict__ g) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const uint8_t *a = bgrL + 3 * i, *b = bgrR + 3 * i;
        g[i] = (float)((double)a[0] * 0.114 + (double)a[1] * 0.587 + (double)a[2] * 0.299);
        g[n + i] = (float)((double)b[0] * 0.114 + (double)b[1] * 0.587 + (double)b[2] * 0.299);
    }
}
__global__ void ad_volume_kernel(const float *__restrict__ g, int W, int H, int D, float tau,
                                 float *__restrict__ out) {
    const long hw = (long)W * H, total = hw * D;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long d = i / hw, r = i - d * hw;
        const long x = r % W;
        out[i] = x < D - 1 ? 0.f : fminf(fabsf(g[r] - g[hw + r - d]), tau);
    }
}

// f32 [D][H][W] copy of the volume for the reference's float* getter (stereo_cost.hpp:42-46).
__global__ void volume_to_f32_kernel(const uint8_t *__restrict__ vol, int W, int H, int D, int pitch, long plane,
                                     float *__restrict__ out) {
    long n = (long)W * H;
    long total = n * D;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        long d = i / n, r = i % n;
        long y = r / W, x = r % W;
        out[i] = (float)vol[d * plane + y * pitch + x];
    }
}

// ----------------------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------------------
void census_launch(Ctx &c, const uint8_t *bgrL, const uint8_t *bgrR, int W, int H, uint32_t *cl, uint32_t *cr) {
    dim3 grid(div_up(W, CT_W), div_up(H, CT_H));
    ProfScope ps(c, "census", 14.0 * W * H);
    hipLaunchKernelGGL(census_kernel, grid, dim3(256), 0, c.stream, bgrL, bgrR, W, H, cl, cr);
    HBP_CHECK(hipGetLastError());
}

void hamming_launch(Ctx &c, Volume &v, bool right) {
    dim3 grid(div_up(v.pitch, HM_SEG), v.H);
    int nwin = HM_SEG + v.D - 1;
    size_t lds = sizeof(uint32_t) * (size_t)nwin * (right ? 2 : 1);
    // algorithmic bytes: census words in (8N), the u8 volume out (D*N), WTA out (4N)
    ProfScope ps(c, right ? "hamming_wta_right" : "hamming_wta", (double)(8 + v.D + 4) * v.W * v.H);
    if (right)
        hipLaunchKernelGGL(hamming_wta_kernel<true>, grid, dim3(HM_THREADS), lds, c.stream, v.cl.p, v.cr.p, v.W, v.H,
                           v.D, v.pitch, v.plane, v.vol.p, v.dwta.p);
    else
        hipLaunchKernelGGL(hamming_wta_kernel<false>, grid, dim3(HM_THREADS), lds, c.stream, v.cl.p, v.cr.p, v.W,
                           v.H, v.D, v.pitch, v.plane, v.vol.p, v.dwta.p);
    HBP_CHECK(hipGetLastError());
}

void volume_build(Ctx &c, Volume &v, const uint8_t *d_bgrL, const uint8_t *d_bgrR, int W, int H, int D) {
    HBP_REQUIRE(W >= 9 && H >= 7, "image must be at least 9x7 (census window)");
    HBP_REQUIRE(D >= 1 && D <= 1024, "max disparity must be in [1, 1024]");
    HBP_REQUIRE(D <= W, "max disparity must not exceed the image width");
    HBP_REQUIRE((long)W * H < (1L << 24), "images are limited to 2^24 pixels (ids are held in floats)");
    HBP_REQUIRE(W < 65536 && H < 65536, "image side must be < 65536");
    v.W = W;
    v.H = H;
    v.D = D;
    v.pitch = (int)round_up(W, 64);
    v.plane = (long)v.pitch * H;
    v.tau = kCensusTau;
    v.right = false;
    long n = (long)W * H;
    v.cl.alloc(c.pool, n);
    v.cr.alloc(c.pool, n);
    v.vol.alloc(c.pool, (size_t)v.plane * D);
    v.dwta.alloc(c.pool, n);
    census_launch(c, d_bgrL, d_bgrR, W, H, v.cl.p, v.cr.p);
    hamming_launch(c, v, false);
}

static int grid_of(const Ctx &c, long n) {
    const long b = (n + 255) / 256, cap = (long)c.num_cu * 16;
    return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

static void f32_shape(Volume &v, int W, int H, int D, float tau) {
    HBP_REQUIRE(W >= 9 && H >= 7, "image must be at least 9x7");
    HBP_REQUIRE(D >= 1 && D <= 1024, "max disparity must be in [1, 1024]");
    HBP_REQUIRE(D <= W, "max disparity must not exceed the image width");
    HBP_REQUIRE((long)W * H < (1L << 24), "images are limited to 2^24 pixels (ids are held in floats)");
    HBP_REQUIRE(W < 65536 && H < 65536, "image side must be < 65536");
    v.W = W;
    v.H = H;
    v.D = D;
    v.pitch = W;
    v.plane = (long)W * H;
    v.tau = tau;
    v.right = false;
    v.f32src = true;
}

// WTA of the f32 volume and the largest |entry| (which bounds the plane costs' fixed point)
static void f32_finish(Ctx &c, Volume &v) {
    const long n = (long)v.W * v.H;
    hipLaunchKernelGGL(f32_wta_kernel, dim3(grid_of(c, n)), dim3(256), 0, c.stream, v.f32.p, v.W, v.H, v.D,
                       v.dwta.p);
    DevBuf<unsigned int> am(c.pool, 1);
    HBP_CHECK(hipMemsetAsync(am.p, 0, sizeof(unsigned int), c.stream));
    hipLaunchKernelGGL(f32_absmax_kernel, dim3(grid_of(c, n * v.D)), dim3(256), 0, c.stream, v.f32.p, n * v.D,
                       am.p);
    HBP_CHECK(hipGetLastError());
    unsigned int bits = 0;
    HBP_CHECK(hipMemcpyAsync(&bits, am.p, sizeof(bits), hipMemcpyDeviceToHost, c.stream));
    HBP_CHECK(hipStreamSynchronize(c.stream));
    std::memcpy(&v.absmax, &bits, 4);
}

void volume_build_f32(Ctx &c, Volume &v, const float *d_raw, int W, int H, int D, float tau) {
    f32_shape(v, W, H, D, tau);
    const long n = (long)W * H;
    v.f32.alloc(c.pool, (size_t)n * D);
    v.dwta.alloc(c.pool, (size_t)n);
    ProfScope ps(c, "volume_f32", 8.0 * n * D + 4.0 * n);
    hipLaunchKernelGGL(f32_volume_kernel, dim3(grid_of(c, n * D)), dim3(256), 0, c.stream, d_raw, W, H, D, tau,
                       v.f32.p);
    f32_finish(c, v);
}

void volume_build_ad(Ctx &c, Volume &v, const uint8_t *d_bgrL, const uint8_t *d_bgrR, int W, int H, int D) {
    f32_shape(v, W, H, D, kAdTau);
    const long n = (long)W * H;
    v.f32.alloc(c.pool, (size_t)n * D);
    v.dwta.alloc(c.pool, (size_t)n);
    DevBuf<float> g(c.pool, 2 * (size_t)n);
    ProfScope ps(c, "volume_ad", 6.0 * n + 4.0 * n * D + 4.0 * n);
    hipLaunchKernelGGL(grey_f32_kernel, dim3(grid_of(c, n)), dim3(256), 0, c.stream, d_bgrL, d_bgrR, n, g.p);
    hipLaunchKernelGGL(ad_volume_kernel, dim3(grid_of(c, n * D)), dim3(256), 0, c.stream, g.p, W, H, D, kAdTau,
                       v.f32.p);
    f32_finish(c, v);
}

void volume_to_right(Ctx &c, Volume &v) {
    if (v.f32src) {
        const long n = (long)v.W * v.H;
        DevBuf<float> r(c.pool, (size_t)n * v.D);
        hipLaunchKernelGGL(f32_to_right_kernel, dim3(grid_of(c, n * v.D)), dim3(256), 0, c.stream, v.f32.p, v.W,
                           v.H, v.D, r.p);
        hipLaunchKernelGGL(f32_wta_kernel, dim3(grid_of(c, n)), dim3(256), 0, c.stream, r.p, v.W, v.H, v.D,
                           v.dwta.p);
        HBP_CHECK(hipGetLastError());
        std::swap(v.f32, r);
        v.right = true;
        return;
    }
    hamming_launch(c, v, true);
    v.right = true;
    v.f32.release();
}

float *volume_f32(Ctx &c, Volume &v) {
    if (!v.f32.p) {
        v.f32.alloc(c.pool, (size_t)v.W * v.H * v.D);
        int blocks = c.num_cu * 8;
        hipLaunchKernelGGL(volume_to_f32_kernel, dim3(blocks), dim3(256), 0, c.stream, v.vol.p, v.W, v.H, v.D,
                           v.pitch, v.plane, v.f32.p);
        HBP_CHECK(hipGetLastError());
    }
    return v.f32.p;
}

}  // namespace hbp
