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
// One 64-lane workgroup per 64 HC columns of one row, lane l on columns HC l .. HC l + HC - 1
// (HC = 4 or 8, a tuning choice).  The census row segment with its D-wide halo (left view:
// CR(x - d); right view: CL(x + d) and the words CR(W-1-d) of the clamped columns) is staged in
// LDS once; disparities then go in blocks of DB = 4, the block's census words in a window of
// HC + DB - 1 registers read from LDS, so each cost is an xor, a popcount and a min.  The WTA
// keeps min(cost << 10 | d) per column in u16 halves (cost <= 25, d < 1024: 15 bits), columns
// (4i, 4i+2) and (4i+1, 4i+3) paired, so one packed min serves two columns.  STORE: also write the
// HC costs of each d (one dword per 4 columns) into the u8 volume (a wavefront writes 64 HC
// contiguous bytes of a disparity plane); without it only the WTA.
// ----------------------------------------------------------------------------------------
constexpr int DB = 4;
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pk_min_u16(uint32_t a, uint32_t b) {
    u16x2 x = __builtin_bit_cast(u16x2, a), y = __builtin_bit_cast(u16x2, b);
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(x, y));
}

// min(popcount(a ^ b), 25) in unsigned arithmetic
__device__ __forceinline__ uint32_t ham25(uint32_t a, uint32_t b) {
    const uint32_t p = (uint32_t)__popc(a ^ b);
    return p < 25u ? p : 25u;
}

// the costs c[HC] of one disparity d: WTA keys, and the volume's dwords when storing
template <int HC, bool STORE>
__device__ __forceinline__ void ham_emit(const uint32_t (&c)[HC], int d, uint32_t (&best)[HC / 2], uint8_t *out,
                                         long plane) {
    const uint32_t dp = (uint32_t)d | ((uint32_t)d << 16);
    uint32_t w[HC / 4];
#pragma unroll
    for (int q = 0; q < HC / 4; ++q) {
        const uint32_t t02 = c[4 * q] | (c[4 * q + 2] << 16), t13 = c[4 * q + 1] | (c[4 * q + 3] << 16);
        best[2 * q] = pk_min_u16(best[2 * q], (t02 << 10) | dp);
        best[2 * q + 1] = pk_min_u16(best[2 * q + 1], (t13 << 10) | dp);
        w[q] = t02 | (t13 << 8);
    }
    if (STORE) {
        if (HC == 4) *(uint32_t *)(out + (long)d * plane) = w[0];
        else *(uint2 *)(out + (long)d * plane) = make_uint2(w[0], w[HC / 4 - 1]);
    }
}

template <int HC, bool RIGHT, bool STORE>
__global__ __launch_bounds__(64) void hamming_wta_kernel(const uint32_t *__restrict__ cl,
                                                         const uint32_t *__restrict__ cr, int W, int H, int D,
                                                         int pitch, long plane, uint8_t *__restrict__ vol,
                                                         float *__restrict__ dwta) {
    static_assert(HC == 4 || HC == 8, "4 or 8 columns per lane");
    constexpr int HS = 64 * HC;
    extern __shared__ uint32_t lds[];
    const int y = blockIdx.y, seg0 = blockIdx.x * HS, lane = threadIdx.x;
    const uint32_t *clr = cl + (long)y * W;
    const uint32_t *crr = cr + (long)y * W;
    const int nwin = HS + D + DB - 1;
    // left: lds[i] = CR(clamp(seg0 - D - (DB-1) + i, 0, W-1)); right: lds[i] = CL(min(seg0 + i,
    // W-1)), then edge[d] = CR(max(W-1-d, 0))
    uint32_t *edge = lds + nwin;
    for (int i = lane; i < nwin; i += 64)
        lds[i] = RIGHT ? clr[min(seg0 + i, W - 1)] : crr[min(max(seg0 - D - (DB - 1) + i, 0), W - 1)];
    if (RIGHT)
        for (int d = lane; d < D; d += 64) edge[d] = crr[max(W - 1 - d, 0)];
    __syncthreads();
    const int x0 = seg0 + lane * HC;
    if (x0 >= W) return;   // padding columns are never read
    uint32_t fix[HC];      // left view: CL(x); right view: CR(x)
#pragma unroll
    for (int j = 0; j < HC; ++j) fix[j] = (RIGHT ? crr : clr)[min(x0 + j, W - 1)];
    const uint32_t cl_last = RIGHT ? clr[W - 1] : 0u;
    uint32_t best[HC / 2];   // u16 keys of columns (4q, 4q+2) and (4q+1, 4q+3)
#pragma unroll
    for (int q = 0; q < HC / 2; ++q) best[q] = 0xFFFFFFFFu;
    uint8_t *out = vol + (long)y * pitch + x0;
    // left: CR(x0 + j - d) = win_base[j - d + DB - 1]; right: CL(x0 + j + d) = win_base[j + d]
    const uint32_t *win_base = lds + lane * HC + (RIGHT ? 0 : D);
    int d0 = 0;
    for (; d0 + DB <= D; d0 += DB) {
        uint32_t win[HC + DB - 1];
#pragma unroll
        for (int m = 0; m < HC + DB - 1; ++m) win[m] = RIGHT ? win_base[d0 + m] : win_base[m - d0];
#pragma unroll
        for (int dd = 0; dd < DB; ++dd) {
            const int d = d0 + dd;
            uint32_t c[HC];
#pragma unroll
            for (int j = 0; j < HC; ++j)
                c[j] = RIGHT ? ham25(win[j + dd], fix[j]) : ham25(fix[j], win[j - dd + DB - 1]);
            if (RIGHT && x0 + HC - 1 + d > W - 1) {
                // columns with x + d > W - 1 take the entry of column W - 1 - d
                const uint32_t ce = ham25(cl_last, edge[d]);
#pragma unroll
                for (int j = 0; j < HC; ++j) c[j] = x0 + j + d > W - 1 ? ce : c[j];
            }
            ham_emit<HC, STORE>(c, d, best, out, plane);
        }
    }
    for (int d = d0; d < D; ++d) {   // D not a multiple of DB: the last disparities one by one
        uint32_t c[HC];
#pragma unroll
        for (int j = 0; j < HC; ++j) {
            const int x = x0 + j;
            uint32_t a, b;
            if (!RIGHT) {
                a = fix[j];
                b = win_base[j - d + DB - 1];                 // CR(x - d)
            } else {
                const bool clamped = x + d > W - 1;
                a = clamped ? cl_last : win_base[j + d];      // CL(min(x + d, W - 1))
                b = clamped ? edge[d] : fix[j];               // CR(max(min(x + d, W - 1) - d, 0))
            }
            c[j] = ham25(a, b);
        }
        ham_emit<HC, STORE>(c, d, best, out, plane);
    }
#pragma unroll
    for (int q = 0; q < HC / 4; ++q) {
        const uint32_t arg[4] = {best[2 * q] & 1023u, best[2 * q + 1] & 1023u, (best[2 * q] >> 16) & 1023u,
                                 (best[2 * q + 1] >> 16) & 1023u};
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (x0 + 4 * q + k < W) dwta[(long)y * W + x0 + 4 * q + k] = (float)arg[k];
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

// store: also write the u8 volume (else the WTA alone)
template <int HC>
static void hamming_launch_hc(Ctx &c, Volume &v, bool right, bool store) {
    dim3 grid((unsigned)div_up(v.W, 64 * HC), (unsigned)v.H);
    const size_t lds = sizeof(uint32_t) * ((size_t)64 * HC + v.D + DB - 1 + (right ? v.D : 0));
    uint8_t *out = store ? v.vol.p : nullptr;
    if (right && store)
        hipLaunchKernelGGL((hamming_wta_kernel<HC, true, true>), grid, dim3(64), lds, c.stream, v.cl.p, v.cr.p, v.W,
                           v.H, v.D, v.pitch, v.plane, out, v.dwta.p);
    else if (right)
        hipLaunchKernelGGL((hamming_wta_kernel<HC, true, false>), grid, dim3(64), lds, c.stream, v.cl.p, v.cr.p, v.W,
                           v.H, v.D, v.pitch, v.plane, out, v.dwta.p);
    else if (store)
        hipLaunchKernelGGL((hamming_wta_kernel<HC, false, true>), grid, dim3(64), lds, c.stream, v.cl.p, v.cr.p, v.W,
                           v.H, v.D, v.pitch, v.plane, out, v.dwta.p);
    else
        hipLaunchKernelGGL((hamming_wta_kernel<HC, false, false>), grid, dim3(64), lds, c.stream, v.cl.p, v.cr.p,
                           v.W, v.H, v.D, v.pitch, v.plane, out, v.dwta.p);
    HBP_CHECK(hipGetLastError());
}

// store: also write the u8 volume (else the WTA alone); the context's hamming_cols picks the
// columns per lane (same results either way)
static void hamming_launch(Ctx &c, Volume &v, bool right, bool store) {
    // algorithmic bytes: census words in (8N), the u8 volume out (D*N) when stored, WTA out (4N)
    const double n = (double)v.W * v.H;
    ProfScope ps(c, std::string(store ? "hamming_volume" : "hamming_wta") + (right ? "_right" : ""),
                 (8.0 + (store ? v.D : 0) + 4.0) * n);
    if (c.hamming_cols == 8) hamming_launch_hc<8>(c, v, right, store);
    else hamming_launch_hc<4>(c, v, right, store);
}

// the u8 volume of a census cost, written on first request: BP reads only the census words and
// the WTA, so the plain, batched and served paths never need it
uint8_t *volume_u8(Ctx &c, Volume &v) {
    HBP_REQUIRE(!v.f32src, "an f32 (AD or MC-CNN) volume has no u8 form");
    if (!v.vol.p) {
        v.vol.alloc(c.pool, (size_t)v.plane * v.D);
        hamming_launch(c, v, v.right, true);
    }
    return v.vol.p;
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
    v.vol.release();
    v.dwta.alloc(c.pool, n);
    census_launch(c, d_bgrL, d_bgrR, W, H, v.cl.p, v.cr.p);
    hamming_launch(c, v, false, false);
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
    hamming_launch(c, v, true, false);
    v.right = true;
    v.vol.release();
    v.f32.release();
}

float *volume_f32(Ctx &c, Volume &v) {
    if (!v.f32.p) {
        const uint8_t *u8 = volume_u8(c, v);
        v.f32.alloc(c.pool, (size_t)v.W * v.H * v.D);
        int blocks = c.num_cu * 8;
        hipLaunchKernelGGL(volume_to_f32_kernel, dim3(blocks), dim3(256), 0, c.stream, u8, v.W, v.H, v.D,
                           v.pitch, v.plane, v.f32.p);
        HBP_CHECK(hipGetLastError());
    }
    return v.f32.p;
}

}  // namespace hbp
