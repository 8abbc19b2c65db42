// Stage (d): left-right consistency check and fill (post_process.cu of the reference, which
// runs it on the host with OpenCV; here it is four small kernels on the device).
#include <mutex>

#include "hbp_internal.h"

namespace hbp {

// doConsistencyCheck (post_process.cu:4-41), both views in one pass
__global__ void lr_consistency_kernel(const float *__restrict__ dL, const float *__restrict__ dR, int W, int H,
                                      uint8_t *__restrict__ fL, uint8_t *__restrict__ fR) {
    long n = (long)W * H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < 2 * n; i += (long)gridDim.x * blockDim.x) {
        int view = i >= n;
        long p = view ? i - n : i;
        int x = (int)(p % W), y = (int)(p / W);
        const float *d = view ? dR : dL;
        const float *o = view ? dL : dR;
        float sign = view ? -1.0f : 1.0f;
        float ds = d[p];
        int rx = (int)((float)x - ds * sign + 0.5f);
        uint8_t f = 0;
        if (rx >= 0 && rx < W) {
            float dsr = o[(long)y * W + rx];
            if ((double)fabsf(dsr - ds) > 1.5) f = 255;
        }
        (view ? fR : fL)[p] = f;
    }
}

// cv::dilate with the default 3x3 kernel (post_process.cu:68-69)
__global__ void dilate3_kernel(const uint8_t *__restrict__ in, int W, int H, uint8_t *__restrict__ out) {
    long n = (long)W * H;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        int x = (int)(p % W), y = (int)(p / W);
        uint8_t m = 0;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int yy = y + dy, xx = x + dx;
                if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
                uint8_t v = in[(long)yy * W + xx];
                m = v > m ? v : m;
            }
        out[p] = m;
    }
}

// nearest column at or left (right) of each pixel whose fail2 is not 255, -1 / W if none
__global__ void nearest_valid_kernel(const uint8_t *__restrict__ f2, int W, int H, int *__restrict__ lft,
                                     int *__restrict__ rgt) {
    for (long y = (long)blockIdx.x * blockDim.x + threadIdx.x; y < H; y += (long)gridDim.x * blockDim.x) {
        const uint8_t *row = f2 + y * W;
        int last = -1;
        for (int x = 0; x < W; ++x) {
            if (row[x] != 255) last = x;
            lft[y * W + x] = last;
        }
        last = W;
        for (int x = W - 1; x >= 0; --x) {
            if (row[x] != 255) last = x;
            rgt[y * W + x] = last;
        }
    }
}

__device__ __forceinline__ float get_z(const float *v, int x, int y) { return v[0] * (float)x + v[1] * (float)y + v[2]; }

// horizontal fill of the left view (post_process.cu:71-109)
__global__ void lr_fill_kernel(const uint8_t *__restrict__ fail, const int *__restrict__ lft,
                               const int *__restrict__ rgt, int W, int H, float *__restrict__ plane3) {
    long n = (long)W * H;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        if (fail[p] == 0) continue;
        int x = (int)(p % W), y = (int)(p / W);
        int xl = lft[p], xr = rgt[p];
        const float *pl = xl >= 0 ? plane3 + 3 * ((long)y * W + xl) : nullptr;
        const float *pr = xr < W ? plane3 + 3 * ((long)y * W + xr) : nullptr;
        const float *src;
        if (!pl && !pr) continue;
        if (!pl) src = pr;
        else if (!pr) src = pl;
        else if (get_z(pl, x, y) < get_z(pr, x, y)) src = pl;
        else src = pr;
        float a = src[0], b = src[1], c = src[2];
        plane3[3 * p] = a;
        plane3[3 * p + 1] = b;
        plane3[3 * p + 2] = c;
    }
}

// plane_label_to_disp (post_process.cu:373-390), unclamped
__global__ void plane_to_disp_kernel(const float *__restrict__ plane3, int W, int H, float *__restrict__ disp) {
    long n = (long)W * H;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        int x = (int)(p % W), y = (int)(p / W);
        disp[p] = get_z(plane3 + 3 * p, x, y);
    }
}

static int grid_for(const Ctx &c, long n, int block = 256) {
    long b = (n + block - 1) / block;
    long cap = (long)c.num_cu * 16;
    return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

// the four kernels; lft / rgt: 2 * W * H ints of scratch, every entry written before it is read
static void lr_launch(Ctx &c, const float *d_dispL, const float *d_dispR, float *d_planeL, int W, int H,
                      uint8_t *d_failL, uint8_t *d_failR, uint8_t *d_fail2, int *lft, int *rgt) {
    const long n = (long)W * H;
    hipLaunchKernelGGL(lr_consistency_kernel, dim3(grid_for(c, 2 * n)), dim3(256), 0, c.stream, d_dispL, d_dispR, W,
                       H, d_failL, d_failR);
    hipLaunchKernelGGL(dilate3_kernel, dim3(grid_for(c, n)), dim3(256), 0, c.stream, d_failL, W, H, d_fail2);
    hipLaunchKernelGGL(nearest_valid_kernel, dim3(grid_for(c, H, 64)), dim3(64), 0, c.stream, d_fail2, W, H, lft, rgt);
    hipLaunchKernelGGL(lr_fill_kernel, dim3(grid_for(c, n)), dim3(256), 0, c.stream, d_failL, lft, rgt, W, H,
                       d_planeL);
    HBP_CHECK(hipGetLastError());
}

void lr_check_fill(Ctx &c, const float *d_dispL, const float *d_dispR, float *d_planeL, int W, int H,
                   uint8_t *d_failL, uint8_t *d_failR, uint8_t *d_fail2) {
    long n = (long)W * H;
    // algorithmic bytes: both disparities in, three masks out, the left planes in and out
    ProfScope ps(c, "lr_check", (8.0 + 3.0 + 24.0) * n);
    DevBuf<int> lr(c.pool, 2 * n);
    lr_launch(c, d_dispL, d_dispR, d_planeL, W, H, d_failL, d_failR, d_fail2, lr.p, lr.p + n);
}

// ---- the allocation-free form (hbp_lr_check_fill_dev_noalloc).  Its scratch is this
// file-scope device array: the row scans' nearest valid columns (2 ints per pixel), then the
// three masks for callers that pass none (3 bytes per pixel), for images of up to
// kNoAllocMaxPixels pixels: 88 MiB that each GPU the library runs on holds until the process
// ends.  Host code never uses the array's name as an address: its device address is looked up
// once per device (lr_scratch_setup, from hbp_ctx_create) and passed to the kernels.  Every call
// on one GPU shares the array, so calls that overlap on one GPU race; no part of it needs
// clearing (each entry a call reads, the call has written first).
constexpr long kNoAllocMaxPixels = 1L << 23;
constexpr size_t kLrScratchBytes = (size_t)kNoAllocMaxPixels * (2 * sizeof(int) + 3);
static __device__ __align__(256) uint8_t scratch[kLrScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint8_t *g_scratch_ptr[kMaxDevices] = {};

void lr_scratch_setup(int device) {   // with `device` current
    HBP_REQUIRE(device >= 0 && device < kMaxDevices, "device index out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    HBP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    HBP_REQUIRE(p, "no device address for the scratch array");
    g_scratch_ptr[device] = (uint8_t *)p;
}

long lr_noalloc_max_pixels() { return kNoAllocMaxPixels; }

void lr_check_fill_noalloc(Ctx &c, const float *d_dispL, const float *d_dispR, float *d_planeL, int W, int H,
                           uint8_t *d_failL, uint8_t *d_failR, uint8_t *d_fail2) {
    const long n = (long)W * H;
    HBP_REQUIRE(n <= kNoAllocMaxPixels, "image larger than the allocation-free scratch holds (" +
                                            std::to_string(kNoAllocMaxPixels) + " pixels)");
    uint8_t *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = c.device >= 0 && c.device < kMaxDevices ? g_scratch_ptr[c.device] : nullptr;
    }
    HBP_REQUIRE(base, "the scratch array of this device is not set up");
    int *lft = (int *)base, *rgt = lft + n;
    uint8_t *masks = base + sizeof(int) * 2 * (size_t)kNoAllocMaxPixels;
    lr_launch(c, d_dispL, d_dispR, d_planeL, W, H, d_failL ? d_failL : masks, d_failR ? d_failR : masks + n,
              d_fail2 ? d_fail2 : masks + 2 * n, lft, rgt);
}

void plane_to_disp(Ctx &c, const float *d_plane3, int W, int H, float *d_disp) {
    hipLaunchKernelGGL(plane_to_disp_kernel, dim3(grid_for(c, (long)W * H)), dim3(256), 0, c.stream, d_plane3, W, H,
                       d_disp);
    HBP_CHECK(hipGetLastError());
}

}  // namespace hbp
