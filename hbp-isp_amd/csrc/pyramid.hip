// Stage (b): CRTrees superpixel pyramid (segment.cu, CRTrees.cu, imutil.cu of the reference).
//
// Per level: one cooperative kernel runs the whole CRTrees iteration with device-side loop
// control (the reference polls a flag from the host after every kernel); then compaction,
// the edge reduction (sort of (s', t') keys, run-length dedupe, boundary-length sums), the
// per-parent sums in a fixed child order and the next level's 1-NN.  The host reads back two
// counts per level (V', E') to size the next level.

#include <hip/hip_cooperative_groups.h>

#include <cmath>

#include "hbp_internal.h"


namespace cg = cooperative_groups;

namespace hbp {

__constant__ double c_srgb_lut[256];

// ----------------------------------------------------------------------------------------
// Lab conversion (imutil.cu:4-40 BGR_to_LAB, img_uchar3_to_float4 170-185) and the level-0
// mean array (init_data_mean, segment.cu:856-861 -> img_struct_to_array imutil.cu:92-114).
// ----------------------------------------------------------------------------------------
__global__ void lab_kernel(const uint8_t *__restrict__ bgr, int W, int H, float4 *__restrict__ lab,
                           float *__restrict__ mean) {
    long n = (long)W * H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double b = c_srgb_lut[bgr[3 * i]], g = c_srgb_lut[bgr[3 * i + 1]], r = c_srgb_lut[bgr[3 * i + 2]];
        double x = (r * 0.412453 + g * 0.357580 + b * 0.180423) / 0.950456;
        double y = (r * 0.212671 + g * 0.715160 + b * 0.072169) / 1.0;
        double z = (r * 0.019334 + g * 0.119193 + b * 0.950227) / 1.088754;
        const double epsilon = 216 / 24389.0;
        const double k = 841 / 108.0;
        const double kb = 16 / 116.0;
        x = (x > epsilon) ? cbrt_newton(x) : (k * x) + kb;
        y = (y > epsilon) ? cbrt_newton(y) : (k * y) + kb;
        z = (z > epsilon) ? cbrt_newton(z) : (k * z) + kb;
        float L = (float)((116 * y) - 16);
        float A = (float)(500 * (x - y));
        float B = (float)(200 * (y - z));
        lab[i] = make_float4(L, A, B, 1.0f);
        mean[i] = L;
        mean[n + i] = A;
        mean[2 * n + i] = B;
        mean[3 * n + i] = (float)(int)(i % W);
        mean[4 * n + i] = (float)(int)(i / W);
        mean[5 * n + i] = 1.0f;
    }
}

// The Gaussian pre-blur for sigma > FLT_MIN (img_CPU_to_GPU, segment.cu:795-809): gauss_blur
// (imutil.cu:399-441) of img_uchar3_to_float4's (B, G, R, 1), then BGR_to_LAB and the level-0
// mean array as in lab_kernel.  A block takes a 32x8 tile of pixels and stages its clamped
// (32+2S)x(8+2S) neighbourhood as floats in LDS; each pixel's channels are left folds in f32
// over r, then c, in [-S, S] of pixel * w[r][c] (the reference's order).  The blurred w channel
// is the same fold of the weights for every pixel (every input w is 1), so the host passes it.
// Where the reference's edge blocks race (several threads write the last column / row, from
// windows centred beyond the image), each pixel here takes the window centred on itself.
constexpr int kBlurTX = 32, kBlurTY = 8, kBlurMaxS = 8;
struct GaussTaps {
    float w[(2 * kBlurMaxS + 1) * (2 * kBlurMaxS + 1)];   // (2S+1)^2 used, row-major
};

__global__ __launch_bounds__(256) void blur_lab_kernel(const uint8_t *__restrict__ bgr, int W, int H, int S,
                                                       GaussTaps taps, float wsum, float4 *__restrict__ lab,
                                                       float *__restrict__ mean) {
    constexpr int kTile = (kBlurTX + 2 * kBlurMaxS) * (kBlurTY + 2 * kBlurMaxS);
    __shared__ float tb[kTile], tg[kTile], tr[kTile];
    __shared__ float sw[(2 * kBlurMaxS + 1) * (2 * kBlurMaxS + 1)];
    const int F = 2 * S + 1;
    const int tw = kBlurTX + 2 * S, th = kBlurTY + 2 * S;
    const int x0 = (int)blockIdx.x * kBlurTX - S, y0 = (int)blockIdx.y * kBlurTY - S;
    for (int k = threadIdx.x; k < F * F; k += blockDim.x) sw[k] = taps.w[k];
    for (int k = threadIdx.x; k < tw * th; k += blockDim.x) {
        int ty = k / tw, tx = k - ty * tw;
        int yy = min(max(y0 + ty, 0), H - 1), xx = min(max(x0 + tx, 0), W - 1);
        const uint8_t *px = bgr + 3 * ((long)yy * W + xx);
        tb[k] = (float)px[0];
        tg[k] = (float)px[1];
        tr[k] = (float)px[2];
    }
    __syncthreads();
    const int lx = (int)threadIdx.x % kBlurTX, ly = (int)threadIdx.x / kBlurTX;
    const int x = (int)blockIdx.x * kBlurTX + lx, y = (int)blockIdx.y * kBlurTY + ly;
    if (x >= W || y >= H) return;
    float fb = 0.f, fg = 0.f, fr = 0.f;
    for (int r = 0; r < F; ++r) {
        const int row = (ly + r) * tw + lx;
        for (int c = 0; c < F; ++c) {
            const float w = sw[r * F + c];
            fb = fb + tb[row + c] * w;
            fg = fg + tg[row + c] * w;
            fr = fr + tr[row + c] * w;
        }
    }
    double b = fb / 255.0, g = fg / 255.0, rr = fr / 255.0;
    rr = (rr > 0.04045) ? pow24_newton((rr + 0.055) / 1.055) : rr / 12.92;
    g = (g > 0.04045) ? pow24_newton((g + 0.055) / 1.055) : g / 12.92;
    b = (b > 0.04045) ? pow24_newton((b + 0.055) / 1.055) : b / 12.92;
    double X = (rr * 0.412453 + g * 0.357580 + b * 0.180423) / 0.950456;
    double Y = (rr * 0.212671 + g * 0.715160 + b * 0.072169) / 1.0;
    double Z = (rr * 0.019334 + g * 0.119193 + b * 0.950227) / 1.088754;
    const double epsilon = 216 / 24389.0;
    const double k = 841 / 108.0;
    const double kb = 16 / 116.0;
    X = (X > epsilon) ? cbrt_newton(X) : (k * X) + kb;
    Y = (Y > epsilon) ? cbrt_newton(Y) : (k * Y) + kb;
    Z = (Z > epsilon) ? cbrt_newton(Z) : (k * Z) + kb;
    const float L = (float)((116 * Y) - 16);
    const float A = (float)(500 * (X - Y));
    const float B = (float)(200 * (Y - Z));
    const long n = (long)W * H, i = (long)y * W + x;
    lab[i] = make_float4(L, A, B, wsum);
    mean[i] = L;
    mean[n + i] = A;
    mean[2 * n + i] = B;
    mean[3 * n + i] = (float)x;
    mean[4 * n + i] = (float)y;
    mean[5 * n + i] = wsum;
}

// segment.cu:14-22
__device__ __forceinline__ float dist_px(float4 s, float4 t) {
    float dx = s.x - t.x, dy = s.y - t.y, dz = s.z - t.z;
    float d = dx * dx + dy * dy + dz * dz;
    float h = dx * 1.f;
    d += h > 0.f ? h : 0.f;
    return d;
}

// compute_1nn_grid_k (segment.cu:57-195) with bdl = 1 (segment.cu:846-850): neighbours in the
// order u, d, l, r (+ ul, ur, dr, dl), first strict minimum, border slots are self-loops.
// edist (Min/Max linkages only, else null): each slot's distance, 0 for a self-loop slot (the
// reduction drops those).
__global__ void nn_grid_kernel(const float4 *__restrict__ lab, int W, int H, int num_nb, int *__restrict__ nn,
                               float *__restrict__ dist_min, int2 *__restrict__ bd, int *__restrict__ bdl,
                               float *__restrict__ edist) {
    long n = (long)W * H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int x = (int)(i % W), y = (int)(i / W);
        float4 v = lab[i];
        float best = FLT_MAX;
        int arg = 0;
        bool ok[8] = {y > 0, y < H - 1, x > 0, x < W - 1, y > 0 && x > 0, y > 0 && x < W - 1,
                      y < H - 1 && x < W - 1, y < H - 1 && x > 0};
        long nb[8] = {i - W, i + W, i - 1, i + 1, i - W - 1, i - W + 1, i + W + 1, i + W - 1};
        for (int k = 0; k < num_nb; ++k) {
            long j = ok[k] ? nb[k] : i;
            float d = 0.f;
            if (ok[k]) {
                d = dist_px(v, lab[j]);
                if (d < best) {
                    best = d;
                    arg = (int)j;
                }
            }
            bd[num_nb * i + k] = make_int2((int)i, (int)j);
            bdl[num_nb * i + k] = 1;
            if (edist) edist[num_nb * i + k] = d;
        }
        nn[i] = arg;
        dist_min[i] = best;
    }
}

// ----------------------------------------------------------------------------------------
// CRTrees::get_clus (CRTrees.cu:315-354).  The loop is the reference's literal iteration
//   A = nnc, B = nnn, M = max_d, C = clus; nnw is the working copy of nn;
// re_link reads a snapshot of A (writes go to B), the double-buffered form of the in-place
// kernel (the oracle checks both forms give the same labels).  Three executions of it:
//  * levels up to kCrtSingleMax nodes: one 1024-thread block, phases separated by
//    __syncthreads;
//  * larger levels: one kernel per phase, launched in a schedule of at most kCrtRounds rounds
//    of at most kCrtJumps pointer-jumping steps and at most kCrtMaxJumps max-jumping steps.
//    The loop flags live in device memory; a kernel whose loop has already ended returns at
//    once (a launch all the same), and the ping-pong buffer of jumping step j is fixed by the
//    parity of j, so the host never waits inside the loop.  A loop that needs more steps than
//    scheduled sets the overflow word.  The schedule of a level is the steps the context's
//    previous levels used plus one of each (CrtSched), so that few dead launches are queued; a
//    level that overflows it runs again with the full schedule;
//  * then (rare) one cooperative kernel with grid barriers recomputes the level with the full
//    bounds (64 rounds, 64 jumps).
// ----------------------------------------------------------------------------------------
struct CRTBuf {
    int *nnw, *A, *B, *M, *C, *mask, *maskc, *indeg, *flags, *status;
};
constexpr int kCrtSingleMax = 8000;
constexpr int kCrtRounds = 5, kCrtJumps = 8, kCrtMaxJumps = 8;
static_assert(CrtSched{}.rounds == kCrtRounds && CrtSched{}.jumps == kCrtJumps && CrtSched{}.max_jumps == kCrtMaxJumps,
              "the default schedule is the full one");
struct CRTFlags {               // zeroed before a scheduled run
    int more[kCrtRounds + 1];   // round r continues (some node has indegree > 1)
    int jl[kCrtRounds][kCrtJumps];
    int jm[kCrtMaxJumps];
    int overflow;               // the loop needed more steps than scheduled
};

template <bool GRID>
__device__ __forceinline__ void crt_sync() {
    if (GRID) cg::this_grid().sync();
    else __syncthreads();
}

// loop flags: a block-wide OR, then one device-scope atomic per block; read back with an
// agent-scope atomic load after the barrier.
__device__ __forceinline__ void flag_or(int *flags, int slot, int local) {
    if (__syncthreads_or(local) && threadIdx.x == 0) atomicOr(&flags[slot], 1);
}
__device__ __forceinline__ int flag_read(const int *flags, int slot) {
    return __hip_atomic_load(&flags[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// sets a loop flag (it only ever goes from 0 to 1): many blocks end with this, so those that
// find it set skip the atomic, whose updates of one word are serialised at the memory side
__device__ __forceinline__ void flag_set(int *flags, int slot) {
    if (!flag_read(flags, slot)) atomicOr(&flags[slot], 1);
}
__device__ __forceinline__ void flag_clear(int *flags, int slot) {
    __hip_atomic_store(&flags[slot], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- the phases (every execution form runs these bodies over its share of the nodes)
__device__ __forceinline__ void crt_delink(const CRTBuf &s, long i) {     // CRTrees.cu:58-74, 304-311
    int m = s.mask[i];
    s.maskc[i] = m;
    int dl = s.indeg[i] > 1;
    s.mask[i] = dl;
    if (dl) s.A[i] = (int)i;
}
// pointer_jumping (CRTrees.cu:314-333), one step from p to q; returns the continue condition
__device__ __forceinline__ int crt_jump(const CRTBuf &s, const int *p, int *q, long i) {
    if (!s.maskc[i]) return 0;
    int a = p[i], b = p[a];
    q[i] = b;
    return (!s.mask[a] && s.mask[b]) ? 1 : 0;
}
__device__ __forceinline__ void crt_relink(const CRTBuf &s, long i) {     // CRTrees.cu:178-201
    int v = s.A[i];
    if (s.mask[i]) {
        int a = s.nnw[i];
        v = a;
        if (!s.mask[a]) v = s.A[a];
    } else if (!s.maskc[i]) {
        int a = s.A[i];
        if (s.maskc[a]) v = s.A[a];
    }
    s.B[i] = v;
}
__device__ __forceinline__ void crt_relink_all(const CRTBuf &s, long i) { // CRTrees.cu:203-225
    int a = s.B[i];
    if (s.mask[i]) s.nnw[i] = a;
    else if (s.mask[a]) s.nnw[i] = s.B[a];
}
// pointer_jumping_max (CRTrees.cu:274-298), one step; returns whether the label changed
__device__ __forceinline__ int crt_max_jump(const int *p, int *q, const int *mx, int *cl, long i) {
    int a = p[i];
    int m = mx[i];
    int c = max(m, mx[a]);
    cl[i] = c;
    q[i] = p[a];
    return c != m ? 1 : 0;
}

// ---- one kernel for the whole loop (single block, or cooperative grid)
template <bool GRID>
__global__ __launch_bounds__(1024) void crtrees_kernel(const int *__restrict__ nn0, int V, CRTBuf s,
                                                       int *__restrict__ label) {
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nth = (long)gridDim.x * blockDim.x;
    int *flags = s.flags;
    int fi = 0;
    for (long i = tid; i < V; i += nth) {
        int v = nn0[i];
        s.nnw[i] = v;
        s.A[i] = v;
        s.mask[i] = 1;
    }
    if (tid < 4) flag_clear(flags, (int)tid);
    crt_sync<GRID>();
    const int kMaxRounds = 64, kMaxJumps = 64;
    int rounds = 0;
    for (;;) {
        // get_in_degree (CRTrees.cu:287-303)
        for (long i = tid; i < V; i += nth) s.indeg[i] = 0;
        crt_sync<GRID>();
        for (long i = tid; i < V; i += nth)
            if (s.mask[i]) atomicAdd(&s.indeg[s.A[i]], 1);
        crt_sync<GRID>();
        int loc = 0;
        for (long i = tid; i < V; i += nth)
            if (s.mask[i] && s.indeg[i] > 1) loc = 1;
        flag_or(flags, fi, loc);
        if (tid == 0) flag_clear(flags, (fi + 1) & 3);
        crt_sync<GRID>();
        bool more = flag_read(flags, fi) != 0;
        fi = (fi + 1) & 3;
        if (!more) break;
        if (++rounds > kMaxRounds) {
            if (tid == 0) *s.status = 1;
            break;
        }
        for (long i = tid; i < V; i += nth) crt_delink(s, i);
        crt_sync<GRID>();
        // ping-pong between A and B; the caller keeps reading buffer A, which holds the
        // result of the last even step
        {
            int *p = s.A, *q = s.B;
            for (int it = 0;; ++it) {
                int jl = 0;
                for (long i = tid; i < V; i += nth) jl |= crt_jump(s, p, q, i);
                flag_or(flags, fi, jl);
                if (tid == 0) flag_clear(flags, (fi + 1) & 3);
                crt_sync<GRID>();
                bool jp = flag_read(flags, fi) != 0;
                fi = (fi + 1) & 3;
                int *t = p;
                p = q;
                q = t;
                if (!jp) break;
                if (it >= kMaxJumps) {
                    if (tid == 0) *s.status = 2;
                    break;
                }
            }
        }
        for (long i = tid; i < V; i += nth) crt_relink(s, i);
        crt_sync<GRID>();
        for (long i = tid; i < V; i += nth) crt_relink_all(s, i);
        crt_sync<GRID>();
        for (long i = tid; i < V; i += nth) s.A[i] = s.nnw[i];   // CRTrees.cu:423
        crt_sync<GRID>();
    }
    for (long i = tid; i < V; i += nth) s.M[i] = s.A[i];
    crt_sync<GRID>();
    int *p = s.A, *q = s.B, *mx = s.M, *cl = s.C;
    for (int it = 0;; ++it) {
        int jl = 0;
        for (long i = tid; i < V; i += nth) jl |= crt_max_jump(p, q, mx, cl, i);
        flag_or(flags, fi, jl);
        if (tid == 0) flag_clear(flags, (fi + 1) & 3);
        crt_sync<GRID>();
        bool jp = flag_read(flags, fi) != 0;
        fi = (fi + 1) & 3;
        int *t = p;
        p = q;
        q = t;
        t = mx;
        mx = cl;
        cl = t;
        if (!jp) break;
        if (it >= kMaxJumps) {
            if (tid == 0) *s.status = 3;
            break;
        }
    }
    for (long i = tid; i < V; i += nth) label[i] = mx[i];
}

int crtrees_coop_blocks(int num_cu) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, crtrees_kernel<true>, 256, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    if (per_cu > 4) per_cu = 4;
    return per_cu * num_cu;
}

// ---- the scheduled form: one kernel per phase; every kernel first reads its gate
#define CRT_NODES(i) for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (long)gridDim.x * blockDim.x)

__global__ void crt_init_kernel(const int *__restrict__ nn0, int V, CRTBuf s) {
    CRT_NODES(i) {
        int v = nn0[i];
        s.nnw[i] = v;
        s.A[i] = v;
        s.mask[i] = 1;
        s.indeg[i] = 0;
    }
}
// round r runs when round r-1 continued
__device__ __forceinline__ bool crt_round_open(const CRTFlags *f, int r) {
    return r == 0 || flag_read(f->more, r - 1) != 0;
}
__global__ void crt_count_kernel(int V, CRTBuf s, const CRTFlags *f, int r) {
    if (!crt_round_open(f, r)) return;
    CRT_NODES(i) if (s.mask[i]) atomicAdd(&s.indeg[s.A[i]], 1);
}
// rounds: the schedule's rounds (the check after the last one overflows if the loop goes on)
__global__ void crt_check_kernel(int V, CRTBuf s, CRTFlags *f, int r, int rounds) {
    if (!crt_round_open(f, r)) return;
    int loc = 0;
    CRT_NODES(i) if (s.mask[i] && s.indeg[i] > 1) loc = 1;
    if (__syncthreads_or(loc) && threadIdx.x == 0) {
        flag_set(f->more, r);
        if (r == rounds) flag_set(&f->overflow, 0);
    }
}
__global__ void crt_delink_kernel(int V, CRTBuf s, const CRTFlags *f, int r) {
    if (!flag_read(f->more, r)) return;
    CRT_NODES(i) crt_delink(s, i);
}
__global__ void crt_jump_kernel(int V, CRTBuf s, CRTFlags *f, int r, int j, int jumps) {
    if (!flag_read(f->more, r) || (j > 0 && !flag_read(f->jl[r], j - 1))) return;
    const int *p = (j & 1) ? s.B : s.A;
    int *q = (j & 1) ? s.A : s.B;
    int jl = 0;
    CRT_NODES(i) jl |= crt_jump(s, p, q, i);
    if (__syncthreads_or(jl) && threadIdx.x == 0) {
        flag_set(f->jl[r], j);
        if (j == jumps - 1) flag_set(&f->overflow, 0);
    }
}
__global__ void crt_relink_kernel(int V, CRTBuf s, const CRTFlags *f, int r) {
    if (!flag_read(f->more, r)) return;
    CRT_NODES(i) crt_relink(s, i);
}
// re_link_all, then the copy A = nnw (CRTrees.cu:423) and the next round's indegree reset
__global__ void crt_relink_all_kernel(int V, CRTBuf s, const CRTFlags *f, int r) {
    if (!flag_read(f->more, r)) return;
    CRT_NODES(i) {
        crt_relink_all(s, i);
        s.A[i] = s.nnw[i];
        s.indeg[i] = 0;
    }
}
__global__ void crt_max_init_kernel(int V, CRTBuf s, const CRTFlags *f) {
    if (flag_read(&f->overflow, 0)) return;
    CRT_NODES(i) s.M[i] = s.A[i];
}
__global__ void crt_max_kernel(int V, CRTBuf s, CRTFlags *f, int j, int max_jumps) {
    if (flag_read(&f->overflow, 0) || (j > 0 && !flag_read(f->jm, j - 1))) return;
    const int *p = (j & 1) ? s.B : s.A;
    int *q = (j & 1) ? s.A : s.B;
    const int *mx = (j & 1) ? s.C : s.M;
    int *cl = (j & 1) ? s.M : s.C;
    int jl = 0;
    CRT_NODES(i) jl |= crt_max_jump(p, q, mx, cl, i);
    if (__syncthreads_or(jl) && threadIdx.x == 0) {
        flag_set(f->jm, j);
        if (j == max_jumps - 1) flag_set(&f->overflow, 0);
    }
}
// the last step changes no label, so M and C hold the same labels.  After an overflow the
// labels are the identity: in range for the compaction that is queued behind this kernel before
// the host sees the overflow word (a rerun then overwrites them).  Thread 0 also leaves in
// status[1..3] the steps the run used: rounds that went on, the most jumping steps of a round
// and the max-jumping steps (a step counts when it ran; the last of a loop finds nothing more)
__global__ void crt_final_kernel(int V, const CRTBuf s, const CRTFlags *f, int *__restrict__ label, int rounds,
                                 int jumps, int max_jumps) {
    const bool ov = flag_read(&f->overflow, 0) != 0;
    CRT_NODES(i) label[i] = ov ? (int)i : s.M[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int ur = 0, uj = 0, um = 1;
        for (int r = 0; r < rounds; ++r) {
            if (!flag_read(f->more, r)) break;
            ++ur;
            int n = 1;
            for (int j = 0; j < jumps; ++j)
                if (flag_read(f->jl[r], j)) n = j + 2;
            uj = max(uj, n);
        }
        for (int j = 0; j < max_jumps; ++j)
            if (flag_read(f->jm, j)) um = j + 2;
        s.status[1] = ur;
        s.status[2] = uj;
        s.status[3] = um;
    }
}

// CRTrees::compact_clus_label (CRTrees.cu:387-402): rank of each label among the labels used.
// Labels from CRTrees name each component by its largest node, so label j is used exactly when
// label[j] == j: the flags need neither a cleared array nor a scatter.  pred[V] = 0.
__global__ void root_flags_kernel(const int *__restrict__ label, int V, int *__restrict__ pred) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i <= V; i += (long)gridDim.x * blockDim.x)
        pred[i] = (i < V && label[i] == (int)i) ? 1 : 0;
}
// any labels in [0, V): pred (cleared, V + 1 entries) marks those used
__global__ void mark_labels_kernel(const int *__restrict__ label, int V, int *__restrict__ pred) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (long)gridDim.x * blockDim.x)
        pred[label[i]] = 1;
}
__global__ void remap_labels_kernel(int *__restrict__ label, int V, const int *__restrict__ pos) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (long)gridDim.x * blockDim.x)
        label[i] = pos[label[i]];
}

// ----------------------------------------------------------------------------------------
// Edge reduction (reduce_boundary_dist, segment.cu:1050-1091)
// ----------------------------------------------------------------------------------------
// map_boundary (303-314): the edge's (clus[s], clus[t]) as one 64-bit key; with Min/Max the
// carried distance travels with it (sort_by_key, 1056-1058), as its float bits
__global__ void map_edges_kernel(const int2 *__restrict__ bd, const float *__restrict__ edist,
                                 const int *__restrict__ clus, long E, int bits,
                                 unsigned long long *__restrict__ keys, unsigned int *__restrict__ dbits) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (long)gridDim.x * blockDim.x) {
        int2 b = bd[e];
        unsigned long long s = (unsigned)clus[b.x], t = (unsigned)clus[b.y];
        keys[e] = (s << bits) | t;
        if (edist) dbits[e] = __float_as_uint(edist[e]);
    }
}
// last element of a run of equal keys with s' != t' (is_diff_bd, segment.cu:320-334)
// (flag[E] = 0, so the exclusive scan's element E is the new edge count).  Also starts the new
// level's E-entry bound: boundary-length sums at 0 and, with Min/Max (dist_out not null), the
// reduced distances at reduce_dist's start value (segment.cu:1070-1071)
__global__ void run_flags_kernel(const unsigned long long *__restrict__ keys, long E, int bits,
                                 int *__restrict__ flag, int *__restrict__ bdl_out, unsigned int *__restrict__ dist_out,
                                 unsigned int dist_start) {
    unsigned long long tmask = (1ull << bits) - 1ull;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e <= E; e += (long)gridDim.x * blockDim.x) {
        if (e == E) {
            flag[e] = 0;
            continue;
        }
        bdl_out[e] = 0;
        if (dist_out) dist_out[e] = dist_start;
        unsigned long long k = keys[e];
        bool self = (k >> bits) == (k & tmask);
        flag[e] = (!self && (e == E - 1 || keys[e + 1] != k)) ? 1 : 0;
    }
}
// reduce_bdl (segment.cu:368-386) + scatter_if (1085): every non-self element at sorted position
// e adds bdl_old[e] -- the bdl stored at that position, which the sort did not move -- to its
// run's slot (integer sums); the run's last element writes the edge.  Min/Max (reduce_dist,
// 338-364): the run's minimum / maximum of the carried distances, which are >= +0 and not NaN
// (squares, a max(0, .) term and non-negative penalties), so their float bits order as
// unsigned integers.
__global__ void scatter_edges_kernel(const unsigned long long *__restrict__ keys, const int *__restrict__ bdl_old,
                                     const unsigned int *__restrict__ dbits, int minlink,
                                     const int *__restrict__ flag, const int *__restrict__ pos, long E, int bits,
                                     int2 *__restrict__ bd_out, int *__restrict__ bdl_out,
                                     unsigned int *__restrict__ dist_out) {
    unsigned long long tmask = (1ull << bits) - 1ull;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (long)gridDim.x * blockDim.x) {
        unsigned long long k = keys[e];
        int s = (int)(k >> bits), t = (int)(k & tmask);
        if (s == t) continue;
        atomicAdd(&bdl_out[pos[e]], bdl_old[e]);
        if (dbits) {
            if (minlink) atomicMin(&dist_out[pos[e]], dbits[e]);
            else atomicMax(&dist_out[pos[e]], dbits[e]);
        }
        if (flag[e]) bd_out[pos[e]] = make_int2(s, t);
    }
}

__global__ void fill_int_kernel(int *__restrict__ p, long n, int v) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void iota_kernel(int *__restrict__ p, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = (int)i;
}
// the sort's input for grouping children by parent: (parent, child) = (clus[i], i)
__global__ void key_iota_kernel(const int *__restrict__ clus, long n, unsigned int *__restrict__ key,
                                int *__restrict__ val) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        key[i] = (unsigned int)clus[i];
        val[i] = (int)i;
    }
}
// offsets of each key's run in a sorted array whose keys cover [0, nk) with every key present
// (compacted labels): the run starting at j writes off[key[j]] = j; off[nk] = n
__global__ void run_starts_kernel(const unsigned int *__restrict__ key, long n, int nk, int *__restrict__ off) {
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        if (j == 0 || key[j] != key[j - 1]) off[key[j]] = (int)j;
        if (j == 0) off[nk] = (int)n;
    }
}
// CSR offsets by source of the edges [0, *en) sorted by (s, t) over nodes [0, nv], where a node
// may have no edge; also stores the edge count in *en_out
__global__ void edge_csr_kernel(const int2 *__restrict__ bd, const int *__restrict__ en, long emax, int nv,
                                int *__restrict__ off, int *__restrict__ en_out) {
    const long n = *en < emax ? *en : emax;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (n > 0 ? n : 1); e += (long)gridDim.x * blockDim.x) {
        if (n == 0) {
            for (int v = 0; v <= nv; ++v) off[v] = 0;
            *en_out = 0;
            continue;
        }
        const int s = bd[e].x;
        const int sp = e > 0 ? bd[e - 1].x : -1;
        for (int v = sp + 1; v <= s; ++v) off[v] = (int)e;
        if (e == n - 1)
            for (int v = s + 1; v <= nv; ++v) off[v] = (int)n;
        if (e == 0) *en_out = (int)n;
    }
}
__global__ void count_keys_kernel(const int *__restrict__ key, long n, int *__restrict__ cnt) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        atomicAdd(&cnt[key[i]], 1);
}

// reduce_data_mean (segment.cu:1134-1153) in the pinned order (children ascending, left fold
// from 0) + sum_to_mean (491-503) + compute_dist_inner_max (1116-1128, max from 0).
__global__ void parent_sums_kernel(const float *__restrict__ mean, int V, const int *__restrict__ child,
                                   const int *__restrict__ child_off, int Vn, const float *__restrict__ dist_min,
                                   float *__restrict__ sums, float *__restrict__ mu, float *__restrict__ dist_max) {
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < Vn; p += (long)gridDim.x * blockDim.x) {
        int b = child_off[p], e = child_off[p + 1];
        float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        float dm = 0.f;
        for (int j = b; j < e; ++j) {
            int c = child[j];
#pragma unroll
            for (int ch = 0; ch < 6; ++ch) acc[ch] += mean[(long)ch * V + c];
            dm = fmaxf(dm, dist_min[c]);
        }
#pragma unroll
        for (int ch = 0; ch < 6; ++ch) sums[(long)ch * Vn + p] = acc[ch];
#pragma unroll
        for (int ch = 0; ch < 5; ++ch) mu[(long)ch * Vn + p] = acc[ch] / acc[5];
        dist_max[p] = dm;
    }
}

// segment.cu:27-37
__device__ __forceinline__ float dist_ch(float s, float n, int c) {
    float d = (s - n) * (s - n);
    if (c == 0) {
        float h = (s - n) * 1.f;
        d += h > 0.f ? h : 0.f;
    }
    return d;
}

// compute_dist_mean_k (508-534) [Ward: ward_weight (561-577), s_w t_w / (s_w + t_w) on the
// pixel counts; Min/Max: the reduced carried distance instead], pernalize_dist_k (582-612),
// compute_1nn (1211-1231): per node over its out-edges (sorted by t): minimum distance,
// ties -> largest t.  Min/Max keep the penalised distance for the next level's reduction.
__global__ void level_nn_kernel(const float *__restrict__ mu, const float *__restrict__ wsum, int link, int Vn,
                                const int2 *__restrict__ bd, const int *__restrict__ off,
                                const float *__restrict__ dist_max, float *__restrict__ edist,
                                int *__restrict__ nn, float *__restrict__ dist_min) {
    const bool use_mean = link == kCentroidLink || link == kWardLink;
    for (long s = (long)blockIdx.x * blockDim.x + threadIdx.x; s < Vn; s += (long)gridDim.x * blockDim.x) {
        float L = mu[s], A = mu[Vn + s], B = mu[2L * Vn + s];
        const float sw = wsum[s];
        float best = FLT_MAX;
        int arg = 0;
        for (int e = off[s]; e < off[s + 1]; ++e) {
            int t = bd[e].y;
            float d;
            if (use_mean) {
                d = dist_ch(L, mu[t], 0) + dist_ch(A, mu[Vn + t], 1) + dist_ch(B, mu[2L * Vn + t], 2);
                if (link == kWardLink) {
                    const float tw = wsum[t];
                    d = sw * tw / (sw + tw) * d;
                }
            } else {
                d = edist[e];
            }
            float delta = d - dist_max[t];
            if (delta > 0) d += delta;
            if (!use_mean) edist[e] = d;
            if (d < best) {
                best = d;
                arg = t;
            } else if (d == best && t > arg) {
                arg = t;
            }
        }
        nn[s] = arg;
        dist_min[s] = best;
    }
}

// ----------------------------------------------------------------------------------------
// get_target_clus (segment.cu:954-1041): re-label a level of Vn < target clusters (whose
// children number V > target) into exactly `target` clusters.  Children are grouped by cluster
// in ascending child index (child / child_off); the pinned orders: each cluster's sum of its
// children's dist_min is a left fold in that order; clusters rank by that sum descending, ties
// by ascending id (a stable sort on the complemented float bits, the sums being >= +0); the
// children of a broken cluster take its values downwards in ascending child index.
// ----------------------------------------------------------------------------------------
__global__ void target_sums_kernel(const int *__restrict__ child, const int *__restrict__ child_off, int Vn,
                                   const float *__restrict__ dist_min, unsigned int *__restrict__ key,
                                   int *__restrict__ id) {
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < Vn; p += (long)gridDim.x * blockDim.x) {
        float acc = 0.f;
        for (int j = child_off[p]; j < child_off[p + 1]; ++j) acc += dist_min[child[j]];
        key[p] = ~__float_as_uint(acc);
        id[p] = (int)p;
    }
}
// the child count of the cluster at each rank (the scan's input; element Vn is 0)
__global__ void target_counts_kernel(const int *__restrict__ id, const int *__restrict__ child_off, int Vn,
                                     int *__restrict__ cnt) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i <= Vn; i += (long)gridDim.x * blockDim.x)
        cnt[i] = i < Vn ? child_off[id[i] + 1] - child_off[id[i]] : 0;
}
// is_break_cluster (segment.cu:647-671): i* = the first rank with cum[i] + Vn - (i+1) >= target,
// cum[i] = excl[i+1]; the condition only grows with i
__global__ void target_istar_kernel(const int *__restrict__ excl, int Vn, int target, int *__restrict__ istar) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < Vn; i += (long)gridDim.x * blockDim.x)
        if ((long)excl[i + 1] + Vn - (i + 1) >= target) atomicMin(istar, (int)i);
}
// pre_clus_label + scatter (segment.cu:675-690, 727-728): the value of each cluster, by id
__global__ void target_pos_kernel(const int *__restrict__ id, const int *__restrict__ excl, int Vn,
                                  const int *__restrict__ istar_p, int *__restrict__ pos) {
    const int is = min(*istar_p, Vn - 1);
    const int nb = excl[is + 1];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < Vn; i += (long)gridDim.x * blockDim.x)
        pos[id[i]] = i <= is ? excl[i + 1] - 1 : (int)i - is + nb - 1;
}
// reset_clus_label (segment.cu:695-718) over the children grouped by cluster: the k-th child of a
// broken cluster takes value pos - k, clamped at `exact`
__global__ void target_label_kernel(const unsigned int *__restrict__ parent_sorted, const int *__restrict__ child,
                                    const int *__restrict__ child_off, long V, int Vn, int target,
                                    const int *__restrict__ excl, const int *__restrict__ istar_p,
                                    const int *__restrict__ pos, int *__restrict__ clus) {
    const int is = min(*istar_p, Vn - 1);
    const int nb = excl[is + 1];
    const int exact = target - Vn + is;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < V; j += (long)gridDim.x * blockDim.x) {
        const int cc = (int)parent_sorted[j];
        const int pre = pos[cc];
        int label;
        if (pre < nb) {
            const int v = pre - (int)(j - child_off[cc]);
            label = v >= exact ? exact : v;
        } else {
            label = pre + exact - nb + 1;
        }
        clus[child[j]] = label;
    }
}

// ----------------------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------------------
static int grid_for(const Ctx &c, long n, int block = 256) {
    long b = (n + block - 1) / block;
    long cap = (long)c.num_cu * 16;
    return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

static int bits_for(long n) {
    int b = 1;
    while ((1L << b) < n) ++b;
    return b;
}

// compute_gauss_kernel (imutil.cu:443-471), host code in the reference too: S = round(3 sigma),
// w[r][c] = exp(-(r^2 + c^2)/(2 sigma^2)) in FP64 (libm), r then c over [-S, S], each times
// 1/sum, rounded to f32; *wsum = the f32 left fold of the weights (the blurred w channel).
static int gauss_taps(double sigma, GaussTaps &t, float *wsum) {
    const int S = (int)std::round(3 * sigma);
    HBP_REQUIRE(S <= kBlurMaxS, "sigma too large for the Gaussian pre-blur: round(3*sigma) must be at most 8 "
                                "(the reference's blur launch needs (16 + 2S)^2 <= 1024 threads)");
    const int F = 2 * S + 1;
    double f[(2 * kBlurMaxS + 1) * (2 * kBlurMaxS + 1)];
    double sum = 0;
    for (int r = -S; r <= S; ++r)
        for (int c = -S; c <= S; ++c) {
            double val = std::exp(-(double)(r * r + c * c) / (2 * sigma * sigma));
            f[(r + S) * F + c + S] = val;
            sum += val;
        }
    sum = 1. / sum;
    float acc = 0.f;
    for (int i = 0; i < F * F; ++i) {
        t.w[i] = (float)(f[i] * sum);
        acc = acc + 1.0f * t.w[i];
    }
    for (int i = F * F; i < (2 * kBlurMaxS + 1) * (2 * kBlurMaxS + 1); ++i) t.w[i] = 0.f;
    *wsum = acc;
    return S;
}

struct SrgbLut {
    double v[256];
    SrgbLut() {
        for (int i = 0; i < 256; ++i) {
            double cc = (double)(float)i / 255.0;
            v[i] = (cc > 0.04045) ? std::pow((cc + 0.055) / 1.055, 2.4) : cc / 12.92;
        }
    }
};
static void upload_srgb_lut(Ctx &c) {
    static const SrgbLut lut;   // built once, thread-safe (several host threads build pyramids)
    HBP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_srgb_lut), lut.v, sizeof(lut.v), 0, hipMemcpyHostToDevice,
                                     c.stream));
}

// device memory of one CRTrees run: 8 node arrays, the loop flags / status words of the
// one-kernel form and the flags of the scheduled form
struct CRTRun {
    DevBuf<int> buf;
    CRTBuf s;
    CRTFlags *f = nullptr;
    int V = 0;
};
static void crt_setup(Ctx &c, CRTRun &r, int V) {
    const size_t words = 8 + (sizeof(CRTFlags) + 3) / 4;
    r.V = V;
    r.buf.alloc(c.pool, (size_t)V * 8 + words);
    CRTBuf &s = r.s;
    s.nnw = r.buf.p;
    s.A = r.buf.p + (size_t)V;
    s.B = r.buf.p + 2 * (size_t)V;
    s.M = r.buf.p + 3 * (size_t)V;
    s.C = r.buf.p + 4 * (size_t)V;
    s.mask = r.buf.p + 5 * (size_t)V;
    s.maskc = r.buf.p + 6 * (size_t)V;
    s.indeg = r.buf.p + 7 * (size_t)V;
    s.flags = r.buf.p + 8 * (size_t)V;
    s.status = s.flags + 4;
    r.f = (CRTFlags *)(s.flags + 8);
    HBP_CHECK(hipMemsetAsync(s.flags, 0, sizeof(int) * words, c.stream));
}
// enqueues the run without waiting: one block for small levels, else the scheduled kernels of
// the schedule sc (each count at most its kCrt* bound)
static void crt_enqueue(Ctx &c, CRTRun &r, const int *d_nn, int *d_label, const CrtSched &sc) {
    const int V = r.V;
    const CRTBuf &s = r.s;
    ProfScope ps(c, "crtrees", 16.0 * V);
    if (V <= kCrtSingleMax) {
        hipLaunchKernelGGL(crtrees_kernel<false>, dim3(1), dim3(1024), 0, c.stream, d_nn, V, s, d_label);
        HBP_CHECK(hipGetLastError());
        return;
    }
    // the phases that end in a loop flag run on 2 blocks per CU (grid-stride over the nodes), so
    // that few blocks set the flag
    const dim3 g(grid_for(c, V)), gf(std::min(grid_for(c, V), 2 * c.num_cu)), b(256);
    CRTFlags *f = r.f;
    HBP_REQUIRE(sc.rounds >= 1 && sc.rounds <= kCrtRounds && sc.jumps >= 1 && sc.jumps <= kCrtJumps &&
                    sc.max_jumps >= 1 && sc.max_jumps <= kCrtMaxJumps,
                "CRTrees schedule out of its bounds");
    hipLaunchKernelGGL(crt_init_kernel, g, b, 0, c.stream, d_nn, V, s);
    for (int rd = 0; rd <= sc.rounds; ++rd) {
        hipLaunchKernelGGL(crt_count_kernel, g, b, 0, c.stream, V, s, f, rd);
        hipLaunchKernelGGL(crt_check_kernel, gf, b, 0, c.stream, V, s, f, rd, sc.rounds);
        if (rd == sc.rounds) break;
        hipLaunchKernelGGL(crt_delink_kernel, g, b, 0, c.stream, V, s, f, rd);
        for (int j = 0; j < sc.jumps; ++j)
            hipLaunchKernelGGL(crt_jump_kernel, gf, b, 0, c.stream, V, s, f, rd, j, sc.jumps);
        hipLaunchKernelGGL(crt_relink_kernel, g, b, 0, c.stream, V, s, f, rd);
        hipLaunchKernelGGL(crt_relink_all_kernel, g, b, 0, c.stream, V, s, f, rd);
    }
    hipLaunchKernelGGL(crt_max_init_kernel, g, b, 0, c.stream, V, s, f);
    for (int j = 0; j < sc.max_jumps; ++j)
        hipLaunchKernelGGL(crt_max_kernel, gf, b, 0, c.stream, V, s, f, j, sc.max_jumps);
    hipLaunchKernelGGL(crt_final_kernel, g, b, 0, c.stream, V, s, f, d_label, sc.rounds, sc.jumps, sc.max_jumps);
    HBP_CHECK(hipGetLastError());
}
// the one-kernel form over a cooperative grid (a scheduled run that overflowed)
static void crt_cooperative(Ctx &c, CRTRun &r, const int *d_nn, int *d_label) {
    CRTBuf s = r.s;
    int V = r.V;
    HBP_CHECK(hipMemsetAsync(s.flags, 0, sizeof(int) * 8, c.stream));
    ProfScope ps(c, "crtrees_cooperative", 16.0 * V);
    int blocks = c.coop_blocks;
    long need = ((long)V + 255) / 256;
    if (need < blocks) blocks = (int)need;
    void *args[] = {(void *)&d_nn, (void *)&V, (void *)&s, (void *)&d_label};
    if (c.coop_forbidden) throw CoopNeeded{};
    if (c.drain_before_coop) HBP_CHECK(hipStreamSynchronize(c.drain_before_coop));
    HBP_CHECK(hipLaunchCooperativeKernel((const void *)crtrees_kernel<true>, dim3(blocks), dim3(256), args, 0,
                                         c.stream));
}
static void crt_require(int status) {
    HBP_REQUIRE(status == 0, "CRTrees iteration did not converge within its bounds");
}

// Device words to the host with one wait and no copies: a one-lane kernel queued after the work
// that produces them stores them into the context's pinned host words.
struct WordPtrs {
    const int *p[8];
};
__global__ void publish_kernel(WordPtrs w, int n, int *__restrict__ out) {
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if (threadIdx.x == 0 && i < n) out[i] = *w.p[i];
}
static void read_words(Ctx &c, const WordPtrs &w, int n, int *out) {
    HBP_REQUIRE(n >= 1 && n <= 8, "one to eight words per read");
    int *hb = c.host_box();
    hipLaunchKernelGGL(publish_kernel, dim3(1), dim3(64), 0, c.stream, w, n, hb);
    HBP_CHECK(hipGetLastError());
    HBP_CHECK(hipStreamSynchronize(c.stream));
    for (int i = 0; i < n; ++i) out[i] = ((volatile int *)hb)[i];
}

void crtrees_labels(Ctx &c, const int *d_nn, int V, int *d_label) {
    CRTRun r;
    crt_setup(c, r, V);
    crt_enqueue(c, r, d_nn, d_label, CrtSched{});
    int w[2] = {0, 0};
    read_words(c, WordPtrs{{r.s.status, &r.f->overflow}}, 2, w);
    if (w[1]) {
        crt_cooperative(c, r, d_nn, d_label);
        read_words(c, WordPtrs{{r.s.status}}, 1, w);
    }
    crt_require(w[0]);
}

// compaction; also reads n_extra (<= 7) device words in the same wait.  roots: the labels come
// from CRTrees (a used label j has label[j] == j)
static int compact_labels_and_read(Ctx &c, int *d_label, int V, const int *const *extra, int *extra_host,
                                   int n_extra, bool roots) {
    DevBuf<int> pred(c.pool, (size_t)V + 1), pos(c.pool, (size_t)V + 1);
    if (roots) {
        hipLaunchKernelGGL(root_flags_kernel, dim3(grid_for(c, (long)V + 1)), dim3(256), 0, c.stream, d_label, V,
                           pred.p);
    } else {
        HBP_CHECK(hipMemsetAsync(pred.p, 0, sizeof(int) * ((size_t)V + 1), c.stream));
        hipLaunchKernelGGL(mark_labels_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream, d_label, V, pred.p);
    }
    scan_exclusive_int(c, pred.p, pos.p, V + 1, nullptr);   // pred[V] = 0, so pos[V] = count
    hipLaunchKernelGGL(remap_labels_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream, d_label, V, pos.p);
    HBP_REQUIRE(n_extra >= 0 && n_extra <= 7, "at most 7 extra words per read");
    WordPtrs w{};
    w.p[0] = pos.p + V;
    for (int i = 0; i < n_extra; ++i) w.p[1 + i] = extra[i];
    int got[8] = {0};
    read_words(c, w, 1 + n_extra, got);
    for (int i = 0; i < n_extra; ++i) extra_host[i] = got[1 + i];
    return got[0];
}
int compact_labels(Ctx &c, int *d_label, int V) {
    return compact_labels_and_read(c, d_label, V, nullptr, nullptr, 0, false);
}

// labels + compaction of one pyramid level with one host wait (more after an overflow); en, when
// not null, is a device word (the previous edge reduction's edge count) read in the same wait.
// The scheduled form runs the context's schedule (c.crt), then, on overflow, the full schedule,
// then the cooperative form; the steps a scheduled run used set the next level's schedule: one
// more of each than used, shrinking by at most one per level.
static int crtrees_level(Ctx &c, const int *d_nn, int V, int *d_label, const int *en, int *en_host) {
    CRTRun r;
    crt_setup(c, r, V);
    const CrtSched full{};
    const CrtSched sc{std::min(std::max(c.crt.rounds, 1), kCrtRounds), std::min(std::max(c.crt.jumps, 1), kCrtJumps),
                      std::min(std::max(c.crt.max_jumps, 1), kCrtMaxJumps)};
    const bool scheduled = V > kCrtSingleMax;
    crt_enqueue(c, r, d_nn, d_label, sc);
    const int *extra[6] = {r.s.status, &r.f->overflow, r.s.status + 1, r.s.status + 2, r.s.status + 3, en};
    int w[6] = {0, 0, 0, 0, 0, 0};
    int Vn = compact_labels_and_read(c, d_label, V, extra, w, en ? 6 : 5, true);
    if (en) *en_host = w[5];
    if (w[1] && scheduled && (sc.rounds < full.rounds || sc.jumps < full.jumps || sc.max_jumps < full.max_jumps)) {
        HBP_CHECK(hipMemsetAsync(r.s.flags, 0, sizeof(int) * (8 + (sizeof(CRTFlags) + 3) / 4), c.stream));
        crt_enqueue(c, r, d_nn, d_label, full);
        Vn = compact_labels_and_read(c, d_label, V, extra, w, 5, true);
    }
    if (w[1]) {
        crt_cooperative(c, r, d_nn, d_label);
        Vn = compact_labels_and_read(c, d_label, V, extra, w, 1, true);
    } else if (scheduled) {
        c.crt.rounds = std::max(w[2] + 1, sc.rounds - 1);
        c.crt.jumps = std::max(w[3] + 1, sc.jumps - 1);
        c.crt.max_jumps = std::max(w[4] + 1, sc.max_jumps - 1);
    }
    crt_require(w[0]);
    return Vn;
}

// get_target_clus on the device; clus holds the compacted labels of V children into Vn < target
// clusters and receives labels in [0, target), every one used (so the reference's following
// compaction is the identity)
static void target_clus_relabel(Ctx &c, int *clus, int V, int Vn, const float *dist_min, int target) {
    const int bits = bits_for(Vn);
    DevBuf<unsigned int> ck(c.pool, V), ck2(c.pool, V);
    DevBuf<int> cv(c.pool, V), cv2(c.pool, V), cnt(c.pool, (size_t)Vn + 1), coff(c.pool, (size_t)Vn + 1);
    HBP_CHECK(hipMemcpyAsync(ck.p, clus, sizeof(int) * (size_t)V, hipMemcpyDeviceToDevice, c.stream));
    hipLaunchKernelGGL(iota_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream, cv.p, (long)V);
    if (sort_pairs_u32(c, ck.p, ck2.p, cv.p, cv2.p, V, bits)) {
        std::swap(ck, ck2);
        std::swap(cv, cv2);
    }
    HBP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(int) * ((size_t)Vn + 1), c.stream));
    hipLaunchKernelGGL(count_keys_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream, clus, (long)V, cnt.p);
    scan_exclusive_int(c, cnt.p, coff.p, Vn + 1, nullptr);
    DevBuf<unsigned int> key(c.pool, Vn), key2(c.pool, Vn);
    DevBuf<int> id(c.pool, Vn), id2(c.pool, Vn), rc(c.pool, (size_t)Vn + 1), excl(c.pool, (size_t)Vn + 1);
    DevBuf<int> istar(c.pool, 1), pos(c.pool, Vn);
    hipLaunchKernelGGL(target_sums_kernel, dim3(grid_for(c, Vn)), dim3(256), 0, c.stream, cv.p, coff.p, Vn, dist_min,
                       key.p, id.p);
    if (sort_pairs_u32(c, key.p, key2.p, id.p, id2.p, Vn, 32)) {
        std::swap(key, key2);
        std::swap(id, id2);
    }
    hipLaunchKernelGGL(target_counts_kernel, dim3(grid_for(c, (long)Vn + 1)), dim3(256), 0, c.stream, id.p, coff.p,
                       Vn, rc.p);
    scan_exclusive_int(c, rc.p, excl.p, Vn + 1, nullptr);
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, c.stream, istar.p, 1L, Vn);
    hipLaunchKernelGGL(target_istar_kernel, dim3(grid_for(c, Vn)), dim3(256), 0, c.stream, excl.p, Vn, target,
                       istar.p);
    hipLaunchKernelGGL(target_pos_kernel, dim3(grid_for(c, Vn)), dim3(256), 0, c.stream, id.p, excl.p, Vn, istar.p,
                       pos.p);
    hipLaunchKernelGGL(target_label_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream, ck.p, cv.p, coff.p,
                       (long)V, Vn, target, excl.p, istar.p, pos.p, clus);
    HBP_CHECK(hipGetLastError());
}

// SegHAC::run_ms (segment.cu:1279-1321) for LabelISPFormat, any linkage; target_clus > 1 ends
// the pyramid at the first level of at most that many clusters, re-labelled to exactly
// target_clus (get_target_clus, then save_output's stop test, segment.cu:1103-1110, 1299-1306).
void pyramid_build(Ctx &c, Pyramid &p, const uint8_t *d_bgr, int W, int H, int num_nb, int link, int target_clus,
                   double sigma) {
    HBP_REQUIRE(num_nb == 4 || num_nb == 8, "num_nb must be 4 or 8");
    HBP_REQUIRE(link >= kMinLink && link <= kWardLink, "linkage must be 0 (Min), 1 (Max), 2 (Centroid) or 3 (Ward)");
    const bool use_mean = link == kCentroidLink || link == kWardLink;
    HBP_REQUIRE(W >= 2 && H >= 2, "image must be at least 2x2");
    long N = (long)W * H;
    HBP_REQUIRE(N < (1L << 24), "images are limited to 2^24 pixels");
    HBP_REQUIRE(target_clus >= 1 && target_clus < N, "target_clus must be in [1, width*height)");
    // the blur only for sigma > FLT_MIN (segment.cu:805)
    const bool blur = sigma > 1.17549435e-38;
    GaussTaps taps;
    float wsum = 1.f;
    const int S = blur ? gauss_taps(sigma, taps, &wsum) : 0;
    HBP_REQUIRE(!blur || (H + kBlurTY - 1) / kBlurTY <= 65535, "image too tall for the pre-blur's grid");
    p.W = W;
    p.H = H;
    p.num_nb = num_nb;
    p.top = target_clus;
    p.V.clear();
    p.E.clear();
    p.voff.clear();
    p.eoff.clear();
    // capacities asserted by the reference (segment.cu:891-892)
    long vcap = 2 * N, ecap = 2L * num_nb * N;
    p.clus.alloc(c.pool, vcap);
    p.mean.alloc(c.pool, 6 * vcap);
    p.bd.alloc(c.pool, ecap);
    p.bdl.alloc(c.pool, ecap);
    upload_srgb_lut(c);

    DevBuf<float4> lab(c.pool, N);
    DevBuf<int> nn(c.pool, N);
    DevBuf<float> dist_min(c.pool, N);
    // Min/Max: the per-edge distances carried from level to level (dist_d, segment.cu:838-844)
    DevBuf<float> edist;
    if (!use_mean) edist.alloc(c.pool, (size_t)num_nb * N);
    if (blur)
        hipLaunchKernelGGL(blur_lab_kernel, dim3((W + kBlurTX - 1) / kBlurTX, (H + kBlurTY - 1) / kBlurTY), dim3(256), 0,
                           c.stream, d_bgr, W, H, S, taps, wsum, lab.p, p.mean.p);
    else
        hipLaunchKernelGGL(lab_kernel, dim3(grid_for(c, N)), dim3(256), 0, c.stream, d_bgr, W, H, lab.p, p.mean.p);
    hipLaunchKernelGGL(nn_grid_kernel, dim3(grid_for(c, N)), dim3(256), 0, c.stream, lab.p, W, H, num_nb, nn.p,
                       dist_min.p, p.bd.p, p.bdl.p, edist.p);
    HBP_CHECK(hipGetLastError());
    lab.release();

    int V = (int)N;
    long E = num_nb * N;
    long vo = 0, eo = 0;
    // en: the edge count of the last edge reduction, read in the next level's wait (E bounds it,
    // so the launches after the reduction are sized by E and need no wait of their own)
    DevBuf<int> en(c.pool, 1);
    bool en_pending = false;
    while (V > 1) {
        int *clus = p.clus.p + vo;
        int En_prev = 0;
        int Vn = crtrees_level(c, nn.p, V, clus, en_pending ? en.p : nullptr, &En_prev);
        if (en_pending) E = En_prev;
        if (Vn < target_clus) {
            target_clus_relabel(c, clus, V, Vn, dist_min.p, target_clus);
            Vn = target_clus;
        }
        p.V.push_back(V);
        p.E.push_back((int)E);
        p.voff.push_back(vo);
        p.eoff.push_back(eo);
        if (Vn == target_clus) break;
        long vo2 = vo + V, eo2 = eo + E;
        HBP_REQUIRE(vo2 + Vn <= vcap && eo2 + E <= ecap, "pyramid capacity exceeded");
        // --- edge reduction
        int bits = bits_for(Vn);
        DevBuf<unsigned long long> keys(c.pool, E), keys2(c.pool, E);
        DevBuf<unsigned int> dbits, dbits2;
        DevBuf<int> flag(c.pool, E + 1), pos(c.pool, E + 1);
        if (!use_mean) {
            dbits.alloc(c.pool, E);
            dbits2.alloc(c.pool, E);
        }
        hipLaunchKernelGGL(map_edges_kernel, dim3(grid_for(c, E)), dim3(256), 0, c.stream, p.bd.p + eo, edist.p,
                           clus, E, bits, keys.p, dbits.p);
        const bool swapped = use_mean ? sort_keys_u64(c, keys.p, keys2.p, E, 2 * bits)
                                      : sort_pairs_u64(c, keys.p, keys2.p, dbits.p, dbits2.p, E, 2 * bits);
        if (swapped) {
            std::swap(keys, keys2);
            std::swap(dbits, dbits2);
        }
        // the new level's sums and (Min/Max) reduced distances start over E entries, which bound
        // its edges: FLT_MAX for Min, 0 for Max
        DevBuf<float> edist2;
        if (!use_mean) edist2.alloc(c.pool, (size_t)E + 1);
        hipLaunchKernelGGL(run_flags_kernel, dim3(grid_for(c, E + 1)), dim3(256), 0, c.stream, keys.p, E, bits,
                           flag.p, p.bdl.p + eo2, (unsigned int *)edist2.p, link == kMinLink ? 0x7f7fffffu : 0u);
        scan_exclusive_int(c, flag.p, pos.p, E + 1, nullptr);   // pos[E]: the new level's edge count
        hipLaunchKernelGGL(scatter_edges_kernel, dim3(grid_for(c, E)), dim3(256), 0, c.stream, keys.p,
                           p.bdl.p + eo, dbits.p, link == kMinLink ? 1 : 0, flag.p, pos.p, E, bits, p.bd.p + eo2,
                           p.bdl.p + eo2, (unsigned int *)edist2.p);
        keys.release();
        keys2.release();
        dbits.release();
        dbits2.release();
        flag.release();
        if (!use_mean) std::swap(edist, edist2);
        edist2.release();
        // --- children grouped by parent, ascending child index (stable sort by parent); every
        // parent has a child (the labels are compacted), so the runs give the offsets
        DevBuf<unsigned int> ck(c.pool, V), ck2(c.pool, V);
        DevBuf<int> cv(c.pool, V), cv2(c.pool, V), coff(c.pool, (size_t)Vn + 1);
        hipLaunchKernelGGL(key_iota_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream, clus, (long)V, ck.p, cv.p);
        if (sort_pairs_u32(c, ck.p, ck2.p, cv.p, cv2.p, V, bits)) {
            std::swap(ck, ck2);
            std::swap(cv, cv2);
        }
        hipLaunchKernelGGL(run_starts_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream, ck.p, (long)V, Vn,
                           coff.p);
        // --- sums, means, dist_max
        DevBuf<float> mu(c.pool, 5L * Vn), dist_max(c.pool, Vn);
        hipLaunchKernelGGL(parent_sums_kernel, dim3(grid_for(c, Vn)), dim3(256), 0, c.stream, p.mean.p + 6 * vo, V,
                           cv.p, coff.p, Vn, dist_min.p, p.mean.p + 6 * vo2, mu.p, dist_max.p);
        // --- CSR of the new edges (sorted by source), then the next level's 1-NN
        hipLaunchKernelGGL(edge_csr_kernel, dim3(grid_for(c, E > 0 ? E : 1)), dim3(256), 0, c.stream, p.bd.p + eo2,
                           pos.p + E, E, Vn, coff.p, en.p);
        pos.release();
        hipLaunchKernelGGL(level_nn_kernel, dim3(grid_for(c, Vn)), dim3(256), 0, c.stream, mu.p,
                           p.mean.p + 6 * vo2 + 5L * Vn, link, Vn, p.bd.p + eo2, coff.p, dist_max.p, edist.p, nn.p,
                           dist_min.p);
        HBP_CHECK(hipGetLastError());
        en_pending = true;
        V = Vn;
        vo = vo2;
        eo = eo2;
    }
    HBP_CHECK(hipStreamSynchronize(c.stream));
}

}  // namespace hbp
