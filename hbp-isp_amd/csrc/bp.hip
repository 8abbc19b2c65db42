// Stage (c): hierarchical plane-particle min-sum BP on the superpixel pyramid
// (hbp_isp.cu / hbp_isp.hpp of the reference).
//
// One BP sweep (generate_plane + compute_plane_cost + belief_propagation, hbp_isp.cu:1302-1308)
// is three kernels on one stream:
//   plane_cost_kernel  per pixel: cost of the 4 candidate planes, summed per superpixel in
//                      int64 fixed point (segmented in LDS, one global atomic per chunk and
//                      cluster), so the sums are independent of the reduction order;
//   edge_kernel        per (edge, target slot): the min-sum message of smooth_msg_k and the
//                      mean removal of normalize_msg_k (16 lanes per edge);
//   node_kernel        per superpixel: belief = cost + incoming messages in edge order, rank
//                      of the 12 beliefs, compaction of planes/costs/beliefs and of the
//                      outgoing messages (select_*), written straight to the reverse edge
//                      (ping_pang_msg), and the next sweep's RANSAC / random candidates.
// Spatial-propagation sweeps add spatial_kernel, which resolves the parent/neighbour/child
// chain of select_child_plane + neighbor_candidate_plane + set_candidate_plane_k per child.
//
// HBM layout per level: planes [V][12] float4, cost/belief [V][12] f32, messages [E][8] f32
// (the 8 slots the next sweep reads), sweep messages [E][12] f32, pixels sorted by superpixel
// as (packed x|y<<16, cluster) pairs, CSR offsets/reverse ids/edge weights per edge.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <cmath>

#include "hbp_internal.h"

namespace cg = cooperative_groups;

namespace hbp {

constexpr int P = kP, K = kK;

// replicas of each fixed-point accumulator: chunk k adds into replica k mod rrep, which spreads
// the atomics of the few large superpixels of the coarse levels over rrep addresses
static inline int acc_replicas(int V) { return V <= 65536 ? 8 : 1; }

// ----------------------------------------------------------------------------------------
// prologue kernels
// ----------------------------------------------------------------------------------------
__global__ void iota_u32_kernel(unsigned int *__restrict__ p, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = (unsigned)i;
}
__global__ void iota_i32_kernel(int *__restrict__ p, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = (int)i;
}
__global__ void compose_kernel(unsigned int *__restrict__ im_clus, const int *__restrict__ clus, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        im_clus[i] = (unsigned)clus[im_clus[i]];
}
__global__ void count_u32_kernel(const unsigned int *__restrict__ key, long n, int *__restrict__ cnt) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        atomicAdd(&cnt[key[i]], 1);
}
__global__ void count_i32_kernel(const int *__restrict__ key, long n, int *__restrict__ cnt) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        atomicAdd(&cnt[key[i]], 1);
}
__global__ void count_src_kernel2(const int2 *__restrict__ bd, long E, int *__restrict__ cnt) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (long)gridDim.x * blockDim.x)
        atomicAdd(&cnt[bd[e].x], 1);
}
// setup_D4 (hbp_isp.cu:871-891): (packed xy, cluster) in (cluster, pixel id) order
__global__ void pix_kernel(const unsigned int *__restrict__ keys, const int *__restrict__ ids, long n, int W,
                           uint2 *__restrict__ pix) {
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        int id = ids[j];
        pix[j] = make_uint2(pack_xy(id % W, id / W), keys[j]);
    }
}
// reverse_bd_id_k (hbp_isp.cu:154-175): first r in t's CSR range with bd[r].y == s
__global__ void rev_kernel(const int2 *__restrict__ bd, const int *__restrict__ off, long E, int *__restrict__ rev) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (long)gridDim.x * blockDim.x) {
        int2 b = bd[e];
        int r_found = -1;
        for (int r = off[b.y]; r < off[b.y + 1]; ++r)
            if (bd[r].y == b.x) {
                r_found = r;
                break;
            }
        rev[e] = r_found;
    }
}
// compute_dist_color_pos (hbp_isp.cu:952-970; kernels 78-126) with smooth_msg_k's per-edge
// factors (hbp_isp.cu:323-328): c0 = max(dc0, 0.01) * bdl, dpos = max(dc1, 1).
__global__ void edge_weights_kernel(const float *__restrict__ mean, int V, const int2 *__restrict__ bd,
                                    const int *__restrict__ bdl, long E, float *__restrict__ c0,
                                    float *__restrict__ dpos) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (long)gridDim.x * blockDim.x) {
        int2 b = bd[e];
        float ws = mean[5L * V + b.x], wt = mean[5L * V + b.y];
        float dcol = 0.f;
#pragma unroll
        for (int c = 0; c < 3; ++c) dcol += fabsf(mean[(long)c * V + b.x] / ws - mean[(long)c * V + b.y] / wt);
        float dp = 0.f;
#pragma unroll
        for (int c = 3; c < 5; ++c) {
            float d = mean[(long)c * V + b.x] / ws - mean[(long)c * V + b.y] / wt;
            dp += d * d;
        }
        float dc0 = (float)exp_cody_waite((double)(-dcol) / 20.0);
        float dc1 = sqrtf(dp);
        float dist_color = fmaxf(dc0, 0.01f);
        c0[e] = dist_color * (float)bdl[e];
        dpos[e] = fmaxf(dc1, 1.0f);
    }
}
__global__ void mean_xy_kernel(const float *__restrict__ mean, int V, float2 *__restrict__ mxy) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (long)gridDim.x * blockDim.x) {
        float w = mean[5L * V + i];
        mxy[i] = make_float2(mean[3L * V + i] / w, mean[4L * V + i] / w);
    }
}

// ----------------------------------------------------------------------------------------
// candidate generators (device restatements of hbp_isp.cu:495-741)
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ void three_diff(int &p0, int &p1, int &p2, int tn) {
    if (tn <= 3) {
        p1 = (p0 + 1) % tn;
        p2 = (p1 + 1) % tn;
        return;
    }
    if (p1 == p0) p1 = tn - 1;
    if (p2 == p0) {
        if (p1 == tn - 2) p2 = tn - 1;
        else p2 = tn - 2;
    } else if (p2 == p1) {
        if (p0 == tn - 2) p2 = tn - 1;
        else p2 = tn - 2;
    }
}

__device__ __forceinline__ float4 fit_plane(float4 A, float4 B, float4 C) {
    float ax = B.x - A.x, ay = B.y - A.y, az = B.z - A.z;
    float bx = C.x - A.x, by = C.y - A.y, bz = C.z - A.z;
    float nx = ay * bz - az * by;
    float ny = az * bx - ax * bz;
    float nz = ax * by - ay * bx;
    nx /= nz;
    ny /= nz;
    nz /= nz;
    if (fabsf(nz - 1.f) > FLT_EPSILON) return make_float4(0.f, 0.f, (A.z + B.z + C.z) / 3.f, 0.f);
    return make_float4(-nx, -ny, nx * A.x + ny * A.y + A.z, 0.f);
}

// ransac_candidate_plane (hbp_isp.cu:592-640) for one superpixel
__device__ void ransac_node(int id, const uint2 *__restrict__ pix, const int *__restrict__ pix_off,
                            const float *__restrict__ dwta, int W, uint32_t *__restrict__ s_ransac,
                            float4 *__restrict__ plane, uint4 *__restrict__ point) {
    int pos = pix_off[id];
    int n = pix_off[id + 1] - pos;
    uint32_t s = s_ransac[id];
    uint32_t q0 = 0, q1 = 0, q2 = 0;
    for (int p = 2 * P; p < 3 * P; ++p) {
        s = lcg_next(s);
        int p0 = (int)umod0(s, n);
        s = lcg_next(s);
        int p1 = (int)umod0(s, n - 1);
        s = lcg_next(s);
        int p2 = (int)umod0(s, n - 2);
        three_diff(p0, p1, p2, n);
        q0 = pix[pos + p0].x;
        q1 = pix[pos + p1].x;
        q2 = pix[pos + p2].x;
        int x0 = unpack_x(q0), y0 = unpack_y(q0), x1 = unpack_x(q1), y1 = unpack_y(q1), x2 = unpack_x(q2),
            y2 = unpack_y(q2);
        float4 A = make_float4((float)x0, (float)y0, dwta[(long)y0 * W + x0], 0.f);
        float4 B = make_float4((float)x1, (float)y1, dwta[(long)y1 * W + x1], 0.f);
        float4 C = make_float4((float)x2, (float)y2, dwta[(long)y2 * W + x2], 0.f);
        plane[(long)id * K + p] = fit_plane(A, B, C);
    }
    point[id] = make_uint4(q0, q1, q2, 0u);
    s_ransac[id] = s;
}

// random_candidate_plane (hbp_isp.cu:641-674) with normalize_plane (495-503)
__device__ void random_node(int id, const float2 *__restrict__ mxy, float max_dz, float max_dn,
                            uint32_t *__restrict__ s_ransac, float4 *__restrict__ plane) {
    uint32_t s = s_ransac[id];
    float2 m = mxy[id];
    float x = m.x, y = m.y;
    for (int p = 2 * P; p < 3 * P; ++p) {
        float4 pl = plane[(long)id * K + p];
        float z = pl.x * x + pl.y * y + pl.z;
        float dz = lcg_uniform(s, max_dz);
        z += dz;
        float dnx = lcg_uniform(s, max_dn);
        float dny = lcg_uniform(s, max_dn);
        float dnz = lcg_uniform(s, max_dn);
        float norm = sqrtf(pl.x * pl.x + pl.y * pl.y + pl.z * pl.z);
        pl.x /= norm;
        pl.y /= norm;
        pl.z /= norm;
        pl.x += dnx;
        pl.y += dny;
        pl.z += dnz;
        float a = -pl.x / pl.z;
        float b = -pl.y / pl.z;
        float c = z - a * x - b * y;
        plane[(long)id * K + p] = make_float4(a, b, c, 0.f);
    }
    s_ransac[id] = s;
}

enum SweepKind { kNone = 0, kRansac = 1, kSpatial = 2, kRandom = 3 };

// Every pointer and constant one sweep reads or writes (level structures, parent level,
// volume, parameters, state).  Passed by value to the kernels.
struct SweepArgs {
    int V, Vh, W, D;
    long E, N;
    const int *clus;
    const int2 *bd;
    const uint2 *pix;
    const int *pix_off;
    const int *off;
    const int *rev;
    const float *c0;
    const float *dpos;
    const float2 *mxy;
    const int *child;
    const int *child_off;
    float4 *plane;
    float *cost;
    float *belief;
    uint4 *point;
    // parent level (spatial propagation, priors)
    const int2 *bd_h;
    const int *off_h;
    const float4 *plane_h;
    // cost volume and parameters
    const uint8_t *vol;
    int pitch;
    long vplane;
    const float *dwta;
    float lambda, tau, tau_s;
    int use_cost_vol, shift;
    // state
    float4 *msg;                 // [E][2] float4
    float *msg_t;                // [E][12]
    unsigned long long *acc;     // [V][rrep][12] fixed-point sums, rrep replicas
    int rrep;
    unsigned int *accnan;        // [V]
    uint32_t *s_ransac, *s_neighbor, *s_child;
};

// the schedule of one sweep as the device reads it
struct SweepDesc {
    int kind, next_kind;
    float ndz, ndn;
};

constexpr int CH12 = 256 * 2, CH4 = 256 * 8;                    // pixels per plane-cost chunk
constexpr int MAXC12 = CH12 / 2 + 2, MAXC4 = CH4 / 2 + 2;        // superpixels a chunk can touch
constexpr size_t PC_LDS_BYTES =
    (MAXC12 * 12 * 8 + MAXC12 * 4) > (MAXC4 * 4 * 8 + MAXC4 * 4) ? (MAXC12 * 12 * 8 + MAXC12 * 4)
                                                                 : (MAXC4 * 4 * 8 + MAXC4 * 4);

__device__ __forceinline__ int64_t acc_sum(const SweepArgs &a, long id, int k) {
    unsigned long long t = 0ull;
    for (int r = 0; r < a.rrep; ++r) t += a.acc[(id * a.rrep + r) * K + k];
    return (int64_t)t;
}
__device__ __forceinline__ void acc_zero(const SweepArgs &a, long id, int k) {
    for (int r = 0; r < a.rrep; ++r) a.acc[(id * a.rrep + r) * K + k] = 0ull;
}

// init_bp (hbp_isp.cu:930-950): copy_prior_plane (56-73), init_rand_seed_k (570-591), plus the
// first sweep's RANSAC candidates.
__device__ void init_node(const SweepArgs &a, long id) {
    int cc = a.clus[id];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        float4 pl = a.plane_h[(long)cc * K + p];
        a.plane[id * K + p] = pl;
        a.plane[id * K + p + P] = pl;
        a.plane[id * K + p + 2 * P] = pl;
    }
    a.s_neighbor[id] = (uint32_t)a.off[id];
    a.s_ransac[id] = (uint32_t)a.pix_off[id];
    if (id < a.Vh) a.s_child[id] = (uint32_t)a.child_off[id];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        acc_zero(a, id, k);
        a.belief[id * K + k] = 0.f;
        a.cost[id * K + k] = 0.f;
    }
    a.accnan[id] = 0u;
    ransac_node((int)id, a.pix, a.pix_off, a.dwta, a.W, a.s_ransac, a.plane, a.point);
}

// costs of slots 0..2P-1 after init (compute_plane_cost(0), (P)); belief = cost
__device__ void init_cost_node(const SweepArgs &a, long id) {
    unsigned int nb = a.accnan[id];
#pragma unroll
    for (int k = 0; k < 2 * P; ++k) {
        float c = from_fixed(acc_sum(a, id, k), (nb >> k) & 1u, a.shift);
        a.cost[id * K + k] = c;
        a.belief[id * K + k] = c;
        acc_zero(a, id, k);
    }
    a.accnan[id] = nb & ~0xffu;
}

// ----------------------------------------------------------------------------------------
// plane cost: read_plane_cost (hbp_isp.cu:177-228) + the per-superpixel sums, for one chunk
// of CH = 256*R consecutive pixels of the sorted list (one 256-thread block).  A lane sums
// its R consecutive pixels while the superpixel stays the same and flushes into LDS int64
// counters indexed by superpixel - first superpixel of the chunk (superpixels at BP levels
// hold >= 2 pixels, so a chunk touches at most CH/2 + 1 of them); the block then adds each
// counter to replica (chunk mod rrep) of the global sum with one atomic.
// ----------------------------------------------------------------------------------------
template <int NS, int R>
__device__ void plane_cost_chunk(const SweepArgs &a, long chunk, int slot0, unsigned long long *sacc,
                                 unsigned int *snan) {
    constexpr int CH = 256 * R;
    constexpr int MAXC = CH / 2 + 2;
    const long base = chunk * CH;
    for (int i = threadIdx.x; i < MAXC * NS; i += 256) sacc[i] = 0ull;
    for (int i = threadIdx.x; i < MAXC; i += 256) snan[i] = 0u;
    const int c_first = (int)a.pix[base].y;
    const long last = (base + CH < a.N ? base + CH : a.N) - 1;
    const int nclus = (int)a.pix[last].y - c_first + 1;
    __syncthreads();

    const long j0 = base + (long)threadIdx.x * R;
    int crun = -1;
    int64_t racc[NS];
    unsigned int rnan = 0u;
    float4 pl[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) racc[k] = 0;
    for (int r = 0; r < R; ++r) {
        long j = j0 + r;
        if (j >= a.N) break;
        uint2 q = a.pix[j];
        int c = (int)q.y;
        if (c != crun) {
            if (crun >= 0) {
                int lc = crun - c_first;
#pragma unroll
                for (int k = 0; k < NS; ++k)
                    if (racc[k]) atomicAdd(&sacc[lc * NS + k], (unsigned long long)racc[k]);
                if (rnan) atomicOr(&snan[lc], rnan);
            }
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                racc[k] = 0;
                pl[k] = a.plane[(long)c * K + slot0 + k];
            }
            rnan = 0u;
            crun = c;
        }
        int x = unpack_x(q.x), y = unpack_y(q.x);
        float fx = (float)x, fy = (float)y;
        long pofs = (long)y * a.pitch + x;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            float df = pl[k].x * fx + pl[k].y * fy + pl[k].z;
            float cost;
            if (a.use_cost_vol) {
                int d = floor_to_int(df);
                if (d < 0 || (long long)d + 1 >= a.D) {
                    cost = a.tau;
                } else {
                    float v0 = (float)a.vol[(long)d * a.vplane + pofs];
                    float v1 = (float)a.vol[(long)(d + 1) * a.vplane + pofs];
                    cost = ((float)(d + 1) - df) * v0 + (df - (float)d) * v1;
                }
            } else {
                float Dw = a.dwta[(long)y * a.W + x];
                cost = fabsf(Dw - df) <= 1.f ? 0.f : 1.f;
            }
            float v = a.lambda * cost;
            if (v != v) rnan |= 1u << (slot0 + k);
            else racc[k] += to_fixed(v, a.shift);
        }
    }
    if (crun >= 0) {
        int lc = crun - c_first;
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (racc[k]) atomicAdd(&sacc[lc * NS + k], (unsigned long long)racc[k]);
        if (rnan) atomicOr(&snan[lc], rnan);
    }
    __syncthreads();
    const int rep = (int)(chunk % a.rrep);
    for (int i = threadIdx.x; i < nclus * NS; i += 256) {
        unsigned long long v = sacc[i];
        int lc = i / NS, k = i % NS;
        if (v) atomicAdd(&a.acc[((long)(c_first + lc) * a.rrep + rep) * K + slot0 + k], v);
    }
    for (int i = threadIdx.x; i < nclus; i += 256)
        if (snan[i]) atomicOr(&a.accnan[c_first + i], snan[i]);
    __syncthreads();
}

// ----------------------------------------------------------------------------------------
// messages for 16 edges (one 256-thread block): passing_msg_k (hbp_isp.cu:292-304),
// smooth_msg_k (306-355), normalize_msg_k (273-290).  Lane y < 12 of edge e = (s, t) computes
// the message into s for s's slot y from t's 8 best + prior particles.
// ----------------------------------------------------------------------------------------
__device__ void edge_group(const SweepArgs &a, long grp) {
    const long e = grp * 16 + (threadIdx.x >> 4);
    const int y = threadIdx.x & 15;
    const bool live = e < a.E && y < K;
    float out = 0.f;
    if (live) {
        int2 b = a.bd[e];
        int s = b.x, t = b.y;
        uint4 ps = a.point[s], pt = a.point[t];
        float X[6], Y[6];
        X[0] = (float)unpack_x(ps.x); Y[0] = (float)unpack_y(ps.x);
        X[1] = (float)unpack_x(ps.y); Y[1] = (float)unpack_y(ps.y);
        X[2] = (float)unpack_x(ps.z); Y[2] = (float)unpack_y(ps.z);
        X[3] = (float)unpack_x(pt.x); Y[3] = (float)unpack_y(pt.x);
        X[4] = (float)unpack_x(pt.y); Y[4] = (float)unpack_y(pt.y);
        X[5] = (float)unpack_x(pt.z); Y[5] = (float)unpack_y(pt.z);
        float c0 = a.c0[e];
        float dpos = a.dpos[e];
        float4 psl = a.plane[(long)s * K + y];
        float4 m0 = a.msg[2 * e], m1 = a.msg[2 * e + 1];
        float mh[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
        float minimum = FLT_MAX;
#pragma unroll
        for (int p = 0; p < 2 * P; ++p) {
            if (p == P) {
                out += minimum;
                minimum = FLT_MAX;
            }
            float m = a.belief[(long)t * K + p] - mh[p];
            float4 ptl = a.plane[(long)t * K + p];
            float dpx = psl.x - ptl.x;
            float dpy = psl.y - ptl.y;
            float dpz = psl.z - ptl.z;
            float phi = fabsf(dpx * X[0] + dpy * Y[0] + dpz);
#pragma unroll
            for (int j = 1; j < 6; ++j) phi = phi + fabsf(dpx * X[j] + dpy * Y[j] + dpz);
            phi = fminf(phi / 3 / dpos, a.tau_s);
            float v = m + c0 * phi;
            if (v < minimum) minimum = v;
        }
        out += minimum;
    }
    // normalize_msg_k: mean over the 12 slots, summed in slot order
    float mean = 0.f;
    const int g = (threadIdx.x & 63) & ~15;
#pragma unroll
    for (int k = 0; k < K; ++k) mean += __shfl(out, g + k, 64);
    mean /= K;
    if (live) a.msg_t[e * K + y] = out - mean;
}

// ----------------------------------------------------------------------------------------
// node update: aggregate_belief (hbp_isp.cu:1048-1058) in edge order, ping_pang_msg
// (1059-1066), get_belief_order (432-450), select_belief (452-472), select_msg (474-493);
// then the parents' seed advance of a spatial sweep and the next sweep's candidates.
// ----------------------------------------------------------------------------------------
__device__ void node_update(const SweepArgs &a, long s, int this_kind, int next_kind, float max_dz, float max_dn) {
    float cost[K], bel[K];
    unsigned int nb = a.accnan[s];
#pragma unroll
    for (int k = 0; k < 2 * P; ++k) cost[k] = a.cost[s * K + k];
#pragma unroll
    for (int k = 2 * P; k < K; ++k) {
        cost[k] = from_fixed(acc_sum(a, s, k), (nb >> k) & 1u, a.shift);
        acc_zero(a, s, k);
    }
    a.accnan[s] = 0u;
#pragma unroll
    for (int k = 0; k < K; ++k) bel[k] = cost[k];
    const int eb = a.off[s], ee = a.off[s + 1];
    for (int e = eb; e < ee; ++e) {
        const float4 *mt = (const float4 *)(a.msg_t + (long)e * K);
        float4 u0 = mt[0], u1 = mt[1], u2 = mt[2];
        bel[0] += u0.x; bel[1] += u0.y; bel[2] += u0.z; bel[3] += u0.w;
        bel[4] += u1.x; bel[5] += u1.y; bel[6] += u1.z; bel[7] += u1.w;
        bel[8] += u2.x; bel[9] += u2.y; bel[10] += u2.z; bel[11] += u2.w;
    }
    int order[K];
#pragma unroll
    for (int k = 0; k < K; ++k) order[k] = 0;
#pragma unroll
    for (int i = 0; i < K - 1; ++i)
#pragma unroll
        for (int j = i + 1; j < K; ++j) {
            if (bel[j] < bel[i]) order[i]++;
            else order[j]++;
        }
    // in-place compaction (select_belief / select_msg): output slot j < nsel takes the j-th
    // selected slot in slot order, slots j >= nsel keep their own value.
    bool sel[K];
    int pos[K];
    int nsel = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        sel[k] = order[k] < P;
        pos[k] = nsel;
        nsel += sel[k] ? 1 : 0;
    }
    float4 *pl = a.plane + s * K;
    float4 pin[K];
#pragma unroll
    for (int k = 0; k < K; ++k) pin[k] = pl[k];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        float cv = cost[j], bv = bel[j];
        float4 pv = pin[j];
#pragma unroll
        for (int k = j; k < K; ++k)
            if (sel[k] && pos[k] == j) {
                cv = cost[k];
                bv = bel[k];
                pv = pin[k];
            }
        pl[j] = pv;
        a.cost[s * K + j] = cv;
        a.belief[s * K + j] = bv;
    }
    // messages over s's particles -> reverse edge, compacted by s's order
    for (int e = eb; e < ee; ++e) {
        const float4 *mt = (const float4 *)(a.msg_t + (long)e * K);
        float4 u0 = mt[0], u1 = mt[1], u2 = mt[2];
        float m[K] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w, u2.x, u2.y, u2.z, u2.w};
        float o[2 * P];
#pragma unroll
        for (int j = 0; j < 2 * P; ++j) {
            float v = m[j];
#pragma unroll
            for (int k = j; k < K; ++k)
                if (sel[k] && pos[k] == j) v = m[k];
            o[j] = v;
        }
        long r = a.rev[e];
        a.msg[2 * r] = make_float4(o[0], o[1], o[2], o[3]);
        a.msg[2 * r + 1] = make_float4(o[4], o[5], o[6], o[7]);
    }
    if (this_kind == kSpatial && s < a.Vh) {
        uint32_t sc = a.s_child[s], sn = a.s_neighbor[s];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            sc = lcg_next(sc);
            sn = lcg_next(sn);
        }
        a.s_child[s] = sc;
        a.s_neighbor[s] = sn;
    }
    if (next_kind == kRansac) ransac_node((int)s, a.pix, a.pix_off, a.dwta, a.W, a.s_ransac, a.plane, a.point);
    else if (next_kind == kRandom) random_node((int)s, a.mxy, max_dz, max_dn, a.s_ransac, a.plane);
}

// ----------------------------------------------------------------------------------------
// spatial propagation (hbp_isp.cu:1225-1243): for child id with parent P, candidate p is the
// slot-0 plane of a random child (select_child_plane) of a random neighbour of P at the parent
// level (neighbor_candidate_plane), each draw being the (p+1)-th LCG step of that parent's seed.
// ----------------------------------------------------------------------------------------
__device__ void spatial_node(const SweepArgs &a, long id) {
    int Pp = a.clus[id];
    int nb_pos = a.off_h[Pp], nb_num = a.off_h[Pp + 1] - nb_pos;
    uint32_t sn = a.s_neighbor[Pp];
    for (int p = 0; p < P; ++p) {
        sn = lcg_next(sn);
        int r = (int)umod0(sn, nb_num);
        int t = a.bd_h[r + nb_pos].y;
        uint32_t sc = a.s_child[t];
        for (int q = 0; q <= p; ++q) sc = lcg_next(sc);
        int cpos = a.child_off[t], ccnt = a.child_off[t + 1] - cpos;
        int r2 = (int)umod0(sc, ccnt);
        int c = a.child[cpos + r2];
        a.plane[id * K + 2 * P + p] = a.plane[(long)c * K];
    }
}

// ----------------------------------------------------------------------------------------
// one kernel per phase (used for per-kernel measurements and step-by-step runs)
// ----------------------------------------------------------------------------------------
#define GRID_STRIDE(i, n) for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

__global__ __launch_bounds__(256) void init_level_kernel(SweepArgs a) { GRID_STRIDE(id, a.V) init_node(a, id); }
__global__ __launch_bounds__(256) void init_cost_kernel(SweepArgs a) { GRID_STRIDE(id, a.V) init_cost_node(a, id); }
__global__ __launch_bounds__(256) void spatial_kernel(SweepArgs a) { GRID_STRIDE(id, a.V) spatial_node(a, id); }
__global__ __launch_bounds__(128) void node_kernel(SweepArgs a, int this_kind, int next_kind, float ndz, float ndn) {
    GRID_STRIDE(s, a.V) node_update(a, s, this_kind, next_kind, ndz, ndn);
}
template <int NS, int R>
__global__ __launch_bounds__(256) void plane_cost_kernel(SweepArgs a, int slot0) {
    extern __shared__ unsigned long long lds_pc[];
    constexpr int MAXC = 256 * R / 2 + 2;
    plane_cost_chunk<NS, R>(a, blockIdx.x, slot0, lds_pc, (unsigned int *)(lds_pc + MAXC * NS));
}
__global__ __launch_bounds__(256) void edge_kernel(SweepArgs a) { edge_group(a, blockIdx.x); }

__global__ void zero_f4_kernel(float4 *__restrict__ p, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ----------------------------------------------------------------------------------------
// all sweeps [s_begin, s_end) of one level in one cooperative launch: per sweep
//   (spatial candidates | grid barrier)  plane-cost chunks + message groups  | barrier
//   node updates (+ next candidates)  | barrier
// Work items are spread over the blocks in grid-stride order.  With first_costed, the first
// sweep's candidate costs were computed by the level init and its plane-cost phase is skipped.
// ----------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bp_level_kernel(SweepArgs a, const SweepDesc *__restrict__ sw, int s_begin,
                                                       int s_end, int first_costed) {
    extern __shared__ unsigned long long lds_pc[];
    auto grid = cg::this_grid();
    const long gtid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long gth = (long)gridDim.x * blockDim.x;
    const long nch4 = (a.N + CH4 - 1) / CH4;
    const long neg = (a.E + 15) / 16;
    for (int i = s_begin; i < s_end; ++i) {
        const SweepDesc d = sw[i];
        if (d.kind == kSpatial) {
            for (long id = gtid; id < a.V; id += gth) spatial_node(a, id);
            grid.sync();
        }
        const long nch = (first_costed && i == s_begin) ? 0 : nch4;
        for (long w = blockIdx.x; w < nch + neg; w += gridDim.x) {
            if (w < nch) plane_cost_chunk<4, 8>(a, w, 2 * P, lds_pc, (unsigned int *)(lds_pc + MAXC4 * 4));
            else edge_group(a, w - nch);
        }
        grid.sync();
        for (long s = gtid; s < a.V; s += gth) node_update(a, s, d.kind, d.next_kind, d.ndz, d.ndn);
        grid.sync();
    }
}

// ----------------------------------------------------------------------------------------
// map_msg (hbp_isp.cu:1081-1105): best of the P kept particles, then per pixel
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ float amin_fold(float stored, float v) {
    uint32_t vb = (v != v) ? 0x7FC00000u : __float_as_uint(v);
    uint32_t sb = __float_as_uint(stored);
    float vv = __uint_as_float(vb);
    uint32_t out;
    if (vv >= 0.0f) out = ((int32_t)vb < (int32_t)sb) ? vb : sb;
    else out = (vb > sb) ? vb : sb;
    return __uint_as_float(out);
}
__global__ void map_best_kernel(int V, const float *__restrict__ belief, const float4 *__restrict__ plane,
                                float4 *__restrict__ best) {
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < V; c += (long)gridDim.x * blockDim.x) {
        float mb = FLT_MAX;
#pragma unroll
        for (int p = 0; p < P; ++p) mb = amin_fold(mb, belief[c * K + p]);
        int bp = 0;
#pragma unroll
        for (int p = P - 1; p >= 0; --p)
            if (belief[c * K + p] == mb) bp = p;
        best[c] = plane[c * K + bp];
    }
}
__global__ void map_paint_kernel(const uint2 *__restrict__ pix, long N, const float4 *__restrict__ best, int W, int D,
                                 float *__restrict__ disp, float *__restrict__ plane3) {
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += (long)gridDim.x * blockDim.x) {
        uint2 q = pix[j];
        int x = unpack_x(q.x), y = unpack_y(q.x);
        float4 pl = best[q.y];
        float d = pl.x * (float)x + pl.y * (float)y + pl.z;
        d = fminf(fmaxf(d, 0.0f), (float)(D - 1));
        long i = (long)y * W + x;
        disp[i] = d;
        plane3[3 * i] = pl.x;
        plane3[3 * i + 1] = pl.y;
        plane3[3 * i + 2] = pl.z;
    }
}

// state readback helpers: [V][12] -> [12][V]
__global__ void transpose_state_kernel(int V, const float4 *__restrict__ plane, const float *__restrict__ cost,
                                       const float *__restrict__ belief, float4 *__restrict__ plane_o,
                                       float *__restrict__ cost_o, float *__restrict__ belief_o) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (long)V * K; i += (long)gridDim.x * blockDim.x) {
        long v = i / K, k = i % K;
        plane_o[k * V + v] = plane[i];
        cost_o[k * V + v] = cost[i];
        belief_o[k * V + v] = belief[i];
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

int fixed_point_shift(float lambda, float tau, int use_cost_vol, long npix) {
    double vmax = std::fabs((double)lambda) * (use_cost_vol ? std::fabs((double)tau) : 1.0) * 1.0001;
    int ev = 0;
    if (vmax > 0.0) std::frexp(vmax, &ev);
    int en = 0;
    std::frexp((double)(npix > 0 ? npix : 1), &en);
    int s = 62 - ev - en;
    if (s > 120) s = 120;
    if (s < -60) s = -60;
    return s;
}

// The sweep list of hbp_g (hbp_isp.cu:1283-1311) for levels L-2 .. min_level.
static std::vector<Sweep> schedule(int L, int min_level, int D) {
    std::vector<Sweep> out;
    for (int l = L - 2; l >= min_level; --l) {
        int iter_bp = 10 + l * 10;
        float max_dz = (float)D / 2.0f * (float)l / (float)L;
        float max_dn = 1.0f * (float)l / (float)L;
        size_t begin = out.size();
        for (int i = 0; i < iter_bp; ++i) {
            for (int r = 0; r < 4; ++r) out.push_back({l, kRansac, 0.f, 0.f, false, false});
            out.push_back({l, kSpatial, 0.f, 0.f, false, false});
            if (i > iter_bp / 4 * 3 && max_dz >= 0.01f) {
                out.push_back({l, kRandom, max_dz, max_dn, false, false});
                max_dz /= 2.0f;
                max_dn /= 2.0f;
            }
        }
        if (out.size() > begin) {
            out[begin].first = true;
            out.back().last = true;
        }
    }
    return out;
}
int bp_total_sweeps(int levels, int min_level, int D) { return (int)schedule(levels, min_level, D).size(); }

static SweepArgs sweep_args(BP &b, int l) {
    BPLevel &B = b.lv[l];
    const Volume &v = *b.vol;
    SweepArgs a{};
    a.V = B.V;
    a.Vh = B.Vh;
    a.W = b.W;
    a.D = b.D;
    a.E = B.E;
    a.N = b.N;
    a.clus = B.clus;
    a.bd = B.bd;
    a.pix = B.pix.p;
    a.pix_off = B.pix_off.p;
    a.off = B.off.p;
    a.rev = B.rev.p;
    a.c0 = B.c0.p;
    a.dpos = B.dpos.p;
    a.mxy = B.mxy.p;
    a.child = B.child.p;
    a.child_off = B.child_off.p;
    a.plane = B.plane.p;
    a.cost = B.cost.p;
    a.belief = B.belief.p;
    a.point = B.point.p;
    if (l + 1 < b.L) {
        BPLevel &Hh = b.lv[l + 1];
        a.bd_h = Hh.bd;
        a.off_h = Hh.off.p;
        a.plane_h = Hh.plane.p;
    }
    a.vol = v.vol.p;
    a.pitch = v.pitch;
    a.vplane = v.plane;
    a.dwta = v.dwta.p;
    a.lambda = b.lambda;
    a.tau = b.tau;
    a.tau_s = b.tau_s;
    a.use_cost_vol = b.use_cost_vol;
    a.shift = b.fx_shift;
    a.msg = b.msg.p;
    a.msg_t = b.msg_t.p;
    a.acc = b.acc.p;
    a.rrep = acc_replicas(B.V);
    a.accnan = b.accnan.p;
    a.s_ransac = b.s_ransac.p;
    a.s_neighbor = b.s_neighbor.p;
    a.s_child = b.s_child.p;
    return a;
}

void bp_setup(Ctx &c, BP &b) {
    const Pyramid &p = *b.pyr;
    const Volume &v = *b.vol;
    HBP_REQUIRE(p.W == v.W && p.H == v.H, "pyramid and cost volume sizes differ");
    b.W = p.W;
    b.H = p.H;
    b.D = v.D;
    b.N = (long)b.W * b.H;
    b.L = p.levels();
    b.tau = v.tau;
    HBP_REQUIRE(b.min_level >= 1 && b.min_level <= b.L - 1, "min_level must be in [1, levels-1]");
    b.fx_shift = fixed_point_shift(b.lambda, b.tau, b.use_cost_vol, b.N);
    b.lv.clear();
    b.lv.resize(b.L);
    long N = b.N;
    DevBuf<unsigned int> im_clus(c.pool, N), keys(c.pool, N), keys2(c.pool, N);
    DevBuf<int> ids(c.pool, N), ids2(c.pool, N);
    hipLaunchKernelGGL(iota_u32_kernel, dim3(grid_for(c, N)), dim3(256), 0, c.stream, im_clus.p, N);
    int maxV = 1;
    long maxE = 1;
    for (int l = 0; l < b.L; ++l) {
        BPLevel &B = b.lv[l];
        B.V = p.V[l];
        B.E = p.E[l];
        B.Vh = (l < b.L - 1) ? p.V[l + 1] : 1;
        B.clus = p.clus.p + p.voff[l];
        B.bd = p.bd.p + p.eoff[l];
        if (l >= b.min_level) {
            int V = B.V;
            long E = B.E;
            if (V > maxV) maxV = V;
            if (E > maxE) maxE = E;
            const float *mean = p.mean.p + 6 * p.voff[l];
            // pixels by (cluster, pixel id)
            HBP_CHECK(hipMemcpyAsync(keys.p, im_clus.p, sizeof(unsigned) * N, hipMemcpyDeviceToDevice, c.stream));
            hipLaunchKernelGGL(iota_i32_kernel, dim3(grid_for(c, N)), dim3(256), 0, c.stream, ids.p, N);
            sort_pairs_u32(c, keys.p, keys2.p, ids.p, ids2.p, N, bits_for(V));
            B.pix.alloc(c.pool, N);
            hipLaunchKernelGGL(pix_kernel, dim3(grid_for(c, N)), dim3(256), 0, c.stream, keys.p, ids.p, N, b.W,
                               B.pix.p);
            DevBuf<int> cnt(c.pool, (size_t)V + 1);
            B.pix_off.alloc(c.pool, (size_t)V + 1);
            HBP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(int) * ((size_t)V + 1), c.stream));
            hipLaunchKernelGGL(count_u32_kernel, dim3(grid_for(c, N)), dim3(256), 0, c.stream, im_clus.p, N, cnt.p);
            scan_exclusive_int(c, cnt.p, B.pix_off.p, V + 1, nullptr);
            // CSR by source, reverse edges, edge weights, mean positions
            B.off.alloc(c.pool, (size_t)V + 1);
            HBP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(int) * ((size_t)V + 1), c.stream));
            hipLaunchKernelGGL(count_src_kernel2, dim3(grid_for(c, E)), dim3(256), 0, c.stream, B.bd, E, cnt.p);
            scan_exclusive_int(c, cnt.p, B.off.p, V + 1, nullptr);
            B.rev.alloc(c.pool, E);
            hipLaunchKernelGGL(rev_kernel, dim3(grid_for(c, E)), dim3(256), 0, c.stream, B.bd, B.off.p, E, B.rev.p);
            B.c0.alloc(c.pool, E);
            B.dpos.alloc(c.pool, E);
            hipLaunchKernelGGL(edge_weights_kernel, dim3(grid_for(c, E)), dim3(256), 0, c.stream, mean, V, B.bd,
                               p.bdl.p + p.eoff[l], E, B.c0.p, B.dpos.p);
            B.mxy.alloc(c.pool, V);
            hipLaunchKernelGGL(mean_xy_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream, mean, V, B.mxy.p);
            // children grouped by parent (sort_clus, hbp_isp.cu:985-1001)
            {
                DevBuf<unsigned int> ck(c.pool, V), ck2(c.pool, V);
                DevBuf<int> cv2(c.pool, V), ccnt(c.pool, (size_t)B.Vh + 1);
                B.child.alloc(c.pool, V);
                B.child_off.alloc(c.pool, (size_t)B.Vh + 1);
                HBP_CHECK(hipMemcpyAsync(ck.p, B.clus, sizeof(int) * (size_t)V, hipMemcpyDeviceToDevice, c.stream));
                hipLaunchKernelGGL(iota_i32_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream, B.child.p, (long)V);
                sort_pairs_u32(c, ck.p, ck2.p, B.child.p, cv2.p, V, bits_for(B.Vh));
                HBP_CHECK(hipMemsetAsync(ccnt.p, 0, sizeof(int) * ((size_t)B.Vh + 1), c.stream));
                hipLaunchKernelGGL(count_i32_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream, B.clus, (long)V,
                                   ccnt.p);
                scan_exclusive_int(c, ccnt.p, B.child_off.p, B.Vh + 1, nullptr);
            }
            B.plane.alloc(c.pool, (size_t)V * K);
            B.cost.alloc(c.pool, (size_t)V * K);
            B.belief.alloc(c.pool, (size_t)V * K);
            B.point.alloc(c.pool, V);
            // the top level's planes are never set by the method; they read as zero
            hipLaunchKernelGGL(zero_f4_kernel, dim3(grid_for(c, (long)V * K)), dim3(256), 0, c.stream, B.plane.p,
                               (long)V * K);
            HBP_CHECK(hipMemsetAsync(B.cost.p, 0, sizeof(float) * (size_t)V * K, c.stream));
            HBP_CHECK(hipMemsetAsync(B.belief.p, 0, sizeof(float) * (size_t)V * K, c.stream));
            HBP_CHECK(hipMemsetAsync(B.point.p, 0, sizeof(uint4) * (size_t)V, c.stream));
        }
        if (l < b.L - 1)
            hipLaunchKernelGGL(compose_kernel, dim3(grid_for(c, N)), dim3(256), 0, c.stream, im_clus.p, B.clus, N);
        HBP_CHECK(hipGetLastError());
    }
    b.msg.alloc(c.pool, (size_t)maxE * 2);
    b.msg_t.alloc(c.pool, (size_t)maxE * K);
    size_t acc_n = 0;
    for (int l = b.min_level; l < b.L; ++l)
        acc_n = std::max(acc_n, (size_t)b.lv[l].V * acc_replicas(b.lv[l].V) * K);
    b.acc.alloc(c.pool, acc_n);
    b.accnan.alloc(c.pool, (size_t)maxV);
    b.s_ransac.alloc(c.pool, maxV);
    b.s_neighbor.alloc(c.pool, maxV);
    b.s_child.alloc(c.pool, maxV);
    b.disp.alloc(c.pool, N);
    b.plane3.alloc(c.pool, (size_t)N * 3);
    b.sched = schedule(b.L, b.min_level, b.D);
    b.next = 0;
    b.cur = b.L - 2;
    b.sweeps = 0;
    b.mapped = false;
    HBP_CHECK(hipStreamSynchronize(c.stream));
}


// algorithmic HBM bytes per launch (pixel id 4 B + 2 cost bytes per plane and pixel; per
// superpixel the planes read and the sums written)
static double plane_cost_bytes(long N, int V, int NS) { return (double)N * (4 + 2 * NS) + (double)V * (16 * NS + 8 + 4 * NS); }
static double edge_bytes(long E, int V) { return 96.0 * E + (16.0 * K + 48) * V; }
static double node_bytes(long E, int V) { return 84.0 * E + (48.0 * K + 16) * V; }

template <int NS, int R>
static void launch_plane_cost(Ctx &c, const SweepArgs &a, int l, int slot0) {
    constexpr int CH = 256 * R;
    ProfScope ps(c, "plane_cost_L" + std::to_string(l), plane_cost_bytes(a.N, a.V, NS));
    hipLaunchKernelGGL((plane_cost_kernel<NS, R>), dim3(div_up(a.N, CH)), dim3(256), PC_LDS_BYTES, c.stream, a,
                       slot0);
}

// one sweep as separate kernels
static void run_sweep_split(Ctx &c, BP &b, const Sweep &sw, int next_kind, float ndz, float ndn) {
    int l = sw.level;
    SweepArgs a = sweep_args(b, l);
    if (sw.first) {
        HBP_CHECK(hipMemsetAsync(b.msg.p, 0, sizeof(float4) * 2 * (size_t)a.E, c.stream));
        hipLaunchKernelGGL(init_level_kernel, dim3(grid_for(c, a.V)), dim3(256), 0, c.stream, a);
        launch_plane_cost<12, 2>(c, a, l, 0);
        hipLaunchKernelGGL(init_cost_kernel, dim3(grid_for(c, a.V)), dim3(256), 0, c.stream, a);
    } else {
        if (sw.kind == kSpatial) {
            ProfScope pss(c, "spatial_L" + std::to_string(l), (16.0 * P + 28) * a.V);
            hipLaunchKernelGGL(spatial_kernel, dim3(grid_for(c, a.V)), dim3(256), 0, c.stream, a);
        }
        launch_plane_cost<4, 8>(c, a, l, 2 * P);
    }
    if (a.E > 0) {
        ProfScope ps(c, "edge_L" + std::to_string(l), edge_bytes(a.E, a.V));
        hipLaunchKernelGGL(edge_kernel, dim3(div_up(a.E, 16)), dim3(256), 0, c.stream, a);
    }
    {
        ProfScope psn(c, "node_L" + std::to_string(l), node_bytes(a.E, a.V));
        hipLaunchKernelGGL(node_kernel, dim3(grid_for(c, a.V, 128)), dim3(128), 0, c.stream, a, sw.kind, next_kind,
                           ndz, ndn);
    }
    HBP_CHECK(hipGetLastError());
}

int bp_level_coop_blocks(int num_cu) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bp_level_kernel, 256, PC_LDS_BYTES) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    if (per_cu > 4) per_cu = 4;
    return per_cu * num_cu;
}

// sweeps [i0, i1) of one level in one cooperative launch
static void run_sweeps_fused(Ctx &c, BP &b, size_t i0, size_t i1) {
    const Sweep &s0 = b.sched[i0];
    int l = s0.level;
    SweepArgs a = sweep_args(b, l);
    // the device copy of the level's schedule
    std::vector<SweepDesc> h(i1 - i0);
    for (size_t i = i0; i < i1; ++i) {
        const Sweep &sw = b.sched[i];
        SweepDesc d{sw.kind, kNone, 0.f, 0.f};
        if (!sw.last && i + 1 < b.sched.size()) {
            d.next_kind = b.sched[i + 1].kind;
            d.ndz = b.sched[i + 1].dz;
            d.ndn = b.sched[i + 1].dn;
        }
        h[i - i0] = d;
    }
    b.sweep_desc.alloc(c.pool, h.size() * sizeof(SweepDesc));
    HBP_CHECK(hipMemcpyAsync(b.sweep_desc.p, h.data(), h.size() * sizeof(SweepDesc), hipMemcpyHostToDevice,
                             c.stream));
    if (s0.first) {
        // level init as separate kernels (its 12-slot plane cost needs more registers)
        HBP_CHECK(hipMemsetAsync(b.msg.p, 0, sizeof(float4) * 2 * (size_t)a.E, c.stream));
        hipLaunchKernelGGL(init_level_kernel, dim3(grid_for(c, a.V)), dim3(256), 0, c.stream, a);
        launch_plane_cost<12, 2>(c, a, l, 0);
        hipLaunchKernelGGL(init_cost_kernel, dim3(grid_for(c, a.V)), dim3(256), 0, c.stream, a);
        HBP_CHECK(hipGetLastError());
    }
    long work = std::max({(long)div_up(a.N, CH4), (long)div_up(a.E, 16), (long)div_up(a.V, 256), 1L});
    int blocks = (int)std::min<long>(work, c.bp_coop_blocks);
    const SweepDesc *dsw = (const SweepDesc *)b.sweep_desc.p;
    int sb = 0, se = (int)(i1 - i0), fc = s0.first ? 1 : 0;
    void *args[] = {(void *)&a, (void *)&dsw, (void *)&sb, (void *)&se, (void *)&fc};
    double bytes = (double)(i1 - i0) * (plane_cost_bytes(a.N, a.V, 4) + edge_bytes(a.E, a.V) + node_bytes(a.E, a.V));
    ProfScope ps(c, "bp_level_L" + std::to_string(l), bytes);
    HBP_CHECK(hipLaunchCooperativeKernel((const void *)bp_level_kernel, dim3(blocks), dim3(256), args, PC_LDS_BYTES,
                                         c.stream));
}

int bp_step(Ctx &c, BP &b, int max_sweeps) {
    int done = 0;
    while (done < max_sweeps && b.next < b.sched.size()) {
        const Sweep &sw = b.sched[b.next];
        if (b.fused) {
            // the rest of this level (or as much of it as max_sweeps allows) in one launch
            size_t i1 = b.next;
            while (i1 < b.sched.size() && b.sched[i1].level == sw.level && (long)(i1 - b.next) < max_sweeps - done &&
                   (i1 == b.next || !b.sched[i1].first))
                ++i1;
            run_sweeps_fused(c, b, b.next, i1);
            int n = (int)(i1 - b.next);
            b.cur = sw.level;
            b.next = i1;
            done += n;
            b.sweeps += n;
            continue;
        }
        int next_kind = kNone;
        float ndz = 0.f, ndn = 0.f;
        if (!sw.last && b.next + 1 < b.sched.size()) {
            const Sweep &nx = b.sched[b.next + 1];
            next_kind = nx.kind;
            ndz = nx.dz;
            ndn = nx.dn;
        }
        run_sweep_split(c, b, sw, next_kind, ndz, ndn);
        b.cur = sw.level;
        b.next++;
        done++;
        b.sweeps++;
    }
    if (b.next >= b.sched.size()) b.cur = b.min_level - 1;
    return done;
}

void bp_map(Ctx &c, BP &b) {
    BPLevel &B = b.lv[b.min_level];
    b.best.alloc(c.pool, B.V);
    hipLaunchKernelGGL(map_best_kernel, dim3(grid_for(c, B.V)), dim3(256), 0, c.stream, B.V, B.belief.p, B.plane.p,
                       b.best.p);
    hipLaunchKernelGGL(map_paint_kernel, dim3(grid_for(c, b.N)), dim3(256), 0, c.stream, B.pix.p, b.N, b.best.p, b.W,
                       b.D, b.disp.p, b.plane3.p);
    HBP_CHECK(hipGetLastError());
    b.mapped = true;
}

void bp_get_state(Ctx &c, BP &b, float *planes, float *cost, float *belief, float *msg) {
    int l = b.cur < b.min_level ? b.min_level : b.cur;
    BPLevel &B = b.lv[l];
    long V = B.V;
    DevBuf<float4> po(c.pool, V * K);
    DevBuf<float> co(c.pool, V * K), bo(c.pool, V * K);
    hipLaunchKernelGGL(transpose_state_kernel, dim3(grid_for(c, V * K)), dim3(256), 0, c.stream, (int)V, B.plane.p,
                       B.cost.p, B.belief.p, po.p, co.p, bo.p);
    HBP_CHECK(hipGetLastError());
    if (planes) HBP_CHECK(hipMemcpyAsync(planes, po.p, sizeof(float4) * V * K, hipMemcpyDeviceToHost, c.stream));
    if (cost) HBP_CHECK(hipMemcpyAsync(cost, co.p, sizeof(float) * V * K, hipMemcpyDeviceToHost, c.stream));
    if (belief) HBP_CHECK(hipMemcpyAsync(belief, bo.p, sizeof(float) * V * K, hipMemcpyDeviceToHost, c.stream));
    if (msg) HBP_CHECK(hipMemcpyAsync(msg, b.msg.p, sizeof(float4) * 2 * B.E, hipMemcpyDeviceToHost, c.stream));
    HBP_CHECK(hipStreamSynchronize(c.stream));
}

}  // namespace hbp
