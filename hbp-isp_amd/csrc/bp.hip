// Stage (c): hierarchical plane-particle min-sum BP on the superpixel pyramid
// (hbp_isp.cu / hbp_isp.hpp of the reference).
//
// One BP sweep (generate_plane + compute_plane_cost + belief_propagation, hbp_isp.cu:1302-1308)
// is two kernels on one stream:
//   cost_edge_kernel  (1) per pixel: the cost of the 4 candidate planes of its superpixel,
//                     summed per superpixel as int64 fixed point, so the sums do not depend
//                     on the reduction order; (2) per (edge, target slot): the min-sum message
//                     of smooth_msg_k and the mean removal of normalize_msg_k.  On spatial
//                     sweeps the candidates of select_child_plane + neighbor_candidate_plane
//                     are resolved inside this kernel.
//   node_kernel       per superpixel, 16 lanes (one per slot): belief = cost + incoming
//                     messages in edge order, rank of the 12 beliefs, compaction of planes /
//                     costs / beliefs and of the outgoing messages (select_*), written to the
//                     reverse edge (ping_pang_msg), and the next sweep's RANSAC / random
//                     candidates.
//
// Plane costs have two layouts.  Coarse levels (>= 32 pixels per superpixel on average) cut
// each superpixel's sorted pixel run into segments of up to 512 pixels; a wavefront owns a
// segment, loads the superpixel's planes once, reads 64 consecutive pixels per step, reduces
// across lanes and adds one int64 per plane to the superpixel's sum.  Fine levels use chunks
// of consecutive sorted pixels per 256-thread block with a segmented wavefront scan and LDS
// counters per superpixel.
//
// HBM layout per level: planes [V][12] float4, cost/belief [V][12] f32, messages [E][8] f32
// (the 8 slots the next sweep reads), sweep messages [E][12] f32, pixels sorted by
// (superpixel, pixel id) as packed x|y<<16 [N] and superpixel id [N], CSR offsets / reverse
// ids / edge weights per edge, fixed-point sums [V][rrep][12] int64.
#include <algorithm>
#include <cmath>

#include <utility>

#include "hbp_internal.h"

namespace hbp {

constexpr int P = kP, K = kK;
// incoming messages a node update keeps in registers; the node-major message and reverse-slot
// arrays carry this many padding rows so that a node's 16 row reads stay in bounds
constexpr int kNodeMaxDeg = 16;
constexpr int SEG_R = 8, SEG = 64 * SEG_R;   // pixels per plane-cost segment (coarse levels)

// replicas of each fixed-point accumulator: work item i adds into replica i mod rrep, which
// spreads the atomics of the few large superpixels of the coarse levels over rrep addresses
// (8 for small levels: measured 0.5 ms per 1472x984 pair faster than 16, whose node update sums
// and clears twice the words, at no cost in the plane-cost launch's atomics)
static inline int acc_replicas(int V) { return V <= 4096 ? 8 : (V <= 65536 ? 4 : 1); }   // <= kMaxReplicas

// ----------------------------------------------------------------------------------------
// prologue kernels
// ----------------------------------------------------------------------------------------
// Reversed, padded right census for the left-view plane cost: row y holds CR(y, max(W-1-j, 0))
// at j = 0 .. S-1, so CR(max(x-d, 0)) and CR(max(x-d-1, 0)) are adjacent words for every
// 0 <= d <= S-W.
__global__ void crx_kernel(const uint32_t *__restrict__ cr, int W, int H, int S, uint32_t *__restrict__ out) {
    const long n = (long)H * S;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int y = (int)(i / S), j = (int)(i % S);
        const int x = W - 1 - j;
        out[i] = cr[(long)y * W + (x > 0 ? x : 0)];
    }
}

__global__ void zero_f4_kernel(float4 *__restrict__ p, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}
__global__ void iota_u32_kernel(unsigned int *__restrict__ p, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = (unsigned)i;
}
__global__ void compose_kernel(unsigned int *__restrict__ im_clus, const int *__restrict__ clus, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        im_clus[i] = (unsigned)clus[im_clus[i]];
}
// off[k] = first j with keys[j] >= k, k in [0, nkeys]; keys ascending (a CSR offset array
// without atomics: every key run writes the offsets of the keys it starts)
__global__ void lower_bound_kernel(const unsigned int *__restrict__ keys, long n, int nkeys, int *__restrict__ off) {
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        long k1 = keys[j];
        long k0 = j == 0 ? 0 : (long)keys[j - 1] + 1;
        for (long k = k0; k <= k1 && k <= nkeys; ++k) off[k] = (int)j;
        if (j == n - 1)
            for (long k = k1 + 1; k <= nkeys; ++k) off[k] = (int)n;
    }
}
// the same over the sources of edges sorted by (s, t) (every pyramid level's edge list is);
// no edges (a top level of one cluster): every offset is 0
__global__ void lower_bound_src_kernel(const int2 *__restrict__ bd, long n, int nkeys, int *__restrict__ off) {
    if (n == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            for (int k = 0; k <= nkeys; ++k) off[k] = 0;
        return;
    }
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        long k1 = bd[j].x;
        long k0 = j == 0 ? 0 : (long)bd[j - 1].x + 1;
        for (long k = k0; k <= k1 && k <= nkeys; ++k) off[k] = (int)j;
        if (j == n - 1)
            for (long k = k1 + 1; k <= nkeys; ++k) off[k] = (int)n;
    }
}
// a sort's input: keys copied, values 0 .. n-1
__global__ void copy_iota_kernel(const unsigned int *__restrict__ src, long n, unsigned int *__restrict__ keys,
                                 int *__restrict__ ids) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        keys[i] = src[i];
        ids[i] = (int)i;
    }
}
__global__ void deg_max_kernel(const int *__restrict__ off, int V, int *__restrict__ dmax) {
    int m = 0;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x)
        m = max(m, off[v + 1] - off[v]);
    if (m > 0) atomicMax(dmax, m);
}
// node-major edge slots: edge e = (s, t) is slot eloc[e] = e - off[s] of s; revl[s][slot] = rev(e)
__global__ void node_major_kernel(const int2 *__restrict__ bd, const int *__restrict__ off,
                                  const int *__restrict__ rev, long E, int maxdeg, int *__restrict__ eloc,
                                  int *__restrict__ revl) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (long)gridDim.x * blockDim.x) {
        int sn = bd[e].x;
        int i = (int)(e - off[sn]);
        eloc[e] = i;
        revl[(long)sn * maxdeg + i] = rev[e];
    }
}
// per edge: the slot of its reverse edge among the reverse's source's edges
__global__ void rloc_kernel(const int *__restrict__ eloc, const int *__restrict__ rev, long E, int *__restrict__ rloc) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (long)gridDim.x * blockDim.x)
        rloc[e] = eloc[rev[e]];
}
// WTA disparity of the sorted pixels
__global__ void pw_kernel(const uint32_t *__restrict__ pxy, long n, const float *__restrict__ dwta, int W,
                          float *__restrict__ pw) {
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        uint32_t q = pxy[j];
        pw[j] = dwta[(long)unpack_y(q) * W + unpack_x(q)];
    }
}
// setup_D4 (hbp_isp.cu:871-891): pixels in (cluster, pixel id) order as packed xy + cluster
__global__ void pix_kernel(const unsigned int *__restrict__ keys, const int *__restrict__ ids, long n, int W,
                           uint32_t *__restrict__ pxy, uint32_t *__restrict__ pcl) {
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        int id = ids[j];
        pxy[j] = pack_xy(id % W, id / W);
        pcl[j] = keys[j];
    }
}
// segments of at most sl <= SEG pixels of one superpixel: count, then (first, end, cluster, index)
__global__ void seg_count_kernel(const int *__restrict__ pix_off, int V, int sl, int *__restrict__ cnt) {
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c <= V; c += (long)gridDim.x * blockDim.x)
        cnt[c] = c < V ? (pix_off[c + 1] - pix_off[c] + sl - 1) / sl : 0;
}
// one thread per segment k < sego[V]: its superpixel is the last c with sego[c] <= k (superpixels
// of the coarse levels hold up to tens of thousands of pixels, so one thread per superpixel left
// a few threads writing a hundred segments each)
__global__ void seg_fill_kernel(const int *__restrict__ pix_off, const int *__restrict__ sego, int V, int sl,
                                long nmax, int4 *__restrict__ seg) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < nmax; k += (long)gridDim.x * blockDim.x) {
        if (k >= sego[V]) break;
        int lo = 0, hi = V;   // sego[lo] <= k < sego[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (sego[mid] <= k) lo = mid;
            else hi = mid;
        }
        const int i = (int)(k - sego[lo]);
        const int b = pix_off[lo], e = pix_off[lo + 1];
        seg[k] = make_int4(b + i * sl, min(b + (i + 1) * sl, e), lo, i);
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

// one plane of ransac_candidate_plane (hbp_isp.cu:592-640): three LCG draws from seed s
// (advanced in place), three distinct pixels of the superpixel, the plane through their WTA
// disparities.  q returns the packed pixels.
__device__ __forceinline__ void ransac_draw(uint32_t &s, int n, int &p0, int &p1, int &p2) {
    s = lcg_next(s);
    p0 = (int)umod0(s, n);
    s = lcg_next(s);
    p1 = (int)umod0(s, n - 1);
    s = lcg_next(s);
    p2 = (int)umod0(s, n - 2);
    three_diff(p0, p1, p2, n);
}
__device__ __forceinline__ float4 ransac_plane(uint32_t &s, const uint32_t *__restrict__ pxy, int pos, int n,
                                               const float *__restrict__ pw, uint4 &q) {
    int p0, p1, p2;
    ransac_draw(s, n, p0, p1, p2);
    q.x = pxy[pos + p0];
    q.y = pxy[pos + p1];
    q.z = pxy[pos + p2];
    const float z0 = pw[pos + p0], z1 = pw[pos + p1], z2 = pw[pos + p2];
    float4 A = make_float4((float)unpack_x(q.x), (float)unpack_y(q.x), z0, 0.f);
    float4 B = make_float4((float)unpack_x(q.y), (float)unpack_y(q.y), z1, 0.f);
    float4 C = make_float4((float)unpack_x(q.z), (float)unpack_y(q.z), z2, 0.f);
    return fit_plane(A, B, C);
}

// ransac_candidate_plane for the candidate slots 2P..3P-1 of one superpixel
__device__ void ransac_node(int id, const uint32_t *__restrict__ pxy, const int *__restrict__ pix_off,
                            const float *__restrict__ pw, uint32_t *__restrict__ s_ransac,
                            float4 *__restrict__ plane, uint4 *__restrict__ point) {
    int pos = pix_off[id];
    int n = pix_off[id + 1] - pos;
    uint32_t s = s_ransac[id];
    uint4 q = make_uint4(0u, 0u, 0u, 0u);
    for (int p = 2 * P; p < 3 * P; ++p) plane[(long)id * K + p] = ransac_plane(s, pxy, pos, n, pw, q);
    point[id] = q;
    s_ransac[id] = s;
}

// one plane of random_candidate_plane (hbp_isp.cu:641-674) with normalize_plane (495-503):
// four draws from seed s (advanced in place) perturb plane pl around the mean position (x, y)
__device__ __forceinline__ float4 random_plane(uint32_t &s, float4 pl, float x, float y, float max_dz, float max_dn) {
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
    return make_float4(a, b, c, 0.f);
}

enum SweepKind { kNone = 0, kRansac = 1, kSpatial = 2, kRandom = 3 };

// Every pointer and constant one sweep reads or writes (level structures, parent level,
// volume, parameters, state).  Passed by value to the kernels.
struct SweepArgs {
    int V, Vh, W, D;
    long E, N;
    const int *clus;
    const int2 *bd;
    const uint32_t *pxy;         // packed x | y << 16 of the sorted pixels [N]
    const uint32_t *pcl;         // superpixel of the sorted pixels [N]
    const int *pix_off;
    const int4 *seg;             // coarse levels: (first, end, superpixel, index) [nseg]
    const int *sego;             // coarse levels: first segment of each superpixel [V+1]
    const int *off;
    const int *rev;
    int maxdeg;                  // largest out-degree of the level
    const int *eloc;             // slot of edge e among its source's edges [E]
    const int *revl;             // reverse edge of each (node, slot) [V][maxdeg]
    const float *pw;             // WTA disparity of the sorted pixels [N]
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
    // census words of the pair [H][W] (the cost volume's entries are recomputed from them),
    // view of the volume, WTA disparities and parameters
    const uint32_t *cl, *cr;
    const uint32_t *crx;         // reversed, padded right census [H][crx_stride] (left view)
    int crx_stride;
    const float *volf;           // a given f32 volume [D][H][W] (MC-CNN), else null
    long vplane;                 // H W
    int right;
    const float *dwta;
    float lambda, tau, tau_s;
    float lam_fx;                // |lambda| 2^shift
    int use_cost_vol, shift;
    // state
    float4 *msg;                 // [E][2] float4
    float *msg_t;                // sweep messages, node-major [V][maxdeg][12]
    unsigned long long *acc;     // [V][rrep][12] fixed-point sums, rrep replicas
    int rrep;
    unsigned int *accnan;        // [V]
    uint32_t *s_ransac, *s_neighbor, *s_child;
    // a spatial sweep of a pair whose level start computed its candidates' superpixels
    // (spatial_ahead): [V][P], else null (the candidates are then resolved from the seeds)
    const int *spc;
};

// fine levels: a 256-thread block costs CHUNK consecutive sorted pixels, 4 per lane
constexpr int CHUNK = 1024;
constexpr int MAXC = CHUNK / 2 + 2;    // superpixels a chunk can touch (each holds >= 2 pixels)
constexpr size_t PC_LDS_BYTES = (size_t)MAXC * P * 16 + (size_t)MAXC * P * 8 + (size_t)MAXC * 4;

// sum / clear of the rrep replicas of one accumulator; unrolled to the largest replica count
// so that all loads are in flight together
constexpr int kMaxReplicas = 8;
// A buffer of slot stride S (K, or P for one sweep's candidates alone) holds slots K-S .. K-1 of
// each node: slot k of node n, replica r is at (n rrep + r) S + k - (K - S).
__device__ __forceinline__ int64_t acc_sum_at(const unsigned long long *__restrict__ acc, int rrep, long id, int k,
                                              int stride = K) {
    const unsigned long long *__restrict__ p = acc + (size_t)id * rrep * stride + (k - (K - stride));
    unsigned long long v[kMaxReplicas];
#pragma unroll
    for (int r = 0; r < kMaxReplicas; ++r) v[r] = r < rrep ? p[r * stride] : 0ull;
    unsigned long long t = 0ull;
#pragma unroll
    for (int r = 0; r < kMaxReplicas; ++r) t += v[r];
    return (int64_t)t;
}
__device__ __forceinline__ void acc_zero_at(unsigned long long *__restrict__ acc, int rrep, long id, int k,
                                            int stride = K) {
    unsigned long long *__restrict__ p = acc + (size_t)id * rrep * stride + (k - (K - stride));
#pragma unroll
    for (int r = 0; r < kMaxReplicas; ++r)
        if (r < rrep) p[r * stride] = 0ull;
}
__device__ __forceinline__ int64_t acc_sum(const SweepArgs &a, long id, int k) { return acc_sum_at(a.acc, a.rrep, id, k); }
__device__ __forceinline__ void acc_zero(const SweepArgs &a, long id, int k) { acc_zero_at(a.acc, a.rrep, id, k); }

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
    ransac_node((int)id, a.pxy, a.pix_off, a.pw, a.s_ransac, a.plane, a.point);
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
// spatial propagation (hbp_isp.cu:1225-1243): candidate p of node id with parent Pp is the
// slot-0 plane of a random child (select_child_plane) of a random neighbour of Pp at the parent
// level (neighbor_candidate_plane); each draw is the (p+1)-th LCG step of that parent's seed.
// Parents at BP levels have at least one neighbour and one child.
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ float4 spatial_candidate(const SweepArgs &a, long id, int p) {
    int Pp = a.clus[id];
    int nb_pos = a.off_h[Pp], nb_num = a.off_h[Pp + 1] - nb_pos;
    uint32_t sn = a.s_neighbor[Pp];
    for (int q = 0; q <= p; ++q) sn = lcg_next(sn);
    int r = (int)umod0(sn, nb_num);
    int t = a.bd_h[r + nb_pos].y;
    uint32_t sc = a.s_child[t];
    for (int q = 0; q <= p; ++q) sc = lcg_next(sc);
    int cpos = a.child_off[t], ccnt = a.child_off[t + 1] - cpos;
    int r2 = (int)umod0(sc, ccnt);
    int c = a.child[cpos + r2];
    return a.plane[(long)c * K];
}
__device__ void spatial_node(const SweepArgs &a, long id) {
    for (int p = 0; p < P; ++p) a.plane[id * K + 2 * P + p] = spatial_candidate(a, id, p);
}
// The superpixel whose best plane is candidate p of node id in the m-th spatial sweep of the
// level: spatial_candidate's draws, from the seeds that sweep reads.  The level init sets the
// seeds of parent q to the level's off[q] (neighbour draws) and child_off[q] (child draws), and
// only the node updates of spatial sweeps change them, by P draws each, so the m-th spatial
// sweep reads them m P draws later.
__device__ int spatial_ahead(const SweepArgs &a, int m, long id, int p) {
    const LcgJump jm = lcg_jump_pow((unsigned long)m * P + p + 1);
    const int Pp = a.clus[id];
    const int nb_pos = a.off_h[Pp], nb_num = a.off_h[Pp + 1] - nb_pos;
    const uint32_t sn = (jm.a * (uint32_t)a.off[Pp] + jm.c) % 2147483648u;
    const int t = a.bd_h[(int)umod0(sn, nb_num) + nb_pos].y;
    const uint32_t sc = (jm.a * (uint32_t)a.child_off[t] + jm.c) % 2147483648u;
    const int cpos = a.child_off[t], ccnt = a.child_off[t + 1] - cpos;
    return a.child[cpos + (int)umod0(sc, ccnt)];
}

// Plane-cost modes: the census cost volume of the left view, of the right view
// (cost_to_right), the WTA indicator of StereoData(use_cost_vol = 0), or a given f32 volume
// (the "MC-CNN" cost, either view: the library holds it in the view's own orientation).
enum CostMode { kCostLeft = 0, kCostRight = 1, kCostWta = 2, kCostF32 = 3 };

// ----------------------------------------------------------------------------------------
// read_plane_cost (hbp_isp.cu:177-228) of one pixel and plane, times lambda, as the term
// rint(|lambda * cost| 2^s) of the cluster's fixed-point sum (an integer valued float below
// 2^52 / N, so per-lane sums of these terms are exact in double).
//
// Entry (d, y, x) of the cost volume, 0 <= d < D, is recomputed from the census words exactly
// as the Hamming kernel stores it: left view min(popc(CL(x) ^ CR(max(x-d, 0))), 25); right
// view (cost_to_right) the left entry at x' = min(x+d, W-1).  The entries at d = floor(df) and
// d+1 are read at d clamped into [0, D-2]; the cost is tau when d is outside [0, D-1).
//
// Left view: the words come from the reversed right census crx, whose row y holds
// CR(y, max(W-1-j, 0)) at j = 0 .. S-1, S = W + max(D-1, 1): CR(max(x-d, 0)) and
// CR(max(x-d-1, 0)) are the adjacent words j = W-1-x+d and j+1, one 8-byte load, no clamp.
//
// Two forms of the same arithmetic:
//  * GEN: any plane.  A NaN disparity counts as d = 0, in range, so its cost is NaN (the
//    caller flags it and adds zero); lanes past the end of the pixels add zero.
//  * fast: every plane of the group has |a|, |b|, |c| <= 2^60 (|df| < 2^77: never NaN or
//    inf), D >= 2 and every lane holds a pixel.  d is in range exactly when floor(df) equals
//    its clamp into [0, D-2].
// Both take (d+1) - df as 1 - (df - d): for 0 <= d < 2^23, df - d is exact (Sterbenz) and so
// is d+1, so both are the one rounding of the same real number (out of range the value is
// not used).  |lambda * cost| 2^s is computed as (|lambda| 2^s) * cost: scaling by 2^s
// commutes with rounding while the products are normal floats, and otherwise both forms
// round to 0 (|lambda| 2^s is a finite normal float, checked on the host).
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ void pin(uint32_t &v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void pin(int &v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void pin(float &v) { asm volatile("" : "+v"(v)); }

struct __attribute__((aligned(4))) Words2 {
    uint32_t lo, hi;
};

__device__ __forceinline__ bool tame_plane(float4 p) {
    const float lim = 0x1p60f;
    return fabsf(p.x) <= lim && fabsf(p.y) <= lim && fabsf(p.z) <= lim;
}

// Loads at a 32-bit byte offset from a kernel-argument base: every image array here is below
// 2^31 bytes, and the offset form needs no 64-bit address arithmetic per lane.
template <typename T>
__device__ __forceinline__ T ld_at(const void *base, uint32_t byte_off) {
    return *(const T *)((const char *)base + byte_off);
}

struct PcPix {
    int x;
    float fx, fy;
    uint32_t clx;   // left view: CL(x)
    uint32_t pb;    // left view: word index of crx (y, W-1-x)
    uint32_t row;   // right view: y W
    float dw;       // WTA: WTA disparity
};
struct PcWords {
    uint32_t u0, u1, w0, w1;
    float tf;   // fast form: df - floor(df)
    bool in;    // fast form: floor(df) is in [0, D-2]
};

// x, y < 2^16 and y W + x < 2^24: 24-bit multiplies are exact
template <int MODE>
__device__ __forceinline__ PcPix pc_pixel(const SweepArgs &a, uint32_t q) {
    PcPix p;
    p.x = unpack_x(q);
    const uint32_t y = (uint32_t)unpack_y(q);
    p.fx = (float)p.x;
    p.fy = (float)y;
    p.row = __umul24(y, (uint32_t)a.W);
    if (MODE == kCostLeft) {
        p.clx = ld_at<uint32_t>(a.cl, (p.row + (uint32_t)p.x) << 2);
        p.pb = __umul24(y, (uint32_t)a.crx_stride) + (uint32_t)(a.W - 1 - p.x);
    }
    if (MODE == kCostWta) p.dw = ld_at<float>(a.dwta, (p.row + (uint32_t)p.x) << 2);
    return p;
}
template <int MODE>
__device__ __forceinline__ void pc_pin(PcPix &p) {
    if (MODE == kCostLeft) pin(p.clx);
    if (MODE == kCostWta) pin(p.dw);
}

// issue the census loads of one pixel and plane (the disparity clamped into [0, D-2])
template <int MODE, bool GEN>
__device__ __forceinline__ void pc_load(const SweepArgs &a, const PcPix &px, float4 pl, float dmax, PcWords &wd) {
    if (MODE == kCostWta) return;
    const float df = pl.x * px.fx + pl.y * px.fy + pl.z;
    const float fl = floorf(df);
    // GEN: NaN -> 0, +-inf -> the ends
    const float dcf = GEN ? fminf(fmaxf(fl, 0.f), dmax) : __builtin_amdgcn_fmed3f(fl, 0.f, dmax);
    const int dc = (int)dcf;
    if (!GEN) {
        wd.tf = df - fl;
        wd.in = fl == dcf;
    }
    if (MODE == kCostLeft) {
        const Words2 w = ld_at<Words2>(a.crx, (px.pb + (uint32_t)dc) << 2);
        wd.u0 = w.lo;
        wd.u1 = w.hi;
    } else if (MODE == kCostF32) {
        // entries (dc, y, x) and (dc+1, y, x) of the f32 volume, as bits
        const float *e = a.volf + (long)dc * a.vplane + px.row + px.x;
        wd.u0 = __float_as_uint(e[0]);
        wd.u1 = __float_as_uint(e[a.vplane]);
    } else {
        const int xa = px.x + dc < a.W - 1 ? px.x + dc : a.W - 1;
        const int xb = px.x + dc + 1 < a.W - 1 ? px.x + dc + 1 : a.W - 1;
        wd.u0 = ld_at<uint32_t>(a.cl, (px.row + (uint32_t)xa) << 2);
        wd.w0 = ld_at<uint32_t>(a.cr, (px.row + (uint32_t)(xa - dc)) << 2);
        wd.u1 = ld_at<uint32_t>(a.cl, (px.row + (uint32_t)xb) << 2);
        wd.w1 = ld_at<uint32_t>(a.cr, (px.row + (uint32_t)(xb - dc - 1)) << 2);
    }
}
template <int MODE, bool GEN>
__device__ __forceinline__ void pc_pin(PcWords &w) {
    if (MODE == kCostWta) return;
    if (!GEN) pin(w.tf);
    pin(w.u0);
    pin(w.u1);
    if (MODE == kCostRight) {
        pin(w.w0);
        pin(w.w1);
    }
}

// the term of one pixel and plane from its loaded words, lam = |lambda| 2^s (0 for a lane
// without a pixel); isn: the cost is NaN
template <int MODE, bool GEN>
__device__ __forceinline__ float pc_term(const SweepArgs &a, const PcPix &px, float4 pl, const PcWords &wd,
                                         float dmax, float dlim, float lam, bool &isn) {
    float cost;
    if (MODE == kCostWta) {
        const float df = pl.x * px.fx + pl.y * px.fy + pl.z;
        cost = fabsf(px.dw - df) <= 1.f ? 0.f : 1.f;
    } else if (!GEN) {
        const bool in = wd.in;
        float v0, v1;
        if (MODE == kCostF32) {
            v0 = __uint_as_float(wd.u0);
            v1 = __uint_as_float(wd.u1);
        } else {
            int c0, c1;
            if (MODE == kCostLeft) {
                c0 = __popc(px.clx ^ wd.u0);
                c1 = __popc(px.clx ^ wd.u1);
            } else {
                c0 = __popc(wd.u0 ^ wd.w0);
                c1 = __popc(wd.u1 ^ wd.w1);
            }
            v0 = (float)(c0 < 25 ? c0 : 25);
            v1 = (float)(c1 < 25 ? c1 : 25);
        }
        const float t1 = wd.tf;
        float ci = (1.f - t1) * v0 + t1 * v1;
        pin(ci);   // a select, not a branch around the arithmetic
        cost = in ? ci : a.tau;
    } else {
        const float df = pl.x * px.fx + pl.y * px.fy + pl.z;
        const float fl = floorf(df);
        bool in;
        if (GEN) in = a.D > 1 && !(fl < 0.f) && !(fl >= dlim);
        else in = fl == __builtin_amdgcn_fmed3f(fl, 0.f, dmax);
        float v0, v1;
        if (MODE == kCostF32) {
            v0 = __uint_as_float(wd.u0);
            v1 = __uint_as_float(wd.u1);
        } else {
            int c0, c1;
            if (MODE == kCostLeft) {
                c0 = __popc(px.clx ^ wd.u0);
                c1 = __popc(px.clx ^ wd.u1);
            } else {
                c0 = __popc(wd.u0 ^ wd.w0);
                c1 = __popc(wd.u1 ^ wd.w1);
            }
            v0 = (float)(c0 < 25 ? c0 : 25);
            v1 = (float)(c1 < 25 ? c1 : 25);
        }
        const float t1 = df - fl;
        float ci = (1.f - t1) * v0 + t1 * v1;
        pin(ci);   // a select, not a branch around the arithmetic
        cost = in ? ci : a.tau;
    }
    const float v = lam * cost;
    isn = v != v;
    return rintf(v);
}

// ----------------------------------------------------------------------------------------
// plane cost, coarse levels: one wavefront per segment (<= seg_len <= SEG pixels of one superpixel).
// Lane l reads pixels first + l + 64 r; the planes are wave-uniform.  A batch of B steps first
// issues the B packed positions, then every census word the B x P plane costs need, pins them
// with an empty asm so that no load is deferred into a branch, and only then computes: two
// memory round trips per batch.  The per-lane sums (exact doubles) are reduced across the
// wavefront and added to replica (segment mod rrep) of the superpixel's int64 sums.
// ----------------------------------------------------------------------------------------
// q: the packed positions of steps r0 .. r0+B-1 (loaded by the caller)
template <int MODE, int B, bool GEN>
__device__ __forceinline__ void pc_batch(const SweepArgs &a, const float4 (&pl)[P], int slot0, int first, int end,
                                         int r0, const uint32_t *q, double (&acc)[P], unsigned int &nanm) {
    const int lane = threadIdx.x & 63;
    const float dmax = (float)(a.D > 2 ? a.D - 2 : 0);
    const float dlim = (float)(a.D - 1);
    PcPix px[B];
    PcWords wd[B][P];
#pragma unroll
    for (int r = 0; r < B; ++r) {
        px[r] = pc_pixel<MODE>(a, q[r]);
#pragma unroll
        for (int k = 0; k < P; ++k) pc_load<MODE, GEN>(a, px[r], pl[k], dmax, wd[r][k]);
    }
#pragma unroll
    for (int r = 0; r < B; ++r) {
        pc_pin<MODE>(px[r]);
#pragma unroll
        for (int k = 0; k < P; ++k) pc_pin<MODE, GEN>(wd[r][k]);
    }
#pragma unroll
    for (int r = 0; r < B; ++r) {
        const bool valid = !GEN || first + (r0 + r) * 64 + lane < end;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            bool isn;
            float t = pc_term<MODE, GEN>(a, px[r], pl[k], wd[r][k], dmax, dlim, a.lam_fx, isn);
            if (GEN) {
                t = (valid && !isn) ? t : 0.f;
                nanm |= (valid && isn) ? (1u << (slot0 + k)) : 0u;
            }
            acc[k] += (double)t;
        }
    }
}

// Sums over the wavefront of integer valued doubles below 2^53 are exact in any order.
// dpp_f64: a DPP exchange of a double (both halves with the same pattern)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    return __hiloint2double(__builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true),
                            __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true));
}
// the P = 4 sums of a group of planes at once: lane l returns the sum over the wavefront of
// v[l mod 4].  Within a quad, lanes l and l^1 split the pairs {v0, v1} and {v2, v3}: each keeps
// the value its bit 0 names and adds the one its partner sends (a select, then one exchange per
// pair); lanes l and l^2 then split the two pair sums the same way, so that every lane holds its
// quad's sum of v[l mod 4].  Rotations by 4 and 8 within rows of 16 keep l mod 4, and two
// exchanges across rows finish.  A value only ever meets values of its own index.
static_assert(P == 4, "four sums per group");
__device__ __forceinline__ double wave_sum4_exact(const double (&v)[4]) {
    const int lane = threadIdx.x & 63;
    const bool b0 = (lane & 1) != 0, b1 = (lane & 2) != 0;
    const double x0 = (b0 ? v[1] : v[0]) + dpp_f64<0xB1>(b0 ? v[0] : v[1]);   // quad_perm [1,0,3,2]
    const double x1 = (b0 ? v[3] : v[2]) + dpp_f64<0xB1>(b0 ? v[2] : v[3]);
    double y = (b1 ? x1 : x0) + dpp_f64<0x4E>(b1 ? x0 : x1);                  // quad_perm [2,3,0,1]
    y += dpp_f64<0x124>(y);   // row_ror 4
    y += dpp_f64<0x128>(y);   // row_ror 8
    y += __shfl_xor(y, 16);
    y += __shfl_xor(y, 32);
    return y;
}
// OR over the wavefront, skipped (one ballot) when every lane holds 0
__device__ __forceinline__ unsigned int wave_or_rare(unsigned int m) {
    if (__builtin_amdgcn_ballot_w64(m != 0u) == 0ull) return 0u;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m |= __shfl_xor(m, o);
    return m;
}

// HB: steps per batch.  4 keeps 16 census gathers in flight per wavefront; 2 halves the registers
// the gathers hold, for more resident wavefronts (beside the plane costs computed ahead on the
// second stream, or when many pairs share a launch).
template <int NS, int MODE, int HB = 4>
__device__ void pc_segment(const SweepArgs &a, int w, int slot0, bool spatial) {
    static_assert(NS % P == 0, "planes are costed in groups of P");
    static_assert(HB == 4 || HB == 2, "batches of 4 or 2 steps");
    const int lane = threadIdx.x & 63;
    int4 sg = a.seg[w];
    pin(sg.x); pin(sg.y); pin(sg.z); pin(sg.w);   // one 16-byte load, not two around the test
    if (sg.y <= sg.x) return;   // past the last segment (the array is padded with empty ones)
    const int c = __builtin_amdgcn_readfirstlane(sg.z);
    const int first = __builtin_amdgcn_readfirstlane(sg.x), end = __builtin_amdgcn_readfirstlane(sg.y);
    const int steps = (end - first + 63) >> 6;
    const int full = (end - first) >> 6;
    // the positions of all SEG_R steps (lanes past the end re-read the last pixel), then the
    // first group's planes: one round trip for both, in straight-line code
    uint32_t q[SEG_R];
#pragma unroll
    for (int r = 0; r < SEG_R; ++r) {
        const int j = first + r * 64 + lane;
        q[r] = ld_at<uint32_t>(a.pxy, (uint32_t)(j < end ? j : end - 1) << 2);
    }
    float4 pl[P];
    if (!spatial) {
#pragma unroll
        for (int k = 0; k < P; ++k) pl[k] = a.plane[(long)c * K + slot0 + k];
    }
    unsigned int nanm = 0u;
    unsigned long long mine = 0ull;
#pragma unroll 1
    for (int g0 = 0; g0 < NS; g0 += P) {
        if (spatial) {
            float4 cand = make_float4(0.f, 0.f, 0.f, 0.f);
            if (lane < P) cand = a.spc ? a.plane[(long)a.spc[(long)c * P + lane] * K] : spatial_candidate(a, c, lane);
            if (sg.w == 0 && lane < P) a.plane[(long)c * K + 2 * P + lane] = cand;
#pragma unroll
            for (int k = 0; k < P; ++k)
                pl[k] = make_float4(__shfl(cand.x, k), __shfl(cand.y, k), __shfl(cand.z, k), 0.f);
        } else if (g0 > 0) {
#pragma unroll
            for (int k = 0; k < P; ++k) pl[k] = a.plane[(long)c * K + slot0 + g0 + k];
        }
        bool tame = a.D >= 2;
#pragma unroll
        for (int k = 0; k < P; ++k) tame = tame && tame_plane(pl[k]);
        const bool fast = __builtin_amdgcn_readfirstlane((int)tame) != 0;
        double acc[P];
#pragma unroll
        for (int k = 0; k < P; ++k) acc[k] = 0.0;
        // SEG_R / HB batches of HB steps: a full batch at once (fast form when the planes
        // allow), a partial one in batches of 2 and 1 in the general form, so that no step past
        // the segment's end is computed
        static_assert(SEG_R % HB == 0, "whole batches");
#pragma unroll
        for (int h = 0; h < SEG_R / HB; ++h) {
            const int r0 = h * HB;
            const int rem = steps - r0;
            if (rem <= 0) break;
            const int sl = slot0 + g0;
            if (rem >= HB) {
                if (fast && r0 + HB <= full) pc_batch<MODE, HB, false>(a, pl, sl, first, end, r0, q + r0, acc, nanm);
                else pc_batch<MODE, HB, true>(a, pl, sl, first, end, r0, q + r0, acc, nanm);
            } else if (rem == 1) {
                pc_batch<MODE, 1, true>(a, pl, sl, first, end, r0, q + r0, acc, nanm);
            } else {
                pc_batch<MODE, 2, true>(a, pl, sl, first, end, r0, q + r0, acc, nanm);
                if (rem == 3) pc_batch<MODE, 1, true>(a, pl, sl, first, end, r0 + 2, q + r0 + 2, acc, nanm);
            }
        }
        // lane g0 + k takes slot slot0 + g0 + k (k = lane mod P); the terms carry the costs'
        // signs and |lambda|, the sum takes lambda's sign
        const long long m = (long long)wave_sum4_exact(acc);
        if ((lane & ~(P - 1)) == g0) mine = (unsigned long long)(a.lambda < 0.f ? -m : m);
    }
    nanm = wave_or_rare(nanm);
    const int rep = w % a.rrep;
    if (lane < NS && mine) atomicAdd(&a.acc[((long)c * a.rrep + rep) * K + slot0 + lane], mine);
    if (lane == 0 && nanm) atomicOr(&a.accnan[c], nanm);
}

// ----------------------------------------------------------------------------------------
// plane cost, fine levels: one 256-thread block per chunk of CHUNK consecutive sorted pixels,
// lane l taking pixels base + 4l .. 4l+3 (16-byte loads of positions and superpixel ids).
// The candidate planes of the chunk's superpixels are staged in LDS; a lane issues the census
// words of its 4 pixels x P planes before using any, sums its terms per superpixel run and
// adds each run into LDS int64 counters indexed by superpixel - first superpixel of the chunk;
// the block then adds each counter to replica (chunk mod rrep) of the global sum.  NS = 12
// runs three groups of P planes.
// ----------------------------------------------------------------------------------------
template <int MODE, bool GEN>
__device__ __forceinline__ void pc_run_group(const SweepArgs &a, const uint32_t (&q)[4], const uint32_t (&cl)[4],
                                             const bool (&valid)[4], int c_first, int slot, const float4 *splane,
                                             unsigned long long *ssum, unsigned int *snan) {
    const float dmax = (float)(a.D > 2 ? a.D - 2 : 0);
    const float dlim = (float)(a.D - 1);
    const bool neg = a.lambda < 0.f;
    PcPix px[4];
    PcWords wd[4][P];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        px[i] = pc_pixel<MODE>(a, q[i]);
        const int lc = (int)cl[i] - c_first;
#pragma unroll
        for (int k = 0; k < P; ++k) pc_load<MODE, GEN>(a, px[i], splane[lc * P + k], dmax, wd[i][k]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        pc_pin<MODE>(px[i]);
#pragma unroll
        for (int k = 0; k < P; ++k) pc_pin<MODE, GEN>(wd[i][k]);
    }
    // per-run sums (|terms|; lambda's sign applied at the flush) and NaN bits
    double acc[P];
#pragma unroll
    for (int k = 0; k < P; ++k) acc[k] = 0.0;
    unsigned int nanm = 0u;
    int crun = (int)cl[0];
    auto flush = [&](int run) {
        const int lr = run - c_first;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            if (acc[k] != 0.0) {
                const long long m = (long long)acc[k];
                atomicAdd(&ssum[lr * P + k], (unsigned long long)(neg ? -m : m));
            }
            acc[k] = 0.0;
        }
        if (nanm) atomicOr(&snan[lr], nanm);
        nanm = 0u;
    };
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (valid[i] && (int)cl[i] != crun) {
            flush(crun);
            crun = (int)cl[i];
        }
        const int lc = (int)cl[i] - c_first;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            bool isn;
            float t = pc_term<MODE, GEN>(a, px[i], splane[lc * P + k], wd[i][k], dmax, dlim, a.lam_fx, isn);
            if (GEN) {
                t = (valid[i] && !isn) ? t : 0.f;
                nanm |= (valid[i] && isn) ? (1u << (slot + k)) : 0u;
            }
            acc[k] += (double)t;
        }
    }
    flush(crun);
}

template <int NS, int MODE>
__device__ void pc_runs(const SweepArgs &a, long chunk, int slot0, unsigned char *lds) {
    float4 *splane = (float4 *)lds;                                       // [MAXC][P]
    unsigned long long *ssum = (unsigned long long *)(lds + (size_t)MAXC * P * 16);   // [MAXC][P]
    unsigned int *snan = (unsigned int *)(lds + (size_t)MAXC * P * 24);   // [MAXC]
    const long base = chunk * CHUNK;
    const long last = (base + CHUNK < a.N ? base + CHUNK : a.N) - 1;
    const bool whole = base + CHUNK <= a.N;
    const int c_first = (int)a.pcl[base];
    const int nclus = (int)a.pcl[last] - c_first + 1;
    const long j0 = base + 4L * threadIdx.x;
    // this lane's pixels
    uint32_t q[4], cl[4];
    bool valid[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        valid[i] = j0 + i <= last;
        const long j = valid[i] ? j0 + i : last;
        q[i] = a.pxy[j];
        cl[i] = a.pcl[j];
    }
    for (int g0 = 0; g0 < NS; g0 += P) {
        int wild = 0;
        for (int i = threadIdx.x; i < nclus * P; i += 256) {
            const float4 p = a.plane[(long)(c_first + i / P) * K + slot0 + g0 + i % P];
            splane[i] = p;
            ssum[i] = 0ull;
            wild |= tame_plane(p) ? 0 : 1;
        }
        for (int i = threadIdx.x; i < nclus; i += 256) snan[i] = 0u;
        const bool fast = __syncthreads_or(wild) == 0 && whole && a.D >= 2;
        if (fast) pc_run_group<MODE, false>(a, q, cl, valid, c_first, slot0 + g0, splane, ssum, snan);
        else pc_run_group<MODE, true>(a, q, cl, valid, c_first, slot0 + g0, splane, ssum, snan);
        __syncthreads();
        const int rep = (int)(chunk % a.rrep);
        for (int i = threadIdx.x; i < nclus * P; i += 256) {
            unsigned long long sm = ssum[i];
            if (sm) atomicAdd(&a.acc[((long)(c_first + i / P) * a.rrep + rep) * K + slot0 + g0 + i % P], sm);
        }
        for (int i = threadIdx.x; i < nclus; i += 256)
            if (snan[i]) atomicOr(&a.accnan[c_first + i], snan[i]);
        __syncthreads();
    }
}

// ----------------------------------------------------------------------------------------
// messages for 5 edges (one wavefront, 12 lanes each): passing_msg_k (hbp_isp.cu:292-304),
// smooth_msg_k (306-355), normalize_msg_k (273-290).  Lane y of edge e = (s, t) computes the
// message into s for s's slot y from t's 8 best + prior particles.  On spatial sweeps whose
// candidates are resolved in this kernel, s's candidate planes are resolved here too.
// ----------------------------------------------------------------------------------------
constexpr int EDGES_PER_WAVE = 5;
__device__ void edge_wave(const SweepArgs &a, long wv, bool spatial) {
    const int lane = threadIdx.x & 63;
    const int grp = lane / K;
    const long e = wv * EDGES_PER_WAVE + grp;
    const int y = lane - grp * K;
    const bool live = grp < EDGES_PER_WAVE && e < a.E;
    float out = 0.f;
    long mslot = 0;
    if (live) {
        int2 b = a.bd[e];
        int s = b.x, t = b.y;
        mslot = (long)s * a.maxdeg + a.eloc[e];
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
        float4 psl = (spatial && y >= 2 * P)
                         ? (a.spc ? a.plane[(long)a.spc[(long)s * P + y - 2 * P] * K] : spatial_candidate(a, s, y - 2 * P))
                         : a.plane[(long)s * K + y];
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
            phi = fminf(div3_nonneg(phi) / dpos, a.tau_s);
            float v = m + c0 * phi;
            if (v < minimum) minimum = v;
        }
        out += minimum;
    }
    // normalize_msg_k: mean over the 12 slots, summed in slot order
    float mean = 0.f;
    const int g = grp * K;
#pragma unroll
    for (int k = 0; k < K; ++k) mean += __shfl(out, g + k, 64);
    mean /= K;
    if (live) a.msg_t[mslot * K + y] = out - mean;
}

// ----------------------------------------------------------------------------------------
// node update with 16 lanes per superpixel s, lane k < 12 owning slot k: aggregate_belief
// (hbp_isp.cu:1048-1058) in edge order, ping_pang_msg (1059-1066), get_belief_order
// (432-450), select_belief (452-472), select_msg (474-493); then the parents' seed advance of
// a spatial sweep and the next sweep's candidates.  Every location a lane stores was read by
// that same lane or reaches it through a cross-lane shuffle, so no lane overwrites data
// another lane has still to read.
// ----------------------------------------------------------------------------------------
// lane J of each 16-lane row, to every lane of the row (DPP row_newbcast)
template <int J>
__device__ __forceinline__ int row_bcast(int v) {
    return __builtin_amdgcn_mov_dpp(v, 0x150 + J, 0xF, 0xF, false);
}
template <int J>
__device__ __forceinline__ float row_bcast(float v) {
    return __int_as_float(row_bcast<J>(__float_as_int(v)));
}
// order[k] = #{j < k: !(bel[k] < bel[j])} + #{j > k: bel[j] < bel[k]} over the K slots
template <int... J>
__device__ __forceinline__ int belief_rank(float bel, int k, std::integer_sequence<int, J...>) {
    int order = 0;
    ((order += ((J < k) & !(bel < row_bcast<J>(bel))) | ((J > k) & (row_bcast<J>(bel) < bel))), ...);
    return order;
}
// the incoming slots i < MAXDEG: their reverse (node, slot) index, from lane i of the row
template <int... I>
__device__ __forceinline__ void row_revl(int rv, int (&r)[sizeof...(I)], std::integer_sequence<int, I...>) {
    ((r[I] = row_bcast<I>(rv)), ...);
}

// Inputs of a node update that are not in place: the state a run of fused sweeps leaves (see
// chain_kernel).  Planes of slots < 2P and their costs from pl / co, the candidates (slots
// >= 2P) from cand, the RANSAC points of those candidates from pt; the update writes its
// outputs to the SweepArgs state and also clears a second set of accumulators.
struct NodeIO {
    const float4 *pl;
    const float *co;
    const float4 *cand;
    const uint4 *pt;
    unsigned long long *acc2;   // or null
    unsigned int *nan2;
    int cs, as;                 // slot strides of cand and of SweepArgs::acc (K or P)
};

template <bool IO>
__device__ void node_group(const SweepArgs &a, long s, int this_kind, int next_kind, float max_dz, float max_dn,
                           const NodeIO &io) {
    constexpr int MAXDEG = kNodeMaxDeg;   // incoming messages held in registers (the rest are re-read)
    const int lane = threadIdx.x & 63;
    const int k = lane & 15;
    const int g = lane & ~15;
    const bool own = k < K;
    const int kk = own ? k : 0;
    // ---- everything that depends on s alone, in one round trip
    const int deg = a.off[s + 1] - a.off[s];
    const int ppos = a.pix_off[s], pn = a.pix_off[s + 1] - ppos;
    const unsigned int nb = a.accnan[s];
    const uint32_t s0 = a.s_ransac[s];
    const float cost_in = IO ? io.co[s * K + kk] : a.cost[s * K + kk];
    const int ast = IO ? io.as : K;
    const int64_t fsum = acc_sum_at(a.acc, a.rrep, s, kk >= 2 * P ? kk : 2 * P, ast);
    const float4 pin = !IO ? a.plane[s * K + kk]
                           : (kk < 2 * P ? io.pl[s * K + kk] : io.cand[s * io.cs + kk - (K - io.cs)]);
    const uint4 pt_in = IO ? io.pt[s] : make_uint4(0u, 0u, 0u, 0u);
    const float2 mxy = a.mxy[s];
    const long nm = (long)s * a.maxdeg;
    // rows nm .. nm+15, past deg into the next node's rows or the arrays' kNodeMaxDeg padding
    // rows (never used), at constant offsets from one address
    const int rv = a.revl[nm + k];
    const float *__restrict__ mtp = a.msg_t + nm * K + kk;
    float mt[MAXDEG];
#pragma unroll
    for (int i = 0; i < MAXDEG; ++i) mt[i] = mtp[i * K];
    const bool seeds = this_kind == kSpatial && s < a.Vh && k == 0;
    const long sh = s < a.Vh ? s : a.Vh - 1;
    uint32_t sc = a.s_child[sh], sn = a.s_neighbor[sh];
    // the next sweep's RANSAC samples depend on the seed alone: lane p in [2P, 3P) draws
    // candidate p from the seed after the draws of candidates 2P..p-1 and issues its loads now
    const bool cand = own && k >= 2 * P;
    uint32_t sk = s0;
    uint4 q = make_uint4(0u, 0u, 0u, 0u);
    float z0 = 0.f, z1 = 0.f, z2 = 0.f;
    if (next_kind == kRansac) {
        sk = lcg_skip<3>(s0, cand ? k - 2 * P : 0);
        int p0, p1, p2;
        ransac_draw(sk, pn, p0, p1, p2);
        q = make_uint4(a.pxy[ppos + p0], a.pxy[ppos + p1], a.pxy[ppos + p2], 0u);
        z0 = a.pw[ppos + p0];
        z1 = a.pw[ppos + p1];
        z2 = a.pw[ppos + p2];
    }
    // ---- sums and the selection
    if (k == 2 * P) a.accnan[s] = 0u;
    if (IO && io.nan2 && k == 2 * P) io.nan2[s] = 0u;
    float cost = cost_in;
    if (kk >= 2 * P) {
        cost = from_fixed(fsum, (nb >> kk) & 1u, a.shift);
        acc_zero_at(a.acc, a.rrep, s, kk, ast);
        if (IO && io.acc2) acc_zero_at(io.acc2, a.rrep, s, kk);
    }
    // belief = cost + incoming messages in edge order
    float bel = cost;
#pragma unroll
    for (int i = 0; i < MAXDEG; ++i) bel = i < deg ? bel + mt[i] : bel;
    for (int i = MAXDEG; i < deg; ++i) bel += a.msg_t[(nm + i) * K + kk];
    const int order = belief_rank(bel, k, std::make_integer_sequence<int, K>{});
    const bool sel = own && order < P;
    const uint32_t m = (uint32_t)(__ballot(sel) >> g) & 0xfffu;
    const int nsel = __popc(m);
    // output slot k takes the k-th selected slot (k < nsel) or keeps its own value
    int src = k;
    if (k < nsel) {
        uint32_t mm = m;
        for (int i = 0; i < k; ++i) mm &= mm - 1u;
        src = __ffs(mm) - 1;
    }
    const float cost_n = __shfl(cost, g + src), bel_n = __shfl(bel, g + src);
    const float4 pl_n = make_float4(__shfl(pin.x, g + src), __shfl(pin.y, g + src), __shfl(pin.z, g + src),
                                    __shfl(pin.w, g + src));
    // messages over s's particles -> reverse edge, compacted by s's order
    float o[MAXDEG];
    int r[MAXDEG];
#pragma unroll
    for (int i = 0; i < MAXDEG; ++i) o[i] = __shfl(mt[i], g + src);
    row_revl(rv, r, std::make_integer_sequence<int, MAXDEG>{});
    if (own) {
        a.plane[s * K + k] = pl_n;
        a.cost[s * K + k] = cost_n;
        a.belief[s * K + k] = bel_n;
    }
    // the points of the candidates just costed, unless a new RANSAC draw replaces them below
    if (IO && k == K - 1 && next_kind != kRansac) a.point[s] = pt_in;
    float *msg = (float *)a.msg;
    if (k < 2 * P) {
#pragma unroll
        for (int i = 0; i < MAXDEG; ++i)
            if (i < deg) msg[(long)r[i] * 8 + k] = o[i];
    }
    for (int i = MAXDEG; i < deg; ++i) {
        const float oi = __shfl(a.msg_t[(nm + i) * K + kk], g + src);
        if (k < 2 * P) msg[(long)a.revl[nm + i] * 8 + k] = oi;
    }
    if (seeds) {
        constexpr LcgJump jp = lcg_jump_coef(P);
        sc = (jp.a * sc + jp.c) % 2147483648u;
        sn = (jp.a * sn + jp.c) % 2147483648u;
        a.s_child[s] = sc;
        a.s_neighbor[s] = sn;
    }
    // next candidates; lane 2P also stores the seed after all of them
    if (next_kind == kRansac || next_kind == kRandom) {
        if (cand) {
            if (next_kind == kRansac) {
                const float4 A = make_float4((float)unpack_x(q.x), (float)unpack_y(q.x), z0, 0.f);
                const float4 B = make_float4((float)unpack_x(q.y), (float)unpack_y(q.y), z1, 0.f);
                const float4 C = make_float4((float)unpack_x(q.z), (float)unpack_y(q.z), z2, 0.f);
                a.plane[s * K + k] = fit_plane(A, B, C);
                if (k == K - 1) a.point[s] = q;
            } else {
                uint32_t sr = lcg_skip<4>(s0, k - 2 * P);
                a.plane[s * K + k] = random_plane(sr, pl_n, mxy.x, mxy.y, max_dz, max_dn);
            }
            if (k == 2 * P) {
                constexpr LcgJump j3 = lcg_jump_coef(3 * P), j4 = lcg_jump_coef(4 * P);
                const LcgJump jf = next_kind == kRansac ? j3 : j4;
                a.s_ransac[s] = (jf.a * s0 + jf.c) % 2147483648u;
            }
        }
    }
}

// ----------------------------------------------------------------------------------------
// Fused sweeps (coarse levels, one pair).  A RANSAC sweep's candidates depend on the seeds
// alone, not on the previous sweep's selection, so a run of sweeps j0, j0+1, .., jL (j0 any
// non-spatial sweep, the others RANSAC sweeps of the same level) runs as one launch per sweep
// instead of two.  Launch u (sweep j = j0+u) holds three kinds of wavefronts:
//   segments   the plane costs of sweep j's candidates (drawn one launch earlier);
//   edges      sweep j's messages.  For u >= 1 the selection that ends sweep j-1 (belief,
//              rank, compaction: node_group's arithmetic) is recomputed inside the wavefront
//              for both ends of each edge from that selection's inputs, which launch u-1 left;
//   nodes      the same selection of every superpixel, written out for launch u+1, the seed
//              advance, and the draw of sweep j+1's RANSAC candidates.
// No wavefront reads what another writes in the same launch: the selection's inputs and
// outputs, the candidates and the sums rotate through separate buffers (ChainIO).  After
// launch L a node update (node_group<true>) writes the standard state, so the rest of the
// schedule (spatial and random sweeps, the map, the next level) reads it unchanged.  Every
// value is computed by the same arithmetic as in the split sweeps, so the bits are the same.
// ----------------------------------------------------------------------------------------
struct ChainIO {
    // inputs of the selection that ends sweep j-1
    const float4 *pl_in;                 // [V][K] slots < 2P: planes after the selection of sweep j-2
    const float *co_in;                  // [V][K] slots < 2P: their costs
    const float4 *cand_in;               // slots >= 2P: the candidates of sweep j-1 (slot stride cs_in)
    const unsigned long long *acc_in;    // their fixed-point sums [V][rrep][as_in]
    const unsigned int *nan_in;          // their NaN bits [V]
    const float *mt_in;                  // sweep j-1's messages, node-major [V][maxdeg][K]
    const int *rloc;                     // per edge: the reverse edge's slot among its source's edges [E]
    // sweep j (its candidates are SweepArgs::plane slots >= 2P)
    const float4 *cand_cur;              // slots >= 2P (slot stride cs_cur)
    const uint4 *pt_cur;                 // their RANSAC points [V]
    int cs_in, as_in, cs_cur;            // slot strides (K or P)
    // node wavefronts
    float4 *pl_out;                      // [V][K] slots < 2P: the selection that ends sweep j-1
    float *co_out;
    float4 *cand_next;                   // [V][K] slots >= 2P: sweep j+1's RANSAC candidates
    uint4 *pt_next;
    unsigned long long *acc_clear;       // sums to clear for a later launch, or null
    unsigned int *nan_clear;
    int materialize;                     // u >= 1: the selection of sweep j-1 is computed here
    int draw;                            // sweep j+1 is a RANSAC sweep of the run
    int prev_spatial;                    // sweep j-1 was a spatial sweep (its update advances the parents' seeds)
};

// node_group's selection of superpixel n from a chain's inputs; lane k of a 16-lane row owns
// slot k.  Loads first (one round trip), then the arithmetic.
struct SelIn {
    int deg;
    unsigned int nb;
    float cost_in;
    int64_t fsum;
    float4 pin;
    float mt[kNodeMaxDeg];
};
struct SelOut {
    float bel, cost;
    float4 pl;
    int src;
};
__device__ __forceinline__ void sel_load(const SweepArgs &a, const ChainIO &io, long n, int kk, SelIn &in) {
    in.deg = a.off[n + 1] - a.off[n];
    in.nb = io.nan_in[n];
    in.cost_in = io.co_in[n * K + kk];
    in.fsum = acc_sum_at(io.acc_in, a.rrep, n, kk >= 2 * P ? kk : 2 * P, io.as_in);
    in.pin = kk < 2 * P ? io.pl_in[n * K + kk] : io.cand_in[n * io.cs_in + kk - (K - io.cs_in)];
    const float *__restrict__ mtp = io.mt_in + n * a.maxdeg * K + kk;
#pragma unroll
    for (int i = 0; i < kNodeMaxDeg; ++i) in.mt[i] = mtp[i * K];
}
// the rank and compaction of a selection from each slot's belief: src = the slot whose values
// output slot k takes
__device__ __forceinline__ SelOut sel_rank(float bel, float cost, float4 pin, int k) {
    const int g = (threadIdx.x & 63) & ~15;
    const bool own = k < K;
    const int order = belief_rank(bel, k, std::make_integer_sequence<int, K>{});
    const bool sel = own && order < P;
    const uint32_t m = (uint32_t)(__ballot(sel) >> g) & 0xfffu;
    const int nsel = __popc(m);
    int src = k;
    if (k < nsel) {
        uint32_t mm = m;
        for (int i = 0; i < k; ++i) mm &= mm - 1u;
        src = __ffs(mm) - 1;
    }
    return SelOut{bel, cost, pin, src};
}
__device__ __forceinline__ SelOut sel_compute(const SweepArgs &a, const ChainIO &io, long n, const SelIn &in, int k,
                                              int kk) {
    const float cost = kk >= 2 * P ? from_fixed(in.fsum, (in.nb >> kk) & 1u, a.shift) : in.cost_in;
    float bel = cost;
#pragma unroll
    for (int i = 0; i < kNodeMaxDeg; ++i) bel = i < in.deg ? bel + in.mt[i] : bel;
    for (int i = kNodeMaxDeg; i < in.deg; ++i) bel += io.mt_in[(n * a.maxdeg + i) * K + kk];
    return sel_rank(bel, cost, in.pin, k);
}
// one step p of the message minimum over t's 2P particles (edge_wave's arithmetic), the
// particle's values broadcast from lane p of the row
template <int p>
__device__ __forceinline__ void fused_msg_step(float tm, float tx, float ty, float tz, float4 psl, const float (&X)[6],
                                               const float (&Y)[6], float c0, float dpos, float tau_s, float &out,
                                               float &minimum) {
    if (p == P) {
        out += minimum;
        minimum = FLT_MAX;
    }
    const float m = row_bcast<p>(tm);
    const float dpx = psl.x - row_bcast<p>(tx);
    const float dpy = psl.y - row_bcast<p>(ty);
    const float dpz = psl.z - row_bcast<p>(tz);
    float phi = fabsf(dpx * X[0] + dpy * Y[0] + dpz);
#pragma unroll
    for (int j = 1; j < 6; ++j) phi = phi + fabsf(dpx * X[j] + dpy * Y[j] + dpz);
    phi = fminf(div3_nonneg(phi) / dpos, tau_s);
    const float v = m + c0 * phi;
    if (v < minimum) minimum = v;
}
template <int... p>
__device__ __forceinline__ void fused_msg(float tm, float tx, float ty, float tz, float4 psl, const float (&X)[6],
                                          const float (&Y)[6], float c0, float dpos, float tau_s, float &out,
                                          float &minimum, std::integer_sequence<int, p...>) {
    (fused_msg_step<p>(tm, tx, ty, tz, psl, X, Y, c0, dpos, tau_s, out, minimum), ...);
}
template <int... J>
__device__ __forceinline__ float row_sum(float v, std::integer_sequence<int, J...>) {
    float s = 0.f;
    ((s += row_bcast<J>(v)), ...);
    return s;
}

// sweep j's messages of FUSED_EDGES edges per wavefront, 16 lanes per edge e = (s, t): lane y
// computes the message into s for s's slot y (edge_wave), with the selections of s and t that
// end sweep j-1 recomputed in the row
constexpr int FUSED_EDGES = 4;
__device__ void edge_wave_fused(const SweepArgs &a, const ChainIO &io, long wv) {
    const int lane = threadIdx.x & 63;
    const int k = lane & 15, g = lane & ~15;
    const int kk = k < K ? k : 0;
    const long e = wv * FUSED_EDGES + (lane >> 4);
    const bool live = e < a.E;
    const long ec = live ? e : a.E - 1;
    const int2 b = a.bd[ec];
    const int eslot = a.eloc[ec], rslot = io.rloc[ec];
    const float c0 = a.c0[ec], dpos = a.dpos[ec];
    SelIn is, it;
    sel_load(a, io, b.x, kk, is);
    sel_load(a, io, b.y, kk, it);
    // t's message from s (on t's slot rslot), the points, s's candidates of this sweep
    const float mrow = io.mt_in[((long)b.y * a.maxdeg + rslot) * K + kk];
    const uint4 ps = io.pt_cur[b.x], pt = io.pt_cur[b.y];
    const float4 cc = io.cand_cur[(long)b.x * io.cs_cur + (kk >= 2 * P ? kk : 2 * P) - (K - io.cs_cur)];
    const SelOut ss = sel_compute(a, io, b.x, is, k, kk);
    const SelOut st = sel_compute(a, io, b.y, it, k, kk);
    // s's planes after its selection (slots < 2P) and this sweep's candidates
    float4 psl = make_float4(__shfl(ss.pl.x, g + ss.src), __shfl(ss.pl.y, g + ss.src), __shfl(ss.pl.z, g + ss.src), 0.f);
    if (k >= 2 * P) psl = cc;
    // lanes p < 2P: t's particle p after its selection, its belief less the message from s
    const float tm = __shfl(st.bel, g + st.src) - __shfl(mrow, g + st.src);
    const float tx = __shfl(st.pl.x, g + st.src), ty = __shfl(st.pl.y, g + st.src), tz = __shfl(st.pl.z, g + st.src);
    float X[6], Y[6];
    X[0] = (float)unpack_x(ps.x); Y[0] = (float)unpack_y(ps.x);
    X[1] = (float)unpack_x(ps.y); Y[1] = (float)unpack_y(ps.y);
    X[2] = (float)unpack_x(ps.z); Y[2] = (float)unpack_y(ps.z);
    X[3] = (float)unpack_x(pt.x); Y[3] = (float)unpack_y(pt.x);
    X[4] = (float)unpack_x(pt.y); Y[4] = (float)unpack_y(pt.y);
    X[5] = (float)unpack_x(pt.z); Y[5] = (float)unpack_y(pt.z);
    float out = 0.f, minimum = FLT_MAX;
    fused_msg(tm, tx, ty, tz, psl, X, Y, c0, dpos, a.tau_s, out, minimum, std::make_integer_sequence<int, 2 * P>{});
    out += minimum;
    // normalize_msg_k: mean over the 12 slots, summed in slot order
    float mean = row_sum(out, std::make_integer_sequence<int, K>{});
    mean /= K;
    if (live && k < K) a.msg_t[((long)b.x * a.maxdeg + eslot) * K + k] = out - mean;
}

// node wavefronts of a fused launch, 16 lanes per superpixel s: the selection that ends sweep
// j-1 (materialize), the seed advance it makes (its next sweep, j, is a RANSAC sweep), and the
// draw of sweep j+1's RANSAC candidates from the advanced seed (node_group's draw)
__device__ void chain_node(const SweepArgs &a, const ChainIO &io, long s) {
    const int lane = threadIdx.x & 63;
    const int k = lane & 15, g = lane & ~15;
    const int kk = k < K ? k : 0;
    const bool cand = k >= 2 * P && k < K;
    const int ppos = a.pix_off[s], pn = a.pix_off[s + 1] - ppos;
    uint32_t s0 = a.s_ransac[s];
    if (io.materialize) {
        SelIn in;
        sel_load(a, io, s, kk, in);
        const SelOut so = sel_compute(a, io, s, in, k, kk);
        const float4 pl_n = make_float4(__shfl(so.pl.x, g + so.src), __shfl(so.pl.y, g + so.src),
                                        __shfl(so.pl.z, g + so.src), __shfl(so.pl.w, g + so.src));
        const float co_n = __shfl(so.cost, g + so.src);
        if (k < 2 * P) {
            io.pl_out[s * K + k] = pl_n;
            io.co_out[s * K + k] = co_n;
        }
        constexpr LcgJump j3 = lcg_jump_coef(3 * P);
        s0 = (j3.a * s0 + j3.c) % 2147483648u;
        if (k == 2 * P) a.s_ransac[s] = s0;
        // after a spatial sweep: the seeds of the parent with this index (node_group's rule)
        if (io.prev_spatial && k == 0 && s < a.Vh) {
            constexpr LcgJump jp = lcg_jump_coef(P);
            a.s_child[s] = (jp.a * a.s_child[s] + jp.c) % 2147483648u;
            a.s_neighbor[s] = (jp.a * a.s_neighbor[s] + jp.c) % 2147483648u;
        }
    }
    if (io.acc_clear && cand) {
        acc_zero_at(io.acc_clear, a.rrep, s, k);
        if (k == 2 * P) io.nan_clear[s] = 0u;
    }
    if (io.draw && cand) {
        uint32_t sk = lcg_skip<3>(s0, k - 2 * P);
        int p0, p1, p2;
        ransac_draw(sk, pn, p0, p1, p2);
        const uint4 q = make_uint4(a.pxy[ppos + p0], a.pxy[ppos + p1], a.pxy[ppos + p2], 0u);
        const float4 A = make_float4((float)unpack_x(q.x), (float)unpack_y(q.x), a.pw[ppos + p0], 0.f);
        const float4 B = make_float4((float)unpack_x(q.y), (float)unpack_y(q.y), a.pw[ppos + p1], 0.f);
        const float4 C = make_float4((float)unpack_x(q.z), (float)unpack_y(q.z), a.pw[ppos + p2], 0.f);
        io.cand_next[s * K + k] = fit_plane(A, B, C);
        if (k == K - 1) io.pt_next[s] = q;
    }
}

// ----------------------------------------------------------------------------------------
// RANSAC plane costs ahead of the BP chain (tuning fuse = 2).  A level's seeds advance by a
// fixed number of LCG draws per node update (3P before a RANSAC sweep, 4P before a random one,
// none before a spatial one), so the seed from which each RANSAC sweep's candidates are drawn
// is an affine map of the seed the level init leaves, and every RANSAC sweep's candidates,
// points and plane costs can be computed before the sweeps run: the candidates in one launch
// at the level's start, the costs by persistent launches on a second stream that run beside
// the chain of message and node launches.  Sweep r of the level's RANSAC list (its first sweep
// excluded, whose candidates the level init draws and costs) keeps its candidates [V][P], its
// points [V], its sums [V][rrep][P] and NaN words [V] in slice r.
// ----------------------------------------------------------------------------------------
struct PipeArgs {
    float4 *cand;                 // [nR][V][P]
    uint4 *pt;                    // [nR][V]
    unsigned long long *acc;      // [nR][V][rrep][P]
    unsigned int *nan;            // [nR][V]
    int rd_first, rd_count;       // the level's random sweeps: first iteration, count
    long cslice, aslice;          // V P, V rrep P
};

// LCG steps from the seed the level init leaves to the seed listed sweep r's candidates are
// drawn from: 3P per RANSAC sweep and 4P per random sweep before it (after the level's first
// sweep).  A level's iterations hold 4 RANSAC sweeps each (listed sweep r is the level's
// (r+1)-th, in iteration (r+1)/4) and its random sweeps end iterations rd_first ..
// rd_first+rd_count-1.
__host__ __device__ inline long ransac_draw_steps(int r, int rd_first, int rd_count) {
    const int i = (r + 1) / 4;
    const int nd = i - rd_first < 0 ? 0 : (i - rd_first > rd_count ? rd_count : i - rd_first);
    return 3L * P * r + 4L * P * nd;
}


// candidates p of every RANSAC sweep r of the list for node n (node_group's draw)
__device__ void draw_ahead(const SweepArgs &a, const PipeArgs &pa, long i) {
    const int p = (int)(i % P);
    const long rn = i / P;
    const int n = (int)(rn % a.V), r = (int)(rn / a.V);
    const int ppos = a.pix_off[n], pn = a.pix_off[n + 1] - ppos;
    const LcgJump m = lcg_jump_pow((unsigned long)ransac_draw_steps(r, pa.rd_first, pa.rd_count));
    const uint32_t s0 = (m.a * a.s_ransac[n] + m.c) % 2147483648u;
    uint32_t sk = lcg_skip<3>(s0, p);
    int p0, p1, p2;
    ransac_draw(sk, pn, p0, p1, p2);
    const uint4 q = make_uint4(a.pxy[ppos + p0], a.pxy[ppos + p1], a.pxy[ppos + p2], 0u);
    const float4 A = make_float4((float)unpack_x(q.x), (float)unpack_y(q.x), a.pw[ppos + p0], 0.f);
    const float4 B = make_float4((float)unpack_x(q.y), (float)unpack_y(q.y), a.pw[ppos + p1], 0.f);
    const float4 C = make_float4((float)unpack_x(q.z), (float)unpack_y(q.z), a.pw[ppos + p2], 0.f);
    pa.cand[r * pa.cslice + (long)n * P + p] = fit_plane(A, B, C);
    if (p == P - 1) pa.pt[(long)r * a.V + n] = q;
}

// plane costs of the candidates of sweeps r0 .. r0+nr-1 (nr <= 64 / P) for one segment: the
// segment's positions are read once, then each sweep's P planes as in pc_segment (the same
// terms and sums); lane P g + k ends with sweep r0+g's sum of slot k, lane g with its NaN bits
template <int MODE, int HB>
__device__ void pc_segment_multi(const SweepArgs &a, const PipeArgs &pa, int w, int r0, int nr) {
    const int lane = threadIdx.x & 63;
    int4 sg = a.seg[w];
    pin(sg.x); pin(sg.y); pin(sg.z); pin(sg.w);
    if (sg.y <= sg.x) return;
    const int c = __builtin_amdgcn_readfirstlane(sg.z);
    const int first = __builtin_amdgcn_readfirstlane(sg.x), end = __builtin_amdgcn_readfirstlane(sg.y);
    const int steps = (end - first + 63) >> 6;
    const int full = (end - first) >> 6;
    uint32_t q[SEG_R];
#pragma unroll
    for (int r = 0; r < SEG_R; ++r) {
        const int j = first + r * 64 + lane;
        q[r] = ld_at<uint32_t>(a.pxy, (uint32_t)(j < end ? j : end - 1) << 2);
    }
    const float4 *__restrict__ cp = pa.cand + r0 * pa.cslice + (long)c * P;
    float4 pl[P], pn[P];
#pragma unroll
    for (int k = 0; k < P; ++k) pn[k] = cp[k];
    unsigned int mynan = 0u;
    unsigned long long mine = 0ull;
#pragma unroll 1
    for (int g = 0; g < nr; ++g) {
        // this sweep's planes; the next sweep's are loaded now, in flight with this one's gathers
#pragma unroll
        for (int k = 0; k < P; ++k) pl[k] = pn[k];
        const int gn = g + 1 < nr ? g + 1 : g;
#pragma unroll
        for (int k = 0; k < P; ++k) pn[k] = cp[gn * pa.cslice + k];
        bool tame = a.D >= 2;
#pragma unroll
        for (int k = 0; k < P; ++k) tame = tame && tame_plane(pl[k]);
        const bool fast = __builtin_amdgcn_readfirstlane((int)tame) != 0;
        double acc[P];
#pragma unroll
        for (int k = 0; k < P; ++k) acc[k] = 0.0;
        unsigned int nanm = 0u;
#pragma unroll
        for (int h = 0; h < SEG_R / HB; ++h) {
            const int rb = h * HB;
            const int rem = steps - rb;
            if (rem <= 0) break;
            if (rem >= HB) {
                if (fast && rb + HB <= full) pc_batch<MODE, HB, false>(a, pl, 2 * P, first, end, rb, q + rb, acc, nanm);
                else pc_batch<MODE, HB, true>(a, pl, 2 * P, first, end, rb, q + rb, acc, nanm);
            } else if (rem == 1) {
                pc_batch<MODE, 1, true>(a, pl, 2 * P, first, end, rb, q + rb, acc, nanm);
            } else {
                pc_batch<MODE, 2, true>(a, pl, 2 * P, first, end, rb, q + rb, acc, nanm);
                if (rem == 3) pc_batch<MODE, 1, true>(a, pl, 2 * P, first, end, rb + 2, q + rb + 2, acc, nanm);
            }
        }
        const long long m = (long long)wave_sum4_exact(acc);   // lane l: slot l mod P
        if ((lane & ~(P - 1)) == g * P) mine = (unsigned long long)(a.lambda < 0.f ? -m : m);
        nanm = wave_or_rare(nanm);
        if (lane == g) mynan = nanm;
    }
    const int rep = w % a.rrep;
    if (lane < nr * P && mine)
        atomicAdd(&pa.acc[(r0 + lane / P) * pa.aslice + ((long)c * a.rrep + rep) * P + lane % P], mine);
    if (lane < nr && mynan) atomicOr(&pa.nan[(long)(r0 + lane) * a.V + c], mynan);
}

// ----------------------------------------------------------------------------------------
// kernels
// ----------------------------------------------------------------------------------------
#define GRID_STRIDE(i, n) for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

// The launches of a pair's chain of sweeps raise their wavefronts' priority (s_setprio): with
// fuse 2 the plane costs computed ahead on the second stream share the SIMDs, and the SIMD's
// issue arbiter then serves the latency-bound chain first and the ahead launch with what is left.
__device__ __forceinline__ void chain_priority() { __builtin_amdgcn_s_setprio(2); }

__global__ __launch_bounds__(256) void init_level_kernel(SweepArgs a) { GRID_STRIDE(id, a.V) init_node(a, id); }
__global__ __launch_bounds__(256) void init_cost_kernel(SweepArgs a) { GRID_STRIDE(id, a.V) init_cost_node(a, id); }
__global__ __launch_bounds__(256) void spatial_kernel(SweepArgs a) { GRID_STRIDE(id, a.V) spatial_node(a, id); }
__global__ __launch_bounds__(256) void node_kernel(SweepArgs a, int this_kind, int next_kind, float ndz, float ndn) {
    chain_priority();
    const long s = ((long)blockIdx.x * 256 + threadIdx.x) >> 4;
    if (s < a.V) node_group<false>(a, s, this_kind, next_kind, ndz, ndn, NodeIO{});
}

// Plane costs of slots [slot0, slot0+NS) and, for NS == P, the sweep's messages.  Coarse
// levels: wavefronts [0, npc) take segments and the rest take 4 edges each.  Fine levels:
// blocks [0, npc) take chunks and the waves of the other blocks take 4 edges each.
// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b mod 8); this maps
// them so that each XCD takes one contiguous range of work items, whose pixels share census
// rows in that XCD's L2.  A bijection on [0, n) for any n.
__device__ __forceinline__ int xcd_contiguous(int b, int n) {
    constexpr int X = 8;
    const int q = n / X, r = n % X, x = b % X;
    return x * q + (x < r ? x : r) + b / X;
}

// the sweep's messages alone, 4 waves of EDGES_PER_WAVE edges per block (fine levels: its
// register budget is the message code's alone, for more resident waves)
__global__ __launch_bounds__(256) void edge_kernel(SweepArgs a) {
    chain_priority();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    edge_wave(a, (long)blockIdx.x * 4 + wave, false);
}

template <int NS, bool COARSE, int MODE, int HB = 4>
__global__ __launch_bounds__(256) void cost_edge_kernel(SweepArgs a, int slot0, int spatial, int npc) {
    extern __shared__ unsigned long long lds_pc[];
    chain_priority();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int b = blockIdx.x;
    // plane-cost blocks first (XCD-contiguous), then the message blocks in dispatch order
    const int npc_blocks = COARSE ? (npc + 3) / 4 : npc;
    if (b < npc_blocks) {
        const int blk = xcd_contiguous(b, npc_blocks);
        if (COARSE) {
            const int w = blk * 4 + wave;
            if (w < npc) pc_segment<NS, MODE, HB>(a, w, slot0, spatial != 0);
        } else {
            pc_runs<NS, MODE>(a, blk, slot0, (unsigned char *)lds_pc);
        }
    } else if (NS == P) {
        edge_wave(a, (long)(b - npc_blocks) * 4 + wave, COARSE && spatial != 0);
    }
}

// One launch of a run of fused sweeps (see ChainIO): segment blocks first (XCD-contiguous),
// then edge blocks, then node blocks of 16 superpixels.  FIRST: the run's first sweep, whose
// messages read the standard state (edge_wave) and whose node wavefronts only draw.
template <bool FIRST, int MODE, int HB>
__global__ __launch_bounds__(256) void chain_kernel(SweepArgs a, ChainIO io, int npc, int nedge_blocks) {
    chain_priority();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int b = blockIdx.x;
    const int npc_blocks = (npc + 3) / 4;
    if (b < npc_blocks) {
        const int w = xcd_contiguous(b, npc_blocks) * 4 + wave;
        if (w < npc) pc_segment<P, MODE, HB>(a, w, 2 * P, false);
    } else if (b < npc_blocks + nedge_blocks) {
        const long wv = (long)(b - npc_blocks) * 4 + wave;
        if (FIRST) edge_wave(a, wv, false);
        else edge_wave_fused(a, io, wv);
    } else {
        const long s = ((long)(b - npc_blocks - nedge_blocks) * 256 + threadIdx.x) >> 4;
        if (s < a.V) chain_node(a, io, s);
    }
}

// a fused launch that costs no planes (its RANSAC sweep costed ahead): edge blocks, then node
// blocks of 16 superpixels; without the plane-cost code of chain_kernel
__global__ __launch_bounds__(256) void chain_msg_kernel(SweepArgs a, ChainIO io, int nedge_blocks) {
    chain_priority();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int b = blockIdx.x;
    if (b < nedge_blocks) {
        edge_wave_fused(a, io, (long)b * 4 + wave);
    } else {
        const long s = ((long)(b - nedge_blocks) * 256 + threadIdx.x) >> 4;
        if (s < a.V) chain_node(a, io, s);
    }
}

// the candidates' superpixels of every spatial sweep of the level (nS sweeps)
__global__ __launch_bounds__(256) void spatial_ahead_kernel(SweepArgs a, int *spc, int nS) {
    GRID_STRIDE(i, (long)nS * a.V * P) {
        const int p = (int)(i % P);
        const long rn = i / P;
        spc[i] = spatial_ahead(a, (int)(rn / a.V), rn % a.V, p);
    }
}

// every RANSAC candidate of the level's list (nR sweeps), P lanes per (sweep, node)
__global__ __launch_bounds__(256) void draw_ahead_kernel(SweepArgs a, PipeArgs pa, int nR) {
    GRID_STRIDE(i, (long)nR * a.V * P) draw_ahead(a, pa, i);
}

// persistent plane costs of sweeps r0 .. r0+nr-1 over the npc segments: the blocks of XCD x
// (block b runs on XCD b mod 8) walk that XCD's contiguous share of the segments (the grid holds
// at least 8 blocks, so every share has a block)
template <int MODE, int HB>
__global__ __launch_bounds__(256) void pc_ahead_kernel(SweepArgs a, PipeArgs pa, int r0, int nr, int npc) {
    constexpr int X = 8;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int x = blockIdx.x % X;
    const int nbx = ((int)gridDim.x - x + X - 1) / X;
    const int qn = npc / X, rn = npc % X;
    const int start = x * qn + (x < rn ? x : rn), len = qn + (x < rn ? 1 : 0);
    for (int i = (int)(blockIdx.x / X) * 4 + wave; i < len; i += nbx * 4)
        pc_segment_multi<MODE, HB>(a, pa, start + i, r0, nr);
}

// ----------------------------------------------------------------------------------------
// Tiny levels (fuse 2): a run's RANSAC sweeps after its first launch and the closing node update
// in one workgroup of 1024 threads.  The RANSAC sweeps' plane costs exist (second stream), so
// what remains per sweep is the selection of every superpixel (phase A, 16 lanes each) and the
// messages of every edge (phase B, 16 lanes each), with a workgroup barrier between them instead
// of a launch.  Everything the run reads besides the first sweep's state is small (V <= RUN_V
// superpixels, E <= RUN_MT edges, at most RUN_NR listed sweeps), so a prologue loads all of it
// into LDS in one memory round trip, together with the first selection's inputs: the listed
// sweeps' candidates, points and costs (their replica sums and NaN words folded into the float
// cost, as the selection computes it), the edges' ends, slots and weights, and the first
// sweep's messages.  After it every phase reads LDS and registers only.  The selection's outputs
// (planes, costs, beliefs and source slots of slots < 2P) and the sweep messages (two node-major
// buffers) live in LDS as well.  The arithmetic is node_group's and edge_wave's: the same bits
// as the fused launches.  Limits: V <= RUN_V, maxdeg <= kNodeMaxDeg, V maxdeg <= RUN_MT (so
// E <= RUN_MT), n <= RUN_NR (the host checks them).
// ----------------------------------------------------------------------------------------
constexpr int RUN_V = 32, RUN_MT = 384, RUN_NR = 4;
constexpr size_t RUN_LDS_BYTES = (size_t)RUN_V * 2 * P * (16 + 4 + 4 + 4)   // sp, sc, sb, ss
                                 + 2 * (size_t)RUN_MT * K * 4                // smt
                                 + (size_t)RUN_NR * RUN_V * P * (16 + 4)     // pcand, pcost
                                 + (size_t)RUN_NR * RUN_V * 16               // ppt
                                 + (size_t)RUN_MT * (16 + 8);                // edges
static_assert(RUN_LDS_BYTES <= 65536, "one workgroup's LDS");
struct RunArgs {
    const unsigned long long *acc0;   // sums of the run's first sweep (slot stride K), cleared at the end
    unsigned int *nan0;
    const float *mt0;                 // its messages (global, node-major)
    const int *rloc;                  // per edge: the reverse edge's slot among its source's edges
    int n;                            // RANSAC sweeps after the first
    int r1;                           // list index of the first of them (the others follow)
    int prev_spatial;                 // the first sweep is a spatial one
    PipeArgs pa;
};

__global__ __launch_bounds__(1024) void run_kernel(SweepArgs a, RunArgs ra) {
    chain_priority();
    extern __shared__ unsigned long long lds_run[];
    float4 *sp = (float4 *)lds_run;                    // [RUN_V][2P] selected planes
    float4 *pcand = sp + RUN_V * 2 * P;                // [RUN_NR][RUN_V][P] listed sweeps' candidates
    uint4 *ppt = (uint4 *)(pcand + RUN_NR * RUN_V * P);   // [RUN_NR][RUN_V] their RANSAC points
    int4 *e4 = (int4 *)(ppt + RUN_NR * RUN_V);         // [RUN_MT] edge (source, target, slot, reverse slot)
    float2 *ecd = (float2 *)(e4 + RUN_MT);             // [RUN_MT] edge (c0, dpos)
    float *pcost = (float *)(ecd + RUN_MT);            // [RUN_NR][RUN_V][P] candidates' costs
    float *sc = pcost + RUN_NR * RUN_V * P;            // [RUN_V][2P]
    float *sb = sc + RUN_V * 2 * P;                    // [RUN_V][2P]
    int *ss = (int *)(sb + RUN_V * 2 * P);             // [RUN_V][2P]
    float *smt = (float *)(ss + RUN_V * 2 * P);        // [2][RUN_MT][K]
    const PipeArgs &pa = ra.pa;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int k = lane & 15, g = lane & ~15;
    const bool own = k < K;
    const int kk = own ? k : 0;
    const int row = wave * 4 + (lane >> 4);            // 64 rows of 16 lanes
    // ---- prologue: the first selection's inputs (row n = row) in registers, the rest in LDS
    const int n = row;
    const bool node = n < a.V;
    const int nc = node ? n : 0;
    const long nm = (long)nc * a.maxdeg;
    const int deg = node ? a.off[nc + 1] - a.off[nc] : 0;
    uint32_t seed = a.s_ransac[nc];
    const bool child_seeds = ra.prev_spatial && k == 0 && node && n < a.Vh;
    uint32_t seed_c = child_seeds ? a.s_child[nc] : 0u, seed_n = child_seeds ? a.s_neighbor[nc] : 0u;
    SelIn in;
    in.deg = deg;
    in.nb = ra.nan0[nc];
    in.cost_in = a.cost[(long)nc * K + kk];
    in.fsum = acc_sum_at(ra.acc0, a.rrep, nc, kk >= 2 * P ? kk : 2 * P);
    in.pin = a.plane[(long)nc * K + kk];
#pragma unroll
    for (int i = 0; i < kNodeMaxDeg; ++i) in.mt[i] = i < deg ? ra.mt0[(nm + i) * K + kk] : 0.f;
    // lane i of the row: the reverse (node, slot) of incoming slot i (for the closing update)
    const int rv = k < deg ? a.revl[nm + k] : 0;
    {
        const int nvp = a.V * P;
        for (int i = tid; i < ra.n * nvp; i += 1024) {
            const int r = i / nvp, v = (i % nvp) / P, p = i % P;
            const int rp = ra.r1 + r;
            const float4 cand = pa.cand[rp * pa.cslice + (long)v * P + p];
            const int64_t fsum = acc_sum_at(pa.acc + rp * pa.aslice, a.rrep, v, 2 * P + p, P);
            const unsigned int nb = pa.nan[(long)rp * a.V + v];
            pcand[(r * RUN_V + v) * P + p] = cand;
            pcost[(r * RUN_V + v) * P + p] = from_fixed(fsum, (nb >> (2 * P + p)) & 1u, a.shift);
        }
        for (int i = tid; i < ra.n * a.V; i += 1024) {
            const int r = i / a.V, v = i % a.V;
            ppt[r * RUN_V + v] = pa.pt[(long)(ra.r1 + r) * a.V + v];
        }
        for (int e = tid; e < a.E; e += 1024) {
            const int2 b = a.bd[e];
            e4[e] = make_int4(b.x, b.y, a.eloc[e], ra.rloc[e]);
            ecd[e] = make_float2(a.c0[e], a.dpos[e]);
        }
        // the first sweep's messages, read by the first sweep's successor as smt buffer 0
        for (int i = tid; i < a.V * a.maxdeg * K; i += 1024) smt[i] = ra.mt0[i];
    }
    // a selection that is not the run's last: its outputs to LDS and the seed advance (its next
    // sweep is a RANSAC sweep)
    auto keep = [&](const SelOut &so, bool first) {
        const float4 pl_n = make_float4(__shfl(so.pl.x, g + so.src), __shfl(so.pl.y, g + so.src),
                                        __shfl(so.pl.z, g + so.src), __shfl(so.pl.w, g + so.src));
        const float co_n = __shfl(so.cost, g + so.src), be_n = __shfl(so.bel, g + so.src);
        if (k < 2 * P) {
            sp[n * 2 * P + k] = pl_n;
            sc[n * 2 * P + k] = co_n;
            sb[n * 2 * P + k] = be_n;
            ss[n * 2 * P + k] = so.src;
        }
        constexpr LcgJump j3 = lcg_jump_coef(3 * P);
        seed = (j3.a * seed + j3.c) % 2147483648u;
        if (k == 2 * P) a.s_ransac[n] = seed;
        if (first && child_seeds) {
            constexpr LcgJump jp = lcg_jump_coef(P);
            a.s_child[n] = (jp.a * seed_c + jp.c) % 2147483648u;
            a.s_neighbor[n] = (jp.a * seed_n + jp.c) % 2147483648u;
        }
    };
    // ---- phase A of the run's first sweep (never its last: n >= 1)
    if (node) keep(sel_compute(a, ChainIO{}, n, in, k, kk), true);
    __syncthreads();
    for (int u = 1; u <= ra.n; ++u) {
        // ---- phase B: the messages of sweep j0+u (listed sweep u-1)
        const float *mtp = smt + (size_t)((u - 1) & 1) * RUN_MT * K;
        float *mto = smt + (size_t)(u & 1) * RUN_MT * K;
        const int rl = (u - 1) * RUN_V;
        for (int e = row; e < a.E; e += 64) {
            const int4 ed = e4[e];
            const int sn = ed.x, tn = ed.y, eslot = ed.z, rslot = ed.w;
            const float2 cd = ecd[e];
            const float c0 = cd.x, dpos = cd.y;
            // slots >= 2P: this sweep's candidates (lanes past the K slots read slot 2P's)
            const int kc = kk >= 2 * P ? kk - 2 * P : 0;
            const float4 psl = k < 2 * P ? sp[sn * 2 * P + k] : pcand[(rl + sn) * P + kc];
            // lanes p < 2P: t's particle p after its selection, its belief less the message from s
            const int pp = k < 2 * P ? k : 0;
            const int tsrc = ss[tn * 2 * P + pp];
            const float tm = sb[tn * 2 * P + pp] - mtp[((long)tn * a.maxdeg + rslot) * K + tsrc];
            const float4 tp = sp[tn * 2 * P + pp];
            const uint4 ps = ppt[rl + sn], pt = ppt[rl + tn];
            float X[6], Y[6];
            X[0] = (float)unpack_x(ps.x); Y[0] = (float)unpack_y(ps.x);
            X[1] = (float)unpack_x(ps.y); Y[1] = (float)unpack_y(ps.y);
            X[2] = (float)unpack_x(ps.z); Y[2] = (float)unpack_y(ps.z);
            X[3] = (float)unpack_x(pt.x); Y[3] = (float)unpack_y(pt.x);
            X[4] = (float)unpack_x(pt.y); Y[4] = (float)unpack_y(pt.y);
            X[5] = (float)unpack_x(pt.z); Y[5] = (float)unpack_y(pt.z);
            float out = 0.f, minimum = FLT_MAX;
            fused_msg(tm, tp.x, tp.y, tp.z, psl, X, Y, c0, dpos, a.tau_s, out, minimum,
                      std::make_integer_sequence<int, 2 * P>{});
            out += minimum;
            float mean = row_sum(out, std::make_integer_sequence<int, K>{});
            mean /= K;
            if (own) mto[((long)sn * a.maxdeg + eslot) * K + k] = out - mean;
        }
        __syncthreads();
        // ---- phase A: the selection that ends sweep j0+u (listed sweep u-1); u = n: the run's
        // last, the closing node update (node_group, next sweep spatial or none) into the
        // standard state
        if (node) {
            const int ps = n * 2 * P + (kk < 2 * P ? kk : 0);
            const int pc = ((u - 1) * RUN_V + n) * P + (kk >= 2 * P ? kk - 2 * P : 0);
            const float cost = kk >= 2 * P ? pcost[pc] : sc[ps];
            const float4 pin = kk < 2 * P ? sp[ps] : pcand[pc];
            float mt[kNodeMaxDeg];
#pragma unroll
            for (int i = 0; i < kNodeMaxDeg; ++i) mt[i] = i < deg ? mto[(nm + i) * K + kk] : 0.f;
            float bel = cost;
#pragma unroll
            for (int i = 0; i < kNodeMaxDeg; ++i) bel = i < deg ? bel + mt[i] : bel;
            const SelOut so = sel_rank(bel, cost, pin, k);
            if (u < ra.n) {
                keep(so, false);
            } else {
                const float4 pl_n = make_float4(__shfl(so.pl.x, g + so.src), __shfl(so.pl.y, g + so.src),
                                                __shfl(so.pl.z, g + so.src), __shfl(so.pl.w, g + so.src));
                const float co_n = __shfl(so.cost, g + so.src), be_n = __shfl(so.bel, g + so.src);
                if (own) {
                    a.plane[(long)n * K + k] = pl_n;
                    a.cost[(long)n * K + k] = co_n;
                    a.belief[(long)n * K + k] = be_n;
                }
                int r[kNodeMaxDeg];
                row_revl(rv, r, std::make_integer_sequence<int, kNodeMaxDeg>{});
                float *msg = (float *)a.msg;
#pragma unroll
                for (int i = 0; i < kNodeMaxDeg; ++i) {
                    const float oi = __shfl(mt[i], g + so.src);
                    if (i < deg && k < 2 * P) msg[(long)r[i] * 8 + k] = oi;
                }
                if (k == K - 1) a.point[n] = ppt[(ra.n - 1) * RUN_V + n];
                if (kk >= 2 * P) acc_zero_at((unsigned long long *)ra.acc0, a.rrep, n, kk);
                if (k == 2 * P) ra.nan0[n] = 0u;
            }
        }
        __syncthreads();
    }
}

// the node update that ends a run of fused sweeps: the run's buffers in, the standard state out
__global__ __launch_bounds__(256) void node_io_kernel(SweepArgs a, NodeIO io, int this_kind, int next_kind, float ndz,
                                                      float ndn) {
    chain_priority();
    const long s = ((long)blockIdx.x * 256 + threadIdx.x) >> 4;
    if (s < a.V) node_group<true>(a, s, this_kind, next_kind, ndz, ndn, io);
}

// ----------------------------------------------------------------------------------------
// batched sweeps: one launch carries one sweep of each of several pairs.  A pair's share of a
// launch is a BatchItem; its blocks are a contiguous range starting at a multiple of 8 (so
// that the XCD-contiguous mapping inside the range holds), found by binary search.
// ----------------------------------------------------------------------------------------
struct BatchItem {
    int args;                   // index of the pair-level SweepArgs
    int kind, next_kind;
    float ndz, ndn;
    int coarse, spatial;        // plane-cost form; spatial candidates resolved in the cost kernel
    int npc;                    // coarse: segment bound; fine: chunks; 0: no plane costs
    int blk_begin, pc_blocks, edge_blocks;   // cost/message launch
    int node_begin, node_blocks;             // node launch
    int sp_begin, sp_blocks;                 // spatial launch (fine levels)
    int msg_begin, msg_blocks;               // message launch (fine levels)
};

// Every pointer of a SweepArgs, for the relative form of the batch table: the host stores each
// pointer as its byte offset from the table's own device address and the kernels add the
// table pointer (a kernel argument) back, so the compiler knows every access is to global
// memory (a pointer loaded from memory would be generic and compile to flat accesses).
#define SWEEP_ARGS_POINTERS(X)                                                                   \
    X(clus) X(bd) X(pxy) X(pcl) X(pix_off) X(seg) X(sego) X(off) X(rev) X(eloc) X(revl) X(pw) X(c0) \
    X(dpos) X(mxy) X(child) X(child_off) X(plane) X(cost) X(belief) X(point) X(bd_h) X(off_h)         \
    X(plane_h) X(cl) X(cr) X(crx) X(volf) X(dwta) X(msg) X(msg_t) X(acc) X(accnan) X(s_ransac) X(s_neighbor) X(s_child)

__device__ __forceinline__ SweepArgs rebased_args(const SweepArgs *__restrict__ args, int i) {
    SweepArgs a = args[i];
#define HBP_REBASE(f) a.f = (decltype(a.f))((const char *)args + (unsigned long long)a.f);
    SWEEP_ARGS_POINTERS(HBP_REBASE)
#undef HBP_REBASE
    return a;
}

template <int FIELD>
__device__ __forceinline__ int batch_find(const BatchItem *__restrict__ it, int n, int b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int v = FIELD == 0   ? it[mid].blk_begin
                      : FIELD == 1 ? it[mid].node_begin
                      : FIELD == 2 ? it[mid].sp_begin
                                   : it[mid].msg_begin;
        if (v <= b) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

template <int MODE, bool COARSE, int HB = 2>
__global__ __launch_bounds__(256) void cost_edge_batch_kernel(const SweepArgs *__restrict__ args,
                                                              const BatchItem *__restrict__ items, int nitems) {
    extern __shared__ unsigned long long lds_pc[];
    const int b = blockIdx.x;
    const BatchItem it = items[batch_find<0>(items, nitems, b)];
    const int local = b - it.blk_begin;
    if (local >= it.pc_blocks + it.edge_blocks) return;   // alignment padding
    const SweepArgs a = rebased_args(args, it.args);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (local < it.pc_blocks) {
        const int blk = xcd_contiguous(local, it.pc_blocks);
        if (COARSE) {
            const int w = blk * 4 + wave;
            if (w < it.npc) pc_segment<P, MODE, HB>(a, w, 2 * P, it.spatial != 0);
        } else {
            pc_runs<P, MODE>(a, blk, 2 * P, (unsigned char *)lds_pc);
        }
    } else {
        edge_wave(a, (long)(local - it.pc_blocks) * 4 + wave, COARSE && it.spatial);
    }
}
__global__ __launch_bounds__(256) void node_batch_kernel(const SweepArgs *__restrict__ args,
                                                         const BatchItem *__restrict__ items, int nitems) {
    const int b = blockIdx.x;
    const BatchItem it = items[batch_find<1>(items, nitems, b)];
    const int local = b - it.node_begin;
    if (local >= it.node_blocks) return;
    const SweepArgs a = rebased_args(args, it.args);
    const long s = ((long)local * 256 + threadIdx.x) >> 4;
    if (s < a.V) node_group<false>(a, s, it.kind, it.next_kind, it.ndz, it.ndn, NodeIO{});
}
__global__ __launch_bounds__(256) void edge_batch_kernel(const SweepArgs *__restrict__ args,
                                                         const BatchItem *__restrict__ items, int nitems) {
    const int b = blockIdx.x;
    const BatchItem it = items[batch_find<3>(items, nitems, b)];
    const int local = b - it.msg_begin;
    if (local >= it.msg_blocks) return;
    const SweepArgs a = rebased_args(args, it.args);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    edge_wave(a, (long)local * 4 + wave, false);
}
__global__ __launch_bounds__(256) void spatial_batch_kernel(const SweepArgs *__restrict__ args,
                                                            const BatchItem *__restrict__ items, int nitems) {
    const int b = blockIdx.x;
    const BatchItem it = items[batch_find<2>(items, nitems, b)];
    const int local = b - it.sp_begin;
    if (local >= it.sp_blocks) return;
    const SweepArgs a = rebased_args(args, it.args);
    const long id = (long)local * 256 + threadIdx.x;
    if (id < a.V) spatial_node(a, id);
}

// a level's particle state from zero: planes, costs, beliefs [V][12], sample points [V]
__global__ void zero_state_kernel(float4 *__restrict__ plane, float *__restrict__ cost, float *__restrict__ belief,
                                  uint4 *__restrict__ point, long nk, long v) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nk; i += (long)gridDim.x * blockDim.x) {
        plane[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        cost[i] = 0.f;
        belief[i] = 0.f;
        if (i < v) point[i] = make_uint4(0u, 0u, 0u, 0u);
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
__global__ void map_paint_kernel(const uint32_t *__restrict__ pxy, const uint32_t *__restrict__ pcl, long N,
                                 const float4 *__restrict__ best, int W, int D, float *__restrict__ disp,
                                 float *__restrict__ plane3) {
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += (long)gridDim.x * blockDim.x) {
        uint32_t q = pxy[j];
        int x = unpack_x(q), y = unpack_y(q);
        float4 pl = best[pcl[j]];
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
    int s = 52 - ev - en;
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
    a.pxy = B.pxy.p;
    a.pcl = B.pcl.p;
    a.pix_off = B.pix_off.p;
    a.seg = B.seg.p;
    a.sego = B.sego.p;
    a.off = B.off.p;
    a.rev = B.rev.p;
    a.maxdeg = B.maxdeg;
    a.eloc = B.eloc.p;
    a.revl = B.revl.p;
    a.pw = B.pw.p;
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
    a.cl = v.cl.p;
    a.cr = v.cr.p;
    a.crx = b.crx.p;
    a.crx_stride = b.crx_stride;
    a.volf = v.f32src ? v.f32.p : nullptr;
    a.vplane = (long)v.W * v.H;
    a.right = v.right ? 1 : 0;
    a.dwta = v.dwta.p;
    a.lambda = b.lambda;
    a.tau = b.tau;
    a.tau_s = b.tau_s;
    a.use_cost_vol = b.use_cost_vol;
    a.shift = b.fx_shift;
    a.lam_fx = std::ldexp(std::fabs(b.lambda), b.fx_shift);
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
    HBP_REQUIRE(std::isfinite(b.lambda) && std::isfinite(b.tau_s), "alpha and tau_smooth must be finite");
    // the bound of the per-pixel costs: tau, or the largest |entry| of a given f32 volume
    // (its entries are truncated at tau from above only)
    const float cmax = std::max(std::fabs(b.tau), v.f32src ? v.absmax : 0.f);
    HBP_REQUIRE(std::isfinite(cmax), "the cost volume holds infinite entries");
    b.fx_shift = fixed_point_shift(b.lambda, cmax, b.use_cost_vol, b.N);
    HBP_REQUIRE(b.fx_shift > -60, "alpha too large for the fixed-point plane-cost sums");
    {
        const float lf = std::ldexp(std::fabs(b.lambda), b.fx_shift);
        HBP_REQUIRE(std::isfinite(lf) && (lf == 0.f || std::isnormal(lf)),
                    "alpha and the cost truncation give no usable fixed-point scale");
    }
    // reversed, padded right census of the left-view plane cost (census volumes)
    if (!v.f32src) {
        b.crx_stride = b.W + std::max(b.D - 1, 1);
        b.crx.alloc(c.pool, (size_t)b.H * b.crx_stride);
        hipLaunchKernelGGL(crx_kernel, dim3(grid_for(c, (long)b.H * b.crx_stride)), dim3(256), 0, c.stream, v.cr.p,
                           b.W, b.H, b.crx_stride, b.crx.p);
        HBP_CHECK(hipGetLastError());
    }
    b.lv.clear();
    b.lv.resize(b.L);
    long N = b.N;
    DevBuf<unsigned int> im_clus(c.pool, N), keys(c.pool, N), keys2(c.pool, N);
    DevBuf<int> ids(c.pool, N), ids2(c.pool, N);
    DevBuf<int> dmax(c.pool, (size_t)b.L);
    HBP_CHECK(hipMemsetAsync(dmax.p, 0, sizeof(int) * (size_t)b.L, c.stream));
    hipLaunchKernelGGL(iota_u32_kernel, dim3(grid_for(c, N)), dim3(256), 0, c.stream, im_clus.p, N);
    int maxV = 1;
    long maxE = 1;
    for (int l = 0; l < b.L; ++l) {
        BPLevel &B = b.lv[l];
        B.V = p.V[l];
        B.E = p.E[l];
        // the top level's parents: one cluster, or target_clus of them (unused by BP, which does
        // not run on the top level, but its children are grouped like every level's)
        B.Vh = (l < b.L - 1) ? p.V[l + 1] : p.top;
        B.clus = p.clus.p + p.voff[l];
        B.bd = p.bd.p + p.eoff[l];
        if (l >= b.min_level) {
            int V = B.V;
            long E = B.E;
            if (V > maxV) maxV = V;
            if (E > maxE) maxE = E;
            const float *mean = p.mean.p + 6 * p.voff[l];
            // pixels by (cluster, pixel id)
            hipLaunchKernelGGL(copy_iota_kernel, dim3(grid_for(c, N)), dim3(256), 0, c.stream, im_clus.p, N, keys.p,
                               ids.p);
            if (sort_pairs_u32(c, keys.p, keys2.p, ids.p, ids2.p, N, bits_for(V))) {
                std::swap(keys, keys2);
                std::swap(ids, ids2);
            }
            B.pxy.alloc(c.pool, N);
            B.pcl.alloc(c.pool, N);
            hipLaunchKernelGGL(pix_kernel, dim3(grid_for(c, N)), dim3(256), 0, c.stream, keys.p, ids.p, N, b.W,
                               B.pxy.p, B.pcl.p);
            B.pix_off.alloc(c.pool, (size_t)V + 1);
            hipLaunchKernelGGL(lower_bound_kernel, dim3(grid_for(c, N)), dim3(256), 0, c.stream, keys.p, N, V,
                               B.pix_off.p);
            // coarse levels: segments of <= seg_len pixels per superpixel
            // (a per-level choice of the context overrides the threshold)
            const int form = l < Ctx::kFormLevels ? c.level_form[l] : -1;
            B.coarse = form < 0 ? N >= (long)c.coarse_min_pix * V : form == 1;
            if (B.coarse) {
                const int sl = c.seg_len;
                HBP_REQUIRE(sl >= 64 && sl <= SEG && sl % 64 == 0, "segment length must be a multiple of 64 in [64, 512]");
                DevBuf<int> scnt(c.pool, (size_t)V + 1);
                B.sego.alloc(c.pool, (size_t)V + 1);
                B.nseg_max = (int)(div_up(N, sl) + V);
                B.seg.alloc(c.pool, (size_t)B.nseg_max);
                HBP_CHECK(hipMemsetAsync(B.seg.p, 0, sizeof(int4) * (size_t)B.nseg_max, c.stream));
                hipLaunchKernelGGL(seg_count_kernel, dim3(grid_for(c, (long)V + 1)), dim3(256), 0, c.stream,
                                   B.pix_off.p, V, sl, scnt.p);
                scan_exclusive_int(c, scnt.p, B.sego.p, V + 1, nullptr);
                hipLaunchKernelGGL(seg_fill_kernel, dim3(grid_for(c, B.nseg_max)), dim3(256), 0, c.stream, B.pix_off.p,
                                   B.sego.p, V, sl, (long)B.nseg_max, B.seg.p);
            }
            // CSR by source, reverse edges, edge weights, mean positions
            B.off.alloc(c.pool, (size_t)V + 1);
            hipLaunchKernelGGL(lower_bound_src_kernel, dim3(grid_for(c, E > 0 ? E : 1)), dim3(256), 0, c.stream, B.bd,
                               E, V, B.off.p);
            B.rev.alloc(c.pool, E);
            hipLaunchKernelGGL(rev_kernel, dim3(grid_for(c, E)), dim3(256), 0, c.stream, B.bd, B.off.p, E, B.rev.p);
            hipLaunchKernelGGL(deg_max_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream, B.off.p, V, dmax.p + l);
            B.pw.alloc(c.pool, N);
            hipLaunchKernelGGL(pw_kernel, dim3(grid_for(c, N)), dim3(256), 0, c.stream, B.pxy.p, N, v.dwta.p, b.W,
                               B.pw.p);
            B.c0.alloc(c.pool, E);
            B.dpos.alloc(c.pool, E);
            hipLaunchKernelGGL(edge_weights_kernel, dim3(grid_for(c, E)), dim3(256), 0, c.stream, mean, V, B.bd,
                               p.bdl.p + p.eoff[l], E, B.c0.p, B.dpos.p);
            B.mxy.alloc(c.pool, V);
            hipLaunchKernelGGL(mean_xy_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream, mean, V, B.mxy.p);
            // children grouped by parent (sort_clus, hbp_isp.cu:985-1001)
            {
                DevBuf<unsigned int> ck(c.pool, V), ck2(c.pool, V);
                DevBuf<int> cv2(c.pool, V);
                B.child.alloc(c.pool, V);
                B.child_off.alloc(c.pool, (size_t)B.Vh + 1);
                hipLaunchKernelGGL(copy_iota_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream,
                                   (const unsigned int *)B.clus, (long)V, ck.p, B.child.p);
                if (sort_pairs_u32(c, ck.p, ck2.p, B.child.p, cv2.p, V, bits_for(B.Vh))) {
                    std::swap(ck, ck2);
                    std::swap(B.child, cv2);
                }
                hipLaunchKernelGGL(lower_bound_kernel, dim3(grid_for(c, V)), dim3(256), 0, c.stream, ck.p, (long)V,
                                   B.Vh, B.child_off.p);
            }
            B.plane.alloc(c.pool, (size_t)V * K);
            B.cost.alloc(c.pool, (size_t)V * K);
            B.belief.alloc(c.pool, (size_t)V * K);
            B.point.alloc(c.pool, V);
            // the top level's planes are never set by the method; they read as zero
            hipLaunchKernelGGL(zero_state_kernel, dim3(grid_for(c, (long)V * K)), dim3(256), 0, c.stream, B.plane.p,
                               B.cost.p, B.belief.p, B.point.p, (long)V * K, (long)V);
        }
        if (l < b.L - 1)
            hipLaunchKernelGGL(compose_kernel, dim3(grid_for(c, N)), dim3(256), 0, c.stream, im_clus.p, B.clus, N);
        HBP_CHECK(hipGetLastError());
    }
    // node-major message slots: the out-degree bound of every BP level
    std::vector<int> hmax(b.L, 0), hseg(b.L, 0);
    HBP_CHECK(hipMemcpyAsync(hmax.data(), dmax.p, sizeof(int) * (size_t)b.L, hipMemcpyDeviceToHost, c.stream));
    // coarse levels: the segment count (the scan's last entry), so that a launch covers the
    // segments alone and no padding lands on one XCD
    for (int l = b.min_level; l < b.L; ++l)
        if (b.lv[l].coarse)
            HBP_CHECK(hipMemcpyAsync(&hseg[l], b.lv[l].sego.p + b.lv[l].V, sizeof(int), hipMemcpyDeviceToHost,
                                     c.stream));
    HBP_CHECK(hipStreamSynchronize(c.stream));
    size_t slots = 1;
    for (int l = b.min_level; l < b.L; ++l) {
        BPLevel &B = b.lv[l];
        B.maxdeg = std::max(1, hmax[l]);
        if (B.coarse) {
            HBP_REQUIRE(hseg[l] >= 1 && hseg[l] <= B.nseg_max, "segment count out of range");
            B.nseg = hseg[l];
        }
        slots = std::max(slots, (size_t)B.V * B.maxdeg);
        B.eloc.alloc(c.pool, std::max<long>(B.E, 1));
        B.revl.alloc(c.pool, (size_t)B.V * B.maxdeg + kNodeMaxDeg);
        if (B.E > 0)
            hipLaunchKernelGGL(node_major_kernel, dim3(grid_for(c, B.E)), dim3(256), 0, c.stream, B.bd, B.off.p,
                               B.rev.p, B.E, B.maxdeg, B.eloc.p, B.revl.p);
        if (B.coarse) {
            B.rloc.alloc(c.pool, std::max<long>(B.E, 1));
            if (B.E > 0)
                hipLaunchKernelGGL(rloc_kernel, dim3(grid_for(c, B.E)), dim3(256), 0, c.stream, B.eloc.p, B.rev.p,
                                   B.E, B.rloc.p);
        }
    }
    b.msg.alloc(c.pool, (size_t)maxE * 2);
    b.msg_t.alloc(c.pool, (slots + kNodeMaxDeg) * K);
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
    // per level, the RANSAC sweeps after its first and the seed maps of their draws: node
    // update j draws sweep j+1's candidates from the seed it reads, then advances it by 3P
    // draws before a RANSAC sweep, 4P before a random one, none before a spatial one, so every
    // map is a number of LCG steps; draw_ahead takes that number from ransac_draw_steps, checked
    // here against the walk over the schedule
    b.sidx.assign(b.sched.size(), -1);
    b.scount.assign(b.L, 0);
    for (size_t j = 0; j < b.sched.size(); ++j)
        if (b.sched[j].kind == kSpatial) b.sidx[j] = b.scount[b.sched[j].level]++;
    b.ridx.assign(b.sched.size(), -1);
    b.rcount.assign(b.L, 0);
    b.rd_first.assign(b.L, 0);
    b.rd_count.assign(b.L, 0);
    {
        int lv = -1, iter = 0;
        for (const Sweep &sw : b.sched) {   // random sweeps: their count, in consecutive iterations
            if (sw.first) {
                lv = sw.level;
                iter = 0;
                b.rd_first[lv] = (10 + lv * 10) / 4 * 3 + 1;
            }
            if (sw.kind == kRandom) {   // it ends the iteration whose spatial sweep was counted
                HBP_REQUIRE(iter - 1 == b.rd_first[lv] + b.rd_count[lv], "random sweeps outside the expected iterations");
                ++b.rd_count[lv];
            }
            if (sw.kind == kSpatial) ++iter;
        }
        long steps = 0;
        for (size_t j = 0; j < b.sched.size(); ++j) {
            const Sweep &sw = b.sched[j];
            if (sw.first) {
                lv = sw.level;
                steps = 0;
                continue;
            }
            if (sw.kind == kRansac) {
                const int r = b.rcount[lv]++;
                b.ridx[j] = r;
                HBP_REQUIRE(steps == ransac_draw_steps(r, b.rd_first[lv], b.rd_count[lv]),
                            "seed map of a RANSAC sweep differs from its closed form");
            }
            steps += sw.kind == kRansac ? 3 * P : (sw.kind == kRandom ? 4 * P : 0);
        }
    }
    b.pp_level = -1;
    b.next = 0;
    b.cur = b.L - 2;
    b.sweeps = 0;
    b.mapped = false;
    HBP_CHECK(hipStreamSynchronize(c.stream));
}


// algorithmic HBM bytes per launch: per pixel its packed position, its left census word and
// one right census word (12 B; the right census row a pixel's planes gather from is read once
// per row); per superpixel the planes read, the int64 sums and the NaN word written; per edge
// the message inputs (6 packed points, weights, 8 prior messages, 8 beliefs and 8 planes of t)
// and 12 outputs
static double plane_cost_bytes(long N, int V, int NS) { return 12.0 * N + (double)V * (24.0 * NS + 4); }
static double edge_bytes(long E, int V) { return 96.0 * E + (16.0 * K + 48) * V; }
static double node_bytes(long E, int V) { return 84.0 * E + (48.0 * K + 16) * V; }

template <int NS, bool COARSE, int HB = 4>
static void launch_cost_edge_mode(Ctx &c, const ProfScope &ps, const SweepArgs &a, int mode, long blocks, size_t lds,
                                  int slot0, int spatial, int npc) {
    if (mode == kCostLeft)
        HBP_LAUNCH(ps, (cost_edge_kernel<NS, COARSE, kCostLeft, HB>), dim3(blocks), dim3(256), lds, c.stream, a, slot0,
                   spatial, npc);
    else if (mode == kCostRight)
        HBP_LAUNCH(ps, (cost_edge_kernel<NS, COARSE, kCostRight, HB>), dim3(blocks), dim3(256), lds, c.stream, a,
                   slot0, spatial, npc);
    else if (mode == kCostF32)
        HBP_LAUNCH(ps, (cost_edge_kernel<NS, COARSE, kCostF32, HB>), dim3(blocks), dim3(256), lds, c.stream, a,
                   slot0, spatial, npc);
    else
        HBP_LAUNCH(ps, (cost_edge_kernel<NS, COARSE, kCostWta, HB>), dim3(blocks), dim3(256), lds, c.stream, a, slot0,
                   spatial, npc);
}

// plane costs of slots [slot0, slot0 + NS) (+ the sweep's messages when with_edges)
template <int NS>
static void launch_cost_edge(Ctx &c, const SweepArgs &a, const BPLevel &B, int l, int slot0, bool spatial,
                             bool costs, bool with_edges) {
    long nedge_w = with_edges ? div_up(a.E, EDGES_PER_WAVE) : 0;
    const int mode = !a.use_cost_vol ? kCostWta : (a.volf ? kCostF32 : (a.right ? kCostRight : kCostLeft));
    // Coarse levels: one launch carries the segments' waves and the message waves (the
    // segments' latency chains and the messages' arithmetic overlap).  Fine levels and
    // message-only launches: the messages run in their own kernel, whose smaller register
    // budget keeps twice the waves resident.
    if (B.coarse && costs) {
        double bytes = (costs ? plane_cost_bytes(a.N, a.V, NS) : 0.0) + (with_edges ? edge_bytes(a.E, a.V) : 0.0);
        int npc = costs ? B.nseg : 0;
        long blocks = div_up(npc, 4) + div_up(nedge_w, 4);
        if (blocks == 0) return;
        ProfScope ps(c, (costs ? "plane_cost_L" : "edge_L") + std::to_string(l), bytes, true);
        if (NS == P && c.single_hb == 2)
            launch_cost_edge_mode<NS, true, 2>(c, ps, a, mode, blocks, 0, slot0, spatial ? 1 : 0, npc);
        else
            launch_cost_edge_mode<NS, true, 4>(c, ps, a, mode, blocks, 0, slot0, spatial ? 1 : 0, npc);
    } else {
        if (costs) {
            ProfScope ps(c, "plane_cost_L" + std::to_string(l), plane_cost_bytes(a.N, a.V, NS), true);
            int npc = (int)div_up(a.N, CHUNK);
            launch_cost_edge_mode<NS, false>(c, ps, a, mode, npc, PC_LDS_BYTES, slot0, 0, npc);
        }
        if (nedge_w > 0) {
            ProfScope ps(c, "edge_L" + std::to_string(l), edge_bytes(a.E, a.V), true);
            HBP_LAUNCH(ps, edge_kernel, dim3(div_up(nedge_w, 4)), dim3(256), 0, c.stream, a);
        }
    }
}

// level init: prior planes, seeds, first RANSAC candidates, costs of all 12 slots
static void run_level_init(Ctx &c, BP &b, const SweepArgs &a, int l) {
    const BPLevel &B = b.lv[l];
    if (a.E > 0)
        hipLaunchKernelGGL(zero_f4_kernel, dim3(grid_for(c, 2 * a.E)), dim3(256), 0, c.stream, b.msg.p, 2 * a.E);
    hipLaunchKernelGGL(init_level_kernel, dim3(grid_for(c, a.V)), dim3(256), 0, c.stream, a);
    launch_cost_edge<K>(c, a, B, l, 0, false, true, false);
    hipLaunchKernelGGL(init_cost_kernel, dim3(grid_for(c, a.V)), dim3(256), 0, c.stream, a);
    HBP_CHECK(hipGetLastError());
}

static void level_start(Ctx &c, BP &b, const SweepArgs &a, int l);
static const int *pipe_spc(const BP &b, const Sweep &sw, size_t j);

// one sweep as separate kernels
static void run_sweep_split(Ctx &c, BP &b, const Sweep &sw, int next_kind, float ndz, float ndn) {
    int l = sw.level;
    const BPLevel &B = b.lv[l];
    SweepArgs a = sweep_args(b, l);
    bool spatial = sw.kind == kSpatial;
    if (spatial && B.coarse) a.spc = pipe_spc(b, sw, (size_t)(&sw - b.sched.data()));
    if (sw.first) {
        level_start(c, b, a, l);
        launch_cost_edge<P>(c, a, B, l, 2 * P, false, false, a.E > 0);
    } else {
        if (spatial && !B.coarse) {
            ProfScope pss(c, "spatial_L" + std::to_string(l), (16.0 * P + 28) * a.V);
            hipLaunchKernelGGL(spatial_kernel, dim3(grid_for(c, a.V)), dim3(256), 0, c.stream, a);
        }
        launch_cost_edge<P>(c, a, B, l, 2 * P, spatial && B.coarse, true, a.E > 0);
    }
    {
        ProfScope psn(c, "node_L" + std::to_string(l), node_bytes(a.E, a.V), true);
        HBP_LAUNCH(psn, node_kernel, dim3(div_up((long)a.V * 16, 256)), dim3(256), 0, c.stream, a, sw.kind, next_kind,
                   ndz, ndn);
    }
    HBP_CHECK(hipGetLastError());
}

template <bool FIRST, int HB>
static void launch_chain_hb(Ctx &c, const ProfScope &ps, const SweepArgs &a, const ChainIO &io, int mode, int blocks,
                            int npc, int nedge_blocks) {
    if (mode == kCostLeft)
        HBP_LAUNCH(ps, (chain_kernel<FIRST, kCostLeft, HB>), dim3(blocks), dim3(256), 0, c.stream, a, io, npc,
                   nedge_blocks);
    else if (mode == kCostRight)
        HBP_LAUNCH(ps, (chain_kernel<FIRST, kCostRight, HB>), dim3(blocks), dim3(256), 0, c.stream, a, io, npc,
                   nedge_blocks);
    else if (mode == kCostF32)
        HBP_LAUNCH(ps, (chain_kernel<FIRST, kCostF32, HB>), dim3(blocks), dim3(256), 0, c.stream, a, io, npc,
                   nedge_blocks);
    else
        HBP_LAUNCH(ps, (chain_kernel<FIRST, kCostWta, HB>), dim3(blocks), dim3(256), 0, c.stream, a, io, npc,
                   nedge_blocks);
}

// one launch of a fused run: plane costs of the sweep's candidates (unless the level init
// costed them), its messages, and the node wavefronts
template <bool FIRST>
static void launch_chain(Ctx &c, const SweepArgs &a, const ChainIO &io, const BPLevel &B, int l, bool costs) {
    const int mode = !a.use_cost_vol ? kCostWta : (a.volf ? kCostF32 : (a.right ? kCostRight : kCostLeft));
    const int npc = costs ? B.nseg : 0;
    const int nedge_w = a.E > 0 ? div_up(a.E, FIRST ? EDGES_PER_WAVE : FUSED_EDGES) : 0;
    const int nedge_blocks = div_up(nedge_w, 4);
    const int node_blocks = div_up((long)a.V * 16, 256);
    const int blocks = div_up(npc, 4) + nedge_blocks + node_blocks;
    const double bytes = (costs ? plane_cost_bytes(a.N, a.V, P) : 0.0) + edge_bytes(a.E, a.V) +
                         (FIRST ? 0.0 : node_bytes(a.E, a.V));
    ProfScope ps(c, (costs ? "plane_cost_L" : "edge_L") + std::to_string(l), bytes, true);
    if (!FIRST && npc == 0) {
        HBP_LAUNCH(ps, chain_msg_kernel, dim3(nedge_blocks + node_blocks), dim3(256), 0, c.stream, a, io, nedge_blocks);
        return;
    }
    if (c.single_hb == 2) launch_chain_hb<FIRST, 2>(c, ps, a, io, mode, blocks, npc, nedge_blocks);
    else launch_chain_hb<FIRST, 4>(c, ps, a, io, mode, blocks, npc, nedge_blocks);
}

// the buffers of the fused runs (ChainIO), sized by the largest segment-form level; made on
// the first run, so that objects that never run one (batched and served pairs) do not hold them
static void chain_alloc(Ctx &c, BP &b) {
    if (b.ch_cand[0].p) return;
    size_t vc = 0, mc = 0, ac = 0;
    for (int l = b.min_level; l < b.L; ++l) {
        const BPLevel &B = b.lv[l];
        if (!B.coarse) continue;
        vc = std::max(vc, (size_t)B.V);
        mc = std::max(mc, (size_t)B.V * B.maxdeg);
        ac = std::max(ac, (size_t)B.V * acc_replicas(B.V) * K);
    }
    HBP_REQUIRE(vc > 0, "no segment-form BP level");
    for (int i = 0; i < 2; ++i) {
        b.ch_pl[i].alloc(c.pool, vc * K);
        b.ch_co[i].alloc(c.pool, vc * K);
        b.ch_pt[i].alloc(c.pool, vc);
        b.ch_acc[i].alloc(c.pool, ac);
        b.ch_nan[i].alloc(c.pool, vc);
        HBP_CHECK(hipMemsetAsync(b.ch_acc[i].p, 0, sizeof(unsigned long long) * ac, c.stream));
        HBP_CHECK(hipMemsetAsync(b.ch_nan[i].p, 0, sizeof(unsigned int) * vc, c.stream));
    }
    for (int i = 0; i < 3; ++i) b.ch_cand[i].alloc(c.pool, vc * K);
    b.ch_mt.alloc(c.pool, (mc + kNodeMaxDeg) * K);
}

// ---- RANSAC plane costs ahead (fuse 2), host side
// algorithmic bytes of one ahead launch: each pixel's position and two census words once
// (the sweeps re-gather the same right census rows), per sweep the planes, sums and NaN words
static double pc_ahead_bytes(long N, int V, int nr) { return 12.0 * N + nr * (V * (24.0 * P + 4)); }

static PipeArgs pipe_args(const BP &b, int l) {
    const BPLevel &B = b.lv[l];
    PipeArgs pa{};
    pa.cand = b.pp_cand.p;
    pa.pt = b.pp_pt.p;
    pa.acc = b.pp_acc.p;
    pa.nan = b.pp_nan.p;
    pa.rd_first = b.rd_first[l];
    pa.rd_count = b.rd_count[l];
    pa.cslice = (long)B.V * P;
    pa.aslice = (long)B.V * acc_replicas(B.V) * P;
    return pa;
}

// the slices of the largest coarse level's list, made on first use
static void pipe_alloc(Ctx &c, BP &b) {
    if (b.pp_cand.p) return;
    size_t nc = 1, na = 1, nv = 1, ns = 1;
    for (int l = b.min_level; l < b.L; ++l) {
        const BPLevel &B = b.lv[l];
        if (!B.coarse) continue;
        const size_t nR = (size_t)b.rcount[l];
        nc = std::max(nc, nR * B.V * P);
        na = std::max(na, nR * B.V * acc_replicas(B.V) * P);
        nv = std::max(nv, nR * B.V);
        ns = std::max(ns, (size_t)b.scount[l] * B.V * P);
    }
    b.pp_spc.alloc(c.pool, ns);
    b.pp_cand.alloc(c.pool, nc);
    b.pp_pt.alloc(c.pool, nv);
    b.pp_acc.alloc(c.pool, na);
    b.pp_nan.alloc(c.pool, nv);
}

// Level start (after the level init, before any node update of the level): every listed
// RANSAC sweep's candidates on the main stream, then on the second stream their sums cleared
// and their plane costs in launches of pipe_batch sweeps, each followed by an event.
// issues the ahead launches [b.pp_issued, upto) of the level b.pp_level on the second stream
static void pipe_issue(Ctx &c, BP &b, int upto) {
    upto = std::min(upto, b.pp_batches);
    if (b.pp_issued >= upto) return;
    const int l = b.pp_level;
    const SweepArgs a = sweep_args(b, l);
    const PipeArgs pa = pipe_args(b, l);
    const int nR = b.rcount[l];
    const int nb = b.pp_batch, f = b.pp_first;
    const int mode = !a.use_cost_vol ? kCostWta : (a.volf ? kCostF32 : (a.right ? kCostRight : kCostLeft));
    const int npc = b.lv[l].nseg;
    // at least one block per XCD: each XCD's blocks cost that XCD's share of the segments
    const int blocks = std::max(8, std::min(c.pipe_blocks, div_up(npc, 4)));
    const bool hb2 = c.pipe_hb == 2;
    hipStream_t s2 = c.side_stream();
    for (int bi = b.pp_issued; bi < upto; ++bi) {
        const int r0 = bi == 0 ? 0 : f + (bi - 1) * nb;
        const int nr = std::min(bi == 0 ? f : nb, nR - r0);
        ProfScope ps(c, "plane_cost_ahead_L" + std::to_string(l), pc_ahead_bytes(a.N, a.V, nr), true);
#define HBP_AHEAD(M)                                                                                             \
    do {                                                                                                         \
        if (hb2) HBP_LAUNCH(ps, (pc_ahead_kernel<M, 2>), dim3(blocks), dim3(256), 0, s2, a, pa, r0, nr, npc);    \
        else HBP_LAUNCH(ps, (pc_ahead_kernel<M, 4>), dim3(blocks), dim3(256), 0, s2, a, pa, r0, nr, npc);        \
    } while (0)
        if (mode == kCostLeft) HBP_AHEAD(kCostLeft);
        else if (mode == kCostRight) HBP_AHEAD(kCostRight);
        else if (mode == kCostF32) HBP_AHEAD(kCostF32);
        else HBP_AHEAD(kCostWta);
#undef HBP_AHEAD
        HBP_CHECK(hipGetLastError());
        HBP_CHECK(hipEventRecord(c.side_event(1 + (size_t)bi), s2));
    }
    b.pp_issued = upto;
}

static void pipe_level_start(Ctx &c, BP &b, const SweepArgs &a, int l) {
    const int nR = b.rcount[l];
    b.pp_level = l;
    b.pp_batches = 0;
    b.pp_issued = 0;
    b.pp_waited = -1;
    pipe_alloc(c, b);
    // the spatial sweeps' candidate superpixels (their seeds advance only in spatial sweeps'
    // node updates) need the level's parent structure
    const int nS = b.scount[l];
    if (nS > 0 && l + 1 < b.L) {
        hipLaunchKernelGGL(spatial_ahead_kernel, dim3(grid_for(c, (long)nS * a.V * P)), dim3(256), 0, c.stream, a,
                           b.pp_spc.p, nS);
        HBP_CHECK(hipGetLastError());
    }
    if (nR == 0) return;
    const PipeArgs pa = pipe_args(b, l);
    hipLaunchKernelGGL(draw_ahead_kernel, dim3(grid_for(c, (long)nR * a.V * P)), dim3(256), 0, c.stream, a, pa, nR);
    HBP_CHECK(hipGetLastError());
    hipStream_t s2 = c.side_stream();
    // the second stream waits for the candidates (and for every earlier use of the slices)
    HBP_CHECK(hipEventRecord(c.side_event(0), c.stream));
    HBP_CHECK(hipStreamWaitEvent(s2, c.side_event(0), 0));
    HBP_CHECK(hipMemsetAsync(pa.acc, 0, sizeof(unsigned long long) * (size_t)nR * pa.aslice, s2));
    HBP_CHECK(hipMemsetAsync(pa.nan, 0, sizeof(unsigned int) * (size_t)nR * a.V, s2));
    // the first launch costs listed sweeps [0, f), f = nb - 1 (1 for nb 1), each later one nb
    // sweeps: with nb 4 a launch holds one iteration's RANSAC sweeps (the level's first sweep,
    // costed by the level init, is not listed), so a run waits for one launch only
    const int nb = c.pipe_batch;
    const int f = nb > 1 ? nb - 1 : 1;
    b.pp_batch = nb;
    b.pp_first = f;
    b.pp_batches = nR <= f ? 1 : 1 + div_up(nR - f, nb);
    // with a window w only the first w + 1 launches are issued here, the rest as the main
    // stream's waits reach them (pipe_wait), so that the main stream's next launches are not
    // queued behind a burst of host-side launches
    pipe_issue(c, b, c.pipe_window > 0 ? c.pipe_window + 1 : b.pp_batches);
}

// a spatial sweep's candidate superpixels when its level start computed them, else null
static const int *pipe_spc(const BP &b, const Sweep &sw, size_t j) {
    if (sw.kind != kSpatial || b.pp_level != sw.level || b.sidx[j] < 0 || !b.pp_spc.p) return nullptr;
    return b.pp_spc.p + (size_t)b.sidx[j] * b.lv[sw.level].V * P;
}

// the main stream waits for the launch that holds listed sweep r's sums
static void pipe_wait(Ctx &c, BP &b, int r) {
    const int bi = r < b.pp_first ? 0 : 1 + (r - b.pp_first) / b.pp_batch;
    if (bi <= b.pp_waited) return;
    pipe_issue(c, b, c.pipe_window > 0 ? bi + 1 + c.pipe_window : b.pp_batches);
    HBP_CHECK(hipStreamWaitEvent(c.stream, c.side_event(1 + (size_t)bi), 0));
    b.pp_waited = bi;
}

// the main stream waits for everything issued on the second stream for this object
static void pipe_drain(Ctx &c, BP &b) {
    pipe_issue(c, b, b.pp_batches);
    if (b.pp_batches > 0 && b.pp_waited < b.pp_batches - 1) {
        HBP_CHECK(hipStreamWaitEvent(c.stream, c.side_event((size_t)b.pp_batches), 0));
        b.pp_waited = b.pp_batches - 1;
    }
}

// level init, and the ahead work of fuse 2
static void level_start(Ctx &c, BP &b, const SweepArgs &a, int l) {
    run_level_init(c, b, a, l);
    if (c.fuse == 2 && b.lv[l].coarse) pipe_level_start(c, b, a, l);
}

// sweeps j0 .. j0+n-1 (n >= 2) of one coarse level as a fused run (see ChainIO).  Launch u
// (sweep j0+u) writes its messages to MT(u), reads the previous launch's from MT(u-1), and
// recomputes the selection that ends sweep j0+u-1, which its node wavefronts write to Q(u-1).
// The candidates, points and sums of sweep j0+u: u = 0, the standard state (candidates and
// points from the previous node update or the level init, sums in the standard buffers);
// with fuse 2, the listed RANSAC sweeps' slices (costed on the second stream); otherwise
// CAND(u), PT(u) (drawn by launch u-1) and ACC(u) (costed by launch u, cleared by launch u+2).
// The closing node update writes the standard state.
static void run_chain(Ctx &c, BP &b, size_t j0, int n) {
    const Sweep &sw0 = b.sched[j0];
    const int l = sw0.level;
    const BPLevel &B = b.lv[l];
    const SweepArgs a = sweep_args(b, l);
    chain_alloc(c, b);
    if (sw0.first) level_start(c, b, a, l);
    const bool pipe = c.fuse == 2 && b.pp_level == l;
    const PipeArgs pa = pipe ? pipe_args(b, l) : PipeArgs{};
    const int uL = n - 1;
    // where sweep j0+u keeps its candidates, points and sums
    struct Loc {
        const float4 *cand;
        const uint4 *pt;
        unsigned long long *acc;
        unsigned int *nan;
        int cstride, astride;   // slot strides of cand and acc: K, or P for a listed sweep's slice
        int r;                  // index in the level's list when costed ahead, else -1
    };
    auto loc = [&](int u) -> Loc {
        const int r = pipe ? b.ridx[j0 + (size_t)u] : -1;
        if (u == 0) {
            if (r >= 0) return Loc{a.plane, a.point, pa.acc + r * pa.aslice, pa.nan + (long)r * a.V, K, P, r};
            return Loc{a.plane, a.point, b.acc.p, b.accnan.p, K, K, -1};
        }
        if (r >= 0) return Loc{pa.cand + r * pa.cslice, pa.pt + (long)r * a.V, pa.acc + r * pa.aslice,
                               pa.nan + (long)r * a.V, P, P, r};
        const int i = u % 3;
        return Loc{b.ch_cand[i].p, b.ch_pt[u % 2].p, i == 0 ? b.acc.p : b.ch_acc[i - 1].p,
                   i == 0 ? b.accnan.p : b.ch_nan[i - 1].p, K, K, -1};
    };
    // a standard-stride sum buffer is cleared by the second launch after the one that read it
    // (in the one launch both ends of an edge and each node may still read it); the slices of
    // the listed sweeps are cleared at the level start
    auto MT = [&](int u) { return u % 2 == 0 ? b.msg_t.p : b.ch_mt.p; };
    // tiny levels (fuse 2): the launches after the first and the closing node update in one
    // workgroup (run_kernel), when the run's sweeps after the first are all costed ahead and
    // the sweep after the run is a spatial one or none
    int next_kind_l = kNone;
    {
        const size_t jl = j0 + (size_t)uL;
        if (!b.sched[jl].last && jl + 1 < b.sched.size()) next_kind_l = b.sched[jl + 1].kind;
    }
    bool tiny = pipe && a.V <= RUN_V && a.E <= c.run_max_e && a.E <= RUN_MT && B.maxdeg <= kNodeMaxDeg &&
                (long)a.V * B.maxdeg <= RUN_MT && uL >= 1 && uL <= RUN_NR &&
                (next_kind_l == kNone || next_kind_l == kSpatial) && loc(0).r < 0;
    for (int u = 1; tiny && u <= uL; ++u) tiny = loc(u).r >= 0 && loc(u).r == loc(1).r + u - 1;
    for (int u = 0; u <= uL; ++u) {
        if (tiny && u == 1) {
            const Loc c0l = loc(0);
            pipe_wait(c, b, loc(uL).r);
            RunArgs ra{};
            ra.acc0 = c0l.acc;
            ra.nan0 = c0l.nan;
            ra.mt0 = MT(0);
            ra.rloc = B.rloc.p;
            ra.n = uL;
            ra.r1 = loc(1).r;
            ra.prev_spatial = sw0.kind == kSpatial ? 1 : 0;
            ra.pa = pa;
            ProfScope psr(c, "run_L" + std::to_string(l), uL * (edge_bytes(a.E, a.V) + node_bytes(a.E, a.V)), true);
            HBP_LAUNCH(psr, run_kernel, dim3(1), dim3(1024), RUN_LDS_BYTES, c.stream, a, ra);
            HBP_CHECK(hipGetLastError());
            return;
        }
        const Loc cu = loc(u);
        SweepArgs ax = a;
        ax.acc = cu.acc;
        ax.accnan = cu.nan;
        ax.msg_t = MT(u);
        ChainIO io{};
        io.rloc = B.rloc.p;
        // fuse 1: launch u draws sweep j0+u+1's candidates (the listed ones exist already)
        const Loc nx = u < uL ? loc(u + 1) : cu;
        io.draw = (u < uL && nx.r < 0) ? 1 : 0;
        io.cand_next = (float4 *)nx.cand;
        io.pt_next = (uint4 *)nx.pt;
        // plane costs in this launch unless costed ahead or by the level init
        const bool costs = cu.r < 0 && !(u == 0 && sw0.first);
        if (u == 0) {
            // a spatial first sweep (fuse 2): its own launch, its node update done by launch 1
            if (sw0.kind == kSpatial) {
                ax.spc = pipe_spc(b, sw0, j0);
                launch_cost_edge<P>(c, ax, B, l, 2 * P, true, true, a.E > 0);
            }
            else launch_chain<true>(c, ax, io, B, l, costs);
            continue;
        }
        io.prev_spatial = (u == 1 && sw0.kind == kSpatial) ? 1 : 0;
        const Loc cp = loc(u - 1);
        if (cp.r >= 0) pipe_wait(c, b, cp.r);
        ax.plane = (float4 *)cu.cand;   // segments (standard stride only)
        io.pl_in = u == 1 ? a.plane : b.ch_pl[u % 2].p;   // Q(u-2)
        io.co_in = u == 1 ? a.cost : b.ch_co[u % 2].p;
        io.cand_in = cp.cand;
        io.cs_in = cp.cstride;
        io.acc_in = cp.acc;
        io.as_in = cp.astride;
        io.nan_in = cp.nan;
        io.mt_in = MT(u - 1);
        io.cand_cur = cu.cand;
        io.cs_cur = cu.cstride;
        io.pt_cur = cu.pt;
        io.pl_out = b.ch_pl[(u - 1) % 2].p;   // Q(u-1)
        io.co_out = b.ch_co[(u - 1) % 2].p;
        io.materialize = 1;
        if (u >= 2) {
            const Loc cc = loc(u - 2);
            if (cc.r < 0) {
                io.acc_clear = cc.acc;
                io.nan_clear = cc.nan;
            }
        }
        launch_chain<false>(c, ax, io, B, l, costs);
    }
    // the node update of the run's last sweep, into the standard state
    const size_t jL = j0 + (size_t)uL;
    const Sweep &swL = b.sched[jL];
    int next_kind = kNone;
    float ndz = 0.f, ndn = 0.f;
    if (!swL.last && jL + 1 < b.sched.size()) {
        const Sweep &nx = b.sched[jL + 1];
        next_kind = nx.kind;
        ndz = nx.dz;
        ndn = nx.dn;
    }
    const Loc cl = loc(uL), cp = loc(uL - 1);
    if (cl.r >= 0) pipe_wait(c, b, cl.r);
    SweepArgs an = a;
    an.acc = cl.acc;
    an.accnan = cl.nan;
    an.msg_t = MT(uL);
    NodeIO nio{b.ch_pl[(uL - 1) % 2].p, b.ch_co[(uL - 1) % 2].p, cl.cand, cl.pt, nullptr, nullptr, cl.cstride,
               cl.astride};
    if (cp.r < 0) {
        nio.acc2 = cp.acc;
        nio.nan2 = cp.nan;
    }
    {
        ProfScope psn(c, "node_L" + std::to_string(l), node_bytes(a.E, a.V), true);
        HBP_LAUNCH(psn, node_io_kernel, dim3(div_up((long)a.V * 16, 256)), dim3(256), 0, c.stream, an, nio, swL.kind,
                   next_kind, ndz, ndn);
    }
    HBP_CHECK(hipGetLastError());
}

int bp_step(Ctx &c, BP &b, int max_sweeps) {
    int done = 0;
    while (done < max_sweeps && b.next < b.sched.size()) {
        const Sweep &sw = b.sched[b.next];
        const BPLevel &B = b.lv[sw.level];
        // a fused run: this sweep and the RANSAC sweeps after it on its level (a spatial first
        // sweep only with fuse 2, whose RANSAC candidates exist before the run)
        int n = 1;
        const bool ahead = c.fuse == 2 && b.pp_level == sw.level;
        if (c.fuse && B.coarse && B.rloc.p && (sw.kind != kSpatial || ahead)) {
            while (done + n < max_sweeps && b.next + n < b.sched.size()) {
                const Sweep &nx = b.sched[b.next + n];
                if (nx.first || nx.level != sw.level || nx.kind != kRansac) break;
                ++n;
            }
        }
        if (n >= 2) {
            run_chain(c, b, b.next, n);
        } else {
            int next_kind = kNone;
            float ndz = 0.f, ndn = 0.f;
            if (!sw.last && b.next + 1 < b.sched.size()) {
                const Sweep &nx = b.sched[b.next + 1];
                next_kind = nx.kind;
                ndz = nx.dz;
                ndn = nx.dn;
            }
            run_sweep_split(c, b, sw, next_kind, ndz, ndn);
        }
        b.cur = sw.level;
        b.next += (size_t)n;
        done += n;
        b.sweeps += n;
    }
    if (b.next >= b.sched.size()) b.cur = b.min_level - 1;
    // nothing of this object stays queued on the second stream past the call
    pipe_drain(c, b);
    return done;
}

// ----------------------------------------------------------------------------------------
// Batched sweeps, host side.  A BatchRunner keeps a device table of SweepArgs (one slot of
// kBatchLevels entries per BP object, pointers stored relative to the table, see
// rebased_args) and runs rounds: one round is one sweep of every BP object given to it, in
// one spatial + two plane-cost/message launches (segment form, run form) + one node launch.
// Per pair the kernels do exactly what the one-pair path does (same work items, same
// fixed-point sums), so every pair's results are bit-identical to its own run.
// ----------------------------------------------------------------------------------------
void BatchRunner::init(Ctx &c, int slots) {
    nslots = slots;
    table.alloc(c.pool, (size_t)slots * kBatchLevels * sizeof(SweepArgs));
    used.assign(slots, 0);
    mode = -1;
    HBP_CHECK(hipHostMalloc((void **)&ring, sizeof(BatchItem) * (size_t)kRing * slots, hipHostMallocDefault));
    for (int i = 0; i < kRing; ++i) HBP_CHECK(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
}

BatchRunner::~BatchRunner() {
    for (int i = 0; i < kRing; ++i)
        if (ev[i]) {
            if (ev_live[i]) (void)hipEventSynchronize(ev[i]);
            (void)hipEventDestroy(ev[i]);
        }
    if (ring) (void)hipHostFree(ring);
}

static int cost_mode(const BP &b) {
    return !b.use_cost_vol ? kCostWta : (b.vol->f32src ? kCostF32 : (b.vol->right ? kCostRight : kCostLeft));
}

int BatchRunner::add(Ctx &c, BP &b) {
    const int m = cost_mode(b);
    int live = 0;
    for (int u : used) live += u;
    HBP_REQUIRE(live == 0 || mode < 0 || m == mode, "a batch must use one cost form (view and use_cost_vol)");
    mode = m;
    HBP_REQUIRE(b.L <= kBatchLevels, "pyramid has more levels than a batch slot holds");
    int slot = -1;
    for (int i = 0; i < nslots; ++i)
        if (!used[i]) {
            slot = i;
            break;
        }
    HBP_REQUIRE(slot >= 0, "no free batch slot");
    used[slot] = 1;
    std::vector<SweepArgs> h(kBatchLevels);
    SweepArgs *dt = (SweepArgs *)table.p;
    for (int l = b.min_level; l < b.L; ++l) {
        SweepArgs a = sweep_args(b, l);
#define HBP_RELATIVE(f) a.f = (decltype(a.f))((uintptr_t)a.f - (uintptr_t)dt);
        SWEEP_ARGS_POINTERS(HBP_RELATIVE)
#undef HBP_RELATIVE
        h[l] = a;
    }
    HBP_CHECK(hipMemcpyAsync(dt + (size_t)slot * kBatchLevels, h.data(), sizeof(SweepArgs) * kBatchLevels,
                             hipMemcpyHostToDevice, c.stream));
    return slot;
}

void BatchRunner::remove(int slot) {
    if (slot >= 0 && slot < nslots) used[slot] = 0;
}

template <int HB>
static void launch_batch_coarse(Ctx &c, int mode, int blocks, const SweepArgs *dargs, const BatchItem *items, int n) {
    if (mode == kCostLeft)
        hipLaunchKernelGGL((cost_edge_batch_kernel<kCostLeft, true, HB>), dim3(blocks), dim3(256), 0, c.stream, dargs,
                           items, n);
    else if (mode == kCostRight)
        hipLaunchKernelGGL((cost_edge_batch_kernel<kCostRight, true, HB>), dim3(blocks), dim3(256), 0, c.stream,
                           dargs, items, n);
    else if (mode == kCostF32)
        hipLaunchKernelGGL((cost_edge_batch_kernel<kCostF32, true, HB>), dim3(blocks), dim3(256), 0, c.stream, dargs,
                           items, n);
    else
        hipLaunchKernelGGL((cost_edge_batch_kernel<kCostWta, true, HB>), dim3(blocks), dim3(256), 0, c.stream, dargs,
                           items, n);
}

bool BatchRunner::plan(const std::vector<std::pair<BP *, int>> &active, RoundPlan &rp) const {
    // items: the segment-form ones first, then the run-form ones; each form is one cost launch
    // with its own block numbering, the node launch covers both
    std::vector<BatchItem> part[2];
    for (const auto &pr : active) {
        BP &b = *pr.first;
        const size_t i = b.next;
        if (i >= b.sched.size()) continue;
        const Sweep &sw = b.sched[i];
        const BPLevel &B = b.lv[sw.level];
        if (sw.first) rp.inits.push_back({&b, sw.level});
        BatchItem it{};
        it.args = pr.second * kBatchLevels + sw.level;
        it.kind = sw.kind;
        it.next_kind = kNone;
        if (!sw.last && i + 1 < b.sched.size()) {
            it.next_kind = b.sched[i + 1].kind;
            it.ndz = b.sched[i + 1].dz;
            it.ndn = b.sched[i + 1].dn;
        }
        it.coarse = B.coarse ? 1 : 0;
        it.spatial = sw.kind == kSpatial ? 1 : 0;
        it.npc = sw.first ? 0 : (B.coarse ? B.nseg : div_up(b.N, CHUNK));
        it.pc_blocks = B.coarse ? div_up(it.npc, 4) : it.npc;
        // coarse levels: the messages share the segments' launch; fine levels: their own
        const int eb = (int)div_up(div_up(B.E, EDGES_PER_WAVE), 4);
        it.edge_blocks = B.coarse ? eb : 0;
        it.msg_blocks = B.coarse ? 0 : eb;
        int &blocks = B.coarse ? rp.coarse_blocks : rp.fine_blocks;
        it.blk_begin = blocks;
        blocks += (int)round_up(it.pc_blocks + it.edge_blocks, 8);
        it.node_blocks = div_up((long)B.V * 16, 256);
        it.sp_blocks = (!B.coarse && sw.kind == kSpatial && !sw.first) ? div_up(B.V, 256) : 0;
        const double pcb = it.npc > 0 ? plane_cost_bytes(b.N, B.V, P) : 0.0;
        const double eb_bytes = B.E > 0 ? edge_bytes(B.E, B.V) : 0.0;
        if (B.coarse) {
            rp.coarse_bytes += pcb + eb_bytes;
        } else {
            rp.fine_bytes += pcb;
            rp.msg_bytes += eb_bytes;
        }
        rp.node_bytes += node_bytes(B.E, B.V);
        if (it.sp_blocks > 0) rp.sp_bytes += (16.0 * P + 28) * B.V;
        part[B.coarse ? 0 : 1].push_back(it);
        b.cur = sw.level;
        ++b.next;
        ++b.sweeps;
        if (b.next >= b.sched.size()) b.cur = b.min_level - 1;
    }
    rp.ncoarse = (int)part[0].size();
    rp.nfine = (int)part[1].size();
    const int count = rp.ncoarse + rp.nfine;
    if (count == 0) return false;
    HBP_REQUIRE(count <= nslots, "more batch items than slots");
    rp.items.reserve(count);
    for (int f = 0; f < 2; ++f)
        for (BatchItem &it : part[f]) {
            it.node_begin = rp.node_blocks;
            rp.node_blocks += it.node_blocks;
            it.msg_begin = rp.msg_blocks;
            rp.msg_blocks += it.msg_blocks;
            it.sp_begin = rp.sp_blocks;
            rp.sp_blocks += it.sp_blocks;
            rp.items.push_back(it);
        }
    return true;
}

void BatchRunner::launch(Ctx &c, const RoundPlan &rp, const BatchItem *d_items) const {
    for (const auto &in : rp.inits) run_level_init(c, *in.first, sweep_args(*in.first, in.second), in.second);
    const int count = rp.ncoarse + rp.nfine;
    const SweepArgs *dargs = (const SweepArgs *)table.p;
    const BatchItem *itc = d_items, *itf = d_items + rp.ncoarse;
    if (rp.sp_blocks > 0) {
        ProfScope ps(c, "batch_spatial", rp.sp_bytes);
        hipLaunchKernelGGL(spatial_batch_kernel, dim3(rp.sp_blocks), dim3(256), 0, c.stream, dargs, itc, count);
    }
    if (rp.coarse_blocks > 0) {
        ProfScope ps(c, "batch_plane_cost_coarse", rp.coarse_bytes);
        if (c.batch_hb == 4) launch_batch_coarse<4>(c, mode, rp.coarse_blocks, dargs, itc, rp.ncoarse);
        else launch_batch_coarse<2>(c, mode, rp.coarse_blocks, dargs, itc, rp.ncoarse);
    }
    if (rp.fine_blocks > 0) {
        ProfScope ps(c, "batch_plane_cost_fine", rp.fine_bytes);
        const dim3 g(rp.fine_blocks), blk(256);
        const int nf = rp.nfine;
        if (mode == kCostLeft)
            hipLaunchKernelGGL((cost_edge_batch_kernel<kCostLeft, false>), g, blk, PC_LDS_BYTES, c.stream, dargs, itf, nf);
        else if (mode == kCostRight)
            hipLaunchKernelGGL((cost_edge_batch_kernel<kCostRight, false>), g, blk, PC_LDS_BYTES, c.stream, dargs, itf, nf);
        else if (mode == kCostF32)
            hipLaunchKernelGGL((cost_edge_batch_kernel<kCostF32, false>), g, blk, PC_LDS_BYTES, c.stream, dargs, itf, nf);
        else
            hipLaunchKernelGGL((cost_edge_batch_kernel<kCostWta, false>), g, blk, PC_LDS_BYTES, c.stream, dargs, itf, nf);
    }
    if (rp.msg_blocks > 0) {
        ProfScope ps(c, "batch_edge", rp.msg_bytes);
        hipLaunchKernelGGL(edge_batch_kernel, dim3(rp.msg_blocks), dim3(256), 0, c.stream, dargs, itc, count);
    }
    {
        ProfScope ps(c, "batch_node", rp.node_bytes);
        hipLaunchKernelGGL(node_batch_kernel, dim3(rp.node_blocks), dim3(256), 0, c.stream, dargs, itc, count);
    }
    HBP_CHECK(hipGetLastError());
}

void BatchRunner::round(Ctx &c, const std::vector<std::pair<BP *, int>> &active) {
    RoundPlan rp;
    if (!plan(active, rp)) return;
    const int count = rp.ncoarse + rp.nfine;
    if (items.n < (size_t)count) items.alloc(c.pool, std::max<size_t>(count, 64));
    // staged in the pinned ring (the copy runs in stream order, after the previous round's
    // kernels have read the device items); the slot's last copy must have run before reuse
    const int slot = ring_pos;
    ring_pos = (ring_pos + 1) % kRing;
    if (ev_live[slot]) HBP_CHECK(hipEventSynchronize(ev[slot]));
    BatchItem *src = ring + (size_t)slot * nslots;
    std::memcpy(src, rp.items.data(), sizeof(BatchItem) * count);
    HBP_CHECK(hipMemcpyAsync(items.p, src, sizeof(BatchItem) * count, hipMemcpyHostToDevice, c.stream));
    HBP_CHECK(hipEventRecord(ev[slot], c.stream));
    ev_live[slot] = true;
    launch(c, rp, items.p);
}

// The whole schedule of rounds captured in one HIP graph and replayed once: every round's items
// go to the device in one copy before the capture, and the graph holds kernel launches only
// (the level inits clear their messages with a kernel, not a memset).  The same launches in
// the same order as round by round, so the same bits.
static void bp_run_batch_graph(Ctx &c, BatchRunner &r, const std::vector<std::pair<BP *, int>> &active) {
    std::vector<RoundPlan> rounds;
    size_t total = 0;
    for (;;) {
        RoundPlan rp;
        if (!r.plan(active, rp)) break;
        total += rp.items.size();
        rounds.push_back(std::move(rp));
    }
    if (rounds.empty()) return;
    std::vector<BatchItem> all;
    all.reserve(total);
    for (const RoundPlan &rp : rounds) all.insert(all.end(), rp.items.begin(), rp.items.end());
    DevBuf<BatchItem> d_all(c.pool, total);
    HBP_CHECK(hipMemcpyAsync(d_all.p, all.data(), sizeof(BatchItem) * total, hipMemcpyHostToDevice, c.stream));
    HBP_CHECK(hipStreamSynchronize(c.stream));
    // the profiler's events stay out of the graph
    const bool prof = c.prof.on;
    c.prof.on = false;
    hipGraph_t g = nullptr;
    HBP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    try {
        size_t off = 0;
        for (const RoundPlan &rp : rounds) {
            r.launch(c, rp, d_all.p + off);
            off += rp.items.size();
        }
    } catch (...) {
        (void)hipStreamEndCapture(c.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        c.prof.on = prof;
        throw;
    }
    c.prof.on = prof;
    HBP_CHECK(hipStreamEndCapture(c.stream, &g));
    hipGraphExec_t ge = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        HBP_CHECK(ie);
    }
    const hipError_t le = hipGraphLaunch(ge, c.stream);
    const hipError_t se = le == hipSuccess ? hipStreamSynchronize(c.stream) : le;
    (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    HBP_CHECK(se);
}

// Runs the remaining sweeps of several BP objects together (round i carries sweep next+i of
// every pair that has one), then the map of each.
void bp_run_batch(Ctx &c, const std::vector<BP *> &bps) {
    const int n = (int)bps.size();
    if (n == 0) return;
    for (BP *bp : bps)
        HBP_REQUIRE(cost_mode(*bp) == cost_mode(*bps[0]), "a batch must use one cost form (view and use_cost_vol)");
    BatchRunner r;
    r.init(c, n);
    std::vector<std::pair<BP *, int>> active;
    for (BP *bp : bps) active.push_back({bp, r.add(c, *bp)});
    if (c.batch_graph) {
        bp_run_batch_graph(c, r, active);
    } else {
        for (;;) {
            bool any = false;
            for (BP *bp : bps) any |= bp->next < bp->sched.size();
            if (!any) break;
            r.round(c, active);
        }
    }
    for (BP *bp : bps) bp_map(c, *bp);
    HBP_CHECK(hipStreamSynchronize(c.stream));
}

void bp_map(Ctx &c, BP &b) {
    BPLevel &B = b.lv[b.min_level];
    b.best.alloc(c.pool, B.V);
    hipLaunchKernelGGL(map_best_kernel, dim3(grid_for(c, B.V)), dim3(256), 0, c.stream, B.V, B.belief.p, B.plane.p,
                       b.best.p);
    hipLaunchKernelGGL(map_paint_kernel, dim3(grid_for(c, b.N)), dim3(256), 0, c.stream, B.pxy.p, B.pcl.p, b.N,
                       b.best.p, b.W, b.D, b.disp.p, b.plane3.p);
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
