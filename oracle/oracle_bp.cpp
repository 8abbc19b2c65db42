// Oracle (test infrastructure only): stage (c), hierarchical plane-particle BP on the pyramid.
// Restates hbp_isp.cu / hbp_isp.hpp of the reference (citations per function), with the
// reference's unordered float reductions replaced by these pinned orders:
//   * plane-cost cluster sums: int64 fixed point (oracle_math.h, fixed_shift), any order;
//   * dc0 channel sum: |dL| + |da| + |db| left to right from 0; dc1: dx^2 + dy^2 from 0;
//   * belief = ((cost + m_e0) + m_e1) + ... over the node's edges in edge (CSR) order;
//   * map_msg's atomicMinFloat: a left fold over p = 0..P-1 (amin_fold).
// Top-level planes (never written by the reference, hbp_isp.hpp:340) are (0, 0, 0, 0).
// Checked (tests/test_oracle_bp_statement.py) against an independent numpy statement of
// hbp_isp.cu after every sweep, bit for bit.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "hbp_oracle.h"
#include "oracle_math.h"
#include "oracle_pyramid.h"

namespace orc {

struct F4 {
    float x, y, z, w;
};

static const int P = 4;       // hbp_isp.hpp:279
static const int K = 3 * P;   // hbp_isp.hpp:280

struct BPLevel {
    int V = 0, Vh = 1;
    long E = 0;
    const Level *lev = nullptr;
    std::vector<float> mu;        // normalised means [5][V]
    std::vector<int> off;         // CSR offsets [V+1]
    std::vector<int> rev;         // reverse edge [E]
    std::vector<float> dc0, dc1;  // [E]
    std::vector<int> px;          // pixel ids sorted by (cluster, id) [N]
    std::vector<int> pix_off;     // [V+1]
    std::vector<int> child;       // this level's nodes sorted by parent [V]
    std::vector<int> child_off;   // [Vh+1]
    std::vector<F4> plane;        // [K][V]
    std::vector<float> cost, belief;  // [K][V]
    std::vector<int> point;       // [V][3] pixel ids of the last RANSAC triple
};

struct BP {
    const Pyramid *pyr;
    const uint8_t *vol;
    const float *volf;   // a given f32 volume (MC-CNN), used instead of vol when set
    const float *dwta;
    int W, H, D, L, min_level, use_cost_vol;
    long N;
    float tau, lambda, tau_s;
    int fx_shift;
    std::vector<BPLevel> lv;
    std::vector<float> msg, msg_t;   // [K][E] of the current level
    std::vector<uint32_t> s_ransac, s_neighbor, s_child;
    // schedule
    int cur = -1;
    bool inited = false;
    int outer = 0, plane_iter = 0, iter_bp = 0;
    float max_dz = 0, max_dn = 0;
    int sweeps = 0;
};

// HBP_ISP ctor + setup_D4 (hbp_isp.hpp:288-386, hbp_isp.cu:871-891) and the per-level
// prologue of run_ms (hbp_isp.cu:1319-1325) for the levels BP reads (min_level .. L-1).
static void prologue(BP &b) {
    const Pyramid &P_ = *b.pyr;
    b.L = (int)P_.levels.size();
    b.lv.assign(b.L, BPLevel());
    std::vector<int> im_clus(b.N);
    std::iota(im_clus.begin(), im_clus.end(), 0);
    for (int l = 0; l < b.L; ++l) {
        const Level &Lv = P_.levels[l];
        BPLevel &B = b.lv[l];
        B.lev = &Lv;
        B.V = Lv.V;
        B.E = Lv.E;
        // the top level's parents: one cluster, or target_clus of them (unused by BP, which
        // does not run on the top level, but its children are grouped like every level's)
        B.Vh = (l < b.L - 1) ? P_.levels[l + 1].V : P_.top;
        if (l >= b.min_level) {
            int V = B.V;
            long E = B.E;
            // set_D4_k + thrust::sort_by_key (stable): pixels by (cluster, pixel id)
            B.pix_off.assign(V + 1, 0);
            for (long i = 0; i < b.N; ++i) B.pix_off[im_clus[i] + 1]++;
            for (int c = 0; c < V; ++c) B.pix_off[c + 1] += B.pix_off[c];
            B.px.assign(b.N, 0);
            std::vector<int> fill(B.pix_off.begin(), B.pix_off.end() - 1);
            for (long i = 0; i < b.N; ++i) B.px[fill[im_clus[i]]++] = (int)i;
            // reverse_bd_id (hbp_isp.cu:1003-1019, 140-175): CSR by source, reverse search
            B.off.assign(V + 1, 0);
            for (long e = 0; e < E; ++e) B.off[Lv.bd[2 * e] + 1]++;
            for (int c = 0; c < V; ++c) B.off[c + 1] += B.off[c];
            B.rev.assign(E, -1);
            for (long e = 0; e < E; ++e) {
                int s = Lv.bd[2 * e], t = Lv.bd[2 * e + 1];
                for (int r = B.off[t]; r < B.off[t + 1]; ++r)
                    if (Lv.bd[2 * r + 1] == s) {
                        B.rev[e] = r;
                        break;
                    }
            }
            // compute_dist_color_pos (hbp_isp.cu:952-970, 78-126): sum_to_mean in place,
            // then dc0 = exp(-(|dL|+|da|+|db|)/gamma) in FP64, dc1 = sqrtf(dx^2 + dy^2)
            B.mu.assign(5L * V, 0.f);
            for (int c = 0; c < 5; ++c)
                for (int i = 0; i < V; ++i)
                    B.mu[(long)c * V + i] = Lv.mean[(long)c * V + i] / Lv.mean[5L * V + i];
            B.dc0.assign(E, 0.f);
            B.dc1.assign(E, 0.f);
            for (long e = 0; e < E; ++e) {
                int s = Lv.bd[2 * e], t = Lv.bd[2 * e + 1];
                float dcol = 0.f;
                for (int c = 0; c < 3; ++c) dcol += std::fabs(B.mu[(long)c * V + s] - B.mu[(long)c * V + t]);
                float dpos = 0.f;
                for (int c = 3; c < 5; ++c) {
                    float d = B.mu[(long)c * V + s] - B.mu[(long)c * V + t];
                    dpos += d * d;
                }
                B.dc0[e] = (float)pin_exp((double)(-dcol) / 20.0);
                B.dc1[e] = std::sqrt(dpos);
            }
            // sort_clus (hbp_isp.cu:985-1001): children grouped by parent, ascending id
            B.child_off.assign(B.Vh + 1, 0);
            for (int i = 0; i < V; ++i) B.child_off[Lv.clus[i] + 1]++;
            for (int c = 0; c < B.Vh; ++c) B.child_off[c + 1] += B.child_off[c];
            B.child.assign(V, 0);
            std::vector<int> cf(B.child_off.begin(), B.child_off.end() - 1);
            for (int i = 0; i < V; ++i) B.child[cf[Lv.clus[i]]++] = i;
            B.plane.assign((long)K * V, F4{0, 0, 0, 0});
            B.cost.assign((long)K * V, 0.f);
            B.belief.assign((long)K * V, 0.f);
            B.point.assign(3L * V, 0);
        }
        if (l < b.L - 1)
            for (long i = 0; i < b.N; ++i) im_clus[i] = Lv.clus[im_clus[i]];
    }
}

static inline float plane_px_cost(const BP &b, const F4 &pl, int pxid) {
    int x = pxid % b.W, y = pxid / b.W;
    float df = pl.x * (float)x + pl.y * (float)y + pl.z;
    if (!b.use_cost_vol) {
        // hbp_isp.cu:189-204
        float Dw = b.dwta[pxid];
        float c = std::fabs(Dw - df) <= 1.f ? 0.f : 1.f;
        return b.lambda * c;
    }
    // hbp_isp.cu:205-227 (P5: NaN -> 0, saturation; d + 1 evaluated without overflow)
    int d = f2i_rd(df);
    float c;
    if (d < 0 || (long long)d + 1 >= b.D) {
        c = b.tau;
    } else {
        float c0 = b.volf ? b.volf[(long)d * b.N + pxid] : (float)b.vol[(long)d * b.N + pxid];
        float c1 = b.volf ? b.volf[(long)(d + 1) * b.N + pxid] : (float)b.vol[(long)(d + 1) * b.N + pxid];
        c = ((float)(d + 1) - df) * c0 + (df - (float)d) * c1;
    }
    return b.lambda * c;
}

// compute_plane_cost (hbp_isp.cu:1159-1174) with read_plane_cost (177-228); cluster sums in
// int64 fixed point.  The sorted pixel list is cut into equal ranges, one per thread; a thread
// folds runs of one cluster locally and adds each run to the cluster's total atomically (the
// sums are exact integers, so the split does not change them).
static void plane_cost(BP &b, int l, int s_c) {
    BPLevel &B = b.lv[l];
    std::vector<int64_t> acc((size_t)B.V * P, 0);
    std::vector<int> nanf((size_t)B.V * P, 0);
    std::vector<int> cl_of(b.N);
    for (int c = 0; c < B.V; ++c)
        for (int j = B.pix_off[c]; j < B.pix_off[c + 1]; ++j) cl_of[j] = c;
    const long N = b.N;
#pragma omp parallel
    {
        int nt = 1, t = 0;
#ifdef _OPENMP
        nt = omp_get_num_threads();
        t = omp_get_thread_num();
#endif
        long j0 = N * t / nt, j1 = N * (t + 1) / nt;
        int crun = -1;
        int64_t racc[P] = {0, 0, 0, 0};
        int rnan[P] = {0, 0, 0, 0};
        F4 pl[P];
        auto flush = [&]() {
            if (crun < 0) return;
            for (int p = 0; p < P; ++p) {
                int64_t v = racc[p];
#pragma omp atomic
                acc[(size_t)crun * P + p] += v;
                if (rnan[p]) {
#pragma omp atomic write
                    nanf[(size_t)crun * P + p] = 1;
                }
            }
        };
        for (long j = j0; j < j1; ++j) {
            int c = cl_of[j];
            if (c != crun) {
                flush();
                crun = c;
                for (int p = 0; p < P; ++p) {
                    racc[p] = 0;
                    rnan[p] = 0;
                    pl[p] = B.plane[(long)(s_c + p) * B.V + c];
                }
            }
            int pxid = B.px[j];
            for (int p = 0; p < P; ++p) {
                float v = plane_px_cost(b, pl[p], pxid);
                if (v != v) rnan[p] = 1;
                else racc[p] += to_fixed(v, b.fx_shift);
            }
        }
        flush();
    }
    for (int c = 0; c < B.V; ++c)
        for (int p = 0; p < P; ++p)
            B.cost[(long)(s_c + p) * B.V + c] = from_fixed(acc[(size_t)c * P + p], nanf[(size_t)c * P + p], b.fx_shift);
}

// init_bp (hbp_isp.cu:930-950) with copy_prior_plane (56-73) and init_rand_seed_k (570-591).
static void init_level(BP &b, int l) {
    BPLevel &B = b.lv[l];
    BPLevel &Hh = b.lv[l + 1];
    b.msg.assign((long)K * B.E, 0.f);
    b.msg_t.assign((long)K * B.E, 0.f);
#pragma omp parallel for schedule(static)
    for (int id = 0; id < B.V; ++id) {
        int cc = B.lev->clus[id];
        for (int p = 0; p < P; ++p) {
            F4 pl = Hh.plane[(long)p * Hh.V + cc];
            B.plane[(long)p * B.V + id] = pl;
            B.plane[(long)(p + P) * B.V + id] = pl;
            B.plane[(long)(p + 2 * P) * B.V + id] = pl;
        }
    }
    plane_cost(b, l, 0);
    plane_cost(b, l, P);
    B.belief = B.cost;
    b.s_ransac.assign(B.V, 0);
    b.s_neighbor.assign(B.V, 0);
    b.s_child.assign(B.Vh, 0);
#pragma omp parallel for schedule(static)
    for (int id = 0; id < B.V; ++id) {
        b.s_neighbor[id] = (uint32_t)B.off[id];
        b.s_ransac[id] = (uint32_t)B.pix_off[id];
    }
    for (int id = 0; id < B.Vh; ++id) b.s_child[id] = (uint32_t)B.child_off[id];
    b.iter_bp = 10 + l * 10;                                      // hbp_isp.cu:1289-1290
    b.max_dz = (float)b.D / 2.0f * (float)l / (float)b.L;         // hbp_isp.cu:1291
    b.max_dn = 1.0f * (float)l / (float)b.L;                      // hbp_isp.cu:1292
    b.outer = 0;
    b.plane_iter = 0;
}

// make_three_diff_id (hbp_isp.cu:504-529)
static inline void three_diff(int &p0, int &p1, int &p2, int tn) {
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

// fit_plane_from_points (hbp_isp.cu:531-548)
static inline F4 fit_plane(const F4 &A, const F4 &Bp, const F4 &C) {
    F4 a{Bp.x - A.x, Bp.y - A.y, Bp.z - A.z, Bp.w - A.w};
    F4 bb{C.x - A.x, C.y - A.y, C.z - A.z, C.w - A.w};
    float nx = a.y * bb.z - a.z * bb.y;
    float ny = a.z * bb.x - a.x * bb.z;
    float nz = a.x * bb.y - a.y * bb.x;
    nx /= nz;
    ny /= nz;
    nz /= nz;
    if (std::fabs(nz - 1.f) > FLT_EPSILON) return F4{0.f, 0.f, (A.z + Bp.z + C.z) / 3.f, 0.f};
    return F4{-nx, -ny, nx * A.x + ny * A.y + A.z, 0.f};
}

// unsigned modulo as the kernels evaluate it; a zero divisor (clusters of 1 or 2 pixels,
// hbp_isp.cu:611-613) yields 0, and make_three_diff_id then overrides the value anyway.
static inline uint32_t umod(uint32_t s, int m) { return (uint32_t)m == 0u ? 0u : s % (uint32_t)m; }

// ransac_candidate_plane (hbp_isp.cu:592-640)
static void ransac(BP &b, int l) {
    BPLevel &B = b.lv[l];
#pragma omp parallel for schedule(static)
    for (int id = 0; id < B.V; ++id) {
        int n = B.pix_off[id + 1] - B.pix_off[id];
        int pos = B.pix_off[id];
        uint32_t s = b.s_ransac[id];
        for (int p = 2 * P; p < 3 * P; ++p) {
            s = lcg(s);
            int p0 = (int)umod(s, n);
            s = lcg(s);
            int p1 = (int)umod(s, n - 1);
            s = lcg(s);
            int p2 = (int)umod(s, n - 2);
            three_diff(p0, p1, p2, n);
            int i0 = B.px[pos + p0], i1 = B.px[pos + p1], i2 = B.px[pos + p2];
            F4 A{(float)(i0 % b.W), (float)(i0 / b.W), b.dwta[i0], 0.f};
            F4 Bq{(float)(i1 % b.W), (float)(i1 / b.W), b.dwta[i1], 0.f};
            F4 C{(float)(i2 % b.W), (float)(i2 / b.W), b.dwta[i2], 0.f};
            B.plane[(long)p * B.V + id] = fit_plane(A, Bq, C);
            if (p == 3 * P - 1) {
                B.point[3L * id] = i0;
                B.point[3L * id + 1] = i1;
                B.point[3L * id + 2] = i2;
            }
        }
        b.s_ransac[id] = s;
    }
}

// select_child_plane (705-726), neighbor_candidate_plane (676-703) on the parent level, the
// copy of plane_buf (1237-1238) and set_candidate_plane_k (727-741), as generate_plane runs
// them (hbp_isp.cu:1225-1243).
static void spatial(BP &b, int l) {
    BPLevel &B = b.lv[l];
    BPLevel &Hh = b.lv[l + 1];
    int Vh = Hh.V;
#pragma omp parallel for schedule(static)
    for (int id = 0; id < Vh; ++id) {
        int cnt = B.child_off[id + 1] - B.child_off[id];
        int pos = B.child_off[id];
        uint32_t s = b.s_child[id];
        for (int p = 0; p < P; ++p) {
            s = lcg(s);
            int r = (int)umod(s, cnt);
            int c = B.child[r + pos];
            Hh.plane[(long)(p + 2 * P) * Vh + id] = B.plane[c];
        }
        b.s_child[id] = s;
    }
    std::vector<F4> buf((long)P * Vh);
#pragma omp parallel for schedule(static)
    for (int id = 0; id < Vh; ++id) {
        int nb_num = Hh.off[id + 1] - Hh.off[id];
        int nb_pos = Hh.off[id];
        uint32_t s = b.s_neighbor[id];
        for (int p = 0; p < P; ++p) {
            s = lcg(s);
            int r = (int)umod(s, nb_num);
            int t = Hh.lev->bd[2L * (r + nb_pos) + 1];
            buf[(long)p * Vh + id] = Hh.plane[(long)(p + 2 * P) * Vh + t];
        }
        b.s_neighbor[id] = s;
    }
    for (int p = 0; p < P; ++p)
        for (int id = 0; id < Vh; ++id) Hh.plane[(long)(p + 2 * P) * Vh + id] = buf[(long)p * Vh + id];
#pragma omp parallel for schedule(static)
    for (int id = 0; id < B.V; ++id) {
        int c = B.lev->clus[id];
        for (int p = 2 * P; p < 3 * P; ++p) B.plane[(long)p * B.V + id] = Hh.plane[(long)p * Vh + c];
    }
}

// random_candidate_plane (hbp_isp.cu:641-674) with normalize_plane (495-503)
static void random_adjust(BP &b, int l) {
    BPLevel &B = b.lv[l];
#pragma omp parallel for schedule(static)
    for (int id = 0; id < B.V; ++id) {
        uint32_t s = b.s_ransac[id];
        float x = B.mu[3L * B.V + id];
        float y = B.mu[4L * B.V + id];
        for (int p = 2 * P; p < 3 * P; ++p) {
            F4 pl = B.plane[(long)p * B.V + id];
            float z = pl.x * x + pl.y * y + pl.z;
            float dz = lcg_uniform(s, b.max_dz);
            z += dz;
            float dnx = lcg_uniform(s, b.max_dn);
            float dny = lcg_uniform(s, b.max_dn);
            float dnz = lcg_uniform(s, b.max_dn);
            float norm = std::sqrt(pl.x * pl.x + pl.y * pl.y + pl.z * pl.z);
            pl.x /= norm;
            pl.y /= norm;
            pl.z /= norm;
            pl.x += dnx;
            pl.y += dny;
            pl.z += dnz;
            float a = -pl.x / pl.z;
            float bb = -pl.y / pl.z;
            float c = z - a * x - bb * y;
            B.plane[(long)p * B.V + id] = F4{a, bb, c, 0.f};
        }
        b.s_ransac[id] = s;
    }
}

// belief_propagation (hbp_isp.cu:1206-1214): passing_msg (1021-1047: passing_msg_k 292-304,
// smooth_msg_k 306-355, normalize_msg_k 273-290), aggregate_belief (1048-1058), ping_pang_msg
// (1059-1066), select_candidate_plane (1067-1079: get_belief_order 432-450, select_belief
// 452-472, select_msg 474-493).
static void bp_step(BP &b, int l) {
    BPLevel &B = b.lv[l];
    const Level &Lv = *B.lev;
    int V = B.V, W = b.W;
    long E = B.E;
    std::vector<float> &msg = b.msg, &msg_t = b.msg_t;
    // passing_msg_k
#pragma omp parallel for schedule(static)
    for (long e = 0; e < E; ++e) {
        int t = Lv.bd[2 * e + 1];
        for (int p = 0; p < 2 * P; ++p) msg[(long)p * E + e] = B.belief[(long)p * V + t] - msg[(long)p * E + e];
    }
    // smooth_msg_k + normalize_msg_k
#pragma omp parallel for schedule(static)
    for (long e = 0; e < E; ++e) {
        int s = Lv.bd[2 * e], t = Lv.bd[2 * e + 1];
        const int *ps = &B.point[3L * s], *pt = &B.point[3L * t];
        float X[6], Y[6];
        for (int j = 0; j < 3; ++j) {
            X[j] = (float)(ps[j] % W);
            Y[j] = (float)(ps[j] / W);
            X[3 + j] = (float)(pt[j] % W);
            Y[3 + j] = (float)(pt[j] / W);
        }
        float dist_color = B.dc0[e];
        float dist_pos = B.dc1[e];
        int bdl = Lv.bdl[e];
        dist_pos = std::fmax(dist_pos, 1.0f);
        dist_color = std::fmax(dist_color, 0.01f);
        float c0 = dist_color * (float)bdl;
        for (int y = 0; y < K; ++y) {
            F4 ps4 = B.plane[(long)y * V + s];
            float minimum = FLT_MAX;
            float out = 0;
            for (int p = 0; p < 2 * P; ++p) {
                if (p == P) {
                    out += minimum;
                    minimum = FLT_MAX;
                }
                F4 pt4 = B.plane[(long)p * V + t];
                float dpx = ps4.x - pt4.x;
                float dpy = ps4.y - pt4.y;
                float dpz = ps4.z - pt4.z;
                float phi = std::fabs(dpx * X[0] + dpy * Y[0] + dpz);
                for (int j = 1; j < 6; ++j) phi = phi + std::fabs(dpx * X[j] + dpy * Y[j] + dpz);
                phi = std::fmin(phi / 3 / dist_pos, b.tau_s);
                float m = msg[(long)p * E + e] + c0 * phi;
                if (m < minimum) minimum = m;
            }
            out += minimum;
            msg_t[(long)y * E + e] = out;
        }
        float mean = 0;
        for (int k = 0; k < K; ++k) mean += msg_t[(long)k * E + e];
        mean /= K;
        for (int k = 0; k < K; ++k) msg_t[(long)k * E + e] -= mean;
    }
    // aggregate_belief: cost + messages in edge order
#pragma omp parallel for schedule(static)
    for (int s = 0; s < V; ++s)
        for (int k = 0; k < K; ++k) {
            float acc = B.cost[(long)k * V + s];
            for (int e = B.off[s]; e < B.off[s + 1]; ++e) acc += msg_t[(long)k * E + e];
            B.belief[(long)k * V + s] = acc;
        }
    // ping_pang_msg_k
#pragma omp parallel for schedule(static)
    for (long e = 0; e < E; ++e)
        for (int k = 0; k < K; ++k) msg[(long)k * E + B.rev[e]] = msg_t[(long)k * E + e];
    // get_belief_order, select_belief
    std::vector<int> order((long)K * V, 0);
#pragma omp parallel for schedule(static)
    for (int id = 0; id < V; ++id) {
        for (int i = 0; i < K - 1; ++i)
            for (int j = i + 1; j < K; ++j) {
                if (B.belief[(long)j * V + id] < B.belief[(long)i * V + id]) order[(long)i * V + id]++;
                else order[(long)j * V + id]++;
            }
        int nb = 0;
        for (int k = 0; k < K; ++k)
            if (order[(long)k * V + id] < P) {
                B.plane[(long)nb * V + id] = B.plane[(long)k * V + id];
                B.cost[(long)nb * V + id] = B.cost[(long)k * V + id];
                B.belief[(long)nb * V + id] = B.belief[(long)k * V + id];
                nb++;
            }
    }
    // select_msg
#pragma omp parallel for schedule(static)
    for (long e = 0; e < E; ++e) {
        int t = Lv.bd[2 * e + 1];
        int nb = 0;
        for (int k = 0; k < K; ++k)
            if (order[(long)k * V + t] < P) {
                msg[(long)nb * E + e] = msg[(long)k * E + e];
                nb++;
            }
    }
}

// generate_plane (hbp_isp.cu:1216-1261)
static bool generate(BP &b, int l, int iter_ransac, int iter_spatial, int iter_random) {
    int &it = b.plane_iter;
    if (it <= iter_ransac) ransac(b, l);
    else if (it <= iter_spatial) spatial(b, l);
    else if (it <= iter_random && b.max_dz >= 0.01f) {
        random_adjust(b, l);
        b.max_dz /= 2.0f;
        b.max_dn /= 2.0f;
    } else return false;
    it++;
    return true;
}

// hbp_g (hbp_isp.cu:1283-1311) and run_ms (1313-1346) as a resumable loop over sweeps.
static int step(BP &b, int max_sweeps) {
    int done = 0;
    while (done < max_sweeps && b.cur >= b.min_level) {
        int l = b.cur;
        if (!b.inited) {
            init_level(b, l);
            b.inited = true;
        }
        if (b.outer >= b.iter_bp) {
            b.cur--;
            b.inited = false;
            continue;
        }
        int iter_ransac = 3, iter_spatial = 1 + iter_ransac, iter_random = 1 + iter_spatial;
        if (b.outer <= b.iter_bp / 4 * 3) iter_random = iter_spatial;
        if (generate(b, l, iter_ransac, iter_spatial, iter_random)) {
            plane_cost(b, l, 2 * P);
            bp_step(b, l);
            done++;
            b.sweeps++;
        } else {
            b.outer++;
            b.plane_iter = 0;
        }
    }
    return done;
}

// map_msg (hbp_isp.cu:1081-1105): map_msg_k (783-794), min_belief_to_plane (796-817),
// plane_to_label (819-841).
static void map(const BP &b, float *disp, float *plane3) {
    int l = b.min_level;
    const BPLevel &B = b.lv[l];
    std::vector<F4> best(B.V);
    for (int c = 0; c < B.V; ++c) {
        float mb = FLT_MAX;
        for (int p = 0; p < P; ++p) mb = amin_fold(mb, B.belief[(long)p * B.V + c]);
        int bp = 0;
        for (int p = 0; p < P; ++p)
            if (B.belief[(long)p * B.V + c] == mb) {
                bp = p;
                break;
            }
        best[c] = B.plane[(long)bp * B.V + c];
    }
    for (int c = 0; c < B.V; ++c)
        for (int j = B.pix_off[c]; j < B.pix_off[c + 1]; ++j) {
            int pxid = B.px[j];
            int x = pxid % b.W, y = pxid / b.W;
            F4 pl = best[c];
            float d = pl.x * (float)x + pl.y * (float)y + pl.z;
            d = std::fmin(std::fmax(d, 0.0f), (float)(b.D - 1));
            if (disp) disp[pxid] = d;
            if (plane3) {
                plane3[3L * pxid] = pl.x;
                plane3[3L * pxid + 1] = pl.y;
                plane3[3L * pxid + 2] = pl.z;
            }
        }
}

}  // namespace orc

extern "C" {

void *orc_bp_new(void *pyr, const uint8_t *vol, const float *dwta, int W, int H, int D, float tau,
                 float alpha, float tau_s, int min_level, int use_cost_vol) {
    auto *b = new orc::BP();
    b->pyr = (const orc::Pyramid *)pyr;
    b->vol = vol;
    b->volf = nullptr;
    b->dwta = dwta;
    b->W = W;
    b->H = H;
    b->D = D;
    b->N = (long)W * H;
    b->tau = tau;
    b->lambda = alpha;
    b->tau_s = tau_s;
    b->use_cost_vol = use_cost_vol;
    int L = (int)b->pyr->levels.size();
    if (min_level < 1 || min_level > L - 1) {
        delete b;
        return nullptr;
    }
    b->min_level = min_level;
    b->fx_shift = orc::fixed_shift(alpha, tau, use_cost_vol, (int)b->N);
    orc::prologue(*b);
    b->cur = b->L - 2;
    return b;
}
// the same on a given f32 volume [D][H][W] (MC-CNN, tau 0.5)
void *orc_bp_new_f32(void *pyr, const float *vol, const float *dwta, int W, int H, int D, float tau, float alpha,
                     float tau_s, int min_level, int use_cost_vol) {
    auto *b = (orc::BP *)orc_bp_new(pyr, nullptr, dwta, W, H, D, tau, alpha, tau_s, min_level, use_cost_vol);
    if (!b) return nullptr;
    b->volf = vol;
    // the bound of the per-pixel costs: tau, or the largest |entry| (truncated from above only)
    float amax = 0.f;
    for (long i = 0; i < (long)D * W * H; ++i) amax = std::fmax(amax, std::fabs(vol[i]));
    const float cmax = std::fmax(std::fabs(tau), amax);
    if (!std::isfinite(cmax)) {
        orc_bp_free(b);
        return nullptr;
    }
    b->fx_shift = orc::fixed_shift(alpha, cmax, use_cost_vol, (int)b->N);
    return b;
}
int orc_bp_step(void *bp, int max_sweeps) { return orc::step(*(orc::BP *)bp, max_sweeps); }
int orc_bp_level(void *bp) { return ((orc::BP *)bp)->cur; }
int orc_bp_sweeps_done(void *bp) { return ((orc::BP *)bp)->sweeps; }
void orc_bp_level_size(void *bp, int level, int *V, int *E) {
    auto &B = ((orc::BP *)bp)->lv[level];
    *V = B.V;
    *E = (int)B.E;
}
void orc_bp_get_state(void *bp, float *planes, float *cost, float *belief, float *msg) {
    auto &b = *(orc::BP *)bp;
    int l = b.cur < b.min_level ? b.min_level : b.cur;
    auto &B = b.lv[l];
    if (planes) std::memcpy(planes, B.plane.data(), sizeof(float) * 4 * orc::K * (long)B.V);
    if (cost) std::memcpy(cost, B.cost.data(), sizeof(float) * orc::K * (long)B.V);
    if (belief) std::memcpy(belief, B.belief.data(), sizeof(float) * orc::K * (long)B.V);
    if (msg && (long)b.msg.size() == (long)orc::K * B.E)
        for (long e = 0; e < B.E; ++e)
            for (int k = 0; k < 2 * orc::P; ++k) msg[e * 2 * orc::P + k] = b.msg[(long)k * B.E + e];
}
void orc_bp_map(void *bp, float *disp, float *plane3) { orc::map(*(orc::BP *)bp, disp, plane3); }
void orc_bp_free(void *bp) { delete (orc::BP *)bp; }

// Sweeps hbp_g performs over levels L-2 .. min_level for a pyramid of `levels` levels.
int orc_bp_total_sweeps(int levels, int min_level, int D) {
    int total = 0;
    for (int l = levels - 2; l >= min_level; --l) {
        int iter_bp = 10 + 10 * l;
        float max_dz = (float)D / 2.0f * (float)l / (float)levels;
        for (int i = 0; i < iter_bp; ++i) {
            total += 5;
            if (i > iter_bp / 4 * 3 && max_dz >= 0.01f) {
                total += 1;
                max_dz /= 2.0f;
            }
        }
    }
    return total;
}

int orc_plain_pair(const uint8_t *bgrL, const uint8_t *bgrR, int W, int H, int D, float alpha,
                   float tau_s, int min_level, int bp_levels, float *disp, float *plane3,
                   int *levels_out) {
    long N = (long)W * H;
    std::vector<uint8_t> vol((long)D * N);
    std::vector<float> dwta(N);
    if (orc_cost_volume(bgrL, bgrR, W, H, D, 0, vol.data(), dwta.data()) != 0) return -1;
    void *pyr = orc_pyramid_build(bgrL, W, H, 4);
    int L = orc_pyr_levels(pyr);
    if (levels_out) *levels_out = L;
    if (bp_levels > 0) min_level = std::max(1, L - 1 - bp_levels);
    void *bp = orc_bp_new(pyr, vol.data(), dwta.data(), W, H, D, 25.f, alpha, tau_s, min_level, 1);
    if (!bp) {
        orc_pyr_free(pyr);
        return -2;
    }
    while (orc_bp_step(bp, 1 << 30) > 0) {
    }
    orc_bp_map(bp, disp, plane3);
    orc_bp_free(bp);
    orc_pyr_free(pyr);
    return min_level;
}

}  // extern "C"
