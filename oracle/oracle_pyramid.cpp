// Oracle (test infrastructure only): stage (b), the CRTrees superpixel pyramid.
// Restates segment.cu, CRTrees.cu and imutil.cu of the reference (citations per function).
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "hbp_oracle.h"
#include "oracle_math.h"
#include "oracle_pyramid.h"

namespace orc {

// sRGB linearisation of one u8 channel (imutil.cu:11-17): c = v/255.0,
// c > 0.04045 ? pow((c+0.055)/1.055, 2.4) : c/12.92, in FP64 with libm pow.
static const double *srgb_lut() {
    static double lut[256];
    static bool init = false;
    if (!init) {
        for (int v = 0; v < 256; ++v) {
            double c = (double)(float)v / 255.0;
            lut[v] = (c > 0.04045) ? std::pow((c + 0.055) / 1.055, 2.4) : c / 12.92;
        }
        init = true;
    }
    return lut;
}

// imutil.cu:4-40 (BGR_to_LAB) with w = 1 (imutil.cu:170-185); the cube root is pin_cbrt.
static void lab_pixel(uint8_t B, uint8_t G, uint8_t R, float out[4]) {
    const double *lut = srgb_lut();
    double b = lut[B], g = lut[G], r = lut[R];
    double x = (r * 0.412453 + g * 0.357580 + b * 0.180423) / 0.950456;
    double y = (r * 0.212671 + g * 0.715160 + b * 0.072169) / 1.0;
    double z = (r * 0.019334 + g * 0.119193 + b * 0.950227) / 1.088754;
    const double epsilon = 216 / 24389.0;
    const double k = 841 / 108.0;
    const double kb = 16 / 116.0;
    x = (x > epsilon) ? pin_cbrt(x) : (k * x) + kb;
    y = (y > epsilon) ? pin_cbrt(y) : (k * y) + kb;
    z = (z > epsilon) ? pin_cbrt(z) : (k * z) + kb;
    out[0] = (float)((116 * y) - 16);
    out[1] = (float)(500 * (x - y));
    out[2] = (float)(200 * (y - z));
    out[3] = 1.0f;
}

// segment.cu:14-22: sum of squared Lab differences + max(0, dL).
static inline float dist_px(const float *s, const float *n) {
    float dx = s[0] - n[0], dy = s[1] - n[1], dz = s[2] - n[2];
    float d = dx * dx + dy * dy + dz * dz;
    float h = dx * 1.f;
    d += h > 0.f ? h : 0.f;
    return d;
}

// segment.cu:27-37: per-channel term, the max(0, ds) part only for channel 0.
static inline float dist_ch(float s, float n, int c) {
    float d = (s - n) * (s - n);
    if (c == 0) {
        float h = (s - n) * 1.f;
        d += h > 0.f ? h : 0.f;
    }
    return d;
}

void lab_image(const uint8_t *bgr, int W, int H, float *lab4) {
    long n = (long)W * H;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) lab_pixel(bgr[3 * i], bgr[3 * i + 1], bgr[3 * i + 2], lab4 + 4 * i);
}

// compute_gauss_kernel (imutil.cu:443-471) with S = round(3 sigma) (img_gauss_blur 481):
// w[r][c] = exp(-(r^2 + c^2)/(2 sigma^2)) in FP64 (libm exp, host code in the reference too),
// r then c over [-S, S], each times 1/sum, rounded to f32.  Returns S.
int gauss_weights(double sigma, std::vector<float> &w) {
    const int S = (int)std::round(3 * sigma);
    const int F = 2 * S + 1;
    std::vector<double> f((size_t)F * F);
    double sum = 0;
    for (int r = -S; r <= S; ++r)
        for (int c = -S; c <= S; ++c) {
            double val = std::exp(-(double)(r * r + c * c) / (2 * sigma * sigma));
            f[(size_t)(r + S) * F + c + S] = val;
            sum += val;
        }
    sum = 1. / sum;
    w.assign((size_t)F * F, 0.f);
    for (size_t i = 0; i < f.size(); ++i) w[i] = (float)(f[i] * sum);
    return S;
}

// gauss_blur (imutil.cu:399-441) of img_uchar3_to_float4's (B, G, R, 1) (imutil.cu:170-185):
// out(y, x) = sum over r, then c, in [-S, S] of in(clamp(y + r), clamp(x + c)) * w[r][c], each
// channel a left fold in f32 from 0.  The reference's edge thread blocks also write the last
// column / row from threads whose windows are centred beyond the image (a write race); the
// pinned value is the window centred on the pixel itself.  The reference's launch needs
// (16 + 2S)^2 <= 1024 threads, so S <= 8; returns S, or -1 (nothing written) above that.
int gauss_blur(const uint8_t *bgr, int W, int H, double sigma, float *out4) {
    std::vector<float> w;
    const int S = gauss_weights(sigma, w);
    if (S > 8) return -1;
    const int F = 2 * S + 1;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float acc[4] = {0.f, 0.f, 0.f, 0.f};
            for (int r = -S; r <= S; ++r) {
                int yy = std::min(std::max(0, y + r), H - 1);
                for (int c = -S; c <= S; ++c) {
                    int xx = std::min(std::max(0, x + c), W - 1);
                    const uint8_t *px = bgr + 3 * ((long)yy * W + xx);
                    float wt = w[(size_t)(r + S) * F + c + S];
                    acc[0] = acc[0] + (float)px[0] * wt;
                    acc[1] = acc[1] + (float)px[1] * wt;
                    acc[2] = acc[2] + (float)px[2] * wt;
                    acc[3] = acc[3] + 1.0f * wt;
                }
            }
            std::memcpy(out4 + 4 * ((long)y * W + x), acc, sizeof(acc));
        }
    return S;
}

// imutil.cu:4-40 (BGR_to_LAB) on a float pixel (the blurred image): the same arithmetic as
// lab_pixel with pow((c + 0.055)/1.055, 2.4) by pin_pow24; w passes through.
static void lab_pixel_f(const float *in, float out[4]) {
    double b = in[0] / 255.0, g = in[1] / 255.0, r = in[2] / 255.0;
    r = (r > 0.04045) ? pin_pow24((r + 0.055) / 1.055) : r / 12.92;
    g = (g > 0.04045) ? pin_pow24((g + 0.055) / 1.055) : g / 12.92;
    b = (b > 0.04045) ? pin_pow24((b + 0.055) / 1.055) : b / 12.92;
    double x = (r * 0.412453 + g * 0.357580 + b * 0.180423) / 0.950456;
    double y = (r * 0.212671 + g * 0.715160 + b * 0.072169) / 1.0;
    double z = (r * 0.019334 + g * 0.119193 + b * 0.950227) / 1.088754;
    const double epsilon = 216 / 24389.0;
    const double k = 841 / 108.0;
    const double kb = 16 / 116.0;
    x = (x > epsilon) ? pin_cbrt(x) : (k * x) + kb;
    y = (y > epsilon) ? pin_cbrt(y) : (k * y) + kb;
    z = (z > epsilon) ? pin_cbrt(z) : (k * z) + kb;
    out[0] = (float)((116 * y) - 16);
    out[1] = (float)(500 * (x - y));
    out[2] = (float)(200 * (y - z));
    out[3] = in[3];
}

void lab_image_f(const float *in4, long n, float *lab4) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) lab_pixel_f(in4 + 4 * i, lab4 + 4 * i);
}

// segment.cu:57-195 (compute_1nn_grid_k): neighbours in the order up, down, left, right
// (then ul, ur, dr, dl for num_nb = 8); first strict minimum wins; border slots become
// self-loops (i, i) that take no part in the minimum.  nn starts at 0, dist_min at FLT_MAX.
// edist (optional, the per-edge distances the Min/Max linkages carry): the distance of each
// in-image slot; a self-loop slot is never read (0 here, unwritten in the reference).
void nn_grid(const float *lab4, int W, int H, int num_nb, int *nn, float *dist_min, int *bd, float *edist) {
    long n = (long)W * H;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        int x = (int)(i % W), y = (int)(i / W);
        float best = FLT_MAX;
        int arg = 0;
        struct Nb { bool ok; long j; };
        Nb nbs[8] = {{y > 0, i - W},
                     {y < H - 1, i + W},
                     {x > 0, i - 1},
                     {x < W - 1, i + 1},
                     {y > 0 && x > 0, i - W - 1},
                     {y > 0 && x < W - 1, i - W + 1},
                     {y < H - 1 && x < W - 1, i + W + 1},
                     {y < H - 1 && x > 0, i + W - 1}};
        for (int k = 0; k < num_nb; ++k) {
            long j = nbs[k].ok ? nbs[k].j : i;
            float d = 0.0f;
            if (nbs[k].ok) {
                d = dist_px(lab4 + 4 * i, lab4 + 4 * j);
                if (d < best) {
                    best = d;
                    arg = (int)j;
                }
            }
            if (edist) edist[num_nb * i + k] = d;
            if (bd) {
                bd[2 * (num_nb * i + k)] = (int)i;
                bd[2 * (num_nb * i + k) + 1] = (int)j;
            }
        }
        nn[i] = arg;
        dist_min[i] = best;
    }
}

// Connected components of the functional graph i -> nn[i]; every node is labelled with the
// largest index in its weakly connected component.  Used by the tests as an independent
// check that CRTrees labels are constant on exactly these components (CRTrees picks a
// different representative: see crtrees_literal).
void cc_max(const int *nn, int V, int *label) {
    std::vector<int> parent(V);
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int a) {
        while (parent[a] != a) {
            parent[a] = parent[parent[a]];
            a = parent[a];
        }
        return a;
    };
    for (int i = 0; i < V; ++i) {
        int a = find(i), b = find(nn[i]);
        if (a == b) continue;
        if (a < b) parent[a] = b; else parent[b] = a;   // the larger root survives
    }
    for (int i = 0; i < V; ++i) label[i] = find(i);
}

// CRTrees::compact_clus_label (CRTrees.cu:387-402 with 356-376): label <- rank of the label
// among all distinct labels (exclusive scan of the "is a label" predicate).  Returns V'.
int compact(int *label, int V) {
    std::vector<int> pred(V, 0), pos(V, 0);
    for (int i = 0; i < V; ++i) pred[label[i]] = 1;
    int acc = 0;
    for (int i = 0; i < V; ++i) {
        pos[i] = acc;
        acc += pred[i];
    }
    for (int i = 0; i < V; ++i) label[i] = pos[label[i]];
    return acc;
}

// Literal restatement of CRTrees::get_clus (CRTrees.cu:315-354) with its kernels (28-188).
// The object's scratch buffers are modelled explicitly: nnc (A), nnn (B), max_d (M) and the
// caller's clus (C).  B and M start with in-range garbage drawn from a seed tied to
// `order` (cudaMalloc leaves them uninitialised).  Each kernel runs as a sequential loop over
// thread ids in the order
// given by `order` (0 ascending, 1 descending, >= 2 a seeded permutation), in place, which is
// one of the interleavings the races inside re_link allow; order < 0 runs ascending with
// re_link reading a snapshot of A taken before the kernel (the double-buffered form).  pointer_jumping and
// pointer_jumping_max ping-pong between A/B (and M/C) exactly as the host loops swap them.
void crtrees_literal(const int *nn_in, int V, int order, int *out) {
    std::vector<int> nn(nn_in, nn_in + V), A(V), B(V), M(V), C(V), mask(V, 1), maskc(V), indeg(V);
    std::vector<int> perm(V);
    std::iota(perm.begin(), perm.end(), 0);
    uint64_t g = 0x9E3779B97F4A7C15ull * (uint64_t)(order + 2);
    auto rnd = [&]() {
        g += 0x9E3779B97F4A7C15ull;
        uint64_t z = g;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    if (order == 1) std::reverse(perm.begin(), perm.end());
    if (order >= 2)
        for (int i = V - 1; i > 0; --i) std::swap(perm[i], perm[rnd() % (uint64_t)(i + 1)]);
    for (int i = 0; i < V; ++i) {
        B[i] = (int)(rnd() % (uint64_t)V);
        M[i] = (int)(rnd() % (uint64_t)V);
    }
    A = nn;
    C = nn;
    for (;;) {
        // get_in_degree (CRTrees.cu:287-303)
        std::fill(indeg.begin(), indeg.end(), 0);
        for (int i : perm)
            if (mask[i]) indeg[A[i]]++;
        bool jump = false;
        for (int i = 0; i < V; ++i)
            if (mask[i] && indeg[i] > 1) jump = true;
        if (!jump) break;
        maskc = mask;
        // de_link_and_compute_mask_h (304-311)
        std::fill(mask.begin(), mask.end(), 0);
        for (int i : perm)
            if (indeg[i] > 1) {
                A[i] = i;
                mask[i] = 1;
            }
        // pointer_jumping (314-333): the member nnc buffer is A whatever the local swaps do
        {
            std::vector<int> *p = &A, *q = &B;
            bool jp;
            do {
                jp = false;
                for (int i : perm) {
                    if (!maskc[i]) continue;
                    int a = (*p)[i], b = (*p)[a];
                    (*q)[i] = b;
                    if (!mask[a] && mask[b]) jp = true;
                }
                std::swap(p, q);
            } while (jp);
        }
        // re_link (178-201), in place on A (order < 0: every thread reads the pre-kernel A)
        {
            std::vector<int> A0;
            if (order < 0) A0 = A;
            const std::vector<int> &Ar = order < 0 ? A0 : A;
            for (int i : perm) {
                if (mask[i]) {
                    int a = nn[i];
                    A[i] = a;
                    if (!mask[a]) A[i] = Ar[a];
                } else if (!maskc[i]) {
                    int a = Ar[i];
                    if (maskc[a]) A[i] = Ar[a];
                }
            }
        }
        // re_link_all (203-225)
        for (int i : perm) {
            int a = A[i];
            if (mask[i]) nn[i] = a;
            else if (mask[a]) nn[i] = A[a];
        }
        A = nn;   // CRTrees.cu:423
    }
    // pointer_jumping_max (274-298); the caller's clus buffer is C
    M = A;
    std::vector<int> *p = &A, *q = &B, *mx = &M, *cl = &C;
    bool jp;
    do {
        jp = false;
        for (int i : perm) {
            int a = (*p)[i];
            int m = (*mx)[i];
            (*cl)[i] = std::max(m, (*mx)[a]);
            (*q)[i] = (*p)[a];
            if ((*cl)[i] != m) jp = true;
        }
        std::swap(p, q);
        std::swap(mx, cl);
    } while (jp);
    for (int i = 0; i < V; ++i) out[i] = C[i];
}

// get_target_clus (segment.cu:954-1041) for num_clus < target < num_clus_p: re-labels the level
// so that it has exactly `target` clusters, breaking the clusters with the largest sums of their
// children's 1-NN distances into their children.
//  * per cluster: the child count and the sum of the children's dist_min (reduce_dist_pos_atomic,
//    a float atomicAdd; pinned: ascending child index, left fold from 0);
//  * clusters ranked by that sum, descending (thrust::sort_by_key with greater; pinned stable:
//    ties keep ascending cluster id); cum = inclusive scan of the counts in rank order;
//  * i* = the first rank with cum[i] + num_clus - (i+1) >= target (is_break_cluster; its
//    write of i* into pos_scan[num_clus-1] races with that element's read in the reference:
//    pinned as read before write); nb = cum[i*], exact = target - num_clus + i*;
//  * rank i <= i*: its children take the values cum[i]-1, cum[i]-2, ... (the atomicSub of
//    reset_clus_label; pinned: ascending child index takes them in that order), each label
//    min(value, exact); rank i > i*: one cluster labelled (i - i* + nb - 1) + exact - nb + 1.
static void target_clus_labels(int *clus, int V, int Vn, const float *dist_min, int target) {
    std::vector<float> dsum(Vn, 0.0f);
    std::vector<int> cnt(Vn, 0);
    for (int i = 0; i < V; ++i) {
        dsum[clus[i]] += dist_min[i];
        cnt[clus[i]] += 1;
    }
    std::vector<int> rank_id(Vn);
    std::iota(rank_id.begin(), rank_id.end(), 0);
    std::stable_sort(rank_id.begin(), rank_id.end(), [&](int a, int b) { return dsum[a] > dsum[b]; });
    std::vector<long> cum(Vn);
    long run = 0;
    for (int i = 0; i < Vn; ++i) cum[i] = run += cnt[rank_id[i]];
    int istar = -1;
    for (int i = 0; i < Vn && istar < 0; ++i)
        if (cum[i] + Vn - (i + 1) >= target) istar = i;
    const long nb = cum[istar];
    const long exact = (long)target - Vn + istar;
    std::vector<long> pos(Vn);
    for (int i = 0; i < Vn; ++i) pos[rank_id[i]] = i <= istar ? cum[i] - 1 : i - istar + nb - 1;
    for (int i = 0; i < V; ++i) {
        const int cc = clus[i];
        const long pre = pos[cc];
        if (pre < nb) {
            const long v = pos[cc]--;
            clus[i] = (int)(v >= exact ? exact : v);
        } else {
            clus[i] = (int)(pre + exact - nb + 1);
        }
    }
}

// SegHAC::run_ms (segment.cu:1279-1321) for LabelISPFormat, pernalize on, with any of the four
// linkages (use_mean = Centroid or Ward, segment.hpp:248).  sigma > FLT_MIN blurs the image
// before the Lab conversion (img_CPU_to_GPU, segment.cu:795-809); the blurred w channel becomes
// the level-0 weight.  target_clus > 1 ends the pyramid at the first level with at most that
// many clusters, re-labelled to exactly target_clus (get_target_clus, then save_output's stop
// test, segment.cu:1103-1110, 1299-1306).  Returns false (P untouched) for S = round(3 sigma) > 8.
bool pyramid_build(const uint8_t *bgr, int W, int H, int num_nb, Pyramid &P, int link, int target_clus,
                   double sigma) {
    const bool use_mean = link == kCentroidLink || link == kWardLink;
    long N = (long)W * H;
    std::vector<float> lab(4 * N);
    if (sigma > FLT_MIN) {
        std::vector<float> blur(4 * N);
        if (gauss_blur(bgr, W, H, sigma, blur.data()) < 0) return false;
        lab_image_f(blur.data(), N, lab.data());
    } else {
        lab_image(bgr, W, H, lab.data());
    }
    P.W = W;
    P.H = H;
    P.top = target_clus;
    P.levels.clear();
    std::vector<int> nn(N);
    std::vector<float> dist_min(N);
    std::vector<int> bd(2 * num_nb * N);
    // the per-edge distances carried from level to level by Min/Max (dist_d, segment.cu:838-844)
    std::vector<float> edist(use_mean ? 0 : num_nb * N);
    nn_grid(lab.data(), W, H, num_nb, nn.data(), dist_min.data(), bd.data(), use_mean ? nullptr : edist.data());
    // init_data_mean (segment.cu:856-861, imutil.cu:92-114): SoA (L, a, b, x, y, w)
    int V = (int)N;
    long E = num_nb * N;
    std::vector<float> mean(6 * N);
    for (long i = 0; i < N; ++i) {
        mean[i] = lab[4 * i];
        mean[N + i] = lab[4 * i + 1];
        mean[2 * N + i] = lab[4 * i + 2];
        mean[3 * N + i] = (float)(int)(i % W);
        mean[4 * N + i] = (float)(int)(i / W);
        mean[5 * N + i] = lab[4 * i + 3];
    }
    std::vector<int> bdl(E, 1);   // segment.cu:846-850

    while (V > 1) {
        std::vector<int> clus(V);
        crtrees_literal(nn.data(), V, 0, clus.data());
        int Vn = compact(clus.data(), V);
        if (Vn <= target_clus && Vn != target_clus) {
            target_clus_labels(clus.data(), V, Vn, dist_min.data(), target_clus);
            Vn = compact(clus.data(), V);
        }
        // save_output -> push_label_isp (segment.cu:1096-1111, 884-905)
        Level L;
        L.V = V;
        L.E = E;
        L.clus = clus;
        L.mean = mean;
        L.bd = bd;
        L.bdl = bdl;
        P.levels.push_back(std::move(L));
        if (Vn == target_clus) break;
        // reduce_boundary_dist (segment.cu:1050-1091): map the edges, sort them (int2
        // lexicographic order, util.h:118-127; with sort_by_key the carried distances move with
        // their edges, 1056-1062), keep the last of each run with s != t.  bdl is not moved by
        // the sort: reduce_bdl (368-386) adds bdl[i] of the *position* i of each run element, so
        // a new edge's bdl is the sum of the previous level's bdl over its run's positions (the
        // boundary length only while the previous bdl are all 1, i.e. for level 1).  Min/Max
        // reduce the carried distances over the run (reduce_dist 338-364, from FLT_MAX / 0).
        std::vector<std::pair<long long, float>> keys(E);
        for (long e = 0; e < E; ++e) {
            int s = clus[bd[2 * e]], t = clus[bd[2 * e + 1]];
            keys[e] = {((long long)s << 32) | (unsigned)t, use_mean ? 0.0f : edist[e]};
        }
        std::stable_sort(keys.begin(), keys.end(),
                         [](const std::pair<long long, float> &a, const std::pair<long long, float> &b) {
                             return a.first < b.first;
                         });
        std::vector<int> nbd, nbdl;
        std::vector<float> ndist;
        for (long e = 0; e < E;) {
            long f = e;
            int sum = 0;
            float red = link == kMinLink ? FLT_MAX : 0.0f;
            for (; f < E && keys[f].first == keys[e].first; ++f) {
                sum += bdl[f];
                if (link == kMinLink) red = std::min(red, keys[f].second);
                if (link == kMaxLink) red = std::max(red, keys[f].second);
            }
            int s = (int)(keys[e].first >> 32), t = (int)(keys[e].first & 0xffffffff);
            if (s != t) {
                nbd.push_back(s);
                nbd.push_back(t);
                nbdl.push_back(sum);
                ndist.push_back(red);
            }
            e = f;
        }
        long En = (long)nbdl.size();
        // compute_dist_inner_max (segment.cu:1116-1128): max over children, from 0
        std::vector<float> dist_max(Vn, 0.0f);
        for (int i = 0; i < V; ++i) dist_max[clus[i]] = std::max(dist_max[clus[i]], dist_min[i]);
        // reduce_data_mean (segment.cu:1134-1153): pinned order = children in ascending
        // index, left fold from 0; mean = sum / w for channels 0..4.
        std::vector<float> sums(6 * (long)Vn, 0.0f);
        for (int c = 0; c < 6; ++c)
            for (int i = 0; i < V; ++i) sums[(long)c * Vn + clus[i]] += mean[(long)c * V + i];
        std::vector<float> mu(5 * (long)Vn);
        for (int c = 0; c < 5; ++c)
            for (int p = 0; p < Vn; ++p) mu[(long)c * Vn + p] = sums[(long)c * Vn + p] / sums[5L * Vn + p];
        // Centroid / Ward: compute_dist_mean_k (segment.cu:508-534), Ward's weight
        // s_w t_w / (s_w + t_w) on the pixel counts (ward_weight 561-577); Min / Max: the
        // reduced carried distance.  Then pernalize_dist_k (582-612), which Min / Max carry on.
        std::vector<float> dist(En);
        for (long e = 0; e < En; ++e) {
            int s = nbd[2 * e], t = nbd[2 * e + 1];
            float d;
            if (use_mean) {
                d = dist_ch(mu[s], mu[t], 0) + dist_ch(mu[Vn + s], mu[Vn + t], 1) +
                    dist_ch(mu[2L * Vn + s], mu[2L * Vn + t], 2);
                if (link == kWardLink) {
                    const float sw = sums[5L * Vn + s], tw = sums[5L * Vn + t];
                    d = sw * tw / (sw + tw) * d;
                }
            } else {
                d = ndist[e];
            }
            float delta = d - dist_max[t];
            if (delta > 0) d += delta;
            dist[e] = d;
        }
        // compute_1nn (segment.cu:1211-1231): min over out-edges, ties -> largest t
        std::vector<int> nn2(Vn, 0);
        std::vector<float> dmin(Vn, FLT_MAX);
        for (long e = 0; e < En; ++e) dmin[nbd[2 * e]] = std::min(dmin[nbd[2 * e]], dist[e]);
        for (long e = 0; e < En; ++e)
            if (dist[e] == dmin[nbd[2 * e]]) nn2[nbd[2 * e]] = std::max(nn2[nbd[2 * e]], nbd[2 * e + 1]);
        nn.swap(nn2);
        dist_min.swap(dmin);
        mean.swap(sums);
        bd.swap(nbd);
        bdl.swap(nbdl);
        if (!use_mean) edist.swap(dist);
        V = Vn;
        E = En;
    }
    return true;
}

}  // namespace orc

extern "C" {

void orc_lab(const uint8_t *bgr, int W, int H, float *lab4) { orc::lab_image(bgr, W, H, lab4); }
void orc_nn_grid(const float *lab4, int W, int H, int num_nb, int *nn, float *dist_min, int *bd) {
    orc::nn_grid(lab4, W, H, num_nb, nn, dist_min, bd);
}
void orc_cc_max(const int *nn, int V, int *label) { orc::cc_max(nn, V, label); }
void orc_crtrees_literal(const int *nn, int V, int order, int *label) {
    orc::crtrees_literal(nn, V, order, label);
}
int orc_compact(int *label, int V) { return orc::compact(label, V); }

void *orc_pyramid_build(const uint8_t *bgr, int W, int H, int num_nb) {
    auto *p = new orc::Pyramid();
    orc::pyramid_build(bgr, W, H, num_nb, *p);
    return p;
}
void *orc_pyramid_build_link(const uint8_t *bgr, int W, int H, int num_nb, int link) {
    if (link < 0 || link > 3) return nullptr;
    auto *p = new orc::Pyramid();
    orc::pyramid_build(bgr, W, H, num_nb, *p, link);
    return p;
}
void *orc_pyramid_build_target(const uint8_t *bgr, int W, int H, int num_nb, int link, int target_clus) {
    return orc_pyramid_build_sigma(bgr, W, H, num_nb, link, target_clus, 0.0);
}
void *orc_pyramid_build_sigma(const uint8_t *bgr, int W, int H, int num_nb, int link, int target_clus, double sigma) {
    if (link < 0 || link > 3 || target_clus < 1 || (long)target_clus >= (long)W * H) return nullptr;
    auto *p = new orc::Pyramid();
    if (!orc::pyramid_build(bgr, W, H, num_nb, *p, link, target_clus, sigma)) {
        delete p;
        return nullptr;
    }
    return p;
}
int orc_gauss_weights(double sigma, float *w) {
    std::vector<float> v;
    int S = orc::gauss_weights(sigma, v);
    if (S > 8) return -1;
    std::memcpy(w, v.data(), sizeof(float) * v.size());
    return S;
}
int orc_gauss_blur(const uint8_t *bgr, int W, int H, double sigma, float *out4) {
    return orc::gauss_blur(bgr, W, H, sigma, out4);
}
void orc_lab_f(const float *in4, long n, float *lab4) { orc::lab_image_f(in4, n, lab4); }
double orc_pin_pow24(double t) { return orc::pin_pow24(t); }
int orc_pyr_levels(void *p) { return (int)((orc::Pyramid *)p)->levels.size(); }
void orc_pyr_level_size(void *p, int l, int *V, int *E) {
    auto &L = ((orc::Pyramid *)p)->levels[l];
    *V = L.V;
    *E = (int)L.E;
}
void orc_pyr_get_level(void *p, int l, int *clus, float *mean, int *bd, int *bdl) {
    auto &L = ((orc::Pyramid *)p)->levels[l];
    if (clus) std::memcpy(clus, L.clus.data(), sizeof(int) * L.V);
    if (mean) std::memcpy(mean, L.mean.data(), sizeof(float) * 6 * (long)L.V);
    if (bd) std::memcpy(bd, L.bd.data(), sizeof(int) * 2 * L.E);
    if (bdl) std::memcpy(bdl, L.bdl.data(), sizeof(int) * L.E);
}
void orc_pyr_free(void *p) { delete (orc::Pyramid *)p; }

double orc_pin_cbrt(double x) { return orc::pin_cbrt(x); }
double orc_pin_exp(double x) { return orc::pin_exp(x); }
float orc_i64_to_f32(int64_t q) { return orc::i64_to_f32(q); }
int orc_f2i_rd(float x) { return orc::f2i_rd(x); }
int orc_fixed_shift(float lambda, float tau, int use_cost_vol, int npix) {
    return orc::fixed_shift(lambda, tau, use_cost_vol, npix);
}
float orc_amin_fold(float stored, float v) { return orc::amin_fold(stored, v); }

}  // extern "C"
