// Internal objects of the HBP-ISP HIP library: cost volume, superpixel pyramid, BP state.
#pragma once
#include "hbp_common.h"

namespace hbp {

// Cost volume of one stereo pair (StereoCost / StereoData of the reference).
struct Volume {
    int W = 0, H = 0, D = 0, pitch = 0;
    long plane = 0;          // bytes between disparity planes (H * pitch)
    float tau = kCensusTau;  // truncation (stereo_cost.cu:433)
    bool right = false;      // cost_to_right applied
    bool f32src = false;     // an f32 volume (MC-CNN given, or AD): f32 holds it, there is no u8 volume
    float absmax = 0.f;      // f32src: the largest |entry| after truncation (bounds the plane costs)
    DevBuf<uint32_t> cl, cr; // census words of the left / right image [H][W]
    DevBuf<uint8_t> vol;     // u8 [D][H][pitch], written on first request (volume_u8)
    DevBuf<float> dwta;      // WTA disparity [H][W]
    DevBuf<float> f32;       // f32 [D][H][W]: the given volume (f32src), else made for the float* getter
};

// Superpixel pyramid in the reference's packed ISP format (segment.cu:884-905): per level l,
// clus[V_l] (id at level l+1), mean[6][V_l] (sums of L, a, b, x, y, w), bd[E_l] (s, t),
// bdl[E_l]; the levels follow each other in the four packed arrays.
struct Pyramid {
    int W = 0, H = 0, num_nb = 4;
    int top = 1;                    // clusters the last level maps to: 1, or target_clus
    std::vector<int> V, E;          // per level
    std::vector<long> voff, eoff;   // element offsets of each level in the packed arrays
    DevBuf<int> clus;
    DevBuf<float> mean;
    DevBuf<int2> bd;
    DevBuf<int> bdl;
    int levels() const { return (int)V.size(); }
};

// Per-level structures BP reads (HBP_ISP ctor + run_ms prologue, hbp_isp.hpp:288-386,
// hbp_isp.cu:1319-1325), and the level's BP state.
struct BPLevel {
    int V = 0, Vh = 1;
    long E = 0;
    const int *clus = nullptr;     // into the pyramid
    const int2 *bd = nullptr;      // into the pyramid
    DevBuf<uint32_t> pxy, pcl;     // pixels sorted by (cluster, pixel id): packed x|y<<16, cluster [N]
    DevBuf<int> pix_off;           // [V+1]
    bool coarse = false;           // plane costs by per-superpixel segments (else by chunks)
    int nseg_max = 0;              // coarse: upper bound of the segment count (N/seg_len + V)
    int nseg = 0;                  // coarse: the segment count (read back at set-up)
    DevBuf<int4> seg;              // coarse: (first pixel, end pixel, cluster, index in cluster)
    DevBuf<int> sego;              // coarse: first segment of each superpixel; sego[V] = count
    DevBuf<int> off;               // CSR offsets by source [V+1]
    DevBuf<int> rev;               // reverse edge [E]
    int maxdeg = 1;                // largest out-degree
    DevBuf<int> eloc;              // slot of each edge among its source's edges [E]
    DevBuf<int> revl;              // reverse edge per (node, slot) [V][maxdeg]
    DevBuf<int> rloc;              // coarse: slot of each edge's reverse among its source's edges [E]
    DevBuf<float> pw;              // WTA disparity of the sorted pixels [N]
    DevBuf<float> c0, dpos;        // per edge: max(dc0, .01) * bdl and max(dc1, 1) [E]
    DevBuf<float2> mxy;            // normalised mean position [V]
    DevBuf<int> child, child_off;  // nodes grouped by parent [V], [Vh+1]
    DevBuf<float4> plane;          // [V][K]
    DevBuf<float> cost, belief;    // [V][K]
    DevBuf<uint4> point;           // packed xy of the last RANSAC triple [V]
};

// one BP sweep of the schedule of hbp_g (hbp_isp.cu:1283-1311)
struct Sweep {
    int level, kind;   // kind: 1 RANSAC, 2 spatial propagation, 3 random adjustment
    float dz, dn;      // max_dz / max_dn of a random sweep
    bool first, last;  // first / last sweep of its level
};

struct BP {
    const Pyramid *pyr = nullptr;
    const Volume *vol = nullptr;
    int W = 0, H = 0, D = 0, L = 0, min_level = 2, use_cost_vol = 1;
    long N = 0;
    float lambda = 1000.1f, tau_s = 0.6f, tau = kCensusTau;
    int fx_shift = 0;
    std::vector<BPLevel> lv;
    DevBuf<float4> msg;            // [E][8] persistent messages of the current level (2 x float4)
    DevBuf<float> msg_t;           // [V][maxdeg][12] messages of the current sweep (node-major)
    DevBuf<uint32_t> crx;          // reversed, padded right census [H][crx_stride]
    int crx_stride = 0;
    DevBuf<unsigned long long> acc;  // [V][12] fixed-point plane-cost sums
    DevBuf<unsigned int> accnan;     // [V] NaN bits per slot
    // runs of fused sweeps on coarse levels (bp.hip, ChainIO): selections [Vc][K], candidates
    // [Vc][K], RANSAC points [Vc], a second sweep-message buffer and a second and third set of
    // sums, Vc the largest coarse level
    DevBuf<float4> ch_pl[2], ch_cand[3];
    DevBuf<float> ch_co[2];
    DevBuf<uint4> ch_pt[2];
    DevBuf<float> ch_mt;
    DevBuf<unsigned long long> ch_acc[2];
    DevBuf<unsigned int> ch_nan[2];
    // RANSAC plane costs ahead (fuse 2): per sweep of sched its index in its level's list of
    // RANSAC sweeps (the level's first sweep excluded) or -1; per level the list's length and
    // the first iteration and count of its random sweeps; the slices of the level pp_level
    // (bp.hip, PipeArgs), its launches, those issued so far and the last one the main stream waited for
    std::vector<int> ridx, rcount, rd_first, rd_count;
    // per sweep its index among its level's spatial sweeps or -1, per level their count; the
    // candidates' superpixels of every spatial sweep of the level pp_level [nS][V][P]
    std::vector<int> sidx, scount;
    DevBuf<int> pp_spc;
    DevBuf<float4> pp_cand;
    DevBuf<uint4> pp_pt;
    DevBuf<unsigned long long> pp_acc;
    DevBuf<unsigned int> pp_nan;
    int pp_level = -1, pp_batches = 0, pp_batch = 4, pp_first = 3, pp_waited = -1, pp_issued = 0;
    DevBuf<uint32_t> s_ransac, s_neighbor, s_child;
    // schedule state (hbp_g, hbp_isp.cu:1283-1311)
    int cur = -1;
    int sweeps = 0;
    bool mapped = false;
    std::vector<Sweep> sched;
    size_t next = 0;
    DevBuf<float4> best;           // per superpixel plane chosen by map_msg
    DevBuf<float> disp, plane3;    // outputs of map_msg [N], [N][3]
};

// cost.hip
void volume_build(Ctx &c, Volume &v, const uint8_t *d_bgrL, const uint8_t *d_bgrR, int W, int H, int D);
void volume_build_f32(Ctx &c, Volume &v, const float *d_raw, int W, int H, int D, float tau);
void volume_build_ad(Ctx &c, Volume &v, const uint8_t *d_bgrL, const uint8_t *d_bgrR, int W, int H, int D);
void volume_to_right(Ctx &c, Volume &v);
float *volume_f32(Ctx &c, Volume &v);
uint8_t *volume_u8(Ctx &c, Volume &v);

// pyramid.hip
void pyramid_build(Ctx &c, Pyramid &p, const uint8_t *d_bgr, int W, int H, int num_nb, int link = kCentroidLink,
                   int target_clus = 1, double sigma = 0.0);
void crtrees_labels(Ctx &c, const int *d_nn, int V, int *d_label);
int compact_labels(Ctx &c, int *d_label, int V);

// bp.hip
void bp_setup(Ctx &c, BP &b);
int bp_step(Ctx &c, BP &b, int max_sweeps);
void bp_map(Ctx &c, BP &b);
void bp_run_batch(Ctx &c, const std::vector<BP *> &bps);

// Rounds of batched sweeps over a changing set of BP objects (one sweep of each per round);
// see bp.hip.  Used by bp_run_batch and by the request server.
struct SweepArgs;
struct BatchItem;
constexpr int kBatchLevels = 24;   // pyramid levels a slot holds (N < 2^24 halves at least per level)
// one round's launches: its items (segment-form ones first), the launches' block counts and
// the level inits that open it
struct RoundPlan {
    std::vector<BatchItem> items;
    int ncoarse = 0, nfine = 0;
    int coarse_blocks = 0, fine_blocks = 0, node_blocks = 0, sp_blocks = 0, msg_blocks = 0;
    // algorithmic HBM bytes of each launch (the sums of the one-pair launches' figures)
    double coarse_bytes = 0.0, fine_bytes = 0.0, node_bytes = 0.0, sp_bytes = 0.0, msg_bytes = 0.0;
    std::vector<std::pair<BP *, int>> inits;   // (object, level)
};
struct BatchRunner {
    int nslots = 0, mode = -1;
    DevBuf<uint8_t> table;         // [slots][kBatchLevels] SweepArgs, pointers relative to the table
    DevBuf<BatchItem> items;       // one round's items (rounds are ordered on the stream)
    std::vector<int> used;
    // the rounds' items staged in pinned host memory, so that queuing a round never waits for
    // the device: a ring of kRing slots of nslots items, a slot reused once the event recorded
    // behind its copy has completed
    static constexpr int kRing = 256;
    BatchItem *ring = nullptr;
    hipEvent_t ev[kRing] = {};
    bool ev_live[kRing] = {};
    int ring_pos = 0;
    BatchRunner() = default;
    BatchRunner(const BatchRunner &) = delete;
    BatchRunner &operator=(const BatchRunner &) = delete;
    ~BatchRunner();
    void init(Ctx &c, int slots);
    int add(Ctx &c, BP &b);        // a free slot holding b's levels (b set up, same cost form)
    void remove(int slot);
    void round(Ctx &c, const std::vector<std::pair<BP *, int>> &active);   // (object, slot)
    // the next round of the objects (advancing their sweep positions); false when none is left
    bool plan(const std::vector<std::pair<BP *, int>> &active, RoundPlan &rp) const;
    // its launches, with the items already on the device
    void launch(Ctx &c, const RoundPlan &rp, const BatchItem *d_items) const;
};
void bp_get_state(Ctx &c, BP &b, float *planes, float *cost, float *belief, float *msg);
int bp_total_sweeps(int levels, int min_level, int D);

// lr.hip
void lr_check_fill(Ctx &c, const float *d_dispL, const float *d_dispR, float *d_planeL, int W, int H,
                   uint8_t *d_failL, uint8_t *d_failR, uint8_t *d_fail2);
void plane_to_disp(Ctx &c, const float *d_plane3, int W, int H, float *d_disp);
// the allocation-free form: scratch in a static device array, set up once per device
void lr_scratch_setup(int device);
long lr_noalloc_max_pixels();
void lr_check_fill_noalloc(Ctx &c, const float *d_dispL, const float *d_dispR, float *d_planeL, int W, int H,
                           uint8_t *d_failL, uint8_t *d_failR, uint8_t *d_fail2);
void weighted_median(Ctx &c, const uint8_t *d_bgr, const uint8_t *d_fail, float *d_plane3, int W, int H);
void median7_filter(Ctx &c, float *d_disp, int W, int H);
void eval_bad_pixels(Ctx &c, const float *d_disp, const float *d_gt, const uint8_t *d_nonocc, int W, int H,
                     double thr, long long *counts);

// primitives.hip
void scan_exclusive_int(Ctx &c, const int *d_in, int *d_out, long n, int *d_total);
// stable radix sorts; true: the sorted data are in the *_tmp buffers (swap them in)
[[nodiscard]] bool sort_pairs_u64(Ctx &c, unsigned long long *keys, unsigned long long *keys_tmp, unsigned int *vals,
                                  unsigned int *vals_tmp, long n, int end_bit);
[[nodiscard]] bool sort_keys_u64(Ctx &c, unsigned long long *keys, unsigned long long *keys_tmp, long n, int end_bit);
[[nodiscard]] bool sort_pairs_u32(Ctx &c, unsigned int *keys, unsigned int *keys_tmp, int *vals, int *vals_tmp,
                                  long n, int end_bit);

}  // namespace hbp
