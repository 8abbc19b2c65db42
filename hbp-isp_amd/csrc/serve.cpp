// Request server: continuous batching of stereo pairs on one context.
//
// A request (two BGR images, their size and disparity range) waits in a FIFO until there is
// a free batch slot and its estimated device memory fits the HBM budget.  Admission runs the
// pair's cost volume, superpixel pyramid and BP set-up on the context's stream; the pair then
// joins the rounds of the BatchRunner, which run one BP sweep of every admitted pair per round
// in shared launches.  A pair whose schedule is done is mapped, its disparity and planes are
// copied to host memory, and its device memory returns to the context's pool.  Pairs enter
// and leave between rounds, so pairs of different sizes and arrival times overlap on the
// device; each pair's result is the one its own hbp_bp_run gives (same kernels, same sums).
//
// Admission runs on a second stream with its own memory pool (the admission context), so that
// a new pair's cost volume and pyramid overlap the BP rounds already queued on the main stream;
// admit() returns once that work is done, and the pair's buffers return to the admission pool
// only after complete() has synchronised the main stream.  Rounds stage their items in pinned
// memory and never wait for the device, so the host queues rounds ahead of it.
//
// The reference has no serving layer (one pair per HBP_ISP::run_ms, main_stereo_hbp_isp.cu:
// 33-76); this is new.  A server is driven by one host thread: hbp_submit queues, and every
// hbp_server_step admits, runs one round and completes finished pairs.
#include <cstring>
#include <deque>
#include <memory>
#include <unordered_map>

#include "../../include/hbp_hip.h"
#include "handles.h"

namespace hbp {

struct Request {
    long long id = 0;
    int W = 0, H = 0, D = 0, min_level = 2;
    std::vector<uint8_t> left, right;
    std::vector<float> disp, plane;
    size_t est = 0;
    int slot = -1;
    int status = 0;         // 0 ok, < 0 failed (message in error)
    std::string error;
    Volume vol;
    Pyramid pyr;
    BP bp;
};

struct Server {
    Ctx *c = nullptr;
    Ctx admit_ctx;   // own stream and pool; see above
    int max_active = 0;
    size_t budget = 0, in_use = 0;
    float alpha = 1000.1f, tau_s = 0.6f;
    BatchRunner runner;
    std::deque<std::unique_ptr<Request>> pending;
    std::vector<std::unique_ptr<Request>> active;
    std::unordered_map<long long, std::unique_ptr<Request>> finished;
    long long next_id = 1;
    ~Server() {
        // the requests' buffers back to their pools, then the admission context's memory
        finished.clear();
        active.clear();
        pending.clear();
        admit_ctx.scratch.release();
        admit_ctx.counters.release();
        admit_ctx.free_host_box();
        admit_ctx.pool.release_all();
        if (admit_ctx.stream) (void)hipStreamDestroy(admit_ctx.stream);
    }
};

// device bytes of one pair: u8 volume, census/WTA, the pyramid's packed arrays (2N nodes and
// 2*4*2N edges of capacity, segment.hpp:276-285), the BP levels (sorted pixels, segments,
// per-level structures and messages) and the padded right census, rounded up
static size_t estimate_bytes(int W, int H, int D) {
    const size_t N = (size_t)W * H;
    return (size_t)D * H * round_up(W, 64) + 16 * N + 152 * N + 160 * N + 4 * (size_t)H * (W + D);
}

static void admit(Server &s, Request &r) {
    Ctx &c = s.admit_ctx;
    c.coarse_min_pix = s.c->coarse_min_pix;   // the forms the main context is tuned to
    c.single_hb = s.c->single_hb;
    c.batch_hb = s.c->batch_hb;
    const size_t n3 = (size_t)r.W * r.H * 3;
    DevBuf<uint8_t> dl(c.pool, n3), dr(c.pool, n3);
    HBP_CHECK(hipMemcpyAsync(dl.p, r.left.data(), n3, hipMemcpyHostToDevice, c.stream));
    HBP_CHECK(hipMemcpyAsync(dr.p, r.right.data(), n3, hipMemcpyHostToDevice, c.stream));
    volume_build(c, r.vol, dl.p, dr.p, r.W, r.H, r.D);
    pyramid_build(c, r.pyr, dl.p, r.W, r.H, 4);
    r.bp.pyr = &r.pyr;
    r.bp.vol = &r.vol;
    r.bp.lambda = s.alpha;
    r.bp.tau_s = s.tau_s;
    r.bp.use_cost_vol = 1;
    const int L = r.pyr.levels();
    HBP_REQUIRE(L >= 2, "image too small for a pyramid of two levels");
    r.bp.min_level = std::max(1, std::min(r.min_level, L - 1));
    bp_setup(c, r.bp);
    HBP_CHECK(hipStreamSynchronize(c.stream));
    r.slot = s.runner.add(*s.c, r.bp);
    std::vector<uint8_t>().swap(r.left);
    std::vector<uint8_t>().swap(r.right);
}

static void complete(Server &s, std::unique_ptr<Request> &rp) {
    Ctx &c = *s.c;
    Request &r = *rp;
    bp_map(c, r.bp);
    const size_t n = (size_t)r.W * r.H;
    r.disp.resize(n);
    r.plane.resize(3 * n);
    HBP_CHECK(hipMemcpyAsync(r.disp.data(), r.bp.disp.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.stream));
    HBP_CHECK(hipMemcpyAsync(r.plane.data(), r.bp.plane3.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c.stream));
    HBP_CHECK(hipStreamSynchronize(c.stream));
    s.runner.remove(r.slot);
    r.slot = -1;
    // device memory back to the pool (stream-ordered: later work runs after this pair's)
    r.bp = BP();
    r.pyr = Pyramid();
    r.vol = Volume();
    s.in_use -= r.est;
    s.finished[r.id] = std::move(rp);
}

static void fail(Server &s, std::unique_ptr<Request> &rp, const std::string &msg) {
    Request &r = *rp;
    if (r.slot >= 0) s.runner.remove(r.slot);
    r.slot = -1;
    r.status = -1;
    r.error = msg;
    r.bp = BP();
    r.pyr = Pyramid();
    r.vol = Volume();
    s.finished[r.id] = std::move(rp);
}

// admit what fits, one round, complete finished pairs
static void step(Server &s, int *n_active, int *n_pending, int *n_done) {
    int done = 0;
    while (!s.pending.empty() && (int)s.active.size() < s.max_active) {
        std::unique_ptr<Request> &r = s.pending.front();
        if (r->est > s.budget) {
            std::unique_ptr<Request> rp = std::move(r);
            s.pending.pop_front();
            fail(s, rp, "pair does not fit the server's HBM budget");
            continue;
        }
        if (s.in_use + r->est > s.budget) break;
        std::unique_ptr<Request> rp = std::move(r);
        s.pending.pop_front();
        s.in_use += rp->est;
        try {
            admit(s, *rp);
            s.active.push_back(std::move(rp));
        } catch (const HipError &e) {
            s.in_use -= rp->est;
            fail(s, rp, e.msg);
        }
    }
    if (!s.active.empty()) {
        std::vector<std::pair<BP *, int>> pairs;
        for (auto &r : s.active) pairs.push_back({&r->bp, r->slot});
        s.runner.round(*s.c, pairs);
        std::vector<std::unique_ptr<Request>> still;
        for (auto &r : s.active) {
            if (r->bp.next >= r->bp.sched.size()) {
                complete(s, r);
                ++done;
            } else {
                still.push_back(std::move(r));
            }
        }
        s.active.swap(still);
    }
    if (n_active) *n_active = (int)s.active.size();
    if (n_pending) *n_pending = (int)s.pending.size();
    if (n_done) *n_done = done;
}

}  // namespace hbp

struct hbp_server {
    hbp::Server s;
};

extern "C" {

int hbp_server_create(hbp_ctx *ctx, int max_active, long long hbm_budget_bytes, float alpha, float tau_smooth,
                      hbp_server **out) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && out, "NULL argument");
    HBP_REQUIRE(max_active >= 1 && max_active <= 4096, "max_active must be in [1, 4096]");
    HBP_REQUIRE(hbm_budget_bytes > 0, "the HBM budget must be positive");
    *out = nullptr;
    HBP_CHECK(hipSetDevice(ctx->c.device));
    auto *srv = new hbp_server();
    srv->s.c = &ctx->c;
    srv->s.max_active = max_active;
    srv->s.budget = (size_t)hbm_budget_bytes;
    srv->s.alpha = alpha;
    srv->s.tau_s = tau_smooth;
    Ctx &ac = srv->s.admit_ctx;
    ac.device = ctx->c.device;
    ac.num_cu = ctx->c.num_cu;
    ac.coop_blocks = ctx->c.coop_blocks;
    ac.drain_before_coop = ctx->c.stream;
    try {
        HBP_CHECK(hipStreamCreateWithFlags(&ac.stream, hipStreamNonBlocking));
        srv->s.runner.init(ctx->c, max_active);
    } catch (...) {
        delete srv;
        throw;
    }
    *out = srv;
    HBP_API_END
}

int hbp_server_destroy(hbp_server *srv) {
    HBP_API_BEGIN
    if (!srv) return 0;
    HBP_CHECK(hipSetDevice(srv->s.c->device));
    HBP_CHECK(hipStreamSynchronize(srv->s.c->stream));
    if (srv->s.admit_ctx.stream) HBP_CHECK(hipStreamSynchronize(srv->s.admit_ctx.stream));
    delete srv;
    HBP_API_END
}

int hbp_submit(hbp_server *srv, const uint8_t *bgr_left, const uint8_t *bgr_right, int width, int height, int labels,
               int min_level, long long *ticket) {
    HBP_API_BEGIN
    HBP_REQUIRE(srv && bgr_left && bgr_right && ticket, "NULL argument");
    HBP_REQUIRE(width >= 9 && height >= 7 && labels >= 1 && labels <= width && labels <= 1024,
                "bad image size or disparity range");
    auto r = std::make_unique<Request>();
    r->id = srv->s.next_id++;
    r->W = width;
    r->H = height;
    r->D = labels;
    r->min_level = min_level;
    const size_t n3 = (size_t)width * height * 3;
    r->left.assign(bgr_left, bgr_left + n3);
    r->right.assign(bgr_right, bgr_right + n3);
    r->est = estimate_bytes(width, height, labels);
    *ticket = r->id;
    srv->s.pending.push_back(std::move(r));
    HBP_API_END
}

int hbp_server_step(hbp_server *srv, int *n_active, int *n_pending, int *n_done) {
    HBP_API_BEGIN
    HBP_REQUIRE(srv, "NULL argument");
    HBP_CHECK(hipSetDevice(srv->s.c->device));
    step(srv->s, n_active, n_pending, n_done);
    HBP_API_END
}

int hbp_poll(hbp_server *srv, long long ticket, float *disp, float *plane3) {
    try {
        HBP_REQUIRE(srv, "NULL argument");
        auto it = srv->s.finished.find(ticket);
        if (it == srv->s.finished.end()) {
            for (auto &r : srv->s.pending)
                if (r->id == ticket) return 1;
            for (auto &r : srv->s.active)
                if (r->id == ticket) return 1;
            throw HipError{"unknown ticket"};
        }
        Request &r = *it->second;
        if (r.status != 0) {
            std::string msg = r.error;
            srv->s.finished.erase(it);
            throw HipError{msg};
        }
        if (disp) std::memcpy(disp, r.disp.data(), sizeof(float) * r.disp.size());
        if (plane3) std::memcpy(plane3, r.plane.data(), sizeof(float) * r.plane.size());
        srv->s.finished.erase(it);
        return 0;
    } catch (const HipError &e) {
        set_error(e.msg);
        return -1;
    }
}

}  // extern "C"
