// Request server: continuous batching of stereo pairs on one context.
//
// A request (two BGR images, their size and disparity range) waits in a FIFO until there is a
// free batch slot and its device memory fits the HBM budget.  It then goes to the admission
// thread, which builds the pair's cost volume, superpixel pyramid and BP set-up on a stream and
// memory pool of its own (the admission context) and drops the u8 cost volume, which BP does not
// read (it recomputes the costs it needs from the census words).  Admitted pairs join the rounds
// of the BatchRunner on the serving thread, which run one BP sweep of every admitted pair per
// round in shared launches.  When a pair's schedule is done its map and the copies of its
// disparity and planes into pinned host memory are queued behind its last round, with an event;
// the pair's device memory returns to the pool once that event has completed.  So neither
// admission nor completion makes the serving thread wait for the device, pairs of different
// sizes and arrival times overlap on the device, and each pair's result is the one its own
// hbp_bp_run gives (same kernels, same sums).
//
// Failures are per request: an error while admitting, mapping or copying out one pair fails that
// pair only; an error in a shared round fails the pairs of that round, after the device has
// finished the work already queued.  A failed, cancelled or finished pair never leaves a null or
// half-released entry behind.
//
// The reference has no serving layer (one pair per HBP_ISP::run_ms, main_stereo_hbp_isp.cu:
// 33-76); this is new.  hbp_submit, hbp_server_step, hbp_poll and hbp_cancel are called from
// one host thread; the admission thread is internal.
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "../../include/hbp_hip.h"
#include "handles.h"

namespace hbp {

struct Request {
    long long id = 0;
    int W = 0, H = 0, D = 0, min_level = 2;
    int coarse_min_pix = 96, single_hb = 2, batch_hb = 2, hamming_cols = 4, seg_len = 512;   // the serving context's forms
    int level_form[Ctx::kFormLevels] = {};
    std::vector<uint8_t> left, right;
    size_t est = 0;           // device bytes held while on the device (without the u8 volume)
    size_t est_admit = 0;     // peak during admission (with it)
    bool counted = false;     // est is in the server's in_use
    int slot = -1;
    int status = 0;           // 0 ok, < 0 failed (message in error)
    bool cancelled = false;
    std::string error;
    hipEvent_t done = nullptr;    // recorded behind the copies out (or behind the last round that read it)
    float *out = nullptr;         // pinned: disp [N], then plane [3N]
    size_t out_bytes = 0;
    Volume vol;
    Pyramid pyr;
    BP bp;
    void release_device() {
        bp = BP();
        pyr = Pyramid();
        vol = Volume();
    }
};

// pinned host blocks for the results, reused by size
struct PinnedCache {
    struct Blk {
        void *p;
        size_t n;
    };
    std::vector<Blk> free_;
    void *get(size_t n) {
        for (size_t i = 0; i < free_.size(); ++i)
            if (free_[i].n >= n && free_[i].n <= 2 * n) {
                void *p = free_[i].p;
                free_.erase(free_.begin() + (long)i);
                return p;
            }
        void *p = nullptr;
        HBP_CHECK(hipHostMalloc(&p, n, hipHostMallocDefault));
        return p;
    }
    void put(void *p, size_t n) {
        if (p) free_.push_back({p, n});
    }
    ~PinnedCache() {
        for (auto &b : free_) (void)hipHostFree(b.p);
    }
};

struct Server {
    Ctx *c = nullptr;
    Ctx admit_ctx;   // own stream and pool, used only by the admission thread
    int max_active = 0;
    int max_pending = 0;      // queued requests beyond which hbp_submit refuses (0: no limit)
    size_t budget = 0, in_use = 0;
    size_t peak_in_use = 0;   // the largest in_use reached (admission accounting)
    long long rejected = 0;   // submissions refused by max_pending
    float alpha = 1000.1f, tau_s = 0.6f;
    BatchRunner runner;
    PinnedCache pinned;
    std::deque<std::unique_ptr<Request>> pending;                 // not yet handed to admission
    std::vector<std::unique_ptr<Request>> active;                 // in the rounds
    std::vector<std::unique_ptr<Request>> draining;               // copies out or release queued
    std::unordered_map<long long, std::unique_ptr<Request>> finished;
    std::unordered_set<long long> in_admission, cancel_ids;   // serving thread only
    long long next_id = 1;
    // the admission thread's queues (guarded by mu)
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::unique_ptr<Request>> to_admit, admitted;
    int admitting = 0;    // requests handed over and not yet back
    bool stop = false;
    std::thread worker;
    ~Server();
};

// device bytes of one pair while it is on the device: census/WTA, the pyramid's packed arrays
// (2N nodes and 2*4*2N edges of capacity, segment.hpp:276-285), the BP levels (sorted pixels,
// segments, per-level structures and messages) and the padded right census, rounded up; plus
// the u8 volume during admission
static size_t estimate_bytes(int W, int H, int D) {
    const size_t N = (size_t)W * H;
    return 16 * N + 152 * N + 160 * N + 4 * (size_t)H * (W + D);
}
static size_t volume_bytes(int W, int H, int D) { return (size_t)D * H * round_up(W, 64); }

// cost volume, pyramid and BP set-up of one pair on the admission context (admission thread)
static void admit_work(Server &s, Request &r) {
    Ctx &c = s.admit_ctx;
    c.coarse_min_pix = r.coarse_min_pix;
    c.single_hb = r.single_hb;
    c.batch_hb = r.batch_hb;
    c.hamming_cols = r.hamming_cols;
    c.seg_len = r.seg_len;
    std::memcpy(c.level_form, r.level_form, sizeof(c.level_form));
    const size_t n3 = (size_t)r.W * r.H * 3;
    {
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
        // min_level < 0: BP on -min_level levels below the top (min_level = L - 1 + min_level)
        const int ml = r.min_level < 0 ? L - 1 + r.min_level : r.min_level;
        r.bp.min_level = std::max(1, std::min(ml, L - 1));
        bp_setup(c, r.bp);
        HBP_CHECK(hipStreamSynchronize(c.stream));
    }
    r.vol.vol.release();   // BP reads the census words and the WTA only
    std::vector<uint8_t>().swap(r.left);
    std::vector<uint8_t>().swap(r.right);
}

static void admission_loop(Server *sp) {
    Server &s = *sp;
    (void)hipSetDevice(s.admit_ctx.device);
    for (;;) {
        std::unique_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(s.mu);
            s.cv_work.wait(lk, [&] { return s.stop || !s.to_admit.empty(); });
            if (s.to_admit.empty()) return;   // stop, nothing left
            r = std::move(s.to_admit.front());
            s.to_admit.pop_front();
        }
        try {
            admit_work(s, *r);
        } catch (const HipError &e) {
            (void)hipStreamSynchronize(s.admit_ctx.stream);
            r->status = -1;
            r->error = e.msg;
        } catch (const std::bad_alloc &) {
            (void)hipStreamSynchronize(s.admit_ctx.stream);
            r->status = -1;
            r->error = "host allocation failed";
        }
        {
            std::lock_guard<std::mutex> lk(s.mu);
            s.admitted.push_back(std::move(r));
        }
        s.cv_done.notify_all();
    }
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
        for (auto &r : to_admit) {   // never started: nothing on the device
            r->status = -1;
            r->error = "server destroyed";
        }
        to_admit.clear();
    }
    cv_work.notify_all();
    if (worker.joinable()) worker.join();
    if (c && c->stream) (void)hipStreamSynchronize(c->stream);
    if (admit_ctx.stream) (void)hipStreamSynchronize(admit_ctx.stream);
    auto drop = [&](std::unique_ptr<Request> &r) {
        if (!r) return;
        if (r->done) (void)hipEventDestroy(r->done);
        if (r->out) pinned.put(r->out, r->out_bytes);
        r->release_device();
    };
    for (auto &r : admitted) drop(r);
    for (auto &r : active) drop(r);
    for (auto &r : draining) drop(r);
    for (auto &kv : finished) drop(kv.second);
    admitted.clear();
    active.clear();
    draining.clear();
    finished.clear();
    pending.clear();
    admit_ctx.scratch.release();
    admit_ctx.counters.release();
    admit_ctx.free_host_box();
    admit_ctx.free_side();
    admit_ctx.pool.release_all();
    if (admit_ctx.stream) (void)hipStreamDestroy(admit_ctx.stream);
}

// a request leaves for `finished`, failed; the device no longer reads its buffers (callers
// make sure of that: the stream that read them was synchronised, or nothing ran)
static void finish_failed(Server &s, std::unique_ptr<Request> rp, const std::string &msg) {
    Request &r = *rp;
    if (r.slot >= 0) s.runner.remove(r.slot);
    r.slot = -1;
    r.status = -1;
    r.error = msg;
    r.release_device();
    if (r.done) (void)hipEventDestroy(r.done);
    r.done = nullptr;
    if (r.out) s.pinned.put(r.out, r.out_bytes);
    r.out = nullptr;
    if (r.counted) s.in_use -= r.est;
    r.counted = false;
    s.finished[r.id] = std::move(rp);
}

// queue the map and the copies out behind the pair's last round (serving thread)
static void queue_completion(Server &s, Request &r) {
    Ctx &c = *s.c;
    bp_map(c, r.bp);
    const size_t n = (size_t)r.W * r.H;
    r.out_bytes = sizeof(float) * 4 * n;
    r.out = (float *)s.pinned.get(r.out_bytes);
    HBP_CHECK(hipMemcpyAsync(r.out, r.bp.disp.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.stream));
    HBP_CHECK(hipMemcpyAsync(r.out + n, r.bp.plane3.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c.stream));
    HBP_CHECK(hipEventCreateWithFlags(&r.done, hipEventDisableTiming));
    HBP_CHECK(hipEventRecord(r.done, c.stream));
}

// admit what fits, collect admitted pairs, one round, queue the completions, retire what the
// device has finished
static void step(Server &s, int *n_active, int *n_pending, int *n_done) {
    int done = 0;
    // 1. retire: pairs whose copies (or last reads) the device has finished
    {
        std::vector<std::unique_ptr<Request>> keep;
        for (auto &rp : s.draining) {
            hipError_t q = hipEventQuery(rp->done);
            if (q == hipErrorNotReady) {
                keep.push_back(std::move(rp));
                continue;
            }
            (void)hipEventDestroy(rp->done);
            rp->done = nullptr;
            if (q != hipSuccess) {
                (void)hipStreamSynchronize(s.c->stream);
                finish_failed(s, std::move(rp), std::string("copying the result out failed: ") + hipGetErrorString(q));
                continue;
            }
            rp->release_device();
            if (rp->counted) s.in_use -= rp->est;
            rp->counted = false;
            if (rp->cancelled) {
                finish_failed(s, std::move(rp), "cancelled");
                continue;
            }
            ++done;
            s.finished[rp->id] = std::move(rp);
        }
        s.draining.swap(keep);
    }
    // 2. hand pending pairs to the admission thread while slots and budget allow
    {
        int in_flight;
        {
            std::lock_guard<std::mutex> lk(s.mu);
            in_flight = s.admitting;
        }
        while (!s.pending.empty() && (int)s.active.size() + in_flight < s.max_active) {
            std::unique_ptr<Request> &r = s.pending.front();
            if (r->est_admit > s.budget) {
                std::unique_ptr<Request> rp = std::move(r);
                s.pending.pop_front();
                finish_failed(s, std::move(rp), "pair does not fit the server's HBM budget");
                continue;
            }
            // the volume is held only during admission, one admission at a time
            if (s.in_use + r->est_admit > s.budget) break;
            std::unique_ptr<Request> rp = std::move(r);
            s.pending.pop_front();
            s.in_use += rp->est;
            s.peak_in_use = std::max(s.peak_in_use, s.in_use + rp->est_admit - rp->est);
            rp->counted = true;
            s.in_admission.insert(rp->id);
            rp->coarse_min_pix = s.c->coarse_min_pix;
            rp->single_hb = s.c->single_hb;
            rp->batch_hb = s.c->batch_hb;
            rp->hamming_cols = s.c->hamming_cols;
            rp->seg_len = s.c->seg_len;
            std::memcpy(rp->level_form, s.c->level_form, sizeof(rp->level_form));
            {
                std::lock_guard<std::mutex> lk(s.mu);
                s.to_admit.push_back(std::move(rp));
                ++s.admitting;
                ++in_flight;
            }
            s.cv_work.notify_one();
        }
    }
    // 3. admitted pairs join the rounds
    {
        std::deque<std::unique_ptr<Request>> got;
        {
            std::unique_lock<std::mutex> lk(s.mu);
            // nothing else to do: wait a little for an admission instead of spinning
            if (s.admitted.empty() && s.active.empty() && s.draining.empty() && s.admitting > 0)
                s.cv_done.wait_for(lk, std::chrono::milliseconds(2), [&] { return !s.admitted.empty(); });
            got.swap(s.admitted);
            s.admitting -= (int)got.size();
        }
        for (auto &rp : got) {
            s.in_admission.erase(rp->id);
            if (rp->status != 0) {
                std::string msg = rp->error;
                finish_failed(s, std::move(rp), msg);
                continue;
            }
            if (s.cancel_ids.erase(rp->id)) {
                finish_failed(s, std::move(rp), "cancelled");
                continue;
            }
            try {
                rp->slot = s.runner.add(*s.c, rp->bp);
                s.active.push_back(std::move(rp));
            } catch (const HipError &e) {
                (void)hipStreamSynchronize(s.c->stream);
                finish_failed(s, std::move(rp), e.msg);
            }
        }
    }
    // 4. one round of every active pair
    if (!s.active.empty()) {
        std::vector<std::pair<BP *, int>> pairs;
        for (auto &r : s.active) pairs.push_back({&r->bp, r->slot});
        bool round_ok = true;
        std::string msg;
        try {
            fault_point(kFaultRound, "batched round");
            s.runner.round(*s.c, pairs);
        } catch (const HipError &e) {
            round_ok = false;
            msg = e.msg;
        }
        std::vector<std::unique_ptr<Request>> still;
        if (!round_ok) {
            // the pairs' states are undefined after a partial round: fail them all once the
            // device has stopped reading their buffers
            (void)hipStreamSynchronize(s.c->stream);
            for (auto &rp : s.active) finish_failed(s, std::move(rp), "batched round failed: " + msg);
        } else {
            for (auto &rp : s.active) {
                if (rp->bp.next < rp->bp.sched.size()) {
                    still.push_back(std::move(rp));
                    continue;
                }
                s.runner.remove(rp->slot);
                rp->slot = -1;
                try {
                    queue_completion(s, *rp);
                    s.draining.push_back(std::move(rp));
                } catch (const HipError &e) {
                    (void)hipStreamSynchronize(s.c->stream);
                    finish_failed(s, std::move(rp), e.msg);
                }
            }
        }
        s.active.swap(still);
    }
    if (n_active) *n_active = (int)s.active.size();
    if (n_pending) *n_pending = (int)(s.pending.size() + s.in_admission.size());
    if (n_done) *n_done = done;
}

// hbp_cancel: 0 if the ticket was queued, running or finished (it then polls as "cancelled",
// or is gone if it had finished), 1 if it is unknown
static int cancel(Server &s, long long id) {
    for (auto it = s.pending.begin(); it != s.pending.end(); ++it)
        if ((*it)->id == id) {
            std::unique_ptr<Request> rp = std::move(*it);
            s.pending.erase(it);
            finish_failed(s, std::move(rp), "cancelled");
            return 0;
        }
    if (s.in_admission.count(id)) {   // leaves as cancelled when admission hands it back
        s.cancel_ids.insert(id);
        return 0;
    }
    for (auto it = s.active.begin(); it != s.active.end(); ++it)
        if ((*it)->id == id) {
            // leaves the rounds now; its buffers return once the queued rounds have run
            std::unique_ptr<Request> rp = std::move(*it);
            s.active.erase(it);
            s.runner.remove(rp->slot);
            rp->slot = -1;
            rp->cancelled = true;
            try {
                HBP_CHECK(hipEventCreateWithFlags(&rp->done, hipEventDisableTiming));
                HBP_CHECK(hipEventRecord(rp->done, s.c->stream));
                s.draining.push_back(std::move(rp));
            } catch (const HipError &) {
                (void)hipStreamSynchronize(s.c->stream);
                finish_failed(s, std::move(rp), "cancelled");
            }
            return 0;
        }
    for (auto &r : s.draining)
        if (r->id == id) {
            r->cancelled = true;
            return 0;
        }
    auto f = s.finished.find(id);
    if (f != s.finished.end()) {
        if (f->second->out) s.pinned.put(f->second->out, f->second->out_bytes);
        s.finished.erase(f);
        return 0;
    }
    return 1;
}

}  // namespace hbp

struct hbp_server {
    hbp::Server s;
    hbp_ctx *owner = nullptr;
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
        srv->s.worker = std::thread(admission_loop, &srv->s);
    } catch (...) {
        delete srv;
        throw;
    }
    srv->owner = ctx;
    ++ctx->live;
    *out = srv;
    HBP_API_END
}

int hbp_server_destroy(hbp_server *srv) {
    HBP_API_BEGIN
    if (!srv) return 0;
    HBP_CHECK(hipSetDevice(srv->s.c->device));
    hbp_ctx *o = srv->owner;
    delete srv;   // stops the admission thread, waits for both streams, returns the memory
    if (o) --o->live;
    HBP_API_END
}

int hbp_submit(hbp_server *srv, const uint8_t *bgr_left, const uint8_t *bgr_right, int width, int height, int labels,
               int min_level, long long *ticket) {
    HBP_API_BEGIN
    HBP_REQUIRE(srv && bgr_left && bgr_right && ticket, "NULL argument");
    HBP_REQUIRE(width >= 9 && height >= 7 && labels >= 1 && labels <= width && labels <= 1024,
                "bad image size or disparity range");
    HBP_REQUIRE((long)width * height < (1L << 24) && width < 65536 && height < 65536, "image too large");
    if (srv->s.max_pending > 0 &&
        (long)(srv->s.pending.size() + srv->s.in_admission.size()) >= srv->s.max_pending) {
        ++srv->s.rejected;
        throw HipError{"server queue full: request refused"};
    }
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
    r->est_admit = r->est + volume_bytes(width, height, labels);
    *ticket = r->id;
    srv->s.pending.push_back(std::move(r));
    HBP_API_END
}

int hbp_server_set_max_pending(hbp_server *srv, int max_pending) {
    HBP_API_BEGIN
    HBP_REQUIRE(srv && max_pending >= 0, "bad argument");
    srv->s.max_pending = max_pending;
    HBP_API_END
}

int hbp_server_stats(hbp_server *srv, long long *bytes_in_use, long long *peak_bytes, long long *rejected) {
    HBP_API_BEGIN
    HBP_REQUIRE(srv, "NULL argument");
    if (bytes_in_use) *bytes_in_use = (long long)srv->s.in_use;
    if (peak_bytes) *peak_bytes = (long long)srv->s.peak_in_use;
    if (rejected) *rejected = srv->s.rejected;
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
        Server &s = srv->s;
        auto it = s.finished.find(ticket);
        if (it == s.finished.end()) {
            if (ticket >= 1 && ticket < s.next_id) {
                for (auto &r : s.pending)
                    if (r->id == ticket) return 1;
                for (auto &r : s.active)
                    if (r->id == ticket) return 1;
                for (auto &r : s.draining)
                    if (r->id == ticket) return 1;
                if (s.in_admission.count(ticket)) return 1;
            }
            throw HipError{"unknown ticket"};
        }
        Request &r = *it->second;
        if (r.status != 0) {
            std::string msg = r.error;
            s.finished.erase(it);
            throw HipError{msg};
        }
        const size_t n = (size_t)r.W * r.H;
        if (disp) std::memcpy(disp, r.out, sizeof(float) * n);
        if (plane3) std::memcpy(plane3, r.out + n, sizeof(float) * 3 * n);
        s.pinned.put(r.out, r.out_bytes);
        r.out = nullptr;
        s.finished.erase(it);
        return 0;
    } catch (const HipError &e) {
        set_error(e.msg);
        return -1;
    }
}

int hbp_cancel(hbp_server *srv, long long ticket) {
    HBP_API_BEGIN
    HBP_REQUIRE(srv, "NULL argument");
    HBP_CHECK(hipSetDevice(srv->s.c->device));
    int r = cancel(srv->s, ticket);
    HBP_REQUIRE(r == 0, "unknown ticket");
    HBP_API_END
}

}  // extern "C"
