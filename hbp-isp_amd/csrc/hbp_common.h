// Shared declarations of the HBP-ISP HIP library (gfx950 / CDNA4).
//
// All device arithmetic in this library is compiled with -ffp-contract=off and IEEE division
// and square root, so every float expression rounds exactly as written; the scalar rules
// below (cube root, exp, floor-to-int, fixed-point sums, LCG) are the pinned forms of the
// reference's FP64 libm calls and unordered atomics.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace hbp {

// ---------------------------------------------------------------------------------------
// error handling: every C-ABI entry point returns 0 or a negative code and records a message
// ---------------------------------------------------------------------------------------
void set_error(const std::string &msg);
struct HipError {
    std::string msg;
};
// thrown instead of a cooperative launch on a context that forbids them (a worker whose stream
// runs beside other streams' kernels, so the grid's co-residency could not be guaranteed); the
// caller redoes the work where a cooperative launch is safe
struct CoopNeeded {};
#define HBP_CHECK(expr)                                                                    \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            throw ::hbp::HipError{std::string(#expr) + ": " + hipGetErrorString(_e) + " (" + \
                                  __FILE__ + ":" + std::to_string(__LINE__) + ")"};        \
    } while (0)
#define HBP_REQUIRE(cond, msg)                          \
    do {                                                \
        if (!(cond)) throw ::hbp::HipError{std::string(msg)}; \
    } while (0)

// ---------------------------------------------------------------------------------------
// pinned scalar rules (device + host)
// ---------------------------------------------------------------------------------------

// Cube root on the Lab f(t) domain (216/24389, 1]: Newton y <- (2y + x/y^2)/3 from
// 0.45 + 0.55x, eight steps.
__host__ __device__ inline double cbrt_newton(double x) {
    double y = 0.45 + 0.55 * x;
#pragma unroll
    for (int i = 0; i < 8; ++i) y = (2.0 * y + x / (y * y)) / 3.0;
    return y;
}

// t^2.4 on the sRGB linearisation domain ((0.04045 + 0.055)/1.055, 1 + 2^-20]: t^2 times the
// fifth root of t^2, the root by Newton's step y <- y + (a/y^4 - y)/5 from 0.6 + 0.4a, eight
// steps.
__host__ __device__ inline double pow24_newton(double t) {
    double a = t * t;
    double y = 0.6 + 0.4 * a;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        double y2 = y * y;
        y = y + (a / (y2 * y2) - y) / 5.0;
    }
    return a * y;
}

// exp(x), x <= 0: Cody-Waite split of ln2 (high part with 32 trailing zero bits), degree-13
// Taylor polynomial on |r| <= ln2/2 (Horner), scaled by 2^n.
__host__ __device__ inline double exp_cody_waite(double x) {
    if (x < -745.0) return 0.0;
    const double inv_ln2 = 1.4426950408889634;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double n = rint(x * inv_ln2);
    double r = (x - n * ln2_hi) - n * ln2_lo;
    const double c[14] = {1.0,
                          1.0,
                          0.5,
                          0.16666666666666666,
                          0.041666666666666664,
                          0.008333333333333333,
                          0.001388888888888889,
                          0.0001984126984126984,
                          2.48015873015873e-05,
                          2.7557319223985893e-06,
                          2.755731922398589e-07,
                          2.505210838544172e-08,
                          2.08767569878681e-09,
                          1.6059043836821613e-10};
    double p = c[13];
#pragma unroll
    for (int k = 12; k >= 0; --k) p = c[k] + r * p;
    return ldexp(p, (int)n);
}

// floor to int with NaN -> 0 and saturation (the semantics of __float2int_rd).
__host__ __device__ inline int floor_to_int(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT_MAX;
    if (x < -2147483648.0f) return INT_MIN;
    return (int)floorf(x);
}

// floor_to_int with selects only (the conversion of an in-range integer is exact)
__host__ __device__ inline int floor_to_int_sel(float x) {
    float f = fminf(fmaxf(floorf(x), -2147483648.0f), 2147483520.0f);
    int i = (int)f;
    i = x >= 2147483648.0f ? INT_MAX : i;
    return x != x ? 0 : i;
}

// rn(x / 3) for x >= +0 from a multiplication by rn(1/3) and one FMA correction.  Checked
// exhaustively against the IEEE quotient over all 2^32 floats: equal for every finite x
// except -0 (gives +0); +-inf give NaN.  The caller's x is a sum of absolute values (+0,
// finite, inf or NaN) and goes through fminf(x / 3 / dpos, tau_s), where inf and NaN both
// give tau_s, so the result is identical to the division.
__host__ __device__ inline float div3_nonneg(float x) {
    const float y = 0.333333343267440796f;
    const float q0 = x * y;
    const float r = fmaf(-q0, 3.0f, x);
    return fmaf(r, y, q0);
}

// int64 -> float, round to nearest even, from integer operations only.
__host__ __device__ inline float int64_to_float_rne(int64_t q) {
    if (q == 0) return 0.0f;
    bool neg = q < 0;
    uint64_t u = neg ? (uint64_t)0 - (uint64_t)q : (uint64_t)q;
#ifdef __HIP_DEVICE_COMPILE__
    int msb = 63 - __clzll((long long)u);
#else
    int msb = 63 - __builtin_clzll(u);
#endif
    float f;
    if (msb <= 23) {
        f = (float)(uint32_t)u;
    } else {
        int sh = msb - 23;
        uint64_t m = u >> sh;
        uint64_t rem = u & ((1ull << sh) - 1ull);
        uint64_t half = 1ull << (sh - 1);
        if (rem > half || (rem == half && (m & 1ull))) m += 1ull;
        f = ldexpf((float)(uint32_t)m, sh);
    }
    return neg ? -f : f;
}

// Cluster sums of per-pixel plane costs are exact sums of q = rint(|v| 2^s) (below 2^52, so
// exact in double per lane and in int64 across lanes), with lambda's sign; the cost is
// int64_to_float_rne(sum) * 2^-s, or the quiet NaN when any term was NaN.
int fixed_point_shift(float lambda, float tau, int use_cost_vol, long npix);
__host__ __device__ inline float from_fixed(int64_t q, bool nan, int s) {
    if (nan) return __int_as_float(0x7FC00000);
    return ldexpf(int64_to_float_rne(q), -s);
}

// BSD LCG: s' = (1103515245 s + 12345) mod 2^31 in uint32 arithmetic.
__host__ __device__ inline uint32_t lcg_next(uint32_t s) { return (1103515245u * s + 12345u) % 2147483648u; }
// n steps of lcg_next are one affine map mod 2^31: s -> (a s + c) mod 2^31 with a = A^n and
// c = C (1 + A + ... + A^(n-1)), both taken mod 2^32 (2^31 divides 2^32, so reducing once at the
// end gives the stepwise value)
struct LcgJump {
    uint32_t a, c;
};
__host__ __device__ constexpr LcgJump lcg_jump_coef(int n) {
    uint32_t a = 1u, c = 0u;
    for (int i = 0; i < n; ++i) {
        c = 1103515245u * c + 12345u;
        a = 1103515245u * a;
    }
    return LcgJump{a, c};
}
// STEP * j steps of lcg_next from s, j in [0, 3] (per lane), without a loop
template <int STEP>
__host__ __device__ inline uint32_t lcg_skip(uint32_t s, int j) {
    constexpr LcgJump j1 = lcg_jump_coef(STEP), j2 = lcg_jump_coef(2 * STEP), j3 = lcg_jump_coef(3 * STEP);
    const uint32_t a = j == 1 ? j1.a : (j == 2 ? j2.a : j3.a);
    const uint32_t c = j == 1 ? j1.c : (j == 2 ? j2.c : j3.c);
    return j == 0 ? s : (a * s + c) % 2147483648u;
}
// n steps of lcg_next as one affine map, by squaring (all such maps commute)
__host__ __device__ inline LcgJump lcg_jump_pow(unsigned long n) {
    LcgJump r{1u, 0u}, b{1103515245u, 12345u};
    while (n) {
        if (n & 1u) r = LcgJump{b.a * r.a, b.a * r.c + b.c};
        b = LcgJump{b.a * b.a, b.a * b.c + b.c};
        n >>= 1;
    }
    return r;
}
__host__ __device__ inline float lcg_uniform(uint32_t &s, float f) {
    s = lcg_next(s);
    return (float)s / 2147483648.0f * f;
}
// unsigned modulo with a zero divisor mapped to 0
__host__ __device__ inline uint32_t umod0(uint32_t s, int m) { return (uint32_t)m == 0u ? 0u : s % (uint32_t)m; }

// ---------------------------------------------------------------------------------------
// constants of the method (hbp_isp.hpp:265-283)
// ---------------------------------------------------------------------------------------
constexpr int kP = 4;         // particles kept per superpixel
constexpr int kK = 3 * kP;    // best, prior and candidate slots
constexpr float kCensusTau = 25.0f;
// SegHAC linkages (segment.hpp:13)
constexpr int kMinLink = 0, kMaxLink = 1, kCentroidLink = 2, kWardLink = 3;
constexpr float kMcCnnTau = 0.5f;   // compute_stereo_cost("MC-CNN"), stereo_cost.cu:437-438
constexpr float kAdTau = 30.0f;     // compute_stereo_cost("AD"), stereo_cost.cu:426-427

// packed pixel coordinate (x | y << 16)
__host__ __device__ inline uint32_t pack_xy(int x, int y) { return (uint32_t)x | ((uint32_t)y << 16); }
__host__ __device__ inline int unpack_x(uint32_t p) { return (int)(p & 0xffffu); }
__host__ __device__ inline int unpack_y(uint32_t p) { return (int)(p >> 16); }

// ---------------------------------------------------------------------------------------
// device memory: a per-context caching pool (blocks are reused by exact size)
// ---------------------------------------------------------------------------------------
// (thread-safe: a server's admission thread allocates from its pool while the serving thread
// returns finished pairs' blocks to it)
class Pool {
public:
    void *get(size_t bytes);
    void put(void *p);
    void release_all();
    ~Pool() { release_all(); }
    size_t bytes_in_use() const { return in_use_; }
    size_t bytes_reserved() const { return reserved_; }

private:
    struct Blk {
        void *p;
        size_t n;
    };
    std::mutex mu_;
    std::vector<Blk> free_, used_;
    size_t in_use_ = 0, reserved_ = 0;
};

// Host-forced failures for tests of the error paths (hbp_inject_fault): the n-th next device
// allocation from any pool, or the n-th next batched round, throws instead of running.  No
// kernel is ever started in a failing state.
enum FaultKind { kFaultAlloc = 0, kFaultRound = 1, kFaultKinds = 2 };
void fault_arm(int kind, long long after);
void fault_point(int kind, const char *what);   // throws when the armed count reaches this call

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    Pool *pool = nullptr;
    DevBuf() = default;
    DevBuf(Pool &pl, size_t count) { alloc(pl, count); }
    void alloc(Pool &pl, size_t count) {
        release();
        pool = &pl;
        n = count;
        p = count ? (T *)pl.get(count * sizeof(T)) : nullptr;
    }
    void release() {
        if (p && pool) pool->put(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n), pool(o.pool) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            pool = o.pool;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
};

// Per-kernel timing with HIP events recorded on the context's stream around selected
// launches, with the launch's algorithmic HBM bytes; off unless enabled through the C ABI.
struct KernelStat {
    double ms = 0.0, bytes = 0.0;
    long long launches = 0;
};
struct Profiler {
    bool on = false;
    std::string only;   // when set, only launches whose name starts with it are timed
    int every = 1;      // of those, one launch in `every` is timed
    long long seen = 0;
    std::vector<std::string> names;
    std::vector<KernelStat> stats;
    struct Pending {
        int id;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    int id_of(const std::string &n) {
        for (size_t i = 0; i < names.size(); ++i)
            if (names[i] == n) return (int)i;
        names.push_back(n);
        stats.emplace_back();
        return (int)names.size() - 1;
    }
    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    void flush() {
        for (auto &p : pending) {
            float ms = 0.f;
            if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                stats[p.id].ms += ms;
                stats[p.id].bytes += p.bytes;
                stats[p.id].launches += 1;
            }
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    void reset() {
        flush();
        for (auto &s : stats) s = KernelStat();
    }
    ~Profiler() {
        for (auto &p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// a CRTrees schedule of the scheduled form: rounds, pointer-jumping steps per round,
// max-jumping steps (the defaults are the full bounds)
struct CrtSched {
    int rounds = 5, jumps = 8, max_jumps = 8;
};

struct Ctx {
    int device = 0;
    Profiler prof;
    hipStream_t stream = nullptr;
    Pool pool;
    int num_cu = 256;
    int coop_blocks = 0;     // co-resident blocks of the CRTrees cooperative kernel
    // kernel forms, chosen by a tuner; every choice gives the same bits (the plane-cost sums
    // are exact in any order): the average pixels per superpixel from which a BP level takes
    // the segment form of the plane cost (below it a segment wavefront holds one small
    // superpixel; measured crossover 64-128), and the steps per batch of its wavefronts in
    // single-pair launches and in batched ones (2: fewer registers, so that with the plane
    // costs computed ahead on the second stream two or three of a spatial sweep's wavefronts
    // fit beside the ahead kernel's on a SIMD, and more resident wavefronts when many pairs
    // share a launch; 4: more gathers in flight per wavefront)
    int coarse_min_pix = 96, single_hb = 2, batch_hb = 2;
    // columns per lane of the Hamming cost kernel (4 or 8; the wider form writes 512 instead of
    // 256 contiguous bytes per wavefront store)
    int hamming_cols = 4;
    // pixels per segment wavefront of the plane-cost segment form (a multiple of 64 up to 512):
    // shorter segments mean more, shorter wavefronts per launch
    int seg_len = 512;
    // batched BP: 1 captures the whole schedule of rounds in one HIP graph and replays it, 0
    // queues the rounds one by one from the host (the same launches, the same bits)
    int batch_graph = 0;
    // one-pair BP on coarse levels: 1 runs each RANSAC sweep that follows another non-spatial
    // sweep of its level as one launch (the previous sweep's node update recomputed where it is
    // read), 0 as two launches like every other sweep; 2 as 1, with every RANSAC sweep's plane
    // costs computed ahead on a second stream (the same bits in every form)
    int fuse = 2;
    // fuse 2: RANSAC sweeps per plane-cost launch of the second stream, the blocks of such a
    // persistent launch (fewer leave room on every CU for the chain's launches) and the steps
    // per batch of its segment wavefronts (2 or 4)
    int pipe_batch = 4, pipe_blocks = 512, pipe_hb = 2;
    // fuse 2: plane-cost launches issued ahead of the one the main stream waits for (0: every
    // launch of a level at its start)
    int pipe_window = 4;
    // the CRTrees schedule of the next scheduled pyramid level (learned from the levels before)
    CrtSched crt;
    // fuse 2: BP levels with at most this many directed edges (and at most 64 superpixels) run a
    // fused run's sweeps after its first in one workgroup (0: never)
    int run_max_e = 64;
    // the second stream (default priority) and events on it: [0] marks the main stream's
    // level start, [1 + i] the end of the i-th plane-cost launch of a level; made on first use
    hipStream_t side = nullptr;
    std::vector<hipEvent_t> side_ev;
    hipStream_t side_stream() {
        if (!side) HBP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        return side;
    }
    hipEvent_t side_event(size_t i) {
        while (side_ev.size() <= i) {
            hipEvent_t e;
            HBP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            side_ev.push_back(e);
        }
        return side_ev[i];
    }
    void free_side() {
        if (side) (void)hipStreamSynchronize(side);
        for (hipEvent_t e : side_ev) (void)hipEventDestroy(e);
        side_ev.clear();
        if (side) (void)hipStreamDestroy(side);
        side = nullptr;
    }
    // per BP level: -1 the threshold above decides, 0 chunks of sorted pixels, 1 per-superpixel
    // segments (a schedule of forms the tuner can set level by level; the bits are the same)
    static constexpr int kFormLevels = 24;
    int level_form[kFormLevels] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,
                                   -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
    // a stream whose queued work must finish before a cooperative launch on this context (the
    // server's admission context, whose work runs beside the BP rounds of the main stream:
    // a cooperative grid needs every one of its blocks resident at once)
    hipStream_t drain_before_coop = nullptr;
    // no cooperative launch on this context: CoopNeeded is thrown instead
    bool coop_forbidden = false;
    // pinned host words that kernels write counts into, read after one stream wait (no copies)
    int *hbox = nullptr;
    int *host_box() {
        if (!hbox) HBP_CHECK(hipHostMalloc((void **)&hbox, 64 * sizeof(int), hipHostMallocDefault));
        return hbox;
    }
    void free_host_box() {
        if (hbox) (void)hipHostFree(hbox);
        hbox = nullptr;
    }
    // scratch reused by every call on this context
    DevBuf<uint8_t> scratch;
    DevBuf<int> counters;  // small device counters (sizes, flags)
    void *scratch_bytes(size_t bytes) {
        if (scratch.n < bytes) scratch.alloc(pool, bytes + (bytes >> 3));
        return scratch.p;
    }
};

// Times the launches issued while it is alive (when profiling is on).  With kernel_events the
// scope records nothing itself: its one launch takes start_event() / stop_event(), which the
// dispatch stamps with the kernel's own start and end (hipExtLaunchKernelGGL; see HBP_LAUNCH),
// so the time excludes the queue's dispatch latency, as a kernel trace does.
class ProfScope {
public:
    ProfScope(Ctx &c, const std::string &name, double bytes, bool kernel_events = false) : c_(c), kev_(kernel_events) {
        if (!c.prof.on) return;
        if (!c.prof.only.empty() && name.compare(0, c.prof.only.size(), c.prof.only) != 0) return;
        if (c.prof.every > 1 && (c.prof.seen++ % c.prof.every) != 0) return;
        id_ = c.prof.id_of(name);
        a_ = c.prof.take();
        b_ = c.prof.take();
        bytes_ = bytes;
        if (a_ && b_ && !kev_) (void)hipEventRecord(a_, c.stream);
    }
    ~ProfScope() {
        if (!a_ || !b_) return;
        if (!kev_) (void)hipEventRecord(b_, c_.stream);
        c_.prof.pending.push_back({id_, a_, b_, bytes_});
        if (c_.prof.pending.size() > 4096) c_.prof.flush();
    }
    hipEvent_t start_event() const { return kev_ && a_ && b_ ? a_ : nullptr; }
    hipEvent_t stop_event() const { return kev_ && a_ && b_ ? b_ : nullptr; }

private:
    Ctx &c_;
    bool kev_ = false;
    int id_ = -1;
    hipEvent_t a_ = nullptr, b_ = nullptr;
    double bytes_ = 0.0;
};

// one kernel launch under a kernel_events ProfScope: stamped by the dispatch when the scope
// times it, a plain launch otherwise
#define HBP_LAUNCH(ps, kernel, grid, block, lds, stream, ...)                                                 \
    do {                                                                                                     \
        if ((ps).start_event())                                                                              \
            hipExtLaunchKernelGGL(kernel, grid, block, lds, stream, (ps).start_event(), (ps).stop_event(), 0, \
                                  __VA_ARGS__);                                                              \
        else                                                                                                 \
            hipLaunchKernelGGL(kernel, grid, block, lds, stream, __VA_ARGS__);                               \
    } while (0)

inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }
inline long round_up(long a, long b) { return (a + b - 1) / b * b; }

}  // namespace hbp
