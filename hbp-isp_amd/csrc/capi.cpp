// extern "C" entry points of libhbp_hip.so (declared in include/hbp_hip.h).
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/hbp_hip.h"
#include "hbp_internal.h"

namespace hbp {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

static std::atomic<long long> g_fault[kFaultKinds];
void fault_arm(int kind, long long after) {
    HBP_REQUIRE(kind >= 0 && kind < kFaultKinds, "unknown fault kind");
    g_fault[kind].store(after > 0 ? after : 0);
}
void fault_point(int kind, const char *what) {
    long long v = g_fault[kind].load(std::memory_order_relaxed);
    while (v > 0 && !g_fault[kind].compare_exchange_weak(v, v - 1)) {
    }
    if (v == 1) throw HipError{std::string("injected failure of a ") + what};
}

void *Pool::get(size_t bytes) {
    fault_point(kFaultAlloc, "device allocation");
    std::lock_guard<std::mutex> lk(mu_);
    size_t n = (bytes + 255) & ~(size_t)255;
    size_t best = (size_t)-1, bi = 0;
    for (size_t i = 0; i < free_.size(); ++i)
        if (free_[i].n >= n && free_[i].n <= 2 * n && free_[i].n < best) {
            best = free_[i].n;
            bi = i;
        }
    Blk b;
    if (best != (size_t)-1) {
        b = free_[bi];
        free_.erase(free_.begin() + (long)bi);
    } else {
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            // give cached blocks back and retry once
            for (auto &f : free_) (void)hipFree(f.p);
            for (auto &f : free_) reserved_ -= f.n;
            free_.clear();
            e = hipMalloc(&p, n);
            if (e != hipSuccess)
                throw HipError{"device allocation of " + std::to_string(n) + " bytes failed: " + hipGetErrorString(e)};
        }
        b = Blk{p, n};
        reserved_ += n;
    }
    used_.push_back(b);
    in_use_ += b.n;
    return b.p;
}

void Pool::put(void *p) {
    std::lock_guard<std::mutex> lk(mu_);
    for (size_t i = 0; i < used_.size(); ++i)
        if (used_[i].p == p) {
            in_use_ -= used_[i].n;
            free_.push_back(used_[i]);
            used_.erase(used_.begin() + (long)i);
            return;
        }
}

void Pool::release_all() {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto &f : free_) (void)hipFree(f.p);
    for (auto &u : used_) (void)hipFree(u.p);
    free_.clear();
    used_.clear();
    in_use_ = reserved_ = 0;
}

int crtrees_coop_blocks(int num_cu);

}  // namespace hbp

#include "handles.h"

static void use_device(hbp_ctx *ctx) { HBP_CHECK(hipSetDevice(ctx->c.device)); }

// a copy of known size: each of its launches reads and writes exactly `n` float4 once, which is
// what the hardware byte counters of a profiler are scaled by
__global__ void calib_copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" {

const char *hbp_last_error(void) { return hbp::g_last_error.c_str(); }

int hbp_inject_fault(const char *what, long long after) {
    HBP_API_BEGIN
    HBP_REQUIRE(what, "NULL argument");
    if (std::strcmp(what, "alloc") == 0) fault_arm(kFaultAlloc, after);
    else if (std::strcmp(what, "round") == 0) fault_arm(kFaultRound, after);
    else throw HipError{"fault kinds are \"alloc\" and \"round\""};
    HBP_API_END
}

// a context's stream, pools and pinned words, after its queued work
static void ctx_teardown(Ctx &c) {
    if (c.stream) HBP_CHECK(hipStreamSynchronize(c.stream));
    c.free_side();
    c.scratch.release();
    c.counters.release();
    c.free_host_box();
    c.pool.release_all();
    if (c.stream) HBP_CHECK(hipStreamDestroy(c.stream));
    c.stream = nullptr;
}

int hbp_ctx_create(int device, hbp_ctx **out) {
    HBP_API_BEGIN
    HBP_REQUIRE(out, "out is NULL");
    *out = nullptr;
    int n = 0;
    HBP_CHECK(hipGetDeviceCount(&n));
    HBP_REQUIRE(device >= 0 && device < n, "no such HIP device");
    HBP_CHECK(hipSetDevice(device));
    auto *ctx = new hbp_ctx();
    ctx->c.device = device;
    hipDeviceProp_t prop;
    HBP_CHECK(hipGetDeviceProperties(&prop, device));
    ctx->c.num_cu = prop.multiProcessorCount;
    HBP_CHECK(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
    ctx->c.coop_blocks = crtrees_coop_blocks(ctx->c.num_cu);
    lr_scratch_setup(device);
    *out = ctx;
    HBP_API_END
}

int hbp_ctx_destroy(hbp_ctx *ctx) {
    HBP_API_BEGIN
    if (!ctx) return 0;
    HBP_REQUIRE(ctx->live == 0, std::to_string(ctx->live) +
                                    " volume, pyramid, BP or server handles of this context are still alive: "
                                    "free them before the context");
    use_device(ctx);
    for (auto &h : ctx->helpers) ctx_teardown(*h);
    ctx->helpers.clear();
    ctx_teardown(ctx->c);
    delete ctx;
    HBP_API_END
}

int hbp_ctx_sync(hbp_ctx *ctx) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx, "ctx is NULL");
    HBP_CHECK(hipStreamSynchronize(ctx->c.stream));
    HBP_API_END
}

void *hbp_ctx_stream(hbp_ctx *ctx) { return ctx ? (void *)ctx->c.stream : nullptr; }
long long hbp_ctx_pool_bytes(hbp_ctx *ctx) { return ctx ? (long long)ctx->c.pool.bytes_reserved() : 0; }

int hbp_profile_enable(hbp_ctx *ctx, int on) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx, "ctx is NULL");
    ctx->c.prof.flush();
    ctx->c.prof.on = on != 0;
    HBP_API_END
}

int hbp_profile_reset(hbp_ctx *ctx) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx, "ctx is NULL");
    ctx->c.prof.reset();
    HBP_API_END
}

int hbp_profile_read(hbp_ctx *ctx, const char *kernel, double *ms, long long *launches, double *bytes) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && kernel, "NULL argument");
    ctx->c.prof.flush();
    KernelStat s;
    for (size_t i = 0; i < ctx->c.prof.names.size(); ++i)
        if (ctx->c.prof.names[i] == kernel) s = ctx->c.prof.stats[i];
    if (ms) *ms = s.ms;
    if (launches) *launches = s.launches;
    if (bytes) *bytes = s.bytes;
    HBP_API_END
}

int hbp_profile_copy(hbp_ctx *ctx, long long bytes, int launches) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && bytes >= 16 && bytes % 16 == 0 && launches >= 1, "bad argument");
    use_device(ctx);
    const long n = (long)(bytes / 16);
    DevBuf<float4> a(ctx->c.pool, (size_t)n), b(ctx->c.pool, (size_t)n);
    HBP_CHECK(hipMemsetAsync(a.p, 0, (size_t)bytes, ctx->c.stream));
    for (int i = 0; i < launches; ++i)
        hipLaunchKernelGGL(calib_copy_kernel, dim3(4096), dim3(256), 0, ctx->c.stream, a.p, b.p, n);
    HBP_CHECK(hipGetLastError());
    HBP_CHECK(hipStreamSynchronize(ctx->c.stream));
    HBP_API_END
}

int hbp_primitive_sort(hbp_ctx *ctx, void *keys, int key_bytes, uint32_t *vals, long long n, int end_bit) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && keys && n >= 0, "bad argument");
    HBP_REQUIRE(key_bytes == 4 || key_bytes == 8, "key_bytes must be 4 or 8");
    HBP_REQUIRE(end_bit >= 0 && end_bit <= 8 * key_bytes, "end_bit out of range");
    HBP_REQUIRE(n < (1LL << 31), "sort too long");
    if (n == 0) return 0;
    const int kb = key_bytes;
    const uint64_t high = end_bit >= 64 ? 0ull : ~((1ull << end_bit) - 1ull);
    for (long long i = 0; i < n; ++i) {
        const uint64_t k = kb == 8 ? ((const uint64_t *)keys)[i] : ((const uint32_t *)keys)[i];
        HBP_REQUIRE((k & high) == 0, "keys hold bits at or above end_bit");
    }
    use_device(ctx);
    Ctx &c = ctx->c;
    DevBuf<uint8_t> ka(c.pool, (size_t)n * kb), kt(c.pool, (size_t)n * kb);
    DevBuf<uint32_t> va(c.pool, (size_t)n), vt(c.pool, (size_t)n);
    HBP_CHECK(hipMemcpyAsync(ka.p, keys, (size_t)n * kb, hipMemcpyHostToDevice, c.stream));
    if (vals) HBP_CHECK(hipMemcpyAsync(va.p, vals, sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice, c.stream));
    bool flipped;
    if (kb == 4)
        flipped = sort_pairs_u32(c, (unsigned int *)ka.p, (unsigned int *)kt.p, (int *)va.p, (int *)vt.p, (long)n,
                                 end_bit);
    else if (vals)
        flipped = sort_pairs_u64(c, (unsigned long long *)ka.p, (unsigned long long *)kt.p, va.p, vt.p, (long)n,
                                 end_bit);
    else
        flipped = sort_keys_u64(c, (unsigned long long *)ka.p, (unsigned long long *)kt.p, (long)n, end_bit);
    HBP_CHECK(hipMemcpyAsync(keys, flipped ? kt.p : ka.p, (size_t)n * kb, hipMemcpyDeviceToHost, c.stream));
    if (vals)
        HBP_CHECK(hipMemcpyAsync(vals, flipped ? vt.p : va.p, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost,
                                 c.stream));
    HBP_CHECK(hipStreamSynchronize(c.stream));
    HBP_API_END
}

int hbp_primitive_scan(hbp_ctx *ctx, const int32_t *in, int32_t *out, long long n) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && in && out && n >= 0, "bad argument");
    if (n == 0) return 0;
    use_device(ctx);
    Ctx &c = ctx->c;
    DevBuf<int> a(c.pool, (size_t)n), b(c.pool, (size_t)n);
    HBP_CHECK(hipMemcpyAsync(a.p, in, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, c.stream));
    scan_exclusive_int(c, a.p, b.p, (long)n, nullptr);
    HBP_CHECK(hipMemcpyAsync(out, b.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, c.stream));
    HBP_CHECK(hipStreamSynchronize(c.stream));
    HBP_API_END
}

int hbp_profile_only(hbp_ctx *ctx, const char *prefix, int every) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx, "ctx is NULL");
    ctx->c.prof.flush();
    ctx->c.prof.only = prefix ? prefix : "";
    ctx->c.prof.every = every > 1 ? every : 1;
    ctx->c.prof.seen = 0;
    HBP_API_END
}

// "level_form.<l>", l in [0, kFormLevels)
static int level_form_index(const std::string &n) {
    const std::string t = n.substr(11);
    HBP_REQUIRE(!t.empty() && t.size() <= 2 && t.find_first_not_of("0123456789") == std::string::npos,
                "level forms are named level_form.<level>");
    const int l = std::stoi(t);
    HBP_REQUIRE(l >= 0 && l < Ctx::kFormLevels, "level form index out of range");
    return l;
}

int hbp_ctx_set_tuning(hbp_ctx *ctx, const char *name, int value) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && name, "NULL argument");
    const std::string n = name;
    if (n == "coarse_min_pix") {
        HBP_REQUIRE(value >= 1 && value <= (1 << 24), "coarse_min_pix must be in [1, 2^24]");
        ctx->c.coarse_min_pix = value;
    } else if (n == "single_hb" || n == "batch_hb") {
        HBP_REQUIRE(value == 2 || value == 4, "steps per batch must be 2 or 4");
        (n == "single_hb" ? ctx->c.single_hb : ctx->c.batch_hb) = value;
    } else if (n == "hamming_cols") {
        HBP_REQUIRE(value == 4 || value == 8, "hamming_cols must be 4 or 8");
        ctx->c.hamming_cols = value;
    } else if (n == "seg_len") {
        HBP_REQUIRE(value >= 64 && value <= 512 && value % 64 == 0, "seg_len must be a multiple of 64 in [64, 512]");
        ctx->c.seg_len = value;
    } else if (n == "batch_graph") {
        HBP_REQUIRE(value == 0 || value == 1, "batch_graph must be 0 or 1");
        ctx->c.batch_graph = value;
    } else if (n == "fuse") {
        HBP_REQUIRE(value >= 0 && value <= 2, "fuse must be 0, 1 or 2");
        ctx->c.fuse = value;
    } else if (n == "pipe_batch") {
        HBP_REQUIRE(value >= 1 && value <= 16, "pipe_batch must be in [1, 16]");
        ctx->c.pipe_batch = value;
    } else if (n == "pipe_blocks") {
        HBP_REQUIRE(value >= 8 && value <= 65536, "pipe_blocks must be in [8, 65536]");
        ctx->c.pipe_blocks = value;
    } else if (n == "run_max_e") {
        HBP_REQUIRE(value >= 0 && value <= 4096, "run_max_e must be in [0, 4096]");
        ctx->c.run_max_e = value;
    } else if (n == "pipe_hb") {
        HBP_REQUIRE(value == 2 || value == 4, "pipe_hb must be 2 or 4");
        ctx->c.pipe_hb = value;
    } else if (n == "pipe_window") {
        HBP_REQUIRE(value >= 0 && value <= 1024, "pipe_window must be in [0, 1024]");
        ctx->c.pipe_window = value;
    } else if (n == "crt_rounds") {
        HBP_REQUIRE(value >= 1 && value <= CrtSched{}.rounds, "crt_rounds must be in [1, 5]");
        ctx->c.crt.rounds = value;
    } else if (n == "crt_jumps") {
        HBP_REQUIRE(value >= 1 && value <= CrtSched{}.jumps, "crt_jumps must be in [1, 8]");
        ctx->c.crt.jumps = value;
    } else if (n == "crt_max_jumps") {
        HBP_REQUIRE(value >= 1 && value <= CrtSched{}.max_jumps, "crt_max_jumps must be in [1, 8]");
        ctx->c.crt.max_jumps = value;
    } else if (n.rfind("level_form.", 0) == 0) {
        const int l = level_form_index(n);
        HBP_REQUIRE(value >= -1 && value <= 1, "a level form is -1 (by coarse_min_pix), 0 (chunks) or 1 (segments)");
        ctx->c.level_form[l] = value;
    } else {
        throw HipError{"unknown tuning parameter " + n + " (coarse_min_pix, single_hb, batch_hb, hamming_cols, seg_len, batch_graph, fuse, pipe_batch, pipe_blocks, pipe_hb, pipe_window, run_max_e, crt_rounds, crt_jumps, crt_max_jumps, level_form.<l>)"};
    }
    HBP_API_END
}

int hbp_ctx_get_tuning(hbp_ctx *ctx, const char *name, int *value) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && name && value, "NULL argument");
    const std::string n = name;
    if (n == "coarse_min_pix") *value = ctx->c.coarse_min_pix;
    else if (n == "single_hb") *value = ctx->c.single_hb;
    else if (n == "batch_hb") *value = ctx->c.batch_hb;
    else if (n == "hamming_cols") *value = ctx->c.hamming_cols;
    else if (n == "batch_graph") *value = ctx->c.batch_graph;
    else if (n == "fuse") *value = ctx->c.fuse;
    else if (n == "pipe_batch") *value = ctx->c.pipe_batch;
    else if (n == "pipe_blocks") *value = ctx->c.pipe_blocks;
    else if (n == "pipe_hb") *value = ctx->c.pipe_hb;
    else if (n == "pipe_window") *value = ctx->c.pipe_window;
    else if (n == "crt_rounds") *value = ctx->c.crt.rounds;
    else if (n == "crt_jumps") *value = ctx->c.crt.jumps;
    else if (n == "crt_max_jumps") *value = ctx->c.crt.max_jumps;
    else if (n == "run_max_e") *value = ctx->c.run_max_e;
    else if (n == "seg_len") *value = ctx->c.seg_len;
    else if (n.rfind("level_form.", 0) == 0) *value = ctx->c.level_form[level_form_index(n)];
    else throw HipError{"unknown tuning parameter " + n + " (coarse_min_pix, single_hb, batch_hb, hamming_cols, seg_len, batch_graph, fuse, pipe_batch, pipe_blocks, pipe_hb, pipe_window, run_max_e, crt_rounds, crt_jumps, crt_max_jumps, level_form.<l>)"};
    HBP_API_END
}

int hbp_profile_names(hbp_ctx *ctx, char *buf, int buflen) {
    if (!ctx) return -1;
    std::string all;
    for (size_t i = 0; i < ctx->c.prof.names.size(); ++i) all += (i ? "," : "") + ctx->c.prof.names[i];
    if (buf && buflen > 0) {
        std::strncpy(buf, all.c_str(), (size_t)buflen - 1);
        buf[buflen - 1] = 0;
    }
    return (int)all.size() + 1;
}

// ---------------------------------------------------------------- (a)
// compute_stereo_cost (stereo_cost.cu:421-445) from images: "Census" (u8 volume) or "AD" (f32)
static bool image_algo_is_ad(const char *algo) {
    HBP_REQUIRE(algo && (std::strcmp(algo, "Census") == 0 || std::strcmp(algo, "AD") == 0),
                "the costs from images are \"Census\" and \"AD\" (\"MC-CNN\" takes a given volume)");
    return std::strcmp(algo, "AD") == 0;
}

int hbp_cost_volume_dev(hbp_ctx *ctx, const uint8_t *d_l, const uint8_t *d_r, int W, int H, int D, const char *algo,
                        hbp_volume **out) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && out && d_l && d_r, "NULL argument");
    *out = nullptr;
    const bool ad = image_algo_is_ad(algo);
    use_device(ctx);
    auto *v = new hbp_volume();
    try {
        if (ad) volume_build_ad(ctx->c, v->v, d_l, d_r, W, H, D);
        else volume_build(ctx->c, v->v, d_l, d_r, W, H, D);
    } catch (...) {
        delete v;
        throw;
    }
    *out = adopt_handle(ctx, v);
    HBP_API_END
}

int hbp_cost_volume(hbp_ctx *ctx, const uint8_t *l, const uint8_t *r, int W, int H, int D, const char *algo,
                    hbp_volume **out) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && out && l && r, "NULL argument");
    *out = nullptr;
    const bool ad = image_algo_is_ad(algo);
    HBP_REQUIRE(W > 0 && H > 0, "bad image size");
    use_device(ctx);
    size_t bytes = (size_t)W * H * 3;
    DevBuf<uint8_t> dl(ctx->c.pool, bytes), dr(ctx->c.pool, bytes);
    HBP_CHECK(hipMemcpyAsync(dl.p, l, bytes, hipMemcpyHostToDevice, ctx->c.stream));
    HBP_CHECK(hipMemcpyAsync(dr.p, r, bytes, hipMemcpyHostToDevice, ctx->c.stream));
    auto *v = new hbp_volume();
    try {
        if (ad) volume_build_ad(ctx->c, v->v, dl.p, dr.p, W, H, D);
        else volume_build(ctx->c, v->v, dl.p, dr.p, W, H, D);
        HBP_CHECK(hipStreamSynchronize(ctx->c.stream));
    } catch (...) {
        delete v;
        throw;
    }
    *out = adopt_handle(ctx, v);
    HBP_API_END
}

static float f32_algo_tau(const char *algo) {
    HBP_REQUIRE(algo && std::strcmp(algo, "MC-CNN") == 0, "a given f32 volume is the \"MC-CNN\" cost");
    return kMcCnnTau;
}

int hbp_cost_volume_f32_dev(hbp_ctx *ctx, const float *d_vol, int W, int H, int D, const char *algo,
                            hbp_volume **out) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && out && d_vol, "NULL argument");
    *out = nullptr;
    const float tau = f32_algo_tau(algo);
    use_device(ctx);
    auto *v = new hbp_volume();
    try {
        volume_build_f32(ctx->c, v->v, d_vol, W, H, D, tau);
        HBP_CHECK(hipStreamSynchronize(ctx->c.stream));
    } catch (...) {
        delete v;
        throw;
    }
    *out = adopt_handle(ctx, v);
    HBP_API_END
}

int hbp_cost_volume_f32(hbp_ctx *ctx, const float *vol, int W, int H, int D, const char *algo, hbp_volume **out) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && out && vol, "NULL argument");
    *out = nullptr;
    const float tau = f32_algo_tau(algo);
    HBP_REQUIRE(W > 0 && H > 0 && D > 0 && (long)W * H < (1L << 24), "bad volume size");
    use_device(ctx);
    const size_t n = (size_t)W * H * D;
    DevBuf<float> raw(ctx->c.pool, n);
    HBP_CHECK(hipMemcpyAsync(raw.p, vol, sizeof(float) * n, hipMemcpyHostToDevice, ctx->c.stream));
    auto *v = new hbp_volume();
    try {
        volume_build_f32(ctx->c, v->v, raw.p, W, H, D, tau);
        HBP_CHECK(hipStreamSynchronize(ctx->c.stream));
    } catch (...) {
        delete v;
        throw;
    }
    *out = adopt_handle(ctx, v);
    HBP_API_END
}

int hbp_volume_download_f32(hbp_ctx *ctx, hbp_volume *vol, float *vh, float *wh) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && vol, "NULL argument");
    use_device(ctx);
    Volume &v = vol->v;
    if (vh) {
        const float *p = volume_f32(ctx->c, v);
        HBP_CHECK(hipMemcpyAsync(vh, p, sizeof(float) * (size_t)v.W * v.H * v.D, hipMemcpyDeviceToHost,
                                 ctx->c.stream));
    }
    if (wh)
        HBP_CHECK(hipMemcpyAsync(wh, v.dwta.p, sizeof(float) * (size_t)v.W * v.H, hipMemcpyDeviceToHost,
                                 ctx->c.stream));
    HBP_CHECK(hipStreamSynchronize(ctx->c.stream));
    HBP_API_END
}

int hbp_cost_to_right(hbp_ctx *ctx, hbp_volume *vol) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && vol, "NULL argument");
    HBP_REQUIRE(!vol->v.right, "the volume is already the right view");
    use_device(ctx);
    volume_to_right(ctx->c, vol->v);
    HBP_API_END
}

int hbp_volume_info(const hbp_volume *vol, int *W, int *H, int *D, float *tau, int *pitch) {
    HBP_API_BEGIN
    HBP_REQUIRE(vol, "NULL volume");
    if (W) *W = vol->v.W;
    if (H) *H = vol->v.H;
    if (D) *D = vol->v.D;
    if (tau) *tau = vol->v.tau;
    if (pitch) *pitch = vol->v.pitch;
    HBP_API_END
}

float *hbp_volume_dev_wta(hbp_volume *vol) { return vol ? vol->v.dwta.p : nullptr; }
const uint8_t *hbp_volume_dev_u8(const hbp_volume *vol) {
    try {
        if (!vol || !vol->owner || vol->v.f32src) return nullptr;
        hbp_ctx *ctx = vol->owner;
        use_device(ctx);
        const uint8_t *p = volume_u8(ctx->c, const_cast<hbp_volume *>(vol)->v);
        HBP_CHECK(hipStreamSynchronize(ctx->c.stream));
        return p;
    } catch (const HipError &e) {
        set_error(e.msg);
        return nullptr;
    }
}
float *hbp_volume_dev_f32(hbp_ctx *ctx, hbp_volume *vol) {
    try {
        if (!ctx || !vol) return nullptr;
        use_device(ctx);
        float *p = volume_f32(ctx->c, vol->v);
        HBP_CHECK(hipStreamSynchronize(ctx->c.stream));
        return p;
    } catch (const HipError &e) {
        set_error(e.msg);
        return nullptr;
    }
}

int hbp_volume_download(hbp_ctx *ctx, const hbp_volume *vol, uint8_t *vh, float *wh) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && vol, "NULL argument");
    use_device(ctx);
    Volume &v = const_cast<hbp_volume *>(vol)->v;
    HBP_REQUIRE(!(vh && v.f32src), "an f32 (MC-CNN) volume has no u8 form: use hbp_volume_download_f32");
    if (vh)
        HBP_CHECK(hipMemcpy2DAsync(vh, (size_t)v.W, volume_u8(ctx->c, v), (size_t)v.pitch, (size_t)v.W,
                                   (size_t)v.H * v.D, hipMemcpyDeviceToHost, ctx->c.stream));
    if (wh)
        HBP_CHECK(hipMemcpyAsync(wh, v.dwta.p, sizeof(float) * (size_t)v.W * v.H, hipMemcpyDeviceToHost,
                                 ctx->c.stream));
    HBP_CHECK(hipStreamSynchronize(ctx->c.stream));
    HBP_API_END
}

int hbp_volume_free(hbp_ctx *ctx, hbp_volume *vol) {
    HBP_API_BEGIN
    if (!vol) return 0;
    release_handle(ctx, vol);
    HBP_API_END
}

// ---------------------------------------------------------------- (b)
static void check_seg_args(int link, int num_nb, double sigma) {
    HBP_REQUIRE(link >= 0 && link <= 3, "link must be 0 (MinLink), 1 (MaxLink), 2 (CentroidLink) or 3 (WardLink)");
    HBP_REQUIRE(num_nb == 4 || num_nb == 8, "num_nb must be 4 or 8");
    HBP_REQUIRE(!(sigma > 1.17549435e-38) || std::round(3 * sigma) <= 8,
                "sigma too large for the Gaussian pre-blur: round(3*sigma) must be at most 8");
}

int hbp_pyramid_build_target_dev(hbp_ctx *ctx, const uint8_t *d_bgr, int W, int H, int link, int num_nb,
                                 double sigma, int target_clus, hbp_pyramid **out) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && out && d_bgr, "NULL argument");
    *out = nullptr;
    check_seg_args(link, num_nb, sigma);
    use_device(ctx);
    auto *p = new hbp_pyramid();
    try {
        pyramid_build(ctx->c, p->p, d_bgr, W, H, num_nb, link, target_clus, sigma);
    } catch (...) {
        delete p;
        throw;
    }
    *out = adopt_handle(ctx, p);
    HBP_API_END
}

int hbp_pyramid_build_target(hbp_ctx *ctx, const uint8_t *bgr, int W, int H, int link, int num_nb, double sigma,
                             int target_clus, hbp_pyramid **out) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && out && bgr, "NULL argument");
    *out = nullptr;
    check_seg_args(link, num_nb, sigma);
    HBP_REQUIRE(W > 0 && H > 0, "bad image size");
    use_device(ctx);
    size_t bytes = (size_t)W * H * 3;
    DevBuf<uint8_t> d(ctx->c.pool, bytes);
    HBP_CHECK(hipMemcpyAsync(d.p, bgr, bytes, hipMemcpyHostToDevice, ctx->c.stream));
    auto *p = new hbp_pyramid();
    try {
        pyramid_build(ctx->c, p->p, d.p, W, H, num_nb, link, target_clus, sigma);
    } catch (...) {
        delete p;
        throw;
    }
    *out = adopt_handle(ctx, p);
    HBP_API_END
}

int hbp_pyramid_build_dev(hbp_ctx *ctx, const uint8_t *d_bgr, int W, int H, int link, int num_nb, double sigma,
                          hbp_pyramid **out) {
    return hbp_pyramid_build_target_dev(ctx, d_bgr, W, H, link, num_nb, sigma, 1, out);
}

int hbp_pyramid_build(hbp_ctx *ctx, const uint8_t *bgr, int W, int H, int link, int num_nb, double sigma,
                      hbp_pyramid **out) {
    return hbp_pyramid_build_target(ctx, bgr, W, H, link, num_nb, sigma, 1, out);
}

int hbp_pyramid_levels(const hbp_pyramid *pyr) { return pyr ? pyr->p.levels() : -1; }

int hbp_pyramid_level_size(const hbp_pyramid *pyr, int level, int *V, int *E) {
    HBP_API_BEGIN
    HBP_REQUIRE(pyr && level >= 0 && level < pyr->p.levels(), "bad level");
    if (V) *V = pyr->p.V[level];
    if (E) *E = pyr->p.E[level];
    HBP_API_END
}

int *hbp_pyramid_dev_clus(hbp_pyramid *pyr) { return pyr ? pyr->p.clus.p : nullptr; }
float *hbp_pyramid_dev_mean(hbp_pyramid *pyr) { return pyr ? pyr->p.mean.p : nullptr; }
int *hbp_pyramid_dev_bd(hbp_pyramid *pyr) { return pyr ? (int *)pyr->p.bd.p : nullptr; }
int *hbp_pyramid_dev_bdl(hbp_pyramid *pyr) { return pyr ? pyr->p.bdl.p : nullptr; }

int hbp_pyramid_download_level(hbp_ctx *ctx, const hbp_pyramid *pyr, int level, int *clus, float *mean, int *bd,
                               int *bdl) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && pyr && level >= 0 && level < pyr->p.levels(), "bad argument");
    use_device(ctx);
    const Pyramid &p = pyr->p;
    size_t V = (size_t)p.V[level], E = (size_t)p.E[level];
    hipStream_t s = ctx->c.stream;
    if (clus) HBP_CHECK(hipMemcpyAsync(clus, p.clus.p + p.voff[level], sizeof(int) * V, hipMemcpyDeviceToHost, s));
    if (mean)
        HBP_CHECK(hipMemcpyAsync(mean, p.mean.p + 6 * p.voff[level], sizeof(float) * 6 * V, hipMemcpyDeviceToHost, s));
    if (bd) HBP_CHECK(hipMemcpyAsync(bd, p.bd.p + p.eoff[level], sizeof(int2) * E, hipMemcpyDeviceToHost, s));
    if (bdl) HBP_CHECK(hipMemcpyAsync(bdl, p.bdl.p + p.eoff[level], sizeof(int) * E, hipMemcpyDeviceToHost, s));
    HBP_CHECK(hipStreamSynchronize(s));
    HBP_API_END
}

int hbp_pyramid_free(hbp_ctx *ctx, hbp_pyramid *pyr) {
    HBP_API_BEGIN
    if (!pyr) return 0;
    release_handle(ctx, pyr);
    HBP_API_END
}

int hbp_crtrees_get_clus(hbp_ctx *ctx, int *d_nn, int *d_clus, int n) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && d_nn && d_clus && n > 0, "bad argument");
    use_device(ctx);
    crtrees_labels(ctx->c, d_nn, n, d_clus);
    HBP_API_END
}

int hbp_crtrees_compact(hbp_ctx *ctx, int *d_clus, int n, int *num) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && d_clus && n > 0, "bad argument");
    use_device(ctx);
    int v = compact_labels(ctx->c, d_clus, n);
    if (num) *num = v;
    HBP_API_END
}

// ---------------------------------------------------------------- (c)
int hbp_bp_create(hbp_ctx *ctx, const hbp_pyramid *pyr, const hbp_volume *vol, float alpha, float tau_s,
                  int min_level, int use_cost_vol, hbp_bp **out) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && pyr && vol && out, "NULL argument");
    *out = nullptr;
    HBP_REQUIRE(std::isfinite(alpha) && std::isfinite(tau_s), "alpha and tau_smooth must be finite");
    use_device(ctx);
    auto *b = new hbp_bp();
    b->b.pyr = &pyr->p;
    b->b.vol = &vol->v;
    b->b.lambda = alpha;
    b->b.tau_s = tau_s;
    b->b.min_level = min_level;
    b->b.use_cost_vol = use_cost_vol ? 1 : 0;
    try {
        bp_setup(ctx->c, b->b);
    } catch (...) {
        delete b;
        throw;
    }
    *out = adopt_handle(ctx, b);
    HBP_API_END
}

// one pair's cost volume, pyramid and BP object on context c (its stream synchronised after)
struct PreparedPair {
    std::unique_ptr<hbp_volume> vol;
    std::unique_ptr<hbp_pyramid> pyr;
    std::unique_ptr<hbp_bp> bp;
};
static void prepare_pair(Ctx &c, const hbp_pair_dev &q, float alpha, float tau_s, int use_cost_vol,
                         PreparedPair &out) {
    PreparedPair r;
    r.vol.reset(new hbp_volume());
    volume_build(c, r.vol->v, q.left, q.right, q.width, q.height, q.labels);
    r.pyr.reset(new hbp_pyramid());
    pyramid_build(c, r.pyr->p, q.left, q.width, q.height, q.num_nb, q.linkage, 1, 0.0);
    const int L = r.pyr->p.levels();
    HBP_REQUIRE(L >= 2, "image too small for a pyramid of two levels");
    const int ml = q.min_level < 0 ? L - 1 + q.min_level : q.min_level;
    r.bp.reset(new hbp_bp());
    BP &b = r.bp->b;
    b.pyr = &r.pyr->p;
    b.vol = &r.vol->v;
    b.lambda = alpha;
    b.tau_s = tau_s;
    b.min_level = std::max(1, std::min(ml, L - 1));
    b.use_cost_vol = use_cost_vol ? 1 : 0;
    bp_setup(c, b);
    HBP_CHECK(hipStreamSynchronize(c.stream));
    out = std::move(r);
}

int hbp_pairs_prepare_dev(hbp_ctx *ctx, int n, const hbp_pair_dev *pairs, float alpha, float tau_smooth,
                          int use_cost_vol, int workers, hbp_volume **vols, hbp_pyramid **pyrs, hbp_bp **bps) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && n >= 0 && (n == 0 || (pairs && vols && pyrs && bps)), "NULL argument");
    HBP_REQUIRE(workers >= 1 && workers <= 16, "workers must be in [1, 16]");
    HBP_REQUIRE(std::isfinite(alpha) && std::isfinite(tau_smooth), "alpha and tau_smooth must be finite");
    for (int i = 0; i < n; ++i) {
        vols[i] = nullptr;
        pyrs[i] = nullptr;
        bps[i] = nullptr;
        const hbp_pair_dev &q = pairs[i];
        HBP_REQUIRE(q.left && q.right, "NULL image in a pair");
        HBP_REQUIRE(q.width > 0 && q.height > 0 && q.labels >= 1, "bad pair size");
        check_seg_args(q.linkage, q.num_nb, 0.0);
    }
    if (n == 0) return 0;
    use_device(ctx);
    Ctx &m = ctx->c;
    // the caller's queued work (e.g. the images' copies) first
    HBP_CHECK(hipStreamSynchronize(m.stream));
    const int nw = std::min(workers, n);
    while ((int)ctx->helpers.size() < nw) {
        auto h = std::make_unique<Ctx>();
        h->device = m.device;
        h->num_cu = m.num_cu;
        h->coop_blocks = m.coop_blocks;
        h->coop_forbidden = true;
        ctx->helpers.push_back(std::move(h));
    }
    // The workers' streams live for this call only (their memory pools stay with the context,
    // which owns what the workers build): a process's streams share the device's few hardware
    // queues, and streams kept idle push later ones, such as a server's admission stream, onto
    // the queue of a busy one.
    struct StreamsOf {
        hbp_ctx *ctx;
        ~StreamsOf() {
            for (auto &h : ctx->helpers)
                if (h->stream) {
                    (void)hipStreamSynchronize(h->stream);
                    (void)hipStreamDestroy(h->stream);
                    h->stream = nullptr;
                }
        }
    } streams_of{ctx};
    for (int w = 0; w < nw; ++w)
        if (!ctx->helpers[w]->stream)
            HBP_CHECK(hipStreamCreateWithFlags(&ctx->helpers[w]->stream, hipStreamNonBlocking));
    for (int w = 0; w < nw; ++w) {   // the context's kernel forms
        Ctx &h = *ctx->helpers[w];
        h.coarse_min_pix = m.coarse_min_pix;
        h.single_hb = m.single_hb;
        h.batch_hb = m.batch_hb;
        h.hamming_cols = m.hamming_cols;
        h.seg_len = m.seg_len;
        h.fuse = m.fuse;
        h.pipe_batch = m.pipe_batch;
        h.pipe_blocks = m.pipe_blocks;
        h.pipe_hb = m.pipe_hb;
        h.pipe_window = m.pipe_window;
        h.crt = m.crt;
        h.run_max_e = m.run_max_e;
        std::memcpy(h.level_form, m.level_form, sizeof(h.level_form));
    }
    std::vector<PreparedPair> done(n);
    std::vector<char> coop(n, 0);
    std::vector<std::string> err(nw);
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    auto work = [&](int w) {
        if (hipSetDevice(m.device) != hipSuccess) {
            err[w] = "hipSetDevice failed in a worker";
            failed = true;
            return;
        }
        Ctx &h = *ctx->helpers[w];
        for (int i = next++; i < n && !failed; i = next++) {
            try {
                prepare_pair(h, pairs[i], alpha, tau_smooth, use_cost_vol, done[i]);
            } catch (const CoopNeeded &) {
                (void)hipStreamSynchronize(h.stream);
                coop[i] = 1;   // redone below, where a cooperative launch is safe
            } catch (const HipError &e) {
                (void)hipStreamSynchronize(h.stream);
                err[w] = e.msg;
                failed = true;
            } catch (const std::bad_alloc &) {
                (void)hipStreamSynchronize(h.stream);
                err[w] = "host allocation failed";
                failed = true;
            } catch (...) {
                (void)hipStreamSynchronize(h.stream);
                err[w] = "unknown error in a worker";
                failed = true;
            }
        }
    };
    std::vector<std::thread> th;
    try {
        for (int w = 1; w < nw; ++w) th.emplace_back(work, w);
    } catch (...) {   // a thread could not start: the started ones stop at their next pair
        failed = true;
        for (auto &t : th) t.join();
        throw HipError{"could not start the worker threads"};
    }
    work(0);
    for (auto &t : th) t.join();
    for (int w = 0; w < nw; ++w) HBP_REQUIRE(err[w].empty(), err[w]);   // done[] frees what was built
    // pairs whose CRTrees needed the cooperative form: on the caller's context, nothing else queued
    for (int i = 0; i < n; ++i)
        if (coop[i]) prepare_pair(m, pairs[i], alpha, tau_smooth, use_cost_vol, done[i]);
    for (int i = 0; i < n; ++i) {
        vols[i] = adopt_handle(ctx, done[i].vol.release());
        pyrs[i] = adopt_handle(ctx, done[i].pyr.release());
        bps[i] = adopt_handle(ctx, done[i].bp.release());
    }
    HBP_API_END
}

int hbp_bp_step(hbp_ctx *ctx, hbp_bp *bp, int max_sweeps, int *done) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && bp, "NULL argument");
    use_device(ctx);
    int d = bp_step(ctx->c, bp->b, max_sweeps);
    if (done) *done = d;
    HBP_API_END
}

int hbp_bp_total_sweeps(const hbp_bp *bp) { return bp ? (int)bp->b.sched.size() : -1; }
int hbp_bp_level(const hbp_bp *bp) { return bp ? bp->b.cur : -1; }

int hbp_bp_level_size(const hbp_bp *bp, int level, int *V, int *E) {
    HBP_API_BEGIN
    HBP_REQUIRE(bp && level >= 0 && level < bp->b.L, "bad level");
    if (V) *V = bp->b.lv[level].V;
    if (E) *E = (int)bp->b.lv[level].E;
    HBP_API_END
}

int hbp_bp_run(hbp_ctx *ctx, hbp_bp *bp, float *disp, float *plane3) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && bp, "NULL argument");
    use_device(ctx);
    BP &b = bp->b;
    bp_step(ctx->c, b, 1 << 30);
    bp_map(ctx->c, b);
    size_t n = (size_t)b.N;
    if (disp) HBP_CHECK(hipMemcpyAsync(disp, b.disp.p, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->c.stream));
    if (plane3)
        HBP_CHECK(hipMemcpyAsync(plane3, b.plane3.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, ctx->c.stream));
    if (disp || plane3) HBP_CHECK(hipStreamSynchronize(ctx->c.stream));
    HBP_API_END
}

int hbp_bp_run_batch(hbp_ctx *ctx, int n, hbp_bp *const *bps) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && n >= 0 && (n == 0 || bps), "NULL argument");
    use_device(ctx);
    std::vector<BP *> v;
    for (int i = 0; i < n; ++i) {
        HBP_REQUIRE(bps[i], "NULL BP object in the batch");
        for (BP *q : v) HBP_REQUIRE(q != &bps[i]->b, "a BP object appears twice in the batch");
        v.push_back(&bps[i]->b);
    }
    bp_run_batch(ctx->c, v);
    HBP_API_END
}

float *hbp_bp_dev_disp(hbp_bp *bp) { return bp ? bp->b.disp.p : nullptr; }
float *hbp_bp_dev_plane(hbp_bp *bp) { return bp ? bp->b.plane3.p : nullptr; }

int hbp_bp_get_state(hbp_ctx *ctx, hbp_bp *bp, float *planes, float *cost, float *belief, float *msg) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && bp, "NULL argument");
    use_device(ctx);
    bp_get_state(ctx->c, bp->b, planes, cost, belief, msg);
    HBP_API_END
}

int hbp_bp_free(hbp_ctx *ctx, hbp_bp *bp) {
    HBP_API_BEGIN
    if (!bp) return 0;
    release_handle(ctx, bp);
    HBP_API_END
}

int hbp_schedule_sweeps(int levels, int min_level, int labels) { return bp_total_sweeps(levels, min_level, labels); }

// ---------------------------------------------------------------- (d)
int hbp_lr_check_fill_dev(hbp_ctx *ctx, const float *dl, const float *dr, float *pl, int W, int H, uint8_t *fl,
                          uint8_t *fr, uint8_t *f2) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && dl && dr && pl && W > 0 && H > 0, "bad argument");
    use_device(ctx);
    size_t n = (size_t)W * H;
    DevBuf<uint8_t> tmp(ctx->c.pool, 3 * n);
    lr_check_fill(ctx->c, dl, dr, pl, W, H, fl ? fl : tmp.p, fr ? fr : tmp.p + n, f2 ? f2 : tmp.p + 2 * n);
    HBP_API_END
}

int hbp_lr_check_fill_dev_noalloc(hbp_ctx *ctx, const float *dl, const float *dr, float *pl, int W, int H,
                                  uint8_t *fl, uint8_t *fr, uint8_t *f2) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && dl && dr && pl && W > 0 && H > 0, "bad argument");
    use_device(ctx);
    lr_check_fill_noalloc(ctx->c, dl, dr, pl, W, H, fl, fr, f2);
    HBP_API_END
}

int hbp_lr_check_fill(hbp_ctx *ctx, const float *dl, const float *dr, float *pl, int W, int H, uint8_t *fl,
                      uint8_t *fr, uint8_t *f2) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && dl && dr && pl && W > 0 && H > 0, "bad argument");
    use_device(ctx);
    size_t n = (size_t)W * H;
    hipStream_t s = ctx->c.stream;
    DevBuf<float> d(ctx->c.pool, 5 * n);
    DevBuf<uint8_t> f(ctx->c.pool, 3 * n);
    HBP_CHECK(hipMemcpyAsync(d.p, dl, sizeof(float) * n, hipMemcpyHostToDevice, s));
    HBP_CHECK(hipMemcpyAsync(d.p + n, dr, sizeof(float) * n, hipMemcpyHostToDevice, s));
    HBP_CHECK(hipMemcpyAsync(d.p + 2 * n, pl, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
    lr_check_fill(ctx->c, d.p, d.p + n, d.p + 2 * n, W, H, f.p, f.p + n, f.p + 2 * n);
    HBP_CHECK(hipMemcpyAsync(pl, d.p + 2 * n, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
    if (fl) HBP_CHECK(hipMemcpyAsync(fl, f.p, n, hipMemcpyDeviceToHost, s));
    if (fr) HBP_CHECK(hipMemcpyAsync(fr, f.p + n, n, hipMemcpyDeviceToHost, s));
    if (f2) HBP_CHECK(hipMemcpyAsync(f2, f.p + 2 * n, n, hipMemcpyDeviceToHost, s));
    HBP_CHECK(hipStreamSynchronize(s));
    HBP_API_END
}

int hbp_weighted_median_dev(hbp_ctx *ctx, const uint8_t *d_bgr, const uint8_t *d_fail, float *d_plane3, int W,
                            int H) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && d_bgr && d_fail && d_plane3 && W > 0 && H > 0, "bad argument");
    HBP_REQUIRE((long)W * H < (1L << 31), "image too large");
    use_device(ctx);
    weighted_median(ctx->c, d_bgr, d_fail, d_plane3, W, H);
    HBP_API_END
}

int hbp_weighted_median(hbp_ctx *ctx, const uint8_t *bgr, const uint8_t *fail, float *plane3, int W, int H) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && bgr && fail && plane3 && W > 0 && H > 0, "bad argument");
    HBP_REQUIRE((long)W * H < (1L << 31), "image too large");
    use_device(ctx);
    size_t n = (size_t)W * H;
    hipStream_t s = ctx->c.stream;
    DevBuf<float> d(ctx->c.pool, 3 * n);
    DevBuf<uint8_t> b(ctx->c.pool, 4 * n);
    HBP_CHECK(hipMemcpyAsync(b.p, bgr, 3 * n, hipMemcpyHostToDevice, s));
    HBP_CHECK(hipMemcpyAsync(b.p + 3 * n, fail, n, hipMemcpyHostToDevice, s));
    HBP_CHECK(hipMemcpyAsync(d.p, plane3, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
    weighted_median(ctx->c, b.p, b.p + 3 * n, d.p, W, H);
    HBP_CHECK(hipMemcpyAsync(plane3, d.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
    HBP_CHECK(hipStreamSynchronize(s));
    HBP_API_END
}

int hbp_eval_bad_pixels(hbp_ctx *ctx, const float *disp, const float *gt, const uint8_t *nonocc, int W, int H,
                        double threshold, long long *counts) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && disp && gt && nonocc && counts && W > 0 && H > 0, "bad argument");
    use_device(ctx);
    size_t n = (size_t)W * H;
    hipStream_t s = ctx->c.stream;
    DevBuf<float> d(ctx->c.pool, 2 * n);
    DevBuf<uint8_t> m(ctx->c.pool, n);
    HBP_CHECK(hipMemcpyAsync(d.p, disp, sizeof(float) * n, hipMemcpyHostToDevice, s));
    HBP_CHECK(hipMemcpyAsync(d.p + n, gt, sizeof(float) * n, hipMemcpyHostToDevice, s));
    HBP_CHECK(hipMemcpyAsync(m.p, nonocc, n, hipMemcpyHostToDevice, s));
    eval_bad_pixels(ctx->c, d.p, d.p + n, m.p, W, H, threshold, counts);
    HBP_API_END
}

int hbp_median7_dev(hbp_ctx *ctx, float *d_disp, int W, int H) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && d_disp && W > 0 && H > 0, "bad argument");
    use_device(ctx);
    median7_filter(ctx->c, d_disp, W, H);
    HBP_API_END
}

int hbp_median7(hbp_ctx *ctx, float *disp, int W, int H) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && disp && W > 0 && H > 0, "bad argument");
    use_device(ctx);
    size_t n = (size_t)W * H;
    hipStream_t s = ctx->c.stream;
    DevBuf<float> d(ctx->c.pool, n);
    HBP_CHECK(hipMemcpyAsync(d.p, disp, sizeof(float) * n, hipMemcpyHostToDevice, s));
    median7_filter(ctx->c, d.p, W, H);
    HBP_CHECK(hipMemcpyAsync(disp, d.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    HBP_CHECK(hipStreamSynchronize(s));
    HBP_API_END
}

int hbp_plane_to_disp(hbp_ctx *ctx, const float *plane3, int W, int H, float *disp) {
    HBP_API_BEGIN
    HBP_REQUIRE(ctx && plane3 && disp && W > 0 && H > 0, "bad argument");
    use_device(ctx);
    size_t n = (size_t)W * H;
    hipStream_t s = ctx->c.stream;
    DevBuf<float> d(ctx->c.pool, 4 * n);
    HBP_CHECK(hipMemcpyAsync(d.p, plane3, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
    plane_to_disp(ctx->c, d.p, W, H, d.p + 3 * n);
    HBP_CHECK(hipMemcpyAsync(disp, d.p + 3 * n, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    HBP_CHECK(hipStreamSynchronize(s));
    HBP_API_END
}

}  // extern "C"
