// Opaque handle types of the C ABI and the error-catching wrapper of every entry point.
#pragma once
#include <memory>
#include <new>
#include <vector>

#include "hbp_internal.h"

struct hbp_ctx {
    hbp::Ctx c;
    int live = 0;   // volumes, pyramids, BP objects and servers made on it and not yet freed
    // worker contexts of hbp_pairs_prepare_dev (a stream and memory pool each), made on first
    // use; the objects they build belong to this context but their memory stays in the
    // workers' pools, so the workers live until the context is destroyed
    std::vector<std::unique_ptr<hbp::Ctx>> helpers;
};
// every handle remembers the context whose pool holds its device memory
struct hbp_volume {
    hbp::Volume v;
    hbp_ctx *owner = nullptr;
};
struct hbp_pyramid {
    hbp::Pyramid p;
    hbp_ctx *owner = nullptr;
};
struct hbp_bp {
    hbp::BP b;
    hbp_ctx *owner = nullptr;
};

using namespace hbp;

// a new handle joins its context's live count (hbp_ctx_destroy refuses while any remain)
template <class Handle>
Handle *adopt_handle(hbp_ctx *ctx, Handle *h) {
    h->owner = ctx;
    ++ctx->live;
    return h;
}
// frees a handle after its context's queued work; ctx may be NULL or must be the owner
template <class Handle>
void release_handle(hbp_ctx *ctx, Handle *h) {
    hbp_ctx *o = h->owner;
    HBP_REQUIRE(!ctx || ctx == o, "the handle belongs to another context");
    HBP_CHECK(hipSetDevice(o->c.device));
    HBP_CHECK(hipStreamSynchronize(o->c.stream));
    --o->live;
    delete h;
}

#define HBP_API_BEGIN try {
#define HBP_API_END                       \
    }                                     \
    catch (const HipError &e) {           \
        set_error(e.msg);                 \
        return -1;                        \
    }                                     \
    catch (const std::bad_alloc &) {      \
        set_error("host allocation failed"); \
        return -2;                        \
    }                                     \
    catch (...) {                         \
        set_error("unknown error");       \
        return -3;                        \
    }                                     \
    return 0;

