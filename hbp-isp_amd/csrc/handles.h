// Opaque handle types of the C ABI and the error-catching wrapper of every entry point.
#pragma once
#include <new>

#include "hbp_internal.h"

struct hbp_ctx {
    hbp::Ctx c;
};
struct hbp_volume {
    hbp::Volume v;
};
struct hbp_pyramid {
    hbp::Pyramid p;
};
struct hbp_bp {
    hbp::BP b;
};

using namespace hbp;

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

