// Shared plumbing of the C++ drop-in classes: the process-wide context they run on (the
// reference uses device 0 and the legacy default stream, main_stereo_hbp_isp.cu:31) and
// the conversion of C-ABI status codes into exceptions.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>

#include "cvmat.hpp"
#include "hbp_hip.h"

namespace hbp_api {

hbp_ctx *default_ctx();

inline void check(int rc, const char *what) {
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + hbp_last_error());
}

struct VolumeHandle {
    hbp_volume *v = nullptr;
    ~VolumeHandle() { hbp_volume_free(default_ctx(), v); }
};
struct PyramidHandle {
    hbp_pyramid *p = nullptr;
    ~PyramidHandle() { hbp_pyramid_free(default_ctx(), p); }
};

}  // namespace hbp_api
