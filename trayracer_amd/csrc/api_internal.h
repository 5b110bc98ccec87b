// Internal helpers shared by the C-ABI translation units (not part of include/trt.h).
#pragma once
#include <string>

#include "../../include/trt.h"

// Records `msg` as the context's last error and returns `s`.
trt_status trt_internal_fail(trt_context* ctx, trt_status s, const std::string& msg);
