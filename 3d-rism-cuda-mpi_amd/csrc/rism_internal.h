// rism_internal.h -- shared host-side helpers of librism3d.
#pragma once
#include <string>

#define RISM_MAX_SITES 16

namespace rism {
// Records the message for rism_last_error() and returns code.
int set_error(int code, const std::string &msg);
}  // namespace rism
