// capi_common.cpp -- per-thread last-error string of the C ABI.
#include "capi_common.hpp"

#include "../../include/cs3d.h"

namespace cs3d {
static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}
}  // namespace cs3d

extern "C" {
const char *cs3d_last_error(void) { return cs3d::g_last_error.c_str(); }
const char *cs3d_version(void) { return "cs3d-mi355x 0.1 (gfx950)"; }
}
