// capi_common.hpp -- error reporting shared by the C ABI translation units.
#pragma once

#include <string>

namespace cs3d {
// records msg as this thread's last error and returns code
int set_error(int code, const std::string &msg);
}  // namespace cs3d
