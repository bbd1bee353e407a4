// pt_common.h -- host helpers shared by the library's translation units.
#pragma once
#include <string>

namespace ptk_host {
// sets the calling thread's last error (pt_last_error) and returns `code`
int fail(int code, const std::string &msg);
}  // namespace ptk_host
