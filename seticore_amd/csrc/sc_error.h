// Thread-local error message shared by the C-ABI entry points.
#pragma once

#include <string>

namespace sc {
void set_last_error(const std::string& msg);
const char* last_error();
}  // namespace sc
