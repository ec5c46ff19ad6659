// Internal helpers shared by the plan and driver translation units.
#pragma once
#include <string>

namespace evr {
// Record the message returned by evr_last_error() and return `code`.
int set_error(int code, const std::string &msg);
}  // namespace evr
