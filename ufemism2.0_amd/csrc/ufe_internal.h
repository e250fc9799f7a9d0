// Internal declarations shared by the engine's translation units.
#pragma once
#include <string>

namespace ufe {
void set_global_error(const std::string &msg);
const char *global_error();
}  // namespace ufe
