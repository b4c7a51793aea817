// Library-wide C-ABI entry points: error reporting and version.
#include <string>

#include "cxa_common.hpp"

namespace cxa {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
const std::string &get_error() { return g_last_error; }

}  // namespace cxa

extern "C" const char *cxa_last_error(void) { return cxa::get_error().c_str(); }

extern "C" int cxa_abi_version(void) { return 5; }
