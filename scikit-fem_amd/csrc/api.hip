// Library-wide state of the C ABI: the thread-local error message.
#include "common.h"

namespace femhip {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace femhip

extern "C" const char* femhip_last_error(void) { return femhip::g_last_error.c_str(); }

extern "C" int femhip_version(void) { return 1; }
