// Thread-local error text behind prt_last_error().
#include <string>

#include "common.hpp"

namespace prt {
namespace {
thread_local std::string g_last_error;
}

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

void clear_error() { g_last_error.clear(); }

}  // namespace prt

extern "C" const char* prt_last_error(void) { return prt::g_last_error.c_str(); }
extern "C" int prt_abi_version(void) { return PRT_ABI_VERSION; }
