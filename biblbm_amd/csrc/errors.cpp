// Thread-local last-error text behind bl_last_error().
#include <string>

#include "../../include/biblbm.h"

namespace bl {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

}  // namespace bl

extern "C" const char* bl_last_error(void) { return bl::g_last_error.c_str(); }
