#include "common.hpp"

namespace cam {
namespace {
thread_local std::string g_err;
}

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

void clear_error() { g_err.clear(); }

}  // namespace cam

extern "C" const char *cam_last_error(void) { return cam::g_err.c_str(); }
extern "C" const char *cam_version(void) { return "caracal_amd 0.1 (gfx950)"; }
