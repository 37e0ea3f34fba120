// fp_error.cpp — the library's error reporting: a thread-local message set by
// fprt::fail and read through fp_hip_last_error.  Plain C++ (no HIP), so the
// host-side table builder links it in its sanitizer unit test too.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "fast_pauli_hip.h"
#include "fp_tiles_host.hpp"

namespace fprt
{

thread_local std::string g_last_error;

fp_status fail(fp_status code, char const *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

} // namespace fprt

extern "C" size_t fp_hip_last_error(char *buf, size_t len)
{
    if (buf && len)
    {
        size_t const n = std::min(len - 1, fprt::g_last_error.size());
        std::memcpy(buf, fprt::g_last_error.data(), n);
        buf[n] = 0;
    }
    return fprt::g_last_error.size();
}
