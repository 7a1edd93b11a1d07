// Shared host-side helpers of libshx: the per-thread error message behind shx_last_error().
#pragma once
#include <string>

namespace shx {
extern thread_local std::string g_err;
// Record a formatted message and return `code` (a negative SHX_E* status).
int fail(int code, const char* fmt, ...);
}  // namespace shx
