// The calling thread's last error message (rc_last_error), shared by the
// host and device halves of libraycer; the counterpart of the message
// CLManager::checkError throws (src/OpenCL/CLManager.cpp:220-226).
#include <algorithm>
#include <cstring>
#include <string>

#include "raycer_hip.h"
#include "rc_host_internal.hpp"

namespace {
thread_local std::string g_last_error;
}

int rc_host_fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

extern "C" int rc_last_error(char* buf, size_t len) {
    if (!buf || len == 0) return RC_ERR_INVALID;
    size_t n = std::min(len - 1, g_last_error.size());
    std::memcpy(buf, g_last_error.data(), n);
    buf[n] = 0;
    return RC_OK;
}
