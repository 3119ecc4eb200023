// Error reporting shared by the host and device halves of libraycer.
#pragma once

#include <string>

// Records msg as the calling thread's last error and returns code.
int rc_host_fail(int code, const std::string& msg);
