// Internal helpers shared by the device (lisa_device.hip) and host
// (lisa_host.cpp) halves of liblisa.so.  Not part of the C ABI.
#pragma once
#include <string>

#include "lisa.h"

namespace lisa {
// Message for lisa_last_error(NULL): last failure of a call made without a context.
void set_global_error(const std::string& msg);
lisa_status global_fail(lisa_status code, const std::string& msg);
const char* global_error();
}  // namespace lisa
