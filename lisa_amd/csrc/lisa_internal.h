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
// `steps` Adam steps of the radiance cache on the same device-resident
// samples (feature rows of feat_stride floats, the first 12 used; targets 3
// per row, non-finite targets ignored), enqueued back to back; *loss is the
// last step's.
lisa_status cache_train_device(lisa_cache* c, const float* d_feat, int feat_stride,
                               const float* d_tgt, int64_t n, float* loss, int steps = 1);
// Network outputs (3 per row) for n device-resident feature rows (12 floats).
lisa_status cache_infer_device(lisa_cache* c, const float* d_feat, int64_t n, float* d_out);
}  // namespace lisa
