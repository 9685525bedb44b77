// kp_errors.h -- status and message of the last failing call (per thread),
// shared by every translation unit behind the C ABI.  Not part of the ABI.
#pragma once

#include "../../include/kp_image.h"

#include <string>

namespace kp_internal {

void clear_error();
kp_status fail(kp_status status, const std::string& message);

}  // namespace kp_internal
