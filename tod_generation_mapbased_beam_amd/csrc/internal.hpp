// Functions shared by the library's translation units (not part of the C-ABI).
#pragma once
#include <stdint.h>

#include <string>

#include "healpix_ring.hpp"

namespace todg_internal {

// Record msg as the thread's last error (read by todg_last_error); returns code.
int set_error(int code, const std::string& msg);

// Device table of per-ring constants for nside (built once per device).
int ring_table(int64_t nside, const todg::RingEntry** out);

}  // namespace todg_internal
