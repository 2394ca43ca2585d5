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

// Packed-slot layout of a map entry (C, c_q, c_u): spin-2 flag, source
// channel and tod row of each slot, slot count and packed stride CS.
int map_slots(int C, int c_q, int c_u, bool* qu, int perm[4], int row[4], int* nslots, int* cs);

// Compute units of the current device.
int n_cu();

}  // namespace todg_internal
