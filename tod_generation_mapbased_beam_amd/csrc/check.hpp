// Index checks of the checked build.
//
// The release library relies on construction (and a clamp on every map
// gather) to keep each index in range.  Compiled with -DTODG_CHECKED
// (libtodgpu_checked.so), every map gather, weight-grid deposit and ring
// slot lookup also tests its index and counts a violation, per kind, in a
// device counter of its module; todg_check_counts sums and clears them.  The
// clamps stay in place, so a violation is counted, never turned into an
// out-of-bounds access.  In the release build TODG_CHECK expands to nothing,
// so its code is unchanged.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace todg_check {

enum Kind {
    kMapPixel = 0,  // map pixel index outside [0, npix)
    kGridCell = 1,  // weight-grid deposit outside the grid's cells
    kRingSlot = 2,  // slot outside [0, n) of its ring, or cell outside its ring's window
    kKinds = 3
};

}  // namespace todg_check

#ifdef TODG_CHECKED

#include <vector>

namespace todg_check {

// This module's counters (one copy per translation unit and device).
static __device__ unsigned long long g_counts[kKinds];

__device__ __forceinline__ void hit(int kind) { atomicAdd(&g_counts[kind], 1ull); }

// Host side: every module registers a reader of its counters at load time
// (no HIP call happens then); todg_check_counts runs them all.
using Reader = int (*)(unsigned long long* acc, bool reset);
inline std::vector<Reader>& readers() {
    static std::vector<Reader> v;
    return v;
}

// Adds this module's counters to acc[0..kKinds) and clears them if reset.
static int read_module(unsigned long long* acc, bool reset) {
    unsigned long long c[kKinds];
    hipError_t e = hipMemcpyFromSymbol(c, HIP_SYMBOL(g_counts), sizeof c, 0, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return (int)e;
    for (int k = 0; k < kKinds; ++k) acc[k] += c[k];
    if (reset) {
        const unsigned long long z[kKinds] = {};
        e = hipMemcpyToSymbol(HIP_SYMBOL(g_counts), z, sizeof z, 0, hipMemcpyHostToDevice);
        if (e != hipSuccess) return (int)e;
    }
    return 0;
}

[[maybe_unused]] static const bool g_registered = (readers().push_back(&read_module), true);

}  // namespace todg_check

#define TODG_CHECK(cond, kind)                           \
    do {                                                 \
        if (!(cond)) todg_check::hit(todg_check::kind); \
    } while (0)

#else

#define TODG_CHECK(cond, kind) ((void)0)

#endif
