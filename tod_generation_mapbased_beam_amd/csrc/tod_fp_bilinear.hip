// Launchers of the bilinear (mode 0) kernels: the template instantiations of
// tod_device.hpp for this mode live in this translation unit, so that the
// modes compile in parallel.
#include "tod_device.hpp"

namespace todg_fp {

hipError_t launch_m0(const TodArgs& a, bool qu, int CS, int variant, hipStream_t st) {
    return launch_mode<0>(a, qu, CS, variant, st);
}

hipError_t launch_batch_m0(const BatchArgs& ba, bool qu, int CS, int variant, hipStream_t st) {
    return launch_batch_mode<0>(ba, qu, CS, variant, st);
}

hipError_t launch_dump_m0(const TodArgs& a, bool qu, hipStream_t st) {
    return qu ? launch_fp_cs<0, true, 4, 4, kClaim, kP2Cells, true>(a, st)
              : launch_fp_cs<0, false, 1, 4, kClaim, kP2Cells, true>(a, st);
}

}  // namespace todg_fp
