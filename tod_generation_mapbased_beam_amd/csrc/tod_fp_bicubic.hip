// Launchers of the bicubic judge (mode 2) kernels: the template instantiations of
// tod_device.hpp for this mode live in this translation unit, so that the
// modes compile in parallel.
#include "tod_device.hpp"

namespace todg_fp {

hipError_t launch_m2(const TodArgs& a, bool qu, int CS, int variant, hipStream_t st) {
    return launch_mode<2>(a, qu, CS, variant, st);
}

hipError_t launch_batch_m2(const BatchArgs& ba, bool qu, int CS, int variant, hipStream_t st) {
    return launch_batch_mode<2>(ba, qu, CS, variant, st);
}

}  // namespace todg_fp
