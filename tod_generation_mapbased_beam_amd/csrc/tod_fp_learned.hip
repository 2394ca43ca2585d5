// Launchers of the learned corrector (mode 3) kernels: the template instantiations of
// tod_device.hpp for this mode live in this translation unit, so that the
// modes compile in parallel.
#include "tod_device.hpp"

namespace todg_fp {

hipError_t launch_m3(const TodArgs& a, bool qu, int CS, hipStream_t st) {
    return launch_mode<3>(a, qu, CS, 0, st);
}

}  // namespace todg_fp
