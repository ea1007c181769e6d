// PatchMatch kernels for the 7x7 window (see mvs_pm_impl.h).
#include "mvs_pm_impl.h"

namespace amvs {

int launch_pm_t7(int mode, int32_t srcl, int32_t split, bool pinhole, bool persist,
                 const PmArgs &a, int32_t grid, hipStream_t st) {
    return PmDispatch<7, false, 1, 2, 4>::run(mode, srcl, split, pinhole, persist, a, grid, st);
}

}  // namespace amvs
