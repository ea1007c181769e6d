// PatchMatch kernels for the 7x7 window, fast arithmetic (see mvs_pm_impl.h).
#include "mvs_pm_impl.h"

namespace amvs {

int launch_pm_t7f(int mode, int32_t srcl, int32_t split, bool pinhole, bool persist,
                  const PmArgs &a, int32_t grid, hipStream_t st) {
    return PmDispatch<7, true, 1, 2, 4>::run(mode, srcl, split, pinhole, persist, a, grid, st);
}

}  // namespace amvs
