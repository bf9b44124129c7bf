// swd_fstep_born.hip -- the fused step for the Born forward pairs (background and scattered field of one shot).
#include "swd_fstep.h"

hipError_t launch_fstep_born(const KGrid &g, const KStep *insts, int n, int it, int first, int cfg, hipStream_t s)
{
    return launch_fstep_variant<true, false, true>(g, insts, n, it, first, cfg, s);
}
