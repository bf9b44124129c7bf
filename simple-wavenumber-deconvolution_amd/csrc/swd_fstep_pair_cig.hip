// swd_fstep_pair_cig.hip -- the fused step for the paired RTM backward that also bins the angle gathers.
#include "swd_fstep.h"

hipError_t launch_fstep_pair_cig(const KGrid &g, const KStep *insts, int n, int it, int first, int cfg, hipStream_t s)
{
    return launch_fstep_variant<true, true, false>(g, insts, n, it, first, cfg, s);
}
