// swd_fstep_pair.hip -- the fused step for the paired RTM backward (SRC + RCV of one shot per workgroup).
#include "swd_fstep.h"

hipError_t launch_fstep_pair(const KGrid &g, const KStep *insts, int n, int it, int first, int cfg, hipStream_t s)
{
    return launch_fstep_variant<true, false, false>(g, insts, n, it, first, cfg, s);
}
