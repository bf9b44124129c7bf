// swd_fstep.hip -- host entry of the fused leapfrog step: picks the kernel
// variant (forward, paired backward, paired backward with angle gathers, Born
// pairs) compiled in swd_fstep_*.hip from swd_fstep.h.
#include "swd_fstep.h"
#include "swd_launch.h"

hipError_t launch_fstep(const KGrid &g, const KStep *insts, int ninst, int it, bool first, int tx, hipStream_t s)
{
    return launch_fstep(g, insts, ninst, it, first, tx, false, s);
}

hipError_t launch_fstep(const KGrid &g, const KStep *insts, int ninst, int it, bool first, int tx, bool pair,
                        hipStream_t s)
{
    return launch_fstep(g, insts, ninst, it, first, tx, pair, false, s);
}

hipError_t launch_fstep(const KGrid &g, const KStep *insts, int ninst, int it, bool first, int tx, bool pair, bool cig,
                        hipStream_t s)
{
    return launch_fstep(g, insts, ninst, it, first, tx, pair, cig, false, s);
}

hipError_t launch_fstep(const KGrid &g, const KStep *insts, int ninst, int it, bool first, int tx, bool pair, bool cig,
                        bool born, hipStream_t s)
{
    const int n = pair ? ninst / 2 : ninst;
    const int f = first ? 1 : 0;
    if (born && (!pair || cig)) return hipErrorInvalidValue;
    if (pair && cig) return launch_fstep_pair_cig(g, insts, n, it, f, tx, s);
    if (pair && born) return launch_fstep_born(g, insts, n, it, f, tx, s);
    if (pair) return launch_fstep_pair(g, insts, n, it, f, tx, s);
    return launch_fstep_fwd(g, insts, n, it, f, tx, s);
}
