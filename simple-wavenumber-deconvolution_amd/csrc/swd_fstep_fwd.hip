// swd_fstep_fwd.hip -- the fused step for single forward wavefields (FWD) and
// the multi-grid forward launch.
#include "swd_fstep.h"

hipError_t launch_fstep_fwd(const KGrid &g, const KStep *insts, int n, int it, int first, int tx, hipStream_t s)
{
    return launch_fstep_variant<false, false, false>(g, insts, n, it, first, tx, s);
}

hipError_t launch_fstep_multi(int Lc, const KStep *insts, const int2 *btab, int nblk, int it, hipStream_t s)
{
    const int nblocks = (nblk + 7) / 8 * 8;
    DISPATCH_LC(Lc, hipLaunchKernelGGL((k_fstep_multi<L>), dim3(nblocks), dim3(256), 0, s, insts, btab, nblk, it));
    return hipGetLastError();
}

