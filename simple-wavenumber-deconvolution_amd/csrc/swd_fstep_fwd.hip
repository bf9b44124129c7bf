// swd_fstep_fwd.hip -- the fused step for single forward wavefields (FWD) and
// the multi-grid forward launch.
#include "swd_fstep.h"
#include "swd_launch.h"

hipError_t launch_fstep_fwd(const KGrid &g, const KStep *insts, int n, int it, int first, int cfg, hipStream_t s)
{
    return launch_fstep_variant<false, false, false>(g, insts, n, it, first, cfg, s);
}

hipError_t launch_fstep_multi(int Lc, int cfg, const KStep *insts, const int2 *btab, int nblk, int it, hipStream_t s)
{
    const int nblocks = (nblk + 7) / 8 * 8;
    const int tq = fstep_cfg_tq(cfg), wb = fstep_cfg_wb(cfg);
    if (fstep_cfg_tx(cfg) != 16) return hipErrorInvalidValue;
#define LM(LV, TQV, WBV)                                                                                             \
    hipLaunchKernelGGL((k_fstep_multi<LV, TQV, WBV>), dim3(nblocks), dim3(FGeo<LV, 16, TQV>::NT), 0, s, insts, btab, \
                       nblk, it)
    if (tq == 16 && wb == 0) {
        DISPATCH_LC(Lc, LM(L, 16, 0));
        return hipGetLastError();
    }
    if (Lc != 4) return hipErrorInvalidValue;
    switch (tq * 10 + wb) {
    case 80: LM(4, 8, 0); break;
    case 85: LM(4, 8, 5); break;
    case 165: LM(4, 16, 5); break;
    case 320: LM(4, 32, 0); break;
    case 325: LM(4, 32, 5); break;
    default: return hipErrorInvalidValue;
    }
#undef LM
    return hipGetLastError();
}
