// swd_fstep2.hip -- host launcher of the two-step fused kernel (swd_fstep2.h).
#include "swd_fstep2.h"
#include "swd_launch.h"

bool fstep2_shape_supported(int tx, int tq)
{
    return (tx == 32 && tq == 16) || (tx == 64 && tq == 8) || (tx == 16 && tq == 16) || (tx == 32 && tq == 8);
}

std::vector<int> fstep2_candidates()
{
    return {fstep_cfg2(32, 8), fstep_cfg2(16, 16), fstep_cfg2(64, 8), fstep_cfg2(32, 16)};
}

hipError_t launch_fstep2(const KGrid &g, const KStep *insts, int ninst, int it, int ph, int cfg, bool pair, hipStream_t s)
{
    const int tx = fstep_cfg_tx(cfg), tq = fstep_cfg_tq(cfg);
    if (g.Lc != 4 || !fstep_cfg_two(cfg) || !fstep2_shape_supported(tx, tq) || (ph != 0 && ph != 1))
        return hipErrorInvalidValue;
    const int n = pair ? ninst / 2 : ninst;
    const int nzb = (g.ntz + 4 * tq - 1) / (4 * tq), nxb = (g.ntx + tx - 1) / tx;
    const int nblocks = (n * nzb * nxb + 7) / 8 * 8;
#define L2K(TXV, TQV)                                                                                                  \
    {                                                                                                                  \
        if (pair)                                                                                                      \
            hipLaunchKernelGGL((k_fstep2<TXV, TQV, true>), dim3(nblocks), dim3(TXV * TQV), 0, s, g, insts, n, nzb, nxb, \
                               it, ph);                                                                                \
        else                                                                                                           \
            hipLaunchKernelGGL((k_fstep2<TXV, TQV, false>), dim3(nblocks), dim3(TXV * TQV), 0, s, g, insts, n, nzb,    \
                               nxb, it, ph);                                                                           \
    }
    switch (tq * 1000 + tx) {
    case 16032: L2K(32, 16); break;
    case 8064: L2K(64, 8); break;
    case 16016: L2K(16, 16); break;
    case 8032: L2K(32, 8); break;
    default: return hipErrorInvalidValue;
    }
#undef L2K
    return hipGetLastError();
}
