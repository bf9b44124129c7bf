// swd_fstep2.hip -- host launcher of the two-step fused kernel (swd_fstep2.h).
#include "swd_fstep2.h"
#include "swd_launch.h"

std::vector<int> fstep2_candidates()
{
    return {fstep_cfg2(32, 8, 4), fstep_cfg2(32, 8, 4) | kFstepSplit, fstep_cfg2(64, 8, 8) | kFstepSplit,
            fstep_cfg2(64, 8, 8), fstep_cfg2(16, 16, 4), fstep_cfg2(32, 16, 8) | kFstepSplit};
}

bool fstep2_cfg_supported(int cfg)
{
    for (int c : fstep2_candidates())
        if (c == cfg) return true;
    return false;
}

int fstep2_cfg_of_shape(int tx, int tq)
{
    for (int c : fstep2_candidates())
        if (fstep_cfg_tx(c) == tx && fstep_cfg_tq(c) == tq) return c;
    return 0;
}

// One launch over every tile, or (split) the fast tiles then the frame tiles:
// both read slot sa and write the other, so their order does not matter.
hipError_t launch_fstep2(const KGrid &g, const KStep *insts, int ninst, int it, int ph, int cfg, bool pair, hipStream_t s,
                         bool mid, bool born)
{
    const int tx = fstep_cfg_tx(cfg), tq = fstep_cfg_tq(cfg), nw = fstep_cfg_wb(cfg);
    if (g.Lc != 4 || !fstep2_cfg_supported(cfg) || (ph != 0 && ph != 1) || g.ldz >= (1 << 23))
        return hipErrorInvalidValue;
    if (mid || born) {
        // the middle-level and Born variants: the paired kernel in the default configuration only
        if (!pair || cfg != kFstep2Default || (mid && born)) return hipErrorInvalidValue;
        const int nzb = (g.ntz + 4 * 8 - 1) / (4 * 8), nxb = (g.ntx + 32 - 1) / 32;
        const int4 fr = f2_fast_rect(g, 32, 8, nxb, nzb);
        const int nb = (ninst / 2 * nzb * nxb + 7) / 8 * 8;
        if (mid)
            hipLaunchKernelGGL((k_fstep2<32, 8, 4, true, 0, true>), dim3(nb), dim3(64 * 4), 0, s, g, insts, ninst / 2, nzb,
                               nxb, fr, it, ph);
        else
            hipLaunchKernelGGL((k_fstep2<32, 8, 4, true, 0, false, true>), dim3(nb), dim3(64 * 4), 0, s, g, insts,
                               ninst / 2, nzb, nxb, fr, it, ph);
        return hipGetLastError();
    }
    const bool split = (cfg & kFstepSplit) != 0;
    const int n = pair ? ninst / 2 : ninst;
    const int nzb = (g.ntz + 4 * tq - 1) / (4 * tq), nxb = (g.ntx + tx - 1) / tx;
    const int4 fr = f2_fast_rect(g, tx, tq, nxb, nzb);
    const int nfast = fr.y * fr.w, nframe = nzb * nxb - nfast;
    // single launch: the forward gives each XCD a share of the frame tiles and one of the fast
    // tiles (k_fstep2, SEL 0), the paired backward whole tile columns
    const int b0 = pair ? (n * nzb * nxb + 7) / 8 * 8 : 8 * ((n * nframe + 7) / 8 + (n * nfast + 7) / 8);
    const int b1 = (n * nfast + 7) / 8 * 8, b2 = (n * nframe + 7) / 8 * 8;
#define L2K(TXV, TQV, NWV, PV, SV, NB)                                                                                 \
    hipLaunchKernelGGL((k_fstep2<TXV, TQV, NWV, PV, SV>), dim3(NB), dim3(64 * NWV), 0, s, g, insts, n, nzb, nxb, fr, it, ph)
#define L2P(TXV, TQV, NWV, PV)                                                                                         \
    {                                                                                                                  \
        if (!split) {                                                                                                  \
            L2K(TXV, TQV, NWV, PV, 0, b0);                                                                             \
        } else {                                                                                                       \
            if (b1) L2K(TXV, TQV, NWV, PV, 1, b1);                                                                     \
            if (b2) L2K(TXV, TQV, NWV, PV, 2, b2);                                                                     \
        }                                                                                                              \
    }
#define L2S(TXV, TQV, NWV)                                                                                             \
    {                                                                                                                  \
        if (pair) L2P(TXV, TQV, NWV, true) else L2P(TXV, TQV, NWV, false)                                              \
    }
    switch (nw * 100000 + tq * 1000 + tx) {
    case 408032: L2S(32, 8, 4); break;
    case 416016: L2S(16, 16, 4); break;
    case 808064: L2S(64, 8, 8); break;
    case 816032: L2S(32, 16, 8); break;
    default: return hipErrorInvalidValue;
    }
#undef L2S
#undef L2P
#undef L2K
    return hipGetLastError();
}

int4 fstep2_fast_rect(const KGrid &g, int tx, int tq, int nxb, int nzb) { return f2_fast_rect(g, tx, tq, nxb, nzb); }

hipError_t launch_fstep2_multi(const KStep *insts, const int4 *btab, int nblk, int it, int ph, hipStream_t s)
{
    if (ph != 0 && ph != 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fstep2_multi, dim3(nblk), dim3(256), 0, s, insts, btab, nblk, it, ph);
    return hipGetLastError();
}
