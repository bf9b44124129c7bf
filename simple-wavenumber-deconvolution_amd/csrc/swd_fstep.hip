// swd_fstep.hip -- host entry of the fused leapfrog step: picks the kernel
// variant (forward, paired backward, paired backward with angle gathers, Born
// pairs) compiled in swd_fstep_*.hip from swd_fstep.h.
#include "swd_fstep.h"
#include "swd_launch.h"

bool fstep_cfg_supported(int Lc, bool pair, bool cig, bool born, bool multi, int cfg)
{
    const int tx = fstep_cfg_tx(cfg), tq = fstep_cfg_tq(cfg), wb = fstep_cfg_wb(cfg);
    if (cfg != fstep_cfg(tx, tq, wb)) return false;
    if (multi) return tx == 16 && (tq == 16 && wb == 0 || (Lc == 4 && (tq == 8 || tq == 16 || tq == 32) && (wb == 0 || wb == 5)));
    if (tq == 16 && wb == 0) return tx == 16 || tx == 32 || tx == 64;
    if (Lc != 4 || cig || born || (wb != 0 && wb != 5)) return false;
    if (tq == 8 || tq == 16) return tx == 16 || tx == 32 || tx == 64;
    return tq == 32 && (tx == 16 || tx == 32);
}

std::vector<int> fstep_candidates(int Lc, bool pair, bool cig, bool born, bool multi)
{
    std::vector<int> v{kFstepDefault};
    for (int tq : {16, 8, 32})
        for (int wb : {0, 5})
            for (int tx : {16, 32, 64}) {
                const int c = fstep_cfg(tx, tq, wb);
                if (c != kFstepDefault && fstep_cfg_supported(Lc, pair, cig, born, multi, c)) v.push_back(c);
            }
    return v;
}

hipError_t launch_fstep(const KGrid &g, const KStep *insts, int ninst, int it, bool first, int cfg, bool pair, bool cig,
                        bool born, hipStream_t s)
{
    const int n = pair ? ninst / 2 : ninst;
    const int f = first ? 1 : 0;
    if (born && (!pair || cig)) return hipErrorInvalidValue;
    if (pair && cig) return launch_fstep_pair_cig(g, insts, n, it, f, cfg, s);
    if (pair && born) return launch_fstep_born(g, insts, n, it, f, cfg, s);
    if (pair) return launch_fstep_pair(g, insts, n, it, f, cfg, s);
    return launch_fstep_fwd(g, insts, n, it, f, cfg, s);
}
