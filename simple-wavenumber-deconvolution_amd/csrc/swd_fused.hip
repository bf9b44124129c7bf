// swd_fused.hip -- batched time loops.
#include "swd_fused.h"
#include "swd_launch.h"
#include "swd.h"
#include "swd_internal.h"

#define HIP_TRY(expr)                                                                                            \
    do {                                                                                                         \
        hipError_t e_ = (expr);                                                                                  \
        if (e_ != hipSuccess) return swd_fail(SWD_E_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

int fused_forward_loop(const KGrid &g, const KInst *d_inst, const KInst *h_inst, int n, hipStream_t s,
                       long long *launches)
{
    (void)h_inst;
    for (int it = 0; it < g.nt; it++) {
        HIP_TRY(launch_v(g, d_inst, n, false, 3, s));
        HIP_TRY(launch_p(g, d_inst, n, false, it, s));
        *launches += 2;
    }
    return SWD_OK;
}

int fused_backward_loop(const KGrid &g, const KInst *dS, const KInst *dR, const KInst *hS, const KInst *hR, int n,
                        hipStream_t s, long long *launches)
{
    (void)hS;
    (void)hR;
    for (int it = g.nt - 2; it >= 0; it--) {
        HIP_TRY(launch_src_p(g, dS, n, it, s));
        HIP_TRY(launch_src_v(g, dS, n, 3, s));
        HIP_TRY(launch_v(g, dR, n, true, 3, s));
        HIP_TRY(launch_p(g, dR, n, true, it, s));
        *launches += 4;
    }
    return SWD_OK;
}
