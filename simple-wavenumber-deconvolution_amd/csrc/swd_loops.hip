// swd_loops.hip -- batched time loops: every shot of a batch in each launch.
#include "swd_loops.h"
#include "swd_launch.h"
#include "swd.h"
#include "swd_internal.h"

#define HIP_TRY(expr)                                                                                            \
    do {                                                                                                         \
        hipError_t e_ = (expr);                                                                                  \
        if (e_ != hipSuccess) return swd_fail(SWD_E_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

int batched_forward_steps(const KGrid &g, const KInst *d_inst, int n, int it0, int it1, hipStream_t s,
                          long long *launches)
{
    for (int it = it0; it < it1; it++) {
        HIP_TRY(launch_v(g, d_inst, n, false, 3, s));
        HIP_TRY(launch_p(g, d_inst, n, false, it, s));
        *launches += 2;
    }
    return SWD_OK;
}

int batched_backward_loop(const KGrid &g, const KInst *dS, const KInst *dR, int n, hipStream_t s, long long *launches)
{
    for (int it = g.nt - 2; it >= 0; it--) {
        HIP_TRY(launch_src_p(g, dS, n, it, s));
        HIP_TRY(launch_src_v(g, dS, n, 3, s));
        HIP_TRY(launch_v(g, dR, n, true, 3, s));
        HIP_TRY(launch_p(g, dR, n, true, it, s));
        *launches += 4;
    }
    return SWD_OK;
}
