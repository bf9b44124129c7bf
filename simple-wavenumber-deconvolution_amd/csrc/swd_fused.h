// swd_fused.h -- batched time loops (all instances of a batch in each launch).
#pragma once
#include <hip/hip_runtime.h>
#include "swd_kernels.h"

// Forward time loop of n instances (d_inst on device, h_inst its host copy).
int fused_forward_loop(const KGrid &g, const KInst *d_inst, const KInst *h_inst, int n, hipStream_t s,
                       long long *launches);

// RTM backward time loop (it = nt-2 .. 0) of n shots: source-side instances
// dS (reconstruction) and receiver-side instances dR (back-propagation + imaging).
int fused_backward_loop(const KGrid &g, const KInst *dS, const KInst *dR, const KInst *hS, const KInst *hR, int n,
                        hipStream_t s, long long *launches);
