// swd_loops.h -- batched time loops (every shot of a batch in each launch).
#pragma once
#include <hip/hip_runtime.h>
#include "swd_kernels.h"

// Forward steps it0 <= it < it1 of n instances: velocity kernel then pressure kernel.
int batched_forward_steps(const KGrid &g, const KInst *d_inst, int n, int it0, int it1, hipStream_t s,
                          long long *launches);

// RTM backward time loop (it = nt-2 .. 0) of n shots: source-side instances
// dS (reconstruction) and receiver-side instances dR (back-propagation + imaging).
int batched_backward_loop(const KGrid &g, const KInst *dS, const KInst *dR, int n, hipStream_t s, long long *launches);
