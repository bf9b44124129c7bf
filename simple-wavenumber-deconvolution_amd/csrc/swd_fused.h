// swd_fused.h -- fused, batched time-step kernels (one launch per leapfrog step).
#pragma once
#include <cstddef>
#include <hip/hip_runtime.h>
#include "swd_kernels.h"

// Device bytes of the step table of nshot shots.
size_t fused_table_bytes(int nshot);

// Builds and uploads the step table of n instances grouped per_shot (1 or 2)
// wavefields per shot; instances need both buffer sets (p..psz and p2..psz2).
int fused_build_table(const KInst *h_inst, int n, int per_shot, void *d_table, hipStream_t s);

// Launches steps it0 <= it < it1 (step it reads buffer set it&1, writes the other).
int fused_forward_steps(const KGrid &g, const void *d_table, int nshot, int per_shot, int it0, int it1,
                        hipStream_t s, long long *launches);

// RTM backward time loop (it = nt-2 .. 0) of n shots: source-side instances
// dS (reconstruction) and receiver-side instances dR (back-propagation + imaging).
int fused_backward_loop(const KGrid &g, const KInst *dS, const KInst *dR, const KInst *hS, const KInst *hR, int n,
                        hipStream_t s, long long *launches);
