// swd_launch.h -- host-side launchers of the kernels in swd_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "swd_kernels.h"

// plain kernels (one sub-step per launch)
hipError_t launch_v(const KGrid &g, const KInst *insts, int ninst, bool neg, int comp, hipStream_t s);
hipError_t launch_p(const KGrid &g, const KInst *insts, int ninst, bool neg, int it, hipStream_t s);
hipError_t launch_src_p(const KGrid &g, const KInst *insts, int ninst, int it, hipStream_t s);
hipError_t launch_src_v(const KGrid &g, const KInst *insts, int ninst, int comp, hipStream_t s);
hipError_t launch_adjoint_step(const KGrid &g, const KInst *insts, int ninst, int it, bool first, hipStream_t s);
// The same step in two tiled launches (the instances' ex, ez, e2, e3 hold the
// second slot of their CPML memories, all zero before the first step).
hipError_t launch_adjoint_step_tiled(const KGrid &g, const KInst *insts, int ninst, int it, bool first, hipStream_t s);
hipError_t launch_copy_rows(float *gather, const int2 *pairs, int npairs, int nt, hipStream_t s);
hipError_t launch_sum_rows(float *dst, const float *src, const int2 *pairs, int npairs, int nt, hipStream_t s);
hipError_t launch_sum_shots(float *dst, const float *src, int nshots, size_t n, bool accumulate, hipStream_t s);
// fused step (swd_fstep.hip): tx = tile width in columns (16, 32 or 64);
// pair: instances are (SRC, RCV) pairs of the RTM backward, ninst counts both
hipError_t launch_fstep(const KGrid &g, const KStep *insts, int ninst, int it, bool first, int tx, hipStream_t s);
hipError_t launch_fstep(const KGrid &g, const KStep *insts, int ninst, int it, bool first, int tx, bool pair,
                        hipStream_t s);
// cig: the paired backward also bins the angle gathers of step it+1 (RCV instances' cig)
hipError_t launch_fstep(const KGrid &g, const KStep *insts, int ninst, int it, bool first, int tx, bool pair, bool cig,
                        hipStream_t s);
// Born forward pairs (born, with pair and without cig): instances 2k and 2k+1
// are the background and scattered fields of shot k.
hipError_t launch_fstep(const KGrid &g, const KStep *insts, int ninst, int it, bool first, int tx, bool pair, bool cig,
                        bool born, hipStream_t s);
// multi-grid forward step: btab[nblk] = (instance, tile), tile width 16
hipError_t launch_fstep_multi(int Lc, const KStep *insts, const int2 *btab, int nblk, int it, hipStream_t s);
hipError_t launch_kzsep(const float2 *capture, int ncap, int nfftz, const float2 *tw, const float *coef,
                        const KCapCol *cols, const int2 *traces, float *up, float *down, int nt, hipStream_t s);
// up/down imaging of step it for nshot shots (KUpDown each) over the interior columns
hipError_t launch_updown(const KGrid &g, const KUpDown *shots, int nshot, int nfftz, int it, hipStream_t s);
// angle-gather update of step it for nshot shots (KCig each) over the interior cells;
// w1, w2: the bin weights exp(-1/18), exp(-4/18) rounded to float; thr: the
// cosine below which no bin is in reach (cig_bins)
hipError_t launch_cig(const KGrid &g, const KCig *shots, int nshot, int thetan, int delt, float w1, float w2, float thr,
                      int it, hipStream_t s);
