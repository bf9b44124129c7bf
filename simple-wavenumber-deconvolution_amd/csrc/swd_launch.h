// swd_launch.h -- host-side launchers of the kernels in swd_kernels.hip.
#pragma once
#include <vector>

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
// Tile configuration of the fused step: width tx (columns: 16, 32, 64),
// height tq (float4 quads per column: 4 tq rows; 8, 16, 32) and a bound of
// wb wavefronts per SIMD (0: none; 5).
constexpr int fstep_cfg(int tx, int tq, int wb) { return tx | (tq << 8) | (wb << 16); }
constexpr int fstep_cfg_tx(int cfg) { return cfg & 0xff; }
constexpr int fstep_cfg_tq(int cfg) { return (cfg >> 8) & 0xff; }
constexpr int fstep_cfg_wb(int cfg) { return (cfg >> 16) & 0xff; }
constexpr int kFstepDefault = fstep_cfg(16, 16, 0);
// Two time steps per launch (swd_fstep2.h): bit 24 of a configuration, whose
// third field is then the wavefronts per workgroup (64 nw threads for tx tq
// own quads).
constexpr int kFstepTwo = 1 << 24;
// Bit 25 of a two-step configuration: the fast tiles (stencils free of damping
// and frame cells) and the others in two launches of kernels compiled apart.
constexpr int kFstepSplit = 1 << 25;
constexpr int fstep_cfg2(int tx, int tq, int nw) { return fstep_cfg(tx, tq, nw) | kFstepTwo; }
constexpr bool fstep_cfg_two(int cfg) { return (cfg & kFstepTwo) != 0; }
constexpr int kFstep2Default = fstep_cfg2(32, 8, 4);
// Threads per workgroup of a one-step configuration (one own quad per thread, 128 to 512).
constexpr int fstep_threads_of(int cfg)
{
    return fstep_cfg_tx(cfg) * fstep_cfg_tq(cfg) < 128   ? 128
           : fstep_cfg_tx(cfg) * fstep_cfg_tq(cfg) > 512 ? 512
                                                          : fstep_cfg_tx(cfg) * fstep_cfg_tq(cfg);
}
// The two-step kernel's configurations (tx columns x 4 tq rows, nw
// wavefronts), the tuner's candidates (the default first) and the first
// candidate of a tile shape (0: none).
bool fstep2_cfg_supported(int cfg);
std::vector<int> fstep2_candidates();
int fstep2_cfg_of_shape(int tx, int tq);
// Two steps (it, it+1 forward; it, it-1 for the paired backward) of every
// instance in one launch, Lc 4, with the column capture, dp/dt and first-level
// outputs but no angle binning; ph: the launch's place in its pair (launches
// come in pairs, see swd_fstep2.h).  Not for the first reconstruction step.
// mid: the paired kernel's middle-level variant (kFstep2Default only), which also
// stores both fields' first level (pmid, vxmid, vzmid) for the angle gathers.
// born: the Born pair (instances 2k and 2k+1: the background and the scattered
// field of shot k, FWD; kFstep2Default only), the background's sums kept in
// registers for the scattered field's Born source.
hipError_t launch_fstep2(const KGrid &g, const KStep *insts, int ninst, int it, int ph, int cfg, bool pair, hipStream_t s,
                         bool mid = false, bool born = false);
// The configurations a fused launch supports: 64-row tiles 16, 32 or 64 wide
// for every variant and Lc; at Lc 4, for the forward and the paired RTM
// backward (not the angle-gather or Born variants), also 32- and 128-row tiles
// (128 rows up to 32 wide) and the five-wavefront bound.  multi: the
// multi-grid launch, whose tiles are 16 wide.
bool fstep_cfg_supported(int Lc, bool pair, bool cig, bool born, bool multi, int cfg);
// The configurations the tuner times, the default (16 x 64 rows, no bound) first.
std::vector<int> fstep_candidates(int Lc, bool pair, bool cig, bool born, bool multi);

// fused step (swd_fstep.hip) with tile configuration cfg; pair: instances are
// (SRC, RCV) pairs of the RTM backward, ninst counts both; cig: the paired
// backward also bins the angle gathers of step it+1 (RCV instances' cig);
// born (with pair, without cig): instances 2k and 2k+1 are the background and
// scattered fields of the Born forward of shot k.
hipError_t launch_fstep(const KGrid &g, const KStep *insts, int ninst, int it, bool first, int cfg, bool pair, bool cig,
                        bool born, hipStream_t s);
// multi-grid forward step: btab[nblk] = (instance, tile numbered for cfg's tiles)
hipError_t launch_fstep_multi(int Lc, int cfg, const KStep *insts, const int2 *btab, int nblk, int it, hipStream_t s);
// Two steps of the multi-grid forward (Lc 4, every grid's column stride below 2^23):
// btab[nblk] = (instance, tile column, tile row, fast) of 32 x 32-row tiles, the
// fast flag from fstep2_fast_rect (the rectangle of tile columns x .. x+y-1 and
// tile rows z .. z+w-1 whose staged regions avoid damping and frame cells).
int4 fstep2_fast_rect(const KGrid &g, int tx, int tq, int nxb, int nzb);
hipError_t launch_fstep2_multi(const KStep *insts, const int4 *btab, int nblk, int it, int ph, hipStream_t s);
hipError_t launch_kzsep(const float2 *capture, int ncap, int nfftz, const float2 *tw, const float *coef,
                        const KCapCol *cols, const int2 *traces, float *up, float *down, int nt, hipStream_t s);
// up/down imaging of step it for nshot shots (KUpDown each) over the interior columns
hipError_t launch_updown(const KGrid &g, const KUpDown *shots, int nshot, int nfftz, int it, hipStream_t s);
// the same for steps ita then itb in one launch (descriptors a and b of one
// shot share its images; the interior rows at most nfftz)
hipError_t launch_updown2(const KGrid &g, const KUpDown *a, const KUpDown *b, int nshot, int nfftz, int ita, int itb,
                          hipStream_t s);
// the kernel launches launch_updown2 makes (two from 2048 points)
constexpr int updown2_launches(int nfftz) { return nfftz >= 2048 ? 2 : 1; }
// angle-gather update of step it for nshot shots (KCig each) over the interior cells;
// w1, w2: the bin weights exp(-1/18), exp(-4/18) rounded to float; thr: the
// cosine below which no bin is in reach (cig_bins)
hipError_t launch_cig(const KGrid &g, const KCig *shots, int nshot, int thetan, int delt, float w1, float w2, float thr,
                      int it, hipStream_t s);
// delt 1: steps ita (descriptors a) then itb (descriptors b) in one launch, the
// order of two launch_cig calls
hipError_t launch_cig2(const KGrid &g, const KCig *a, const KCig *b, int nshot, int thetan, float w1, float w2, float thr,
                       int ita, int itb, hipStream_t s);
