// swd_kernels.h -- kernel-side data structures shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include "swd.h"

// Grid constants of a plan, passed by value to every kernel.
struct KGrid {
    int ntx, ntz, ldz;      // padded grid; ldz = column stride in floats (>= ntz, multiple of 32)
    int Lc, pmlc, nx, nz, nt, nframe;
    float idx, idz;         // float(1/dx), float(1/dz) as RK:47
    float rc[SWD_MAX_LC];
    const float *ax, *axh, *bx, *bxh;   // [ntx]
    const float *az, *azh, *bz, *bzh;   // [ntz]
    const float *Bx, *Bz, *K, *K0;      // [ntx][ldz]: dt*buoyancy (x, z half points), dt*rho*vp^2, rho*vp^2
    int fx0, fx1, fz0, fz1;             // interior box free of CPML damping (all a == 0), inclusive
};

// One propagated wavefield ("instance") of a batch and what the kernels do with it.
struct KInst {
    float *p, *vx, *vz;          // state [ntx][ldz]
    float *psx, *psz;            // CPML memory of dp/dx, dp/dz (on vx, vz)
    float *pvx, *pvz;            // CPML memory of dvx/dx, dvz/dz (on p)
    const float *wavelet;        // additive source [nt] (forward) or subtracted (reconstruction); may be null
    int s_ix, s_iz;
    const int *colmap;           // [ntx]: offset of the column's row map in rowmap, or -1
    const int *rowmap;           // per receiver column [ntz]: trace index or -1
    float *gather;               // record [ntr][nt] (forward) or null
    const float *adj;            // hard replacement [ntr][nt] (receiver back-propagation) or null
    float *strips;               // interior frame [nt][nframe] (save in forward, restore in reconstruction)
    float2 *capture;             // AFor column capture [nt][ncap][nfftz] or null
    const int *colcap;           // [ntx]: capture column index or -1
    int capcomp, ncap, nfftz;
    const float *bvx, *bvz;      // Born: background velocities of the same shot
    const float *born_m;         // Born: dK/K on the interior [nx][nz]
    float *dpdt;                 // reconstruction: written; receiver side: read for imaging
    float *image, *illum_s, *illum_r;
    const float *s0;             // Born adjoint: background sums [slot][nx][nz]
    float *s0_out;               // Born adjoint recompute: store background sums [slot][nx][nz]
    int s0_t0;                   // time step of slot 0 of s0 / s0_out
    float *grad;                 // Born adjoint: gradient accumulator [nx][nz]
    float *ex, *ez;              // Born adjoint scratch [ntx][ldz]; the tiled adjoint's second slot of pvx, pvz
    float *e2, *e3;              // the tiled adjoint's second slot of psx, psz [ntx][ldz]
    const float *dsum;           // Born adjoint: data summed per receiver node [ntr][nt]
};

// One wavefield of the fused step kernel (velocity and pressure of a time
// step in one launch).  State arrays are double-buffered by step parity: a
// launch reads slot (it+1)&1 of p and writes slot it&1, so no workgroup
// overwrites what a neighbour still reads as its halo.  The source
// reconstruction runs its velocity half one step behind, so its velocity
// slots are swapped (see swd_fstep.hip).
enum KStepMode { KSTEP_FWD = 0, KSTEP_SRC = 1, KSTEP_RCV = 2 };
struct KStep {
    int mode;                    // KStepMode
    int s_ix, s_iz;
    const KGrid *grid;           // own grid (multi-grid launches only)
    float *p[2], *vx[2], *vz[2];
    float *psx[2], *psz[2];      // CPML memory of dp/dx, dp/dz (FWD, RCV)
    float *pvx, *pvz;            // CPML memory of dvx/dx, dvz/dz, own cells only (FWD, RCV)
    float *pvxb, *pvzb;          // their second slots (the two-step kernel reads one slot, writes the other)
    const float *wavelet;        // FWD: added at it; SRC: subtracted at it+1
    const int *colmap, *rowmap;  // receiver lookup (FWD records, RCV replaces)
    float *gather;               // FWD
    const float *adj;            // RCV
    float *strips;               // FWD saves, SRC restores
    float2 *capture;             // FWD (AFor)
    const int *colcap;
    int capcomp, ncap, nfftz;
    float *image, *illum_s, *illum_r;   // RCV (paired with the SRC of the same shot)
    float *dpdt;                 // SRC: dp/dt of step it stored on the interior [ntx][ldz], or null
    float *dpdt_b;               // SRC, two steps per launch: dp/dt of the second level (dpdt holds the first)
    float *pmid;                 // RCV, two steps per launch: p of the first level, own cells, or null (and
                                 //   SRC in the middle-level variant: p of the first level, own cells)
    float *vxmid, *vzmid;        // SRC, RCV in the middle-level variant of the two-step kernel (angle
                                 //   gathers): the velocities of the first level, own cells
    float *cig;                  // RCV, angle gathers in the paired launch: [nx*nz][thetan], or null
    int thetan;
    float cw1, cw2;              // bin weights exp(-1/18), exp(-4/18) rounded to float
    float cthr;                  // cosine below which no bin is in reach
    const float *born_m;         // Born scattered field (paired after its background): dK/K [nx][nz], or null
    float *s0_out;               // FWD: the divergence sum S of the pressure update on the interior cells,
    int s0_t0;                   //   [it - s0_t0][nx][nz] (the Born adjoint's background sums), or null
};

// One shot of the angle-gather update that follows a backward step: the
// reconstruction's and the receiver field's slots (both parities) and the
// gather storage.
struct KCig {
    const float *ps[2], *vxs[2], *vzs[2];   // source reconstruction [ntx][ldz]: p at level L in slot L&1, v likewise
    const float *pr[2], *vxr[2], *vzr[2];   // receiver field [ntx][ldz]
    float *ring;                            // [2][delt][nx*nz] time-averaging rings (delt > 1), or null
    float *cig;                             // [nx*nz][thetan], interior cells
};

// One shot of the up/down imaging that follows a backward step: the two
// reconstructions' dp/dt (the wavelet's and the Hilbert-transformed wavelet's)
// and the two receiver pressures (adjoint source and its Hilbert transform)
// are split along kz per interior column and cross-correlated.
struct KUpDown {
    const float *ds, *dh;        // dp/dt of the two reconstructions [ntx][ld], zero off the interior
    const float *pr[2], *ph[2];  // receiver pressures, both slots [ntx][ld]
    float *du, *ud, *uu, *dd;    // images [ntx][ld] (interior cells); uu and dd may be null
    const float2 *tw;            // e^{-2 pi i k / nfftz}, k < nfftz/2
    const float *coef;           // kz taper [nfftz]
    int ld;                      // column stride (floats) of all the arrays above
};

// Column descriptor of the kz separation: one receiver column of one shot.
struct KCapCol {
    int tstart, tcount;          // range in the trace list
};
