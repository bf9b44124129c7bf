/*
 * swd.h -- C-ABI of the MI355X (gfx950) 2D acoustic staggered-grid CPML FDTD /
 * RTM library (libswd.so).
 *
 * This is the drop-in boundary for the reference's extern "C" seam between its
 * host drivers and its device code.  Each entry point names the reference
 * function it replaces (RC = acoustic_HLVSPRTMC.cpp, RK = acoustic_HLVSPRTMLc.cu,
 * FC = forward/acoustic_hilbert_fftkz_HLForC.cpp,
 * FK = forward/acoustic_hilbert_fftkz_HLForLc.cu, H = headrtm.h,
 * DC = kkdecon_acoustic/kk_decon_interp_psf_minewindow.cpp).
 *
 * Conventions
 *  - Every call returns an int status: 0 = ok, < 0 = error (see SWD_E_*); the
 *    message of the last error of the calling thread is in swd_last_error().
 *  - Plain C types only.  Host buffers are caller-owned and touched only during
 *    the call.  *_dev entry points take device pointers (HIP device memory on the
 *    plan's device) and run on the plan's stream; they return when the results
 *    are complete.
 *  - Grid layout (RC:78-83, RK:45): pmlc = pml+Lc, ntx = nx+2*pmlc,
 *    ntz = nz+2*pmlc; padded arrays are [ntx][ntz] floats (z contiguous, cell
 *    (ix,iz) at ix*ntz+iz); images are padded [ntx][ntz] too.  Interior .bin
 *    files are [nx][nz].  Gathers are trace-major [ntrace][nt].
 *  - Source/receiver coordinates are padded-grid indices, as Source.s_ix/s_iz
 *    and r_iz in the reference (H:9-13).
 *  - A plan is bound to one device; its work runs on its own stream (and, for
 *    alternate RTM shot groups, on its other streams between a fork and a
 *    join).  It is not thread-safe, but independent plans may be used from
 *    independent threads.
 */
#ifndef SWD_H
#define SWD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWD_OK 0
#define SWD_E_ARG (-1)      /* invalid argument / shape */
#define SWD_E_DEVICE (-2)   /* no device or HIP runtime error */
#define SWD_E_NOMEM (-3)    /* allocation failed */
#define SWD_E_IO (-4)       /* file missing, short or unreadable */
#define SWD_E_STATE (-5)    /* call order (e.g. model not set) */
#define SWD_E_RANGE (-6)    /* configuration outside the limits (layerz, CFL, 32-bit field offsets, ...) */

#define SWD_MAX_LC 8

typedef struct swd_grid {
    int nx, nz;        /* interior cells */
    int pml, Lc;       /* CPML cells as in parameter.txt, half stencil order (1..8) */
    float dx, dz, dt;
    int nt;            /* itmax */
} swd_grid;

/* FD and CPML coefficient vectors (host arrays), as produced by swd_cal_xishu
 * and swd_cpml: rc[Lc], *_x[ntx], *_z[ntz]. */
typedef struct swd_coefs {
    const float *rc;
    const float *a_x, *a_x_half, *b_x, *b_x_half;
    const float *a_z, *a_z_half, *b_z, *b_z_half;
} swd_coefs;

/* One shot: source node and receiver list (padded-grid indices), replacing
 * struct Source (H:9-13).  The reference's two-borehole geometry is the list
 * {(pmlc, r_iz[i])} followed by {(ntx-pmlc-1, r_iz[i])}. */
typedef struct swd_shot {
    int s_ix, s_iz;
    int nrec;
    const int *rec_ix;
    const int *rec_iz;
} swd_shot;

/* parameter.txt contents, in file order (RC:1931-2035). */
typedef struct swd_params {
    int nx, nz, pml, Lc;
    float dx, dz;
    int itmax;
    float dt, f0;
    int ns;
    float sx0, shotdx, shotdep, recdix, offsetz;
    int scatter_flag, swin, window;
} swd_params;

/* Execution path of a plan. */
#define SWD_PATH_PLAIN 0   /* one kernel per sub-step (vx, vz, p), one shot per launch */
#define SWD_PATH_BATCHED 1 /* velocity kernel + pressure kernel per step, whole batch per launch */
#define SWD_PATH_FUSED 2   /* one fused kernel per step for the whole batch (default); Born uses BATCHED */

typedef struct swd_stats {
    double fwd_ms;          /* event time of the last forward time loop (all shots of the call) */
    double bwd_ms;          /* event time of the last backward time loop */
    double fft_ms;          /* event time of the kz separation kernels */
    long long fwd_launches; /* kernel launches inside the last forward loop */
    long long bwd_launches;
    long long fwd_cells;    /* cell-updates of the last forward loop (per wavefield, padded grid) */
    long long bwd_cells;
    long long ws_bytes;     /* device workspace held by the plan */
    long long ckpt_bytes;   /* last Born adjoint: snapshot + one segment of background sums */
    long long full_bytes;   /* last Born adjoint: background sums of all nt steps (full storage) */
    long long waves;        /* shot waves of the last call (see swd_plan_set_memory_limit) */
    int tile_x, tile_rows, tile_waves; /* tile of the last fused loop: columns, rows, wavefronts-per-SIMD bound (0: none) */
    int steps_per_launch;   /* time steps per launch of the last fused loop (1 or 2) */
    int tile_group_waves;   /* wavefronts per workgroup of the last fused loop (negative: two-step split launches) */
    int fwd_tile_x, fwd_tile_rows, fwd_tile_group_waves, fwd_steps_per_launch; /* the same of the last fused forward loop */
} swd_stats;

/* ---------------- errors and devices ---------------- */
int swd_last_error(char *buf, size_t n);
int swd_get_device_count(int *n);                              /* getdevice (RK:1839) */

/* ---------------- host setup, reference semantics ---------------- */
int swd_read_parameters(const char *path, swd_params *out);   /* input_parameters (RC:1931-2035) */
int swd_cal_xishu(int Lc, float *rc);                          /* cal_xishu (RC:1858-1878) */
int swd_ricker(float *rick, int nt, float f0, float t0, float dt); /* ricker_wave flag 2 (RC:1284-1299) */
int swd_cpml(int ntx, int ntz, int pml_arg, float dx, float dz, float dt,
             float *a_x, float *a_x_half, float *b_x, float *b_x_half,
             float *a_z, float *a_z_half, float *b_z, float *b_z_half); /* get_absorbing_parameters (RC:891-1129) */
int swd_model_in_pml(float *vp, int ntz, int ntx, int pml);    /* model_in_pml (RC:1428-1468) */
int swd_read_model(const char *path, float *vp, float *rho, int ntx, int ntz, int pml); /* get_acc_model (RC:1136-1204) */
int swd_smooth_model(float *vp, int ntz, int ntx, int pml, int window); /* ini_model_mine (RC:1473-1569) */
/* Largest stable velocity of the scheme, 1/(sum|rc| dt sqrt(1/dx^2 + 1/dz^2)), the
 * limit the reference prints (RC:110-118, dx/(sum|rc| dt sqrt 2) for dx = dz);
 * rc = NULL takes swd_cal_xishu(Lc). */
int swd_cfl_vmax(const swd_grid *g, const float *rc, double *vmax);
/* The checks swd_plan_set_model makes on a padded model [ntx][ntz] (no
 * reference counterpart: the reference only prints the limit, RC:117):
 * SWD_E_ARG for a non-positive or non-finite vp or rho, SWD_E_RANGE for a vp
 * above swd_cfl_vmax. */
int swd_check_model(const swd_grid *g, const float *rc, const float *vp, const float *rho);
int swd_nfft(int n);                                           /* 2^ceil(log2 n) (FK:595-596) */
int swd_kz_taper(float *coef, int nfftz, float dz);            /* kz taper (FC:354-377); returns layerz */
int swd_hilbert1d(const float *rick, float *hrick, int nt);    /* hilbert_transform1d (FK:1257-1324) */
int swd_geometry(int nx, int nz, int pml, int Lc, float dx, float dz, float sx0, float shotdx,
                 float shotdep, float offsetz, int is, int *s_ix, int *s_iz, int *r_iz, int max_rn);
                                                               /* shot/receiver geometry (RC:150-200); returns r_n */
int swd_adjoint_source(const float *d, float *rms, int ntr, int nt, float dt); /* RC:603-637 */
int swd_scatterers(float *rho, const float *vp, int ntx, int ntz, int pmlc, int swin); /* FC:697-705 */
int swd_image_normalize(float *image, const float *illum, int ntp);              /* RC:663-679 */
int swd_image_laplacian(const float *image, const float *vp, float *out, float *acc,
                        int ntx, int ntz, int pmlc, float dx, float dz);          /* RC:722-735 */
int swd_edge_zero(float *mig, int nx, int nz, int Lc);                           /* RC:754-768 */

/* ---------------- plans ---------------- */
/* variables_malloc (RK:1613, FK:1425).  SWD_E_RANGE when a padded field
 * (ntx x ntz rounded up to 32 rows) reaches 2^30 floats. */
int swd_plan_create(const swd_grid *g, int device, void **plan);
int swd_plan_destroy(void *plan);                                /* variables_free (RK:1732, FK:1565) */
int swd_plan_set_coefs(void *plan, const swd_coefs *c);          /* H2D of rc and CPML vectors (RK:278-297) */
/* H2D of vp, rho (RK:281-282), after swd_check_model (the plan's rc if its
 * coefficients are set). */
int swd_plan_set_model(void *plan, const float *vp, const float *rho);
int swd_plan_set_path(void *plan, int path);
/* Tile of the fused path's workgroups (no reference counterpart: the reference
 * fixes 16 x 16 thread blocks, RK:7, forward/headrtm.h:3): tile_x columns
 * (16, 32, 64) by tile_rows rows (64; at Lc 4 also 32, and 128 up to 32
 * columns) with a bound of waves wavefronts per SIMD (0: none; at Lc 4 also
 * 5).  All zero (the default): the tuner times every configuration the loop's
 * kernel variant supports on the first steps of a loop, keeps the default 16 x
 * 64 unless another is more than 10% faster, and remembers the choice for the
 * plan's later loops with as many instances per launch and on disk (the file
 * SWD_TUNE_CACHE, else $XDG_CACHE_HOME or $HOME/.cache/swd_tiles.txt) keyed by
 * device, grid, Lc, instances per launch, variant and streams.  Variants that
 * lack an asked height or bound run the asked width at 64 rows.  For two steps
 * per launch (swd_plan_set_steps_per_launch) waves is the wavefronts per
 * workgroup instead, negative for split launches (the tiles free of damping
 * and frame cells, then the others): 32 x 32 rows with 4 or -4, 64 x 32 with 8
 * or -8, 16 x 64 with 4, 32 x 64 with -8 (0: the first of the shape).
 * Results do not
 * depend on the tile. */
int swd_plan_set_tile_config(void *plan, int tile_x, int tile_rows, int waves);
/* swd_plan_set_tile_config(plan, tile_x, 64, 0), or the tuner for tile_x = 0. */
int swd_plan_set_tile(void *plan, int tile_x);
/* Time steps per launch of the fused forward (swd_forward, swd_afor, the
 * forward of every RTM) and the paired RTM backward at Lc 4 (no reference counterpart: the
 * reference launches three kernels per step, RK:1000-1045): 1, 2 (two steps
 * of every wavefield per launch, temporal blocking: the state, CPML memories
 * and image accumulators are read and written once per two steps), or 0 (the
 * default): two steps per launch where they apply, the tuner timing their
 * tile shapes and keeping 32 x 32 rows unless another is more than 10% faster
 * (a loop too short to time them keeps 32 x 32).
 * With 2 and an asked tile configuration, its shape runs two steps per launch
 * where the two-step kernel has it (32 x 32, 16 x 64, 64 x 32, 32 x 64 rows),
 * else 32 x 32.  Two steps per launch need Lc 4; they cover the forward
 * loops (with the Born adjoint's background passes and swd_multi_forward,
 * whose tile is then 32 x 32 rows) and the paired backward
 * of the RTM, the up/down RTM and the angle-gather RTM; the Born forward pair
 * keeps one step per launch.  Results do not depend on it. */
int swd_plan_set_steps_per_launch(void *plan, int steps);
/* Group sizes of the fused loops: the RTM forward and backward run their
 * batch in groups of at most fwd_shots and bwd_shots shots, and swd_forward,
 * swd_afor and the Born forward theirs in groups of at most fwd_shots
 * wavefields, a shot's wavefields kept together (0, the default: sized so a
 * group's working set stays in the last-level cache).
 * The group count is raised to a multiple of the streams where the batch
 * allows.  Results do not depend on them. */
int swd_plan_set_groups(void *plan, int fwd_shots, int bwd_shots);
/* Streams the shot groups of the fused loops run on at once, 1 to 4 (default 2):
 * consecutive groups take the plan's stream and its other streams (default
 * priority) in turn, so one group's launch gaps and tails overlap the others'
 * work, and the cache budget of a group is divided by the count.  Results do
 * not depend on it. */
int swd_plan_set_streams(void *plan, int streams);
/* Replay fused time loops from captured hipGraphs (on by default): the
 * launches of a loop are recorded once per configuration and replayed on
 * later calls with the same batch layout. */
int swd_plan_set_graphs(void *plan, int on);
/* Device memory a call may use for its workspace (0, the default: 85% of the
 * free device memory).  swd_rtm*, swd_born*, swd_forward and swd_afor run larger
 * batches in consecutive waves of shots that fit; results are the same. */
int swd_plan_set_memory_limit(void *plan, size_t bytes);
int swd_plan_get_stats(void *plan, swd_stats *st);
/* The plan's HIP stream (a hipStream_t), for callers that order their own
 * work or timing events with the plan's queued work. */
int swd_plan_get_stream(void *plan, void **stream);

/* ---------------- operators ----------------
 * nshots shots of one plan run as one batch; wavelets are shared [nt];
 * gathers of the batch are concatenated in shot order ([sum nrec][nt]). */

/* Forward modelling, pressure gather (fdtd_2d_GPU_forward with inv_flag=0, RK:239-432). */
int swd_forward(void *plan, int nshots, const swd_shot *shots, const float *wavelet, float *gather);

/* AFor forward: p (wavelet) and htp (hwavelet) wavefields, gather of p and kz
 * up/down separated gathers (FK:576-863 with FK:52-138, cuFFT FK:762-769).
 * kzcoef is the taper [nfftz], nfftz a power of two >= ntz. */
int swd_afor(void *plan, int nshots, const swd_shot *shots, const float *wavelet,
             const float *hwavelet, const float *kzcoef, int nfftz,
             float *gather, float *up, float *down);
/* Same as swd_afor with gather, up and down in device memory of the plan's
 * device ([sum nrec][nt] each). */
int swd_afor_dev(void *plan, int nshots, const swd_shot *shots, const float *wavelet,
                 const float *hwavelet, const float *kzcoef, int nfftz,
                 float *gather, float *up, float *down);

/* RTM / PSF imaging of a batch: forward with boundary strips (inv_flag=1,
 * RK:239-432), then source reconstruction, receiver back-propagation of
 * adj[sum nrec][nt] and cross-correlation imaging (fdtd_2d_GPU_backward,
 * RK:946-1215).  image/illum_src/illum_rcv are [nshots][ntx*ntz];
 * gather_syn (optional) receives the forward gathers. */
int swd_rtm(void *plan, int nshots, const swd_shot *shots, const float *wavelet,
            const float *adj, float *image, float *illum_src, float *illum_rcv, float *gather_syn);

/* Same as swd_rtm with device pointers (adj, image, illum_*, gather_syn on the
 * plan's device; gather_syn may be NULL). */
int swd_rtm_dev(void *plan, int nshots, const swd_shot *shots, const float *wavelet,
                const float *adj, float *image, float *illum_src, float *illum_rcv, float *gather_syn);

/* RTM with angle-domain common-image gathers (sum_image_GPU_kernel_imageCIG,
 * RK:819-905, whose launch RK:1149-1157 is commented out; thetan = 10 and
 * delt = 1 in RC:235).  Same as swd_rtm, plus cig [nshots][ntx*ntz][thetan]
 * (the reference's d_image_ppcig layout): after every backward step each
 * interior cell bins half the angle between the source and receiver Poynting
 * vectors (in degrees, averaged over the last delt steps) into bins
 * ceil(angle + ii - 0.5), ii = -2..2, in [2, thetan), with weights
 * exp(-ii^2/18) times p_s p_r.  The receiver vector is +p v of the
 * time-reversed field, where RK:866-867 writes -p v (which would give 90
 * degrees minus the reflection angle).  thetan in [1, 1024], delt in [1, 64].
 * Runs on the fused kernels for every path. */
int swd_rtm_cig(void *plan, int nshots, const swd_shot *shots, const float *wavelet, const float *adj,
                int thetan, int delt, float *image, float *illum_src, float *illum_rcv, float *cig);
/* Same with adj, image, illum_* and cig in device memory of the plan's device. */
int swd_rtm_cig_dev(void *plan, int nshots, const swd_shot *shots, const float *wavelet, const float *adj,
                    int thetan, int delt, float *image, float *illum_src, float *illum_rcv, float *cig);

/* Up/down-separated RTM (the image_ppdu/ppud and d_image_ppdu/ppud fields of
 * forward/headrtm.h:51,119, whose separation kernel FK:140-165 the reference
 * never launches).
 * The source side propagates the wavelet and hwavelet (its Hilbert transform
 * along time), the receiver side adj and hadj (the adjoint source and its
 * Hilbert transform along time, [sum nrec][nt] each).  At every backward step
 * the analytic columns S = dp/dt + i dp_h/dt (source, zero off the interior)
 * and R = p + i p_h (receiver) of each interior column (the whole padded
 * column when nfftz >= ntz, else its nz interior rows) are zero padded to
 * nfftz, transformed along z, weighted by kzcoef and split as in swd_afor
 * (bins 0..nfftz/2 "up", the rest "down"), transformed back and divided by
 * nfftz; the real parts are cross-correlated on the interior:
 *   du += Sd Ru, ud += Su Rd, uu += Su Ru, dd += Sd Rd.
 * With kzcoef all ones du+ud+uu+dd is the plain image of swd_rtm.
 * nfftz is a power of two in [max(nz, 128), 4096].  Runs on the fused kernels for
 * every path.  Outputs are [nshots][ntx*ntz]; du and ud are required, uu and
 * dd optional together, and image/illum_src/illum_rcv (the plain RTM outputs
 * of the same run) optional together. */
typedef struct swd_updown_images {
    float *du, *ud, *uu, *dd;
    float *image, *illum_src, *illum_rcv;
} swd_updown_images;
int swd_rtm_updown(void *plan, int nshots, const swd_shot *shots, const float *wavelet, const float *hwavelet,
                   const float *adj, const float *hadj, const float *kzcoef, int nfftz, const swd_updown_images *out);
/* Same with adj, hadj and the outputs in device memory of the plan's device. */
int swd_rtm_updown_dev(void *plan, int nshots, const swd_shot *shots, const float *wavelet, const float *hwavelet,
                       const float *adj, const float *hadj, const float *kzcoef, int nfftz,
                       const swd_updown_images *out);

/* One step of that imaging condition on caller-held wavefields, all device
 * pointers [nshots][ntx*ntz]: src/hsrc are the two source fields (dp/dt of
 * the wavelet's and of hwavelet's reconstruction, zero off the interior),
 * rcv/hrcv the two receiver pressures; du, ud (and uu, dd) accumulate on the
 * interior.  Queued on the plan's stream without waiting (no reference
 * counterpart: the per-step building block of swd_rtm_updown for callers
 * running their own time loop). */
int swd_updown_image_dev(void *plan, int nshots, const float *src, const float *hsrc, const float *rcv,
                         const float *hrcv, const float *kzcoef, int nfftz, const swd_updown_images *out);

/* The step of swd_updown_image_dev with no device allocation.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight per-step serving loops, harnesses that forbid
 * hipMalloc).  Its tables (shot descriptors, twiddles, kz taper) go to a fixed
 * 64 KiB array of the library on each GPU instead of the plan's workspace:
 * nshots <= SWD_NOALLOC_MAX_SHOTS (else SWD_E_RANGE), nfftz <= 4096 as for
 * swd_updown_image_dev.  Every call on a GPU shares that GPU's array, which
 * keeps the last call's tables: calls queued on one stream are ordered, calls
 * that overlap on one GPU (other streams or threads) race. */
#define SWD_NOALLOC_MAX_SHOTS 64
int swd_updown_image_dev_noalloc(void *plan, int nshots, const float *src, const float *hsrc, const float *rcv,
                                 const float *hrcv, const float *kzcoef, int nfftz, const swd_updown_images *out);

/* Born modelling d = L m, m = dK/K on the interior [nx][nz] (no reference
 * counterpart: linearisation of the RK:24-221 scheme), and its exact discrete
 * adjoint m_grad = L^T d with sqrt(nt) checkpointing (nsnap <= 0: automatic).
 * d is [sum nrec][nt]; m and m_grad are [nx][nz] per shot, m_grad summed over
 * the batch. */
int swd_born(void *plan, int nshots, const swd_shot *shots, const float *wavelet,
             const float *m, float *d);
int swd_born_adjoint(void *plan, int nshots, const swd_shot *shots, const float *wavelet,
                     const float *d, float *m_grad, int nsnap);

/* Same as swd_born / swd_born_adjoint with m, d and m_grad in device memory of
 * the plan's device (wavelet stays a host array). */
int swd_born_dev(void *plan, int nshots, const swd_shot *shots, const float *wavelet,
                 const float *m, float *d);
int swd_born_adjoint_dev(void *plan, int nshots, const swd_shot *shots, const float *wavelet,
                         const float *d, float *m_grad, int nsnap);

/* Forward modelling of several plans' batches of uneven grid sizes in one
 * fused launch per time step (no reference counterpart: the reference runs one
 * grid per call, FK:576-863).  All plans must be on one device and share Lc and
 * nt; plan i runs nshots[i] shots shots[i] with wavelet wavelets[i][nt] and
 * gathers[i] receives its [sum nrec][nt].  The launches run on plans[0]'s
 * stream, from a captured hipGraph when plans[0] has graphs on. */
int swd_multi_forward(int nplans, void *const *plans, const int *nshots, const swd_shot *const *shots,
                      const float *const *wavelets, float *const *gathers);

/* ---------------- wavenumber-domain PSF deconvolution ----------------
 * The windowed Wiener deconvolution of a migrated image by the nearest PSF of
 * a point-scatterer lattice (decon2d_mine, DC:17-785), one process.  Images
 * are [nx][nz] float32 (z contiguous), as the raw .bin files (DC:255-261). */

/* kkdecon parameter.txt values used by the deconvolution (DC:1527-1586). */
typedef struct swd_decon_params {
    int nx, nz;
    int psfwinx, psfwinz;        /* PSF lattice spacing / patch (odd) */
    int filterwinx, filterwinz;  /* image window (odd) */
    int snapwinx, snapwinz;      /* taper length of the output window */
    float epsilon;               /* Wiener regularisation, times max |S|^2 */
} swd_decon_params;

/* The file names and flags of the same parameter.txt (DC:1587-1626). */
typedef struct swd_decon_files {
    char vp[256], vp0[256], rho[256], image[256], psf[256], decon[256], inv_psf[256];
    int shift_flag;
    float reg_eps;
} swd_decon_files;

/* Derived sizes (DC:84-129): n = filterwin + psfwin - 1 (square windows
 * only), the output window, the window step and counts, the padded image. */
typedef struct swd_decon_geom {
    int n, outwinx, outwinz, stepwinx, stepwinz, stepnumx, stepnumz, windows, nxp, nzp, cells;
} swd_decon_geom;

typedef struct swd_decon_stats {
    double sums_ms;                                  /* 0: the window sums run inside the window kernel */
    double cells_ms, windows_ms, gather_ms;          /* event times of the last call's kernels */
    double total_ms;
    int windows, cells, n;
    long long ws_bytes;                                        /* device memory held by the plan */
} swd_decon_stats;

int swd_decon_read_parameters(const char *path, swd_decon_params *p, swd_decon_files *files); /* input_parameters (DC:1527-1626) */
int swd_decon_geometry(const swd_decon_params *p, swd_decon_geom *g);  /* DC:84-129 */
int swd_decon_taper(const swd_decon_params *p, float *coefr);         /* coefr[outwinx][outwinz] (DC:355-382) */
/* A deconvolution plan holds the twiddle tables, taper and workspace of one
 * parameter set on one device and stream (the FFTW plans and buffers of DC:449-493). */
int swd_decon_plan_create(const swd_decon_params *p, int device, void **plan);
int swd_decon_plan_destroy(void *plan);
/* Deconvolution of image by the PSF image (DC:543-682): dec receives
 * migration_dudecon (DC:724-731), coef the summed taper weights, migration_coef
 * (DC:702-709; may be NULL).  Host arrays [nx][nz]. */
int swd_decon(void *plan, const float *image, const float *psf, float *dec, float *coef);
/* Same with device arrays on the plan's device (coef required). */
int swd_decon_dev(void *plan, const float *image, const float *psf, float *dec, float *coef);
int swd_decon_get_stats(void *plan, swd_decon_stats *st);

#ifdef __cplusplus
}
#endif
#endif
