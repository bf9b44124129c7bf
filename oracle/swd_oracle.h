/*
 * swd_oracle.h -- CPU oracle for the 2D acoustic staggered-grid CPML FDTD / RTM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, as plain C host loops, the
 * per-cell equations of the reference program qingchenzhang/Simple-wavenumber-
 * deconvolution (files acoustic_HLVSPRTMC.cpp = "RC", acoustic_HLVSPRTMLc.cu = "RK",
 * forward/acoustic_hilbert_fftkz_HLForC.cpp = "FC",
 * forward/acoustic_hilbert_fftkz_HLForLc.cu = "FK").  It is the checker the HIP
 * path is compared against; it is never linked into, called by, or shipped with
 * the product library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.
 *
 * Two twins of every propagation routine exist: *_f32 does all arithmetic in
 * float (the parity reference), *_f64 does the identical sequence in double (the
 * accuracy oracle).  Both take the same float inputs.  The arithmetic contract
 * shared with the HIP kernels ("canonical order") is:
 *   derivative   d = 0; for ic<Lc: d = fma(rc[ic]*(f[+]-f[-]), 1/dx, d)
 *   memory var   psi = fma(b, psi, a*d)
 *   velocity     v = fma(B, d + psi, v),           B = dt * (float)(1/(0.5(rho+rho')))
 *   pressure     p = fma(K, ((dvx+psi_x)+dvz)+psi_z, p),  K = ((dt*rho)*vp)*vp
 * which is the reference expression (RK:55-64, 104-113, 162-171) with the
 * multiply-add contracted the way nvcc contracts it by default.
 *
 * Grid convention (RC:78-83, RK:45): pmlc = pml+Lc, ntx = nx+2*pmlc,
 * ntz = nz+2*pmlc, cell (ix,iz) lives at ip = ix*ntz+iz (z contiguous).
 */
#ifndef SWD_ORACLE_H
#define SWD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_grid {
    int nx, nz;      /* interior cells */
    int pml, Lc;     /* CPML cells (parameter.txt "pml"), half stencil order */
    float dx, dz, dt;
    int nt;          /* itmax */
} or_grid;

typedef struct or_coefs {
    const float *rc;                        /* [Lc]  */
    const float *a_x, *a_x_half, *b_x, *b_x_half;  /* [ntx] */
    const float *a_z, *a_z_half, *b_z, *b_z_half;  /* [ntz] */
} or_coefs;

typedef struct or_shot {
    int s_ix, s_iz;          /* padded-grid source node */
    int nrec;                /* traces */
    const int *rec_ix;       /* [nrec] padded-grid receiver column */
    const int *rec_iz;       /* [nrec] padded-grid receiver row */
} or_shot;

/* ---- setup restatements (mixed precision exactly as the reference) ---- */
void or_cal_xishu(int Lc, float *rc);                                   /* RC:1858-1878 */
void or_ricker(float *rick, int nt, float f0, float t0, float dt);      /* RC:1284-1299 */
void or_cpml(int ntx, int ntz, int pml_arg, float dx, float dz, float dt,
             float *a_x, float *a_x_half, float *b_x, float *b_x_half,
             float *a_z, float *a_z_half, float *b_z, float *b_z_half); /* RC:891-1129 */
void or_model_in_pml(float *vp, int ntz, int ntx, int pml);             /* RC:1428-1468 */
void or_smooth(float *vp, int ntz, int ntx, int pml, int window);       /* RC:1473-1569 */
int  or_nfft(int n);                                                    /* FK:595-596 */
int  or_kz_taper(float *coef, int nfftz, float dz);                     /* FC:354-377 */
void or_hilbert1d(const float *rick, float *hrick, int nt);             /* FK:1257-1324 */
int  or_geometry(int nx, int nz, int pml, int Lc, float dx, float dz,
                 float sx0, float shotdx, float shotdep, float offsetz, int is,
                 int *s_ix, int *s_iz, int *r_iz, int max_rn);          /* RC:150-200 */
void or_adjoint_source(const float *d, float *rms, int ntr, int nt, float dt); /* RC:603-637 */
void or_scatterers(float *rho, const float *vp, int ntx, int ntz, int pmlc, int swin); /* FC:697-705 */
void or_image_normalize(float *image, const float *illum, int ntp);     /* RC:663-679 */
void or_image_laplacian(const float *image, const float *vp, float *out, float *acc,
                        int ntx, int ntz, int pmlc, float dx, float dz); /* RC:722-735 */
void or_edge_zero(float *mig, int nx, int nz, int Lc);                  /* RC:754-768 */

/* ---- propagation, FP32 / FP64 twins ---- */
/* Forward CPML modelling of one shot (RK:239-432 with RK:24-221).  gather is
 * [nrec][nt]; strips (optional) receives the Lc-wide interior frame of p after
 * every step (RK:200-219), layout [nt][nframe] with or_frame_size(); final
 * (optional) receives vx, vz, p at it = nt-1, [3][ntp] (RK:403-415). */
int or_forward_f32(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                   const or_shot *s, const float *wavelet, float *gather, float *strips, float *final_state);
int or_forward_f64(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                   const or_shot *s, const float *wavelet, double *gather, double *strips, double *final_state);

/* RTM backward (RK:946-1215): source reconstruction from final state + strips,
 * time-reversed receiver propagation with hard replacement of adj[nrec][nt],
 * cross-correlation imaging.  image/illum_s/illum_r are [ntp]. */
int or_rtm_backward_f32(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                        const or_shot *s, const float *wavelet, const float *final_state,
                        const float *strips, const float *adj,
                        float *image, float *illum_s, float *illum_r);
int or_rtm_backward_f64(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                        const or_shot *s, const float *wavelet, const double *final_state,
                        const double *strips, const float *adj,
                        double *image, double *illum_s, double *illum_r);

/* The same two runs with per-step histories for the reconstruction tests
 * ([nt][ntp] each, any may be NULL; gather may be NULL too).
 * or_forward_hist: phist[it] = p after the source injection of step it.
 * or_rtm_backward_hist: at backward step it, shist[it] = the reconstructed p
 * after the strip restore, rhist[it] = the receiver p after the hard
 * replacement, dhist[it] = dp/dt of the step (rows it = nt-1 are untouched). */
int or_forward_hist_f32(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                        const or_shot *s, const float *wavelet, float *gather, float *strips, float *final_state,
                        float *phist);
int or_forward_hist_f64(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                        const or_shot *s, const float *wavelet, double *gather, double *strips, double *final_state,
                        double *phist);
int or_rtm_backward_hist_f32(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                             const or_shot *s, const float *wavelet, const float *final_state,
                             const float *strips, const float *adj, float *image, float *illum_s, float *illum_r,
                             float *shist, float *rhist, float *dhist);
int or_rtm_backward_hist_f64(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                             const or_shot *s, const float *wavelet, const double *final_state,
                             const double *strips, const float *adj, double *image, double *illum_s,
                             double *illum_r, double *shist, double *rhist, double *dhist);

/* AFor forward (FK:576-863): p driven by wavelet and htp driven by hwavelet,
 * gather of p, and per-step kz up/down separation of p + i*htp on the receiver
 * columns (FK:705, 762-777, 52-138).  coef is the kz taper [nfftz]. */
int or_afor_f32(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                const or_shot *s, const float *wavelet, const float *hwavelet,
                const float *kzcoef, int nfftz, float *gather, float *up, float *down);
int or_afor_f64(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                const or_shot *s, const float *wavelet, const float *hwavelet,
                const float *kzcoef, int nfftz, double *gather, double *up, double *down);

/* Up/down-separated RTM images of one shot: two reconstructed source fields
 * (wavelet and hwavelet, each with its final state and strips) and two
 * receiver fields (adj and hadj), split along kz per interior column at every
 * backward step; du/ud/uu/dd are [ntp].  Not in the reference's live code
 * (FK:140-165 and forward/headrtm.h:51,119 declare it unused); defined here. */
int or_rtm_updown_f32(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                      const or_shot *s, const float *wavelet, const float *hwavelet,
                      const float *final_state, const float *strips, const float *hfinal_state, const float *hstrips,
                      const float *adj, const float *hadj, const float *kzcoef, int nfftz,
                      float *du, float *ud, float *uu, float *dd);
int or_rtm_updown_f64(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                      const or_shot *s, const float *wavelet, const float *hwavelet,
                      const double *final_state, const double *strips, const double *hfinal_state,
                      const double *hstrips, const float *adj, const float *hadj, const float *kzcoef, int nfftz,
                      double *du, double *ud, double *uu, double *dd);

/* RTM with angle-domain common-image gathers (sum_image_GPU_kernel_imageCIG,
 * RK:819-905, launched nowhere in the reference: RK:1149-1157 is commented out).
 * Same as or_rtm_backward plus cig [ntp][thetan]: Poynting-vector opening
 * half-angles binned in degrees, delt-step time averaging of the vector products. */
int or_rtm_cig_f32(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                   const or_shot *s, const float *wavelet, const float *final_state,
                   const float *strips, const float *adj, int thetan, int delt,
                   float *image, float *illum_s, float *illum_r, float *cig);
int or_rtm_cig_f64(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                   const or_shot *s, const float *wavelet, const double *final_state,
                   const double *strips, const float *adj, int thetan, int delt,
                   double *image, double *illum_s, double *illum_r, double *cig);

/* The angle-gather update of one step on n independent cells (flat arrays of
 * the six fields; rings [delt][n]; cig [n][thetan]): the per-cell part of
 * or_rtm_cig, exposed for tests of the binning. */
void or_cig_cells_f32(int n, const float *ps, const float *vxs, const float *vzs, const float *pr, const float *vxr,
                      const float *vzr, int it, int nt, int delt, int thetan, float *ringF, float *ringG, float *cig);
void or_cig_cells_f64(int n, const double *ps, const double *vxs, const double *vzs, const double *pr,
                      const double *vxr, const double *vzr, int it, int nt, int delt, int thetan, double *ringF,
                      double *ringG, double *cig);

/* Born forward d = L m and its exact discrete adjoint m = L^T d (the linearised
 * CPML scheme with K -> K(1+m) on the interior, linearised in m at the background model). */
int or_born_f32(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                const or_shot *s, const float *wavelet, const float *m, float *d);
int or_born_f64(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                const or_shot *s, const float *wavelet, const float *m, double *d);
int or_born_adjoint_f32(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                        const or_shot *s, const float *wavelet, const float *d, float *mgrad);
int or_born_adjoint_f64(const or_grid *g, const float *vp, const float *rho, const or_coefs *c,
                        const or_shot *s, const float *wavelet, const float *d, double *mgrad);

/* Wavenumber-domain PSF deconvolution (kkdecon_acoustic/kk_decon_interp_psf_minewindow.cpp,
 * DC), one rank.  image, psf, dec, coef are [nx][nz] (z fastest). */
typedef struct or_decon_params {
    int nx, nz, psfwinx, psfwinz, filterwinx, filterwinz, snapwinx, snapwinz;
    float epsilon;
} or_decon_params;
typedef struct or_decon_geom {
    int halfpsfwx, halfpsfwz, halfltwx, halfltwz, nxwpsf, nzwpsf, fullwx, fullwz, midpx, midpz;
    int outwinx, outwinz, halfoutwx, halfoutwz, stepwinx, stepwinz, stepnumx, stepnumz, totalnum, nxp, nzp;
} or_decon_geom;
int or_decon_geometry(const or_decon_params *p, or_decon_geom *g);
void or_decon_coefr(const or_decon_params *p, const or_decon_geom *g, float *coefr);
void or_decon_expand(const float *a, int nx, int nz, float *af, int nxp, int nzp);
void or_twiddles(int n, double *c, double *s);
void or_decon_spectrum(const double *x, int mx, int mz, int offx, int offz, int n, double *re, double *im);
void or_decon_inverse(const double *dr, const double *di, int n, int r0, int nr, int c0, int nc, double *out);
void or_decon_cell(const or_decon_params *p, const or_decon_geom *g, int ix, int iz, int *cx, int *cz);
void or_decon_psf(const or_decon_params *p, const or_decon_geom *g, const float *psff, int cx, int cz, double *sr,
                  double *si, float *tden);
void or_decon_window(const or_decon_params *p, const or_decon_geom *g, const float *imgf, const float *psff,
                     int step, double *out);
int or_decon(const or_decon_params *p, const float *image, const float *psf, float *dec, float *coef);

int or_frame_size(const or_grid *g);
void or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
