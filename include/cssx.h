/*
 * cssx: MI355X-native compact sixth-order derivatives, MHD right-hand side and RK4 stages of
 * the CSS code (orvedahl/CSS-Code), exposed as a plain C ABI.
 *
 * Data layout everywhere is the reference's: a field f(nth, nphi, nr) in Fortran order, i.e.
 * theta fastest, then phi, then r (parallel_mpi.f90:41, rhs_mpi.f90:24).  A state holds nv
 * such fields back to back in the order rho, w=u_theta, v=u_phi, u=u_r, s, A_theta, A_phi,
 * A_r (rhs_mpi.f90:24, the magnetic slots hold the vector potential between stages).
 *
 * All pointers passed to compute entry points are device pointers owned by the caller;
 * `stream` is a hipStream_t (NULL = the null stream).  Calls are stream-ordered and do not
 * synchronize unless stated.  Every entry point returns 0 on success or a CSSX_E* code and
 * never terminates the process (the reference's Graceful_Exit, parallel_mpi.f90:886-889,
 * aborts the whole MPI job instead).
 */
#ifndef CSSX_H
#define CSSX_H

#ifdef __cplusplus
extern "C" {
#endif

#define CSSX_OK 0
#define CSSX_ENOMEM 12      /* device or host allocation failed */
#define CSSX_EINVAL 22      /* bad shape, boundary code, axis or pointer */
#define CSSX_ENAN 33        /* non-finite value detected by a checked call */
#define CSSX_EHIP 100       /* a HIP runtime call failed */
#define CSSX_ESIZE 27       /* problem larger than a no-allocation entry point's fixed scratch */

#define CSSX_PREC_F64 0
#define CSSX_PREC_F32 1

#define CSSX_AXIS_R 0
#define CSSX_AXIS_THETA 1
#define CSSX_AXIS_PHI 2

typedef struct cssx_ctx cssx_ctx;

/* One rank's subdomain (the single-GPU build treats it as the whole domain).
 * The *_wall flags say whether that end is a physical boundary (derivatives.f90:529-544:
 * Determine_Dtype); phi is periodic. */
typedef struct cssx_grid {
    int nr, nth, nphi;
    int r_lo_wall, r_hi_wall;
    int th_lo_wall, th_hi_wall;
    double dri, dti, dpi;   /* 1/dx in index space: derivatives.f90:91-97 */
} cssx_grid;

/* Library version (major*10000 + minor*100 + patch). */
int cssx_version(void);

/* Human-readable text for an error code. */
const char *cssx_strerror(int code);

/* Replaces Initialize_Derivatives (derivatives.f90:39-520): validates the grid and prepares
 * the Thomas/segment plans for every boundary code on the current HIP device. */
int cssx_ctx_create(const cssx_grid *grid, cssx_ctx **out);
void cssx_ctx_destroy(cssx_ctx *ctx);

/* One directional derivative over the whole subdomain.
 * Replaces Dbydr / Dbydt / Dbydp (order 1; derivatives.f90:547, :1010, :1480) and
 * D2bydr2 / D2bydt2 / D2bydp2 (order 2; :799, :1265, :1594).
 *   axis      CSSX_AXIS_R / _THETA / _PHI
 *   bdry      boundary code 0..4 (derivatives.f90:11-21); ignored on phi
 *   prec      CSSX_PREC_F64 (in/out double) or CSSX_PREC_F32 (float)
 *   ghost_lo/ghost_hi  halo buffers in the reference's layout for that axis:
 *             r     (nth, nphi, 4)   binfo(1,side)%bdata
 *             theta (4, nphi, nr)    binfo(2,side)%bdata
 *             phi   (nth, 4, nr)     binfo(3,side)%bdata
 *             At a physical wall slot 0 holds the prescribed physical first derivative
 *             (NULL = 0).  On phi, NULL ghosts select the single-subdomain periodic wrap.
 * The result is f' (or f'') in index units times dri/dti/dpi (squared for order 2), exactly
 * what the reference returns; the caller multiplies by dx/dr on a stretched radial grid. */
int cssx_deriv(cssx_ctx *ctx, int axis, int order, int bdry, int prec, const void *in, void *out,
               const void *ghost_lo, const void *ghost_hi, void *stream);

/* CSS-Python compact_fd6 (CSS-Python/derivatives.py:32-185) on a batch of pencils.
 * Pencil p's point j is arr[j*npencils + p] (pencil index fastest).  b1/b2 hold 4 ghosts
 * per pencil at b[k*npencils + p] (used for dtype 1/3 and 2/3; may be NULL otherwise).
 * darr (may be NULL): on entry rows 0 / n-1 carry the Neumann values f'/dri as in the Python;
 * on exit it holds the assembled right-hand side, as the Python leaves it.
 * result = the derivative (solution * dri), double precision.
 * n >= 5, as in the Python, whose rows reach arr[4] and arr[n-5]; a pencil whose tridiagonal
 * system is singular (n = 5 with Dirichlet-type closures at both walls) returns +-inf, as
 * the reference's Thomas solve does. */
int cssx_compact_fd6(int n, long npencils, double dri, int ibc, int dtype, const double *arr,
                     const double *b1, const double *b2, double *darr, double *result, void *stream);

/* cssx_compact_fd6 that allocates nothing.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and errors as cssx_compact_fd6,
 * bit for bit; its one scratch buffer, the FP64 plan table of the launch, lives in a fixed
 * 64 KiB array per device (8192 doubles; the largest plan, n = 1024, takes 6080) that the
 * library holds until the process ends.  The plan is copied into it on `stream` before the
 * launch, so calls that overlap on different streams of one device race; a plan that would not
 * fit returns CSSX_ESIZE. */
int cssx_compact_fd6_noalloc(int n, long npencils, double dri, int ibc, int dtype, const double *arr,
                             const double *b1, const double *b2, double *darr, double *result, void *stream);

/* ---------------------------------------------------------------------------------------------
 * Compressible MHD right-hand side and RK4 stepping (rhs_mpi.f90, step_rk4.f90, diffusion.f90)
 * ------------------------------------------------------------------------------------------- */

/* Physics constants, switches and boundary codes.  Field names follow the reference's
 * namelist / module variables (constants.f90:17-72, rhs_mpi.f90:10-13, diffusion.f90:8-11).
 * The stratified (not Unstratified), do_radheat=.False. physics is supported, with SLD
 * diffusion, the explicit Laplacian / SLaplacian diffusion, or both; the flags below are the
 * ones that path reads. */
typedef struct cssx_mhd_params {
    double gamma, Cp;            /* Cv=Cp/gamma, pconst=(gamma-1)*Cp (rhs_mpi.f90:59-73) */
    double omega0;               /* frame rotation rate omega_0 */
    double xi_cs;                /* sound-speed reduction: drho/dt *= xi_cs^2 when < 1 */
    double damping_time, T0;     /* top relaxation toward T0 (rhs_mpi.f90:1807-1818) */
    int magnetic;                /* 8 fields with the vector potential in slots 6-8; do_lorentz */
    int top_relax;               /* apply the top relaxation (not Set_Stop/top_inout/...) */
    int bottom_stress_free, top_stress_free;
    int bottom_const_flux, top_const_flux, bottom_const_temp, top_const_temp;
    int bottom_radial_field, top_radial_field;
    int theta_normal_field, theta_perf_cond, theta_stress_free;
    int ibcr[13], ibct[13];      /* boundary codes by variable index 1..12 (ibcr/ibct of
                                    derivatives.f90:50-57: rho,w,v,u,s,T,bt,bp,br,at,ap,ar) */
    /* SLD diffusion (diffusion.f90:8-11; Compute_Diffusion :70-474) */
    int sld_on;                  /* SLDiffusion: Euler diffusion folded into stage 1 */
    double sld_coef;             /* 0 gives an exactly zero diffusion tendency */
    int mass_diffusion, magnetic_diffusion, viscous_heating, ohmic_heating;
    int dstype;                  /* 1: add |u|^2 and B^2/(4 pi rho) to the SLD speed^2 */
    double sld_pr, sld_prm;      /* flux multipliers (1 = off): diffusion.f90:846-854 */
    int bottom_perf_cond, top_perf_cond;   /* A_theta, A_phi ghosts 0 at that r wall in the SLD
                                              fluxes (diffusion.f90:628-643, 691-706) */
    /* explicit diffusion (rhs_mpi.f90:844-1474, :2070-2366).  laplacian: viscosity mur, thermal
     * diffusion ks and resistivity etar; slaplacian: the thermal diffusion alone.  Either one
     * needs the profiles mur..d2xdr2; both combine with sld_on. */
    int laplacian, slaplacian;
    double lumr1, lumr2;         /* luminosities of constant-flux entropy walls under laplacian /
                                    slaplacian: ds/dr = -lumr/(4 pi r^2 ks) (rhs_mpi.f90:586-588,
                                    :618-620) */
} cssx_mhd_params;

/* Radial / latitudinal profiles (host arrays, copied at creation).  Lengths: nr or nth. */
typedef struct cssx_mhd_profiles {
    const double *r;             /* r_ind (nr) */
    const double *dr;            /* dr_ind (nr): cell widths used by diffusion and CFL */
    const double *dxdr;          /* dx/dr (nr); all ones on a uniform grid */
    const double *theta;         /* nth colatitudes */
    const double *gravity;       /* g(r) (nr) */
    const double *eps;           /* heating eps(r) (nr) */
    const double *rho_bar;       /* background density rhor (nr) */
    const double *s_bar;         /* background entropy sr (nr) */
    const double *dsdr1, *dsdr2; /* wall entropy gradients (nth) for constant-flux walls */
    const double *sr1, *sr2;     /* wall entropies (nth) for constant-temperature walls */
    const double *diffusion_speed;/* SLD speed^2 per radius, nr+2 entries (0..nr+1) */
    /* explicit-diffusion profiles (nr), read when laplacian or slaplacian: viscosity mur and
     * d ln mur/dr, thermal diffusivity ks and d ln ks/dr, resistivity etar
     * (mhdsph_mpi.f90:350-354), and d2x/dr2 of the radial map (rhs_mpi.f90:1277-1283; zeros on a
     * uniform grid) */
    const double *mur, *dlnmu, *ks, *dlnks, *etar, *d2xdr2;
} cssx_mhd_profiles;

typedef struct cssx_mhd cssx_mhd;

/* Builds the device-side constants for one subdomain (Initialize_Physics, rhs_mpi.f90:18-73;
 * Initialize_Step_RK4, step_rk4.f90:19-22).  Allocates its own scratch for both precisions. */
int cssx_mhd_create(const cssx_grid *grid, const cssx_mhd_params *params, const cssx_mhd_profiles *prof,
                    cssx_mhd **out);
void cssx_mhd_destroy(cssx_mhd *m);

/* RHS(call_num) (rhs_mpi.f90:279-783) on a state vars1 (nv fields, A in slots 6-8).
 *   rhs   dvars1 (nv fields): d(rho,w,v,u,s)/dt and dA/dt
 *   b_out optional (3 fields): B = curl A with the wall conditions of Compute_B
 *         (derivatives.f90:1830-1892), what the reference leaves in vars1(6:8) for call 1.
 * prec selects the arithmetic (CSSX_PREC_F64: state/rhs double; CSSX_PREC_F32: float).
 * FP32 states are stored in fluctuation form: slot 1 holds rho - rho_bar(r) and slot 5
 * s - s_bar(r) (the profiles given at creation); every other slot, and every rhs slot, is
 * the plain value.  This keeps FP32 gradients of the strongly stratified background exact. */
int cssx_rhs(cssx_mhd *m, int prec, const void *state, void *rhs, void *b_out, void *stream);

/* Bound (rhs_mpi.f90:2395-2491) applied in place to a state (FP32: fluctuation form). */
int cssx_bound(cssx_mhd *m, int prec, void *state, void *stream);

/* nsteps RK4 steps of Step_RK4 (step_rk4.f90:339-493) on the FP64 state y0 (nv fields, updated
 * in place).  With sld_on the weights are the classical 1/6, 1/3, 1/3, 1/6 plus the Euler SLD
 * tendency; without it k1 carries weight 1 as in the reference (vars_save = dvars1, :387), so
 * sld_on with sld_coef = 0 is the plain classical RK4.  prec = CSSX_PREC_F64 runs everything in FP64;
 * CSSX_PREC_F32 evaluates stages, RHS and diffusion in FP32 while the RK4 accumulation and y0
 * stay FP64. */
int cssx_rk4(cssx_mhd *m, int prec, void *y0, double dt, int nsteps, void *stream);

/* nsteps of Step_RK3_TVD (step_rk4.f90:166-335: y1 = y0 + dt k1, y2 = 3/4 y0 + 1/4 y1 + dt/4 k2,
 * y0 <- 1/3 y0 + 2/3 y2 + 2dt/3 k3 [+ dt D]) and of Step_RK2_TVD (:31-162: y1 = y0 + dt k1,
 * y0 <- 1/2 y0 + 1/2 y1 + dt/2 k2 [+ dt D]), Bound after every stage; D = the SLD tendency of
 * y0 when SLDiffusion.  prec as for cssx_rk4 (FP32 stages keep FP64 copies of the stage states
 * for the combinations). */
int cssx_rk3_tvd(cssx_mhd *m, int prec, void *y0, double dt, int nsteps, void *stream);
int cssx_rk2_tvd(cssx_mhd *m, int prec, void *y0, double dt, int nsteps, void *stream);

/* Compute_Diffusion(1) (diffusion.f90:70-474 with Diffusion_R/Th/Phi :476-1495): the SLD
 * tendency dif (nv fields) of a state (A in slots 6-8; FP32 in fluctuation form): Van Albada
 * limited second-order fluxes per variable and direction at speed sld_coef*sqrt(vcc), zero
 * flux through the walls, then viscous heating, mass diffusion and Ohmic heating.
 *   b  optional B = curl A (3 fields) as RHS(1) leaves it in vars1(6:8); NULL computes it.
 * Unlike the reference, the state is not modified (it subtracts the background in place and
 * overwrites the shared temperature array, diffusion.f90:122, :168-175). */
int cssx_diffusion(cssx_mhd *m, int prec, const void *state, const void *b, void *dif, void *stream);

/* Compute_Step_Size (rhs_mpi.f90:92-275): dt = cfl_safety * min over points of the sound,
 * flow and Alfven crossing times, and with laplacian the diffusive limit
 * min(dr, r1 min(min(sin th1, sin thN) dphi, dth))^2 / max(ks, mur/rho_bar, etar) (:194-198).
 * Synchronizes the stream; *dt_out on the host. */
int cssx_cfl(cssx_mhd *m, const void *y0_f64, double cfl_safety, double *dt_out, void *stream);

/* The conservation totals Compute_Step_Size prints (rhs_mpi.f90:156-216, the scalar.dat
 * quantities): out4 = {E_tot, L_tot, m_tot, mrsint} of an FP64 state, E including the magnetic
 * energy |B|^2/(8 pi).  Synchronizes the stream. */
int cssx_conservation(cssx_mhd *m, const void *y0_f64, double *out4, void *stream);

/* Per-phase GPU timing, the counterpart of the reference's Do_Timings buckets (constants.f90:82,
 * timings(1:16) around each phase of rhs_mpi.f90:324-766 and step_rk4.f90:343-491).  While
 * enabled, every kernel launch of the library is bracketed by hipEvents on its stream.
 * cssx_timing_read synchronizes the device and returns, per phase, the summed kernel time
 * (ms), the number of launches and the algorithmic bytes those launches move (one read of
 * every input value and one write of every output value), then resets the table.  Any of the
 * three arrays may be NULL; each has CSSX_NPHASE entries. */
#define CSSX_PHASE_DERIV_R 0
#define CSSX_PHASE_DERIV_THETA 1
#define CSSX_PHASE_DERIV_PHI 2
#define CSSX_PHASE_POINTWISE 3     /* curl assembly and other pointwise kernels */
#define CSSX_PHASE_RK_UPDATE 4     /* stage update + Bound */
#define CSSX_PHASE_OTHER 5
#define CSSX_PHASE_SLD 6           /* SLD diffusion: speed, fluxes, corrections (not its curl) */
#define CSSX_PHASE_RHS_MAIN 7      /* unused: the momentum / entropy / continuity terms are evaluated
                                      inside the RHS_FINAL kernel */
#define CSSX_PHASE_RHS_FINAL 8     /* every pointwise RHS term, the density terms, u x B, and the fused
                                      RK4 stage update */
#define CSSX_NPHASE 9
int cssx_timing_enable(int on);
int cssx_timing_read(double *ms, long *count, double *bytes);

/* The precomputed solve tables for one pencil system, in FP64 on the host (no device work).
 * The counterpart of the module-level Thomas tables gam1/gam2/gam3 that Initialize_Derivatives
 * builds (derivatives.f90:106-518); here they are the segment-local LU factors, interface
 * response vectors and reduced-system factors of the segmented solve (cssx_plan.hpp layout).
 * lo_kind/hi_kind: 0 wall, 1 halo.  variant 0: strided-axis plan (waves = segments, plan_out
 * holds 5*n + 80 doubles); variant 1: contiguous-axis plan (lanes = segments, 5*n + 960
 * doubles).  *nseg receives the number of segments. */
int cssx_plan_tables(int n, int order, int lo_kind, int hi_kind, int bdry, int variant, double *plan_out,
                     int *nseg);

/* ---- batched patches ---------------------------------------------------------------------------
 * Many independent subdomains ("patches") of uneven shape f(nth, nphi, nr), each with walls on
 * all six faces (the wall closures of derivatives.f90 with boundary code bdry on every axis,
 * prescribed wall derivatives zero), differentiated along r, theta and phi in ONE persistent
 * launch per cssx_patchset_run.  Pencil lengths 8..512.  Outputs that are NULL are skipped.
 * Results are in index units times dri/dti/dpi, as cssx_deriv's. */
typedef struct cssx_patch {
    int nr, nth, nphi;
    const void *in;
    void *d_r, *d_th, *d_ph;
    int bdry;
    double dri, dti, dpi;
} cssx_patch;
typedef struct cssx_patchset cssx_patchset;
/* builds every plan and the work-item tables of the four kernel variants (strided segment
 * lengths, contiguous run length, workgroup size) on the current device; the patch pointers
 * are fixed from here on.  CSSX_EINVAL also when a strided axis's (n + 20) rows exceed 2^31
 * bytes (the kernels' 32-bit buffer offsets) */
int cssx_patchset_create(const cssx_patch *patches, int npatch, int prec, cssx_patchset **out);
void cssx_patchset_destroy(cssx_patchset *ps);
/* one persistent launch over every item; variant -1 = the autotuned (or default) one */
int cssx_patchset_run(cssx_patchset *ps, int variant, void *stream);
/* times each variant (reps runs), keeps the fastest; synchronizes.  ms_out (4 floats) may be NULL */
int cssx_patchset_autotune(cssx_patchset *ps, int reps, void *stream, int *best, float *ms_out);
/* n seeded N(0,1) values (SplitMix64 + Box-Muller over the flat index; no reference
 * counterpart): the synthetic initial perturbations generated in place on the device */
int cssx_normal_fill(void *out, int prec, long n, unsigned long long seed, void *stream);
/* seeded N(0,1) field (SplitMix64 + Box-Muller over the flat index) smoothed by [1,2,1]/4 along
 * each axis (edge values repeated); tmp: scratch of the same size */
int cssx_patch_fill(void *out, void *tmp, int prec, int nr, int nth, int nphi, unsigned long long seed,
                    void *stream);

/* ---- learned subgrid closure ------------------------------------------------------------------
 * No counterpart in the reference: a coarse-grid alternative to (or correction on top of) the
 * SLD subgrid diffusion of diffusion.f90, for callers that run a coarse grid in place of a
 * fine one.  A 216 -> h1 -> h2 -> 8 GELU network maps the 3x3x3 stencil of the 8 fields
 * (rho and s as fluctuations about the hydrostatic background) to the 8-field tendency that the
 * coarse RHS misses.  Matrices run on the BF16 matrix cores with FP32 accumulation; the master
 * weights and Adam moments are FP32.  Parameters, flattened (host order):
 *   W1 (216 x h1, row-major), b1 (h1), W2 (h1 x h2), b2 (h2), W3 (h2 x 8), b3 (8)
 * Features of point p (row p): column q = field*27 + (dr+1)*9 + (dphi+1)*3 + (dth+1), the
 * neighbour clamped at r and theta walls and wrapped in phi, standardized as (x - mean)/std.
 * Network inputs are BF16 row-major (rows x 216); outputs/targets FP32 row-major (rows x 8),
 * standardized; the closure tendency is output*y_std + y_mean (y_std = y_mean = 0 switches a
 * field's correction off, e.g. where a validation snapshot shows no gain). */
typedef struct cssx_mlp cssx_mlp;
int cssx_mlp_create(int nin, int h1, int h2, int nout, int max_rows, unsigned long long seed, cssx_mlp **out);
void cssx_mlp_destroy(cssx_mlp *net);
long cssx_mlp_num_params(const cssx_mlp *net);
int cssx_mlp_set_params(cssx_mlp *net, const float *host_params);
/* either pointer may be NULL; grads are those of the last cssx_mlp_grad / _train_step */
int cssx_mlp_get_params(cssx_mlp *net, float *host_params, float *host_grads);
int cssx_mlp_set_norm(cssx_mlp *net, const float *x_mean, const float *x_std, const float *y_mean,
                      const float *y_std);
int cssx_mlp_forward(cssx_mlp *net, const void *x_bf16, int rows, float *y, void *stream);
/* gradient of the mean squared error over rows x 8 outputs; *loss (device float) may be NULL */
int cssx_mlp_grad(cssx_mlp *net, const void *x_bf16, const float *target, int rows, float *loss, void *stream);
/* cssx_mlp_grad + one Adam step (beta 0.9/0.999, eps 1e-8, bias-corrected) */
int cssx_mlp_train_step(cssx_mlp *net, const void *x_bf16, const float *target, int rows, float lr, float *loss,
                        void *stream);
/* C (M x N, FP32, row-major, ldc) = A B with A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn],
 * BF16 inputs; ksplit > 1 splits K over workgroups (requires ldc == N) */
int cssx_gemm_bf16(const void *A, long sam, long sak, const void *B, long sbk, long sbn, int M, int N, int K,
                   int ksplit, float *C, long ldc, void *stream);
/* mean over factor^3 blocks: (nv, nr, nphi, nth) -> FP32 (nv, nr/f, nphi/f, nth/f) */
int cssx_box_filter(const void *in, int prec, int nv, int nr, int nphi, int nth, int factor, float *out,
                    void *stream);
/* features of rows [row0, row0+rows) of a state (8 fields; fluct = 1 when slots 0 and 4 already
 * hold fluctuations, else rho_bar/s_bar (nr, same precision) are subtracted): standardized BF16
 * into x_bf16 (needs net's norm) and/or raw FP32 into x_raw */
int cssx_closure_features(cssx_mlp *net, const void *state, int prec, int fluct, const void *rho_bar,
                          const void *s_bar, int nr, int nphi, int nth, long row0, long rows, void *x_bf16,
                          float *x_raw, void *stream);
/* residual targets: raw[p][iv] = fine_filtered[iv][p] - coarse_rhs[iv][p] (row-major FP32) and
 * target = (raw - y_mean)/y_std; either output may be NULL */
int cssx_closure_targets(const float *fine_filtered, const void *coarse_rhs, int prec, long npts,
                         const float *y_mean, const float *y_std, float *target, float *raw, void *stream);
/* column mean and standard deviation (population) of a row-major FP32 matrix; synchronizes */
int cssx_colstats(const float *X, long rows, int cols, double *mean, double *stddev, void *stream);
/* per-field mean and root mean square of a field-major state (rho and s about the background
 * rho_bar/s_bar (nr) unless fluct); synchronizes */
int cssx_field_stats(const void *state, int prec, int nv, int nr, int nphi, int nth, int fluct, const void *rho_bar,
                     const void *s_bar, double *mean, double *rms, void *stream);
/* (x - mean[c]) / stddev[c] of a row-major FP32 (rows x cols) matrix (mean, stddev: device FP32)
 * into BF16 and/or FP32 outputs (either may be NULL) */
int cssx_standardize_rows(const float *x, long rows, int cols, const float *mean, const float *stddev,
                          void *out_bf16, float *out_f32, void *stream);
/* dst row i = src row idx[i] (idx: device int64), rows of row_bytes bytes (a multiple of 4) */
int cssx_gather_rows(const void *src, long row_bytes, const long *idx, long n, void *dst, void *stream);
/* the closure tendency (8 fields, FP32, field-major like the state) of a whole 8-field state;
 * h1 = h2 = 128 networks run one fused kernel (features, three layers and the scaling with the
 * activations held in registers and the weights in LDS), other widths layer by layer */
int cssx_closure_tendency(cssx_mlp *net, const void *state, int prec, int fluct, const void *rho_bar,
                          const void *s_bar, int nr, int nphi, int nth, float *ext, void *stream);
/* attach (net != NULL) or detach a closure: every later RHS evaluation of m (cssx_rhs and the
 * steppers' stages) adds the network's tendency of its input state.  net is not owned.  Each
 * tendency writes the network's packed weights and inference scratch, so one network must not
 * serve solvers whose RHS evaluations can overlap on different streams (each solver runs its
 * closure on its own side stream). */
int cssx_mhd_set_closure(cssx_mhd *m, cssx_mlp *net);

#ifdef __cplusplus
}
#endif
#endif /* CSSX_H */
