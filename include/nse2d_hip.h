/*
 * nse2d_hip.h - C-ABI of the MI355X (gfx950) implementation of the per-step hot path of the
 * 2-D vorticity-streamfunction channel solver xupingxie/nse2d-channel-gpu.
 *
 * Conventions (the reference's, src/test_api.cu:1-92, src/ops.cuh, src/diag.cuh):
 *   - every array argument is a DEVICE pointer; fields are FP64 row-major [Ny][Nx], x fastest
 *     (common.cuh:127-130); metric vectors have Ny (a_node, y_node, w_node) or Ny-1 entries
 *     (a_edge, dy_edge);
 *   - all buffers are caller-owned, including reduction output scalars; the launchers write
 *     their outputs completely (no memset needed by the caller);
 *   - everything is asynchronous on the given stream; results are visible after a stream sync;
 *   - launch and runtime failures abort the process with a message, like CUDA_CHECK
 *     (common.cuh:116-119); argument errors that the host can see (unsupported sizes, missing
 *     setup) are reported through a return value and nse2d_last_error();
 *   - the reduction launchers use a library-owned device scratch and are not re-entrant across
 *     host threads or concurrent streams of one device, like the reference's function-static
 *     scratch (diag.cu:600-606, 843-855).
 * Citations are file:line in the reference repository.
 */
#ifndef NSE2D_HIP_H
#define NSE2D_HIP_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------
 * 1. The seven C-linkage test wrappers, signatures kept (cudaStream_t -> hipStream_t).
 * ------------------------------------------------------------------------------------------ */
/* test_api.cu:8-18 */
void testapi_uv_from_psi_mapped(const double* d_psi, double* d_u, double* d_v, int Nx, int Ny,
                                double inv_dx, double inv_deta, const double* d_a_node,
                                hipStream_t s);
/* test_api.cu:20-30 */
void testapi_omega_from_psi_interior_mapped(double* d_w, const double* d_psi, int Nx, int Ny,
                                            double inv_dx2, double inv_deta2,
                                            const double* d_a_node, const double* d_a_edge,
                                            hipStream_t s);
/* test_api.cu:32-42 */
void testapi_thom_wall_vorticity_mapped(double* d_w, const double* d_psi, int Nx, int Ny,
                                        const double* d_dy_edge, double psi_bot,
                                        double psi_top, hipStream_t s);
/* test_api.cu:44-53 */
void testapi_arakawa_J_mapped(const double* d_psi, const double* d_w, double* d_J, int Nx,
                              int Ny, double inv_dx, double inv_deta, const double* d_a_node,
                              hipStream_t s);
/* test_api.cu:55-65 */
void testapi_compute_adv_from_uv_mapped(const double* d_u, const double* d_v,
                                        const double* d_omg, double* d_Jref, int Nx, int Ny,
                                        double inv_dx, double inv_deta, const double* d_a_node,
                                        hipStream_t s);
/* test_api.cu:67-77 */
void testapi_compute_adv_flux_mapped(const double* d_u, const double* d_v, const double* d_omg,
                                     double* d_Jref, int Nx, int Ny, double inv_dx,
                                     double inv_deta, const double* d_a_node, hipStream_t s);
/* test_api.cu:79-92; any of J1/J2/J3/Jadv may be NULL */
void testapi_arakawa_J_mapped_split(const double* d_psi, const double* d_omg, double* d_J1,
                                    double* d_J2, double* d_J3, double* d_Jadv, int Nx, int Ny,
                                    double inv_dx, double inv_deta, const double* d_a_node,
                                    hipStream_t s);

/* ------------------------------------------------------------------------------------------
 * 2. Operator launchers (ops.cuh:12-81), same names and parameter order, C linkage.
 * ------------------------------------------------------------------------------------------ */
/* ops.cuh:12-13, ops.cu:64-70 */
void launch_uv_from_psi_mapped(const double* psi, double* u, double* v, int Nx, int Ny,
                               double inv_dx, double inv_deta, const double* d_a_node,
                               hipStream_t s);
/* ops.cuh:18-20, ops.cu:94-101; writes interior rows only */
void launch_omega_from_psi_interior_mapped(double* w, const double* psi, int Nx, int Ny,
                                           double inv_dx2, double inv_deta2,
                                           const double* d_a_node, const double* d_a_edge,
                                           hipStream_t s);
/* ops.cuh:23-24, ops.cu:270-276; wall type from nse2d_set_wall_type() */
void launch_thom_wall_vorticity_mapped(double* w, const double* psi, int Nx, int Ny,
                                       const double* d_dy_edge, double psi_bot, double psi_top,
                                       hipStream_t s);
/* ops.cuh:31-32, ops.cu:328-334 */
void launch_arakawa_J_mapped(const double* psi, const double* omg, double* J, int Nx, int Ny,
                             double inv_dx, double inv_deta, const double* d_a_node,
                             hipStream_t s);
/* ops.cuh:56-60, ops.cu:501-512 */
void launch_laplacian_fd_mapped(const double* d_q, double* d_out, int Nx, int Ny,
                                double inv_dx2, double inv_deta2, const double* d_a_node,
                                const double* d_a_edge, hipStream_t s);
/* ops.cuh:66-69, ops.cu:536-544 */
void launch_rhs_mapped(const double* J, const double* lap, const double* omg, double* rhs,
                       int Nx, int Ny, double nu, double F0, int nforce, double h,
                       double lin_drag, const double* d_y_node, hipStream_t s);
/* ops.cuh:70-75, ops.cu:628-641 */
void launch_mms_add_forcing(double* d_rhs, int Nx, int Ny, double Lx, double h, double t,
                            double nu, double r, double A0, double eps, double Om, int kx_mode,
                            const double* d_y_node, hipStream_t s);
/* ops.cuh:79-81, ops.cu:656-669; launch_ssprk3 allows wnp1 == w (main.cu:530) */
void launch_axpy(double* y, const double* x, const double* r, double dt, size_t N,
                 hipStream_t s);
void launch_ssprk2(double* w2, const double* w, const double* w1, const double* rhs, double dt,
                   size_t N, hipStream_t s);
void launch_ssprk3(double* wnp1, const double* w, const double* w2, const double* rhs,
                   double dt, size_t N, hipStream_t s);

/* Wall closure selection (no reference counterpart: the reference chooses at compile time,
 * ops.cu:108-111 / Makefile WALLS=freeslip). 0 = no-slip Thom (default), 1 = free-slip. */
void nse2d_set_wall_type(int free_slip);
int nse2d_get_wall_type(void);

/* ------------------------------------------------------------------------------------------
 * 3. Poisson solver: x-FFT (hand-written LDS FFT, real data, half spectrum) + y-Thomas.
 *    Replaces struct Poisson2D_Z2Z (poisson.cuh:18-119) and its cuFFT Z2Z plans.
 *    Nx must be a power of two, 4 <= Nx <= 16384; Ny >= 4.
 * ------------------------------------------------------------------------------------------ */
typedef struct nse2d_poisson* nse2d_poisson_t;
/* poisson.cuh:34-55 (constructor); NULL on unsupported sizes (see nse2d_last_error) */
nse2d_poisson_t poisson2d_create(int Nx, int Ny, double Lx, hipStream_t s);
/* poisson.cuh:71-80 (enable_mapped); borrows a_node/a_edge; returns 0 on success */
int poisson2d_enable_mapped(nse2d_poisson_t p, const double* d_a_node, const double* d_a_edge,
                            double inv_deta, hipStream_t s);
/* poisson.cuh:83-118 (solve): -lap psi = omega, psi = psi_bot/psi_top on the walls.
 * dy is unused (API compatibility, as in the reference). d_omega may not alias d_psi. */
void poisson2d_solve(nse2d_poisson_t p, const double* d_omega, double* d_psi, double dy,
                     double psi_bot, double psi_top, hipStream_t s);
/* poisson.cuh:65-67 (set_profile / profile_reset / profile_get): event timing per phase */
void poisson2d_set_profile(nse2d_poisson_t p, int on);
void poisson2d_profile_reset(nse2d_poisson_t p);
void poisson2d_profile_get(nse2d_poisson_t p, float* fft_ms, float* tridiag_ms);
/* poisson.cuh:56-63 (destructor) */
void poisson2d_destroy(nse2d_poisson_t p);

/* x-spectra of u and v, time-averaged over the calls to accumulate: replaces struct Spectra1D
 * (diag.cuh:123-191; kernels diag.cu:869-898) and its cuFFT D2Z plan. Per call, for
 * k = 0..Nx/2: E(k) += wx / (Nx^2 Ny) sum_j |X_j(k)|^2 with wx = 2 for 0 < k < Nx/2, else 1, so
 * that sum_k E_u(k) is the mean of u^2 over the grid points. Same size limits as the Poisson
 * plan. */
typedef struct nse2d_spectra1d* nse2d_spectra1d_t;
/* diag.cuh:133-154 (init) */
nse2d_spectra1d_t nse2d_spectra1d_create(int Nx, int Ny, double Lx, hipStream_t s);
/* diag.cuh:165-174 (accumulate): d_u, d_v are [Ny][Nx] device fields */
void nse2d_spectra1d_accumulate(nse2d_spectra1d_t sp, const double* d_u, const double* d_v,
                                hipStream_t s);
/* diag.cuh:175-190 (write_csv's data): host arrays of Nx/2+1 (any may be NULL): kx = 2 pi k / Lx
 * and the averaged E_u, E_v. Synchronises s; returns the number of samples. */
int nse2d_spectra1d_read(nse2d_spectra1d_t sp, double* kx, double* Eu, double* Ev, hipStream_t s);
/* diag.cuh:156-164 (destructor) */
void nse2d_spectra1d_destroy(nse2d_spectra1d_t sp);

/* ------------------------------------------------------------------------------------------
 * 4. Diagnostics reductions (diag.cuh:12-113), same names and parameter order. Outputs are
 *    the raw sums (the host applies 1/A, nu, 1/2 as main.cu:627-675 does).
 * ------------------------------------------------------------------------------------------ */
/* diag.cuh:12-14, diag.cu:6-38 (interior rows) */
void launch_enstrophy_diss_mapped(const double* omg, int Nx, int Ny, double inv_dx,
                                  double inv_deta, const double* d_a_node,
                                  const double* d_w_node, double* d_sum_w2, double* d_sum_gw2,
                                  hipStream_t s);
/* diag.cuh:20-22, diag.cu:43-101 (all rows, one-sided eta derivative at the walls) */
void launch_enstrophy_diss_mapped_full(const double* omg, int Nx, int Ny, double inv_dx,
                                       double inv_deta, const double* d_a_node,
                                       const double* d_w_node, double* d_sum_w2_full,
                                       double* d_sum_gw2_full, hipStream_t s);
/* diag.cuh:26-27, diag.cu:158-199 */
void launch_enstrophy_sbp_mapped(const double* omg, int Nx, int Ny, double inv_dx,
                                 double inv_deta, const double* d_a_edge,
                                 const double* d_wt_node, double* d_sum_eta_h,
                                 double* d_sum_phi_h, hipStream_t s);
/* diag.cuh:31-32, diag.cu:201-231 */
void launch_kinetic_energy_mapped(const double* psi, int Nx, int Ny, double inv_dx,
                                  double inv_deta, const double* d_a_node,
                                  const double* d_w_node, double* d_sum, hipStream_t s);
/* diag.cuh:33-40, diag.cu:250-348 */
void launch_energy_dissipation_mapped(const double* d_psi, const double* d_omega, int Nx,
                                      int Ny, double inv_dx, double inv_deta,
                                      const double* d_a_node, const double* d_a_edge,
                                      const double* d_w_node, double* d_sum_eps,
                                      hipStream_t s);
/* diag.cuh:45-47, diag.cu:350-376 */
void launch_power_input_mapped(const double* psi, int Nx, int Ny, double inv_deta, double h,
                               double F0, int nforce, const double* d_a_node,
                               const double* d_y_node, const double* d_w_node, double* d_sum,
                               hipStream_t s);
/* diag.cuh:54-57, diag.cu:411-422 */
void launch_enstrophy_input_mapped_full(const double* omega, int Nx, int Ny, double h,
                                        double F0, int nforce, const double* d_y_node,
                                        const double* d_w_node, double* d_sum, hipStream_t s);
/* diag.cuh:64-67, diag.cu:456-467 (unused by the reference driver; kept for completeness) */
void launch_enstrophy_input_mapped(const double* omega, int Nx, int Ny, double h, double F0,
                                   int nforce, const double* d_y_node, const double* d_w_node,
                                   double* d_sum, hipStream_t s);
/* diag.cuh:72-73, diag.cu:493-499 (per-row sums, Ny entries each) */
void launch_row_moments_from_psi_mapped(const double* psi, int Nx, int Ny, double inv_dx,
                                        double inv_deta, const double* d_a_node, double* su,
                                        double* sv, double* su2, double* sv2, double* suv,
                                        hipStream_t s);
/* diag.cuh:75, diag.cu:508-512 (means over x of the two wall rows) */
void launch_mean_wall_omega(const double* omega, int Nx, int Ny, double* d_mean_bot,
                            double* d_mean_top, hipStream_t s);
/* diag.cuh:83-88, diag.cu:571-583 */
void launch_mean_wall_omega_omegay_mapped(const double* omega, int Nx, int Ny,
                                          double inv_deta, const double* d_a_node,
                                          double* d_mean_bot, double* d_mean_top,
                                          hipStream_t s);
/* diag.cuh:89-98, diag.cu:586-625 (unused by the reference driver; kept for completeness) */
void launch_enstrophy_diss_balance_mapped(const double* d_omega, int Nx, int Ny,
                                          double inv_dx2, double inv_deta2,
                                          const double* d_a_node, const double* d_a_edge,
                                          const double* d_w_node, double* d_sum_w2,
                                          double* d_sum_eta, hipStream_t s);
/* diag.cuh:99-105, diag.cu:672-691 */
void launch_omega_rhs_balance_mapped(const double* d_omega, const double* d_rhs, int Nx,
                                     int Ny, const double* d_w_node, double* d_sum_omega_rhs,
                                     hipStream_t stream);
/* diag.cuh:111-113, diag.cu:821-866 (max|u|, max|v|, max|v|*a over interior rows) */
void launch_max_uv_from_psi_mapped(const double* d_psi, int Nx, int Ny, double inv_dx,
                                   double inv_deta, const double* d_a_node, double* d_umax,
                                   double* d_vmax, double* d_vmaxm, hipStream_t stream);
/* One diagnostics row (main.cu:565-692 without --closure_diag; no single reference counterpart):
 * the reductions of launch_max_uv_from_psi_mapped, launch_kinetic_energy_mapped,
 * launch_enstrophy_diss_mapped_full, launch_omega_rhs_balance_mapped, launch_power_input_mapped,
 * launch_energy_dissipation_mapped, launch_enstrophy_input_mapped_full,
 * launch_mean_wall_omega_omegay_mapped and launch_mean_wall_omega, and optionally
 * launch_row_moments_from_psi_mapped (d_moments: [5][Ny] row sums of u, v, u^2, v^2, uv, or
 * NULL), in one pass over psi, omega and rhs (two launches). d_out receives NSE2D_DIAG_NOUT
 * doubles, each equal to the corresponding launcher's output bit for bit. 0, or -1 for Nx < 2,
 * Ny < 4 or a null field or output pointer. */
enum { NSE2D_DIAG_UMAX = 0, NSE2D_DIAG_VMAX = 1, NSE2D_DIAG_VMAXM = 2, NSE2D_DIAG_K = 3,
       NSE2D_DIAG_OMEGA_W2 = 4, NSE2D_DIAG_OMEGA_GRAD2 = 5, NSE2D_DIAG_OMEGA_RHS = 6,
       NSE2D_DIAG_PIN = 7, NSE2D_DIAG_EPS = 8, NSE2D_DIAG_POMEGA = 9,
       NSE2D_DIAG_WALL_OMEGA_OMEGAY_BOT = 10, NSE2D_DIAG_WALL_OMEGA_OMEGAY_TOP = 11,
       NSE2D_DIAG_WALL_OMEGA_BOT = 12, NSE2D_DIAG_WALL_OMEGA_TOP = 13, NSE2D_DIAG_NOUT = 14 };
int nse2d_diag_row(const double* d_psi, const double* d_omega, const double* d_rhs, int Nx, int Ny,
                   double inv_dx, double inv_deta, double h, double F0, int nforce,
                   const double* d_a_node, const double* d_a_edge, const double* d_y_node,
                   const double* d_w_node, double* d_out, double* d_moments, hipStream_t stream);
/* No counterpart in the reference. launch_max_uv_from_psi_mapped for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. Same kernels and results; the row partials live in a fixed device array of
 * 3 x NSE2D_NOALLOC_MAX_ROWS doubles (393 KB) that each GPU holds until the process ends,
 * instead of the library's grow-on-demand scratch. Returns 0, or -1 with nse2d_last_error() when
 * Ny > NSE2D_NOALLOC_MAX_ROWS (nothing is launched then). All calls on one GPU share that array:
 * calls that overlap on one GPU (concurrent streams or host threads) race. */
#define NSE2D_NOALLOC_MAX_ROWS 16385
int launch_max_uv_from_psi_mapped_noalloc(const double* d_psi, int Nx, int Ny, double inv_dx,
                                          double inv_deta, const double* d_a_node, double* d_umax,
                                          double* d_vmax, double* d_vmaxm, hipStream_t stream);

/* ------------------------------------------------------------------------------------------
 * 5. Whole-step entry points: the SSPRK(3,3) step of main.cu:482-538 with each stage's
 *    stencil work (Jacobian, Laplacian, RHS, forcing, update) fused into one kernel, then the
 *    Poisson solve and the Thom closure. No reference counterpart as one call; the sequence
 *    and the arithmetic are main.cu's.
 * ------------------------------------------------------------------------------------------ */
typedef struct {
    int Nx, Ny;
    double Lx, h;
    double nu, F0, lin_drag;
    int nforce;
    double psi_bot, psi_top;        /* main.cu:211: 0 and 2 h Ub */
    int mms;                        /* add the manufactured forcing (ops.cu:559-626) */
    double mms_A0, mms_eps, mms_Om;
    int mms_kx;
    /* adaptive-dt rule of main.cu:451-458, used by nse2d_channel_step_auto */
    double cfl, cvisc, dt_max, dy_min;
    /* borrowed device metric arrays */
    const double *d_a_node, *d_a_edge, *d_y_node, *d_dy_edge, *d_w_node;
    /* wall closure of this channel: 1 = no-slip Thom closure (ops.cu:112-156), 2 = free-slip
     * (omega = 0 on the walls, the FREE_SLIP_WALLS build of ops.cu:108-111), 0 = the library-wide
     * default of nse2d_set_wall_type, read at each step (an ensemble reads it once, at creation) */
    int wall_type;
} nse2d_channel_desc;

typedef struct nse2d_channel* nse2d_channel_t;

enum { NSE2D_FIELD_OMEGA = 0, NSE2D_FIELD_PSI = 1, NSE2D_FIELD_RHS = 2, NSE2D_FIELD_W1 = 3,
       NSE2D_FIELD_W2 = 4 };
enum { NSE2D_STEP_STORE_RHS = 1,     /* keep the stage-3 RHS in the RHS field (diag column) */
       NSE2D_STEP_WALL_BOOKKEEPING = 2 /* accumulate the enstrophy jump of each Thom update */ };

/* Allocates the state (omega, omega1, omega2, psi, rhs) and the Poisson plan. NULL on error. */
nse2d_channel_t nse2d_channel_create(const nse2d_channel_desc* desc, hipStream_t s);
void nse2d_channel_destroy(nse2d_channel_t ch);
/* Device pointer of a state field ([Ny][Nx]); the caller may read or overwrite it. */
double* nse2d_channel_field(nse2d_channel_t ch, int which);
/* Initial solve + closure (main.cu:211-213): psi from omega, then omega's wall rows. */
void nse2d_channel_init(nse2d_channel_t ch);
/* One SSPRK(3,3) step with host-chosen dt (main.cu:482-538). flags: NSE2D_STEP_*. */
void nse2d_channel_step(nse2d_channel_t ch, double t, double dt, int flags);
/* n steps with dt chosen on the device by the rule of main.cu:440-458 (no host sync; the
 * current dt and t live in device memory). t advances from its device value. */
void nse2d_channel_step_auto(nse2d_channel_t ch, int n);
/* The stream the channel was created with (all of its work runs on it). */
hipStream_t nse2d_channel_stream(nse2d_channel_t ch);
/* Device-resident time and dt (read/write through these device pointers). */
double* nse2d_channel_time_ptr(nse2d_channel_t ch);
double* nse2d_channel_dt_ptr(nse2d_channel_t ch);
/* Cumulative enstrophy added by the Thom update of each stage (main.cu:271-297), 3 values. */
void nse2d_channel_wall_bookkeeping(nse2d_channel_t ch, double out3[3], double invA_full);
/* Kernel-class timing with HIP events recorded on the channel's stream around the launches of
 * the classes in class_mask (bit k = class k); resets the accumulators. Replaces the reference's
 * --profile event timing (main.cu:303,436,540; poisson.cuh:88-117). */
enum { NSE2D_PROF_STAGE = 0, NSE2D_PROF_FFT_FWD = 1, NSE2D_PROF_THOMAS = 2, NSE2D_PROF_FFT_INV = 3,
       NSE2D_PROF_THOM = 4, NSE2D_PROF_CFL = 5, NSE2D_PROF_NCLASS = 6 };
void nse2d_channel_profile(nse2d_channel_t ch, int class_mask);
/* Time only the first of every `every` steps (default 1: every step). The events themselves
 * cost time (at 4096x2049, six per step add about 16 us per step), so a sample keeps
 * the measurement from slowing the run it measures. No reference counterpart. */
void nse2d_channel_profile_every(nse2d_channel_t ch, int every);
/* Synchronises the stream and returns the accumulated milliseconds and launch counts per class. */
void nse2d_channel_profile_read(nse2d_channel_t ch, double* ms, long long* count);
/* Optional additive RHS term (no reference counterpart; used by the learned closure of a coarse
 * channel, section 5d): before each SSPRK stage the channel calls
 * fn(ctx, d_omega_s, d_psi_s, d_out, s), which must fill d_out ([Ny][Nx] fp32, owned by the
 * channel) on stream s from the stage state; the stage adds d_out to the RHS of rows 1..Ny-2.
 * Stages with a closure run the generic stage kernel. fn = NULL detaches it. Returns 0. */
typedef void (*nse2d_closure_fn)(void* ctx, const double* d_omega_s, const double* d_psi_s,
                                 float* d_out, hipStream_t s);
int nse2d_channel_set_closure(nse2d_channel_t ch, nse2d_closure_fn fn, void* ctx);

/* ------------------------------------------------------------------------------------------
 * 5b. Ensembles of independent channels (no reference counterpart: the reference advances one
 *     channel per process). Members may differ in size, Re_b, forcing, initial condition and
 *     dt; each advances with its own device-chosen dt (main.cu:440-458) exactly as
 *     nse2d_channel_step_auto would advance it alone under the same kernel shapes. Members that
 *     share Nx, Ny, Lx, metric arrays and wall values form a size class and share one Poisson
 *     plan. Requirements: Nx a power of two in [512, 16384], no manufactured forcing. One step is
 *     one grouped CFL launch over all members, then per size class one grouped launch per stage
 *     and per transform and y-solve, captured once into a hipGraph and replayed. All calls run on
 *     the ensemble's stream (a new stream when s is NULL, since the null stream cannot be
 *     captured); with several size classes, the classes after the first run their stages and
 *     solves on streams the ensemble creates (default priority), forked from and joined back to
 *     the ensemble's stream by events within each step, so work queued on the ensemble's stream
 *     after a step sees the whole step.
 * ------------------------------------------------------------------------------------------ */
typedef struct nse2d_ensemble* nse2d_ensemble_t;
/* n member descriptors (the metric arrays are borrowed); NULL on error (nse2d_last_error). */
nse2d_ensemble_t nse2d_ensemble_create(const nse2d_channel_desc* members, int n, hipStream_t s);
void nse2d_ensemble_destroy(nse2d_ensemble_t e);
int nse2d_ensemble_size(nse2d_ensemble_t e);
/* Device pointer of member m's field (NSE2D_FIELD_OMEGA, _PSI, _W1, _W2), NULL otherwise. */
double* nse2d_ensemble_field(nse2d_ensemble_t e, int member, int which);
/* Device array [n][8]: member m's t, dt, max|u|, max|v|, max|v| a at 8m .. 8m+4. */
double* nse2d_ensemble_scalars(nse2d_ensemble_t e);
hipStream_t nse2d_ensemble_stream(nse2d_ensemble_t e);
/* Initial solve and wall closure of every member (main.cu:211-213). */
void nse2d_ensemble_init(nse2d_ensemble_t e);
/* nsteps steps of every member, each with its own dt (replays of the captured step graph). */
void nse2d_ensemble_step_auto(nse2d_ensemble_t e, int nsteps);
/* on = 0: launch the kernels directly instead of replaying the graph. */
void nse2d_ensemble_set_graph(nse2d_ensemble_t e, int on);
void nse2d_ensemble_sync(nse2d_ensemble_t e);
/* Times whole steps for each stage-kernel shape and, per size class, each row-FFT radix and
 * Thomas column width, keeps the fastest, and restores the state it advanced while timing.
 * Writes a one-line summary of the choice into report. Returns 0. */
int nse2d_ensemble_autotune(nse2d_ensemble_t e, int nsteps, char* report, size_t report_len);

/* ------------------------------------------------------------------------------------------
 * 5c. Learned subfilter closure (no reference counterpart): a CNN of six 3x3 convolutions
 *     (4 -> 64 -> 64 -> 64 -> 64 -> 64 -> 1 channels, GELU after the first five) on images
 *     [batch][H][W] with x periodic and zero padding beyond the walls. bf16 weights and
 *     activations, fp32 accumulation (MFMA), fp32 master weights and Adam moments. The flat
 *     parameter array holds, per layer, W[9][cout][cin] (tap t = 3(dy+1)+(dx+1)) then b[cout];
 *     the first layer's cin is padded from 4 to 32 with weights that stay zero.
 * ------------------------------------------------------------------------------------------ */
typedef struct nse2d_cnn* nse2d_cnn_t;
/* W a multiple of 64, H >= 3; He-initialised weights from seed. NULL on error. */
nse2d_cnn_t nse2d_cnn_create(int H, int W, int max_batch, unsigned long long seed, hipStream_t s);
void nse2d_cnn_destroy(nse2d_cnn_t c);
size_t nse2d_cnn_num_params(nse2d_cnn_t c);
/* host <-> device copies of the fp32 parameters / of the last gradients; set resets Adam */
void nse2d_cnn_get_params(nse2d_cnn_t c, float* host);
void nse2d_cnn_set_params(nse2d_cnn_t c, const float* host);
void nse2d_cnn_get_grads(nse2d_cnn_t c, float* host);
/* d_in: [B][H][W][4] fp32 (normalised inputs), 16-byte aligned; d_out: [B][H][W] fp32 (may be
 * NULL). 0 on success; -1 (nse2d_last_error) for a batch outside [1, max_batch] or a misaligned
 * d_in, in this and the other entry points that take d_in. */
int nse2d_cnn_forward(nse2d_cnn_t c, const float* d_in, float* d_out, int B);
/* The forward of nse2d_cnn_forward without keeping the pre-activations (inference: no backward
 * may follow). 0 on success. */
int nse2d_cnn_infer(nse2d_cnn_t c, const float* d_in, float* d_out, int B);
/* Convolution kernel selection, for cross-checks (all compute the same products in the same
 * order): 0 = the persistent LDS-staged kernel (default), 1 = operands straight from global
 * memory (also NSE2D_CNN_DIRECT=1 at creation), 2 = LDS-staged, one workgroup per row tile. */
int nse2d_cnn_set_direct(nse2d_cnn_t c, int on);
/* Gradients of sum(d_gout * out) for the batch of the preceding forward. 0 on success. */
int nse2d_cnn_backward(nse2d_cnn_t c, const float* d_gout, int B);
/* forward, mean squared error over rows 1..H-2 against d_target [B][H][W], backward, Adam
 * (beta 0.9 / 0.999, eps 1e-8) with learning rate lr; returns the loss (synchronises). */
double nse2d_cnn_train_step(nse2d_cnn_t c, const float* d_in, const float* d_target, int B, float lr);
/* One Adam update from the current gradients. */
int nse2d_cnn_adam_step(nse2d_cnn_t c, float lr);
/* forward and the mean squared error of nse2d_cnn_train_step, without an update (synchronises). */
double nse2d_cnn_eval_loss(nse2d_cnn_t c, const float* d_in, const float* d_target, int B);

/* ------------------------------------------------------------------------------------------
 * 5d. Learned closure of a coarse channel (LES) trained on filtered DNS (no reference
 *     counterpart). Filter F, fine -> coarse: box filter over ry fine cells in mapped y
 *     (trapezoid rule, wall rows kept), then a sharp spectral cut |kx| < Nx_c/2 sampled at every
 *     rx-th fine node. Closure target of a fine state: F(R_f) - R_c(omega_c, psi_c) on the
 *     interior rows, where omega_c = F(omega_f) with Thom wall rows from psi_c = solve(omega_c)
 *     and R = -J + nu lap + f(y) - r omega (ops.cu:515-544). The CNN (section 5c) maps the
 *     normalised coarse (omega, psi, u, v) to the normalised target. FP64 everywhere but the CNN.
 *     Every call runs on the stream given at creation; a coarse channel that uses the closure
 *     must have been created on the same stream.
 * ------------------------------------------------------------------------------------------ */
typedef struct nse2d_les* nse2d_les_t;
/* fine, coarse: channel descriptors of the DNS and LES grids (metric arrays borrowed), with
 * Nx_c a multiple of 64, Nx_f = rx Nx_c, Ny_f - 1 = ry (Ny_c - 1), ry 1 or even, same Lx and h.
 * max_pairs: capacity of the device-resident training set; max_batch: CNN batch capacity. */
nse2d_les_t nse2d_les_create(const nse2d_channel_desc* fine, const nse2d_channel_desc* coarse,
                             int max_pairs, int max_batch, unsigned long long seed, hipStream_t s);
void nse2d_les_destroy(nse2d_les_t L);
/* the closure's CNN (parameters, direct forward/backward) */
nse2d_cnn_t nse2d_les_cnn(nse2d_les_t L);
/* d_coarse [Ny_c][Nx_c] = F(d_fine [Ny_f][Nx_f]). Returns 0. */
int nse2d_les_filter(nse2d_les_t L, const double* d_fine, double* d_coarse);
/* From a fine state (omega_f, psi_f): the coarse state (omega_c with Thom wall rows, psi_c) and
 * the closure target (fp64, 0 on the wall rows); any output may be NULL. append != 0 adds the
 * pair (raw inputs, target) to the training set. 0 on success, -1 when the set is full. */
int nse2d_les_pair(nse2d_les_t L, const double* d_omega_f, const double* d_psi_f,
                   double* d_omega_c, double* d_psi_c, double* d_target, int append);
int nse2d_les_num_pairs(nse2d_les_t L);
void nse2d_les_clear_pairs(nse2d_les_t L);
/* Normalisation: mean and standard deviation of each input over pairs [0, n) and of the target
 * over their interior rows. get/set: 5 values each (omega, psi, u, v, target). */
int nse2d_les_fit_scales(nse2d_les_t L, int n);
void nse2d_les_get_scales(nse2d_les_t L, double* mean5, double* std5);
int nse2d_les_set_scales(nse2d_les_t L, const double* mean5, const double* std5);
/* Adam on mini-batches of `batch` pairs drawn without replacement from [0, n_train), reshuffled
 * each epoch (seeded), learning rate lr0 with cosine decay to lr0/20; losses[step] (may be NULL)
 * receives each step's normalised MSE. Returns the mean loss of the last 10 steps, -1 on error. */
double nse2d_les_train(nse2d_les_t L, int n_train, int steps, int batch, float lr0,
                       unsigned long long seed, float* losses);
/* Normalised MSE over pairs [first, first+count) of the CNN and of a zero closure. 0 on success. */
int nse2d_les_eval(nse2d_les_t L, int first, int count, double* loss_cnn, double* loss_zero);
/* The closure term of one coarse state into d_out ([Ny_c][Nx_c] fp32, 0 on the wall rows). */
int nse2d_les_closure(nse2d_les_t L, const double* d_omega_c, const double* d_psi_c, float* d_out);
/* on = 1: the coarse channel adds the closure of each stage state to its RHS (section 5,
 * nse2d_channel_set_closure); on = 2: the closure of each step's first stage state (the
 * channel's omega^n field) is evaluated once and added in all three stages (a third of the
 * network evaluations; the closure term is then first order in time); on = 0 detaches it.
 * The channel must have been created on the stream given to nse2d_les_create (else -1).
 * nse2d_les_closure always evaluates, whatever is attached. 0 on success. */
int nse2d_les_attach(nse2d_les_t L, nse2d_channel_t coarse_ch, int on);
/* Constraints applied to every later closure evaluation (attached or nse2d_les_closure), as an
 * OR of the flags below; 0 (the default) applies none. In the order they are applied:
 * NSE2D_CLOSURE_MEAN_FREE subtracts each row's x-average from the term, leaving the mean
 * vorticity profile to the resolved dynamics; NSE2D_CLOSURE_DISSIPATIVE sets the term to 0 where
 * omega * C > 0 (backscatter, which adds enstrophy). No counterpart in the reference, whose
 * solver has no closure: these are options of the LES built on it. 0, or -1 on an unknown flag. */
#define NSE2D_CLOSURE_DISSIPATIVE 1
#define NSE2D_CLOSURE_MEAN_FREE 2
int nse2d_les_set_constraints(nse2d_les_t L, int flags);

/* ------------------------------------------------------------------------------------------
 * 6. Runtime helpers (device memory, streams) so that callers without another HIP binding
 *    (ctypes, cgo) can drive the library.
 * ------------------------------------------------------------------------------------------ */
void* nse2d_dev_alloc(size_t bytes);
void nse2d_dev_free(void* p);
void nse2d_memcpy_h2d(void* dst, const void* src, size_t bytes);
void nse2d_memcpy_d2h(void* dst, const void* src, size_t bytes);
void nse2d_memcpy_d2d(void* dst, const void* src, size_t bytes);
void nse2d_device_sync(void);
/* A new stream of default priority, its destruction, and a wait for its work. */
hipStream_t nse2d_stream_create(void);
void nse2d_stream_destroy(hipStream_t s);
void nse2d_stream_sync(hipStream_t s);
int nse2d_device_count(void);
const char* nse2d_last_error(void);
const char* nse2d_version(void);

/* Host-side setup with the reference's definitions: grid metrics (grid.cu:7-94; stretch "none",
 * "uniform", "tanh", or "table:<file>" for the tabulated grid of grid.cu:46-78 with the Ny values
 * of y read as whitespace-separated text in place of the reference's HDF5 dataset /y) and the
 * initial vorticity (init.cu:99-378, plus the seeded benchmark ICs
 * "lamnoise" and "lamband"); argv is the reference command line. Return 0 on success. */
int nse2d_host_metrics(int Ny, double h, const char* stretch, double beta, double* y,
                       double* a_node, double* a_edge, double* dy_edge, double* w_node);
int nse2d_host_init_omega(int argc, char** argv, double* w_out);

/* Host driver with the reference's command line (main.cu:13-816): runs a case, writing
 * diagnostics.csv, profiles_timeavg.csv, progress.log and snapshots into --outdir. */
int nse2d_main(int argc, char** argv);

#ifdef __cplusplus
}
#endif
#endif /* NSE2D_HIP_H */
