// nse2d oracle: FP64 OpenMP C++ restatement of the reference solver's per-step path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP implementation and the CPU
// baseline of bench.py. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it. The product path (nse2d-channel-gpu_amd/) never links or calls it.
//
// Every function restates one reference function; the citation (file:line, relative to the
// reference repository root) is given next to each declaration. Array layout follows the reference:
// fields are row-major [Ny][Nx] with x fastest, metric vectors have Ny or Ny-1 entries.
//
// Pinning (how each part is checked, see tests/test_oracle_*.py):
//   - grid metrics, ICs, the Poisson solve, Thom closure, every diag reduction and the host
//     formulas: row 1 (t = 0) of all 21 fixture diagnostics.csv files (9 significant digits);
//   - SSPRK stages, dt rule, output-time clipping: later fixture rows of short cases;
//   - FFT: against an O(N^2) DFT and numpy.fft; Thomas: against a dense solve of the same
//     spectral-x / FD-y operator; Laplacian/Jacobian: closed forms on polynomials and trig fields;
//     Arakawa: discrete energy/enstrophy invariants.
// Parity unpinned: none of the functions declared here (each has at least one of the checks above).
#pragma once
#include <cstddef>

extern "C" {

// ---- grid.cu:7-94 build_y_metrics. stretch: 0 = none/uniform, 1 = tanh. Returns 0 on success.
int orc_build_y_metrics(int Ny, double h, int stretch, double beta,
                        double* y, double* a_node, double* a_edge, double* dy_edge, double* w_node);
// ---- grid.cu:46-78 tabulated branch: y given (Ny values), metrics by differences in eta.
int orc_table_y_metrics(int Ny, const double* y_in, double* y, double* a_node, double* a_edge,
                        double* dy_edge, double* w_node);

// ---- ops.cu (device operators restated as loops)
void orc_uv_from_psi_mapped(const double* psi, double* u, double* v, int Nx, int Ny,
                            double inv_dx, double inv_deta, const double* a_node);            // ops.cu:47-63
void orc_omega_from_psi_interior_mapped(double* w, const double* psi, int Nx, int Ny,
                            double inv_dx2, double inv_deta2, const double* a_node,
                            const double* a_edge);                                            // ops.cu:72-93
void orc_thom_wall_vorticity_mapped(double* w, const double* psi, int Nx, int Ny,
                            const double* dy_edge, double psi_bot, double psi_top,
                            int free_slip);                                                   // ops.cu:102-158
void orc_arakawa_J_mapped(const double* psi, const double* w, double* J, int Nx, int Ny,
                            double inv_dx, double inv_deta, const double* a_node);            // ops.cu:283-326
void orc_arakawa_J_mapped_split(const double* psi, const double* w, double* J1, double* J2,
                            double* J3, double* Jadv, int Nx, int Ny, double inv_dx,
                            double inv_deta, const double* a_node);                           // ops.cu:415-467
void orc_compute_adv_from_uv_mapped(const double* u, const double* v, const double* w,
                            double* Jadv, int Nx, int Ny, double inv_dx, double inv_deta,
                            const double* a_node);                                            // ops.cu:339-371
void orc_compute_adv_flux_mapped(const double* u, const double* v, const double* w,
                            double* Jadv, int Nx, int Ny, double inv_dx, double inv_deta,
                            const double* a_node);                                            // ops.cu:372-409
void orc_laplacian_fd_mapped(const double* q, double* out, int Nx, int Ny, double inv_dx2,
                            double inv_deta2, const double* a_node, const double* a_edge);    // ops.cu:471-499
void orc_rhs_mapped(const double* J, const double* lap, const double* w, double* rhs,
                            int Nx, int Ny, double nu, double F0, int nforce, double h,
                            double lin_drag, const double* y_node);                           // ops.cu:515-534
void orc_mms_add_forcing(double* rhs, int Nx, int Ny, double Lx, double h, double t, double nu,
                            double r, double A0, double eps, double Om, int kx_mode,
                            const double* y_node);                                            // ops.cu:559-626
void orc_axpy(double* y, const double* x, const double* r, double dt, size_t N);              // ops.cu:644-647
void orc_ssprk2(double* w2, const double* w, const double* w1, const double* rhs, double dt,
                            size_t N);                                                        // ops.cu:648-651
void orc_ssprk3(double* wnp1, const double* w, const double* w2, const double* rhs, double dt,
                            size_t N);                                                        // ops.cu:652-655

// ---- poisson.cu / poisson.cuh
// Unnormalised complex DFT along contiguous rows of length N (interleaved re,im), batched
// `howmany` times: forward exp(-2 pi i k n / N), inverse exp(+2 pi i k n / N) (cuFFT Z2Z
// convention used at poisson.cuh:96,110). Radix-2 for powers of two, direct O(N^2) otherwise.
void orc_fft_z2z(double* data, int N, int howmany, int inverse);
void orc_dft_naive(const double* in, double* out, int N, int inverse);                        // textbook DFT
// Evaluation order of every transform and y elimination of the oracle (orc_fft_z2z, the Poisson
// solves, the driver), bit flags: 1 = FFTs by decimation in frequency instead of in time, 2 = the
// y systems eliminated from the top wall down instead of from the bottom up. A measurement tool
// with no reference counterpart: the same solver evaluated in a second FP64 order shows how
// closely two FP64 implementations of the path can agree.
void orc_set_eval_order(int flags);
int orc_get_eval_order(void);
void orc_fill_kx2(double* kx2, int Nx, double Lx);                                            // poisson.cu:16-22
void orc_precompute_cp_mapped(double* cp, const double* a_node, const double* a_edge,
                            const double* kx2, int Nx, int Ny, double inv_deta2);             // poisson.cu:38-68
// In place on hat ([Ny][Nx] interleaved complex), k-major cp as the reference.
void orc_poisson_tridiag_mapped(double* hat, const double* cp, const double* a_node,
                            const double* a_edge, const double* kx2, int Nx, int Ny,
                            double inv_deta2, double psi_bot, double psi_top);                // poisson.cu:70-133
// Whole solve (poisson.cuh:83-118): pack, FFT, tridiag, inverse FFT, unpack * 1/Nx.
void orc_poisson_solve(const double* omega, double* psi, int Nx, int Ny, double Lx,
                            const double* a_node, const double* a_edge, double inv_deta,
                            double psi_bot, double psi_top);
// Extended-precision solve of the same operator (a measurement tool with no reference
// counterpart): shows the distance of any FP64 evaluation from the operator's exact solution.
void orc_poisson_solve_ld(const double* omega, double* psi, int Nx, int Ny, double Lx,
                          const double* a_node, const double* a_edge, double inv_deta,
                          double psi_bot, double psi_top);

// ---- diag.cuh:165-174 + diag.cu:876-893 (Spectra1D::accumulate): adds one sample of the x-spectra
// of u and v to acc (E_u at [0, Nx/2], E_v at [Nx/2+1, Nx+1]).
void orc_spectra1d_accumulate(const double* u, const double* v, int Nx, int Ny, double* acc);

// ---- diag.cu reductions. Results are the raw sums before the host normalisation, as the
// reference launchers leave them in their device scalars.
void orc_enstrophy_diss_mapped(const double* w, int Nx, int Ny, double inv_dx, double inv_deta,
                            const double* a_node, const double* w_node, double* out2);        // diag.cu:6-38
void orc_enstrophy_diss_mapped_full(const double* w, int Nx, int Ny, double inv_dx,
                            double inv_deta, const double* a_node, const double* w_node,
                            double* out2);                                                    // diag.cu:43-101
void orc_enstrophy_diss_balance_mapped(const double* w, int Nx, int Ny, double inv_dx2,
                            double inv_deta2, const double* a_node, const double* a_edge,
                            const double* w_node, double* out2);                              // diag.cu:103-156,586-625
void orc_enstrophy_sbp_mapped(const double* w, int Nx, int Ny, double inv_dx, double inv_deta,
                            const double* a_edge, const double* wt_node, double* out2);       // diag.cu:158-199
double orc_kinetic_energy_mapped(const double* psi, int Nx, int Ny, double inv_dx,
                            double inv_deta, const double* a_node, const double* w_node);     // diag.cu:201-231
double orc_energy_dissipation_mapped(const double* psi, const double* w, int Nx, int Ny,
                            double inv_dx, double inv_deta, const double* a_node,
                            const double* a_edge, const double* w_node);                      // diag.cu:250-348
double orc_power_input_mapped(const double* psi, int Nx, int Ny, double inv_deta, double h,
                            double F0, int nforce, const double* a_node, const double* y_node,
                            const double* w_node);                                            // diag.cu:350-376
double orc_enstrophy_input_mapped_full(const double* w, int Nx, int Ny, double h, double F0,
                            int nforce, const double* y_node, const double* w_node);          // diag.cu:378-422
double orc_enstrophy_input_mapped(const double* w, int Nx, int Ny, double h, double F0,
                            int nforce, const double* y_node, const double* w_node);          // diag.cu:423-467
void orc_row_moments_from_psi_mapped(const double* psi, int Nx, int Ny, double inv_dx,
                            double inv_deta, const double* a_node, double* su, double* sv,
                            double* su2, double* sv2, double* suv);                           // diag.cu:468-499
void orc_mean_wall_omega(const double* w, int Nx, int Ny, double* out2);                      // diag.cu:500-512
void orc_mean_wall_omega_omegay_mapped(const double* w, int Nx, int Ny, double inv_deta,
                            const double* a_node, double* out2);                              // diag.cu:516-583
double orc_omega_rhs_balance_mapped(const double* w, const double* rhs, int Nx, int Ny,
                            const double* w_node);                                            // diag.cu:628-691
void orc_max_uv_from_psi_mapped(const double* psi, int Nx, int Ny, double inv_dx,
                            double inv_deta, const double* a_node, double* out3);             // diag.cu:702-866

// ---- init.cu:99-378 initial conditions, plus the seeded ICs of the benchmark configurations
// (no reference counterpart; same documented generator as the HIP host code).
// Arguments are passed as a command line ("--Nx 64 --init mix ..."), parsed like parse_args.
int orc_init_omega(int argc, char** argv, double* w_out);

// ---- main.cu driver: full run with the reference's CLI, writing diagnostics.csv,
// profiles_timeavg.csv, snapshots and progress.log into --outdir. Extra oracle-only flags:
//   --max_steps N   stop after N steps (timing and partial runs)
//   --walls freeslip  free-slip wall closure (the reference's FREE_SLIP_WALLS build)
int orc_run_main(int argc, char** argv);

// Timing helper for the CPU baseline: set up the case described by argv, run `warm` steps,
// then `steps` timed steps; returns wall seconds of the timed steps (negative on error).
double orc_time_steps(int argc, char** argv, int warm, int steps);

// Threads actually used by OpenMP regions.
int orc_num_threads(void);
}
