// Internal interface of the Poisson solver (poisson.hip), used by the stepper (channel.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace nse2d {

struct PoissonImpl;

PoissonImpl* poisson_create(int Nx, int Ny, double Lx, hipStream_t s);
// With an explicit row-FFT radix (0 = LDS radix-4 kernels) and Thomas column width; -1 keeps the
// default (or the NSE2D_FFT_RADIX / NSE2D_THOMAS_KC setting). Unavailable choices fall back.
PoissonImpl* poisson_create_with(int Nx, int Ny, double Lx, int fft_radix, int thomas_kc, hipStream_t s);
int poisson_fft_radix(const PoissonImpl* p);
int poisson_thomas_kc(const PoissonImpl* p);
int poisson_enable_mapped(PoissonImpl* p, const double* a_node, const double* a_edge,
                          double inv_deta, hipStream_t s);
// omega -> psi; psi = psi_bot / psi_top on the wall rows. omega and psi must not alias.
void poisson_solve(PoissonImpl* p, const double* omega, double* psi, double psi_bot,
                   double psi_top, hipStream_t s);
void poisson_destroy(PoissonImpl* p);
// The three phases of poisson_solve, for callers that time them separately.
void poisson_forward(PoissonImpl* p, const double* omega, hipStream_t s);
// The forward phase over the interior rows only (the solve never reads the wall rows' spectrum).
void poisson_forward_interior(PoissonImpl* p, const double* omega, hipStream_t s);
void poisson_tridiag(PoissonImpl* p, double psi_bot, double psi_top, hipStream_t s);
void poisson_inverse(PoissonImpl* p, double* psi, hipStream_t s);
// The inverse phase with the Thom closure of omega's wall rows fused in (ops.cu:113-151, or
// omega_wall = 0 with free_slip), launched over the interior rows; psi's wall rows get psi_bot and
// psi_top, which is what their transform gives. Returns false, having launched nothing, when the plan's row FFT
// cannot fuse it (the small-row LDS kernels); the caller then runs poisson_inverse and the
// closure kernel.
bool poisson_fuses_closure(const PoissonImpl* p);
bool poisson_inverse_closure(PoissonImpl* p, double* psi, double* omega, const double* dy_edge, double psi_bot,
                             double psi_top, int free_slip, hipStream_t s);

// Group launches over n equally sized fields (ensemble members) that share plan p's coefficients
// (same Nx, Ny, Lx and metrics); each field has its own spectrum buffer of poisson_hat_bytes(p).
// The pointer tables are device arrays of n pointers. Available when poisson_group_supported(p)
// (the register row FFT, Nx in [512, 16384]). The inverse fuses the wall closure of each field's
// omega (d_omegas may be null: no closure).
size_t poisson_hat_bytes(const PoissonImpl* p);
// The plan's own spectrum buffer ([Ny][Nx/2] complex, slot 0 of a row holds (X_0, X_{Nx/2})):
// written by poisson_forward, read by poisson_inverse.
void* poisson_hat(PoissonImpl* p);
// The register row FFT of plan p, for a kernel that transforms rows it produces itself (the
// fused stage + forward transform): M = Nx/2, the radix R and the twiddle tables (passes >= 1,
// and the real-split factors exp(-2 pi i k / Nx)). False for the small-row LDS plans.
bool poisson_row_fft(const PoissonImpl* p, int* M, int* R, const void** twp, const void** tw);
int poisson_group_supported(const PoissonImpl* p);
void poisson_forward_group(PoissonImpl* p, const double* const* d_omegas, void* const* d_hats, int n,
                           hipStream_t s);
void poisson_tridiag_group(PoissonImpl* p, void* const* d_hats, int n, double psi_bot, double psi_top,
                           hipStream_t s);
void poisson_inverse_group(PoissonImpl* p, void* const* d_hats, double* const* d_psis, double* const* d_omegas,
                           int n, const double* dy_edge, double psi_bot, double psi_top, int free_slip,
                           hipStream_t s);

// Time-averaged x-spectra of u and v (diag.cuh:123-191), on a plan's row FFT.
struct Spectra1DImpl;
Spectra1DImpl* spectra1d_create(int Nx, int Ny, double Lx, hipStream_t s);
void spectra1d_accumulate(Spectra1DImpl* sp, const double* d_u, const double* d_v, hipStream_t s);
int spectra1d_read(Spectra1DImpl* sp, double* kx, double* Eu, double* Ev, hipStream_t s);
void spectra1d_destroy(Spectra1DImpl* sp);

}  // namespace nse2d
