// Internal interface of the Poisson solver (poisson.hip), used by the stepper (channel.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace nse2d {

struct PoissonImpl;

PoissonImpl* poisson_create(int Nx, int Ny, double Lx, hipStream_t s);
int poisson_enable_mapped(PoissonImpl* p, const double* a_node, const double* a_edge,
                          double inv_deta, hipStream_t s);
// omega -> psi; psi = psi_bot / psi_top on the wall rows. omega and psi must not alias.
void poisson_solve(PoissonImpl* p, const double* omega, double* psi, double psi_bot,
                   double psi_top, hipStream_t s);
void poisson_destroy(PoissonImpl* p);
// The three phases of poisson_solve, for callers that time them separately.
void poisson_forward(PoissonImpl* p, const double* omega, hipStream_t s);
void poisson_tridiag(PoissonImpl* p, double psi_bot, double psi_top, hipStream_t s);
void poisson_inverse(PoissonImpl* p, double* psi, hipStream_t s);
// The inverse phase with the Thom closure of omega's wall rows fused in (ops.cu:113-151, or
// omega_wall = 0 with free_slip). Returns false, having launched nothing, when the plan's row FFT
// cannot fuse it (the small-row LDS kernels); the caller then runs poisson_inverse and the
// closure kernel.
bool poisson_fuses_closure(const PoissonImpl* p);
bool poisson_inverse_closure(PoissonImpl* p, double* psi, double* omega, const double* dy_edge, double psi_bot,
                             double psi_top, int free_slip, hipStream_t s);

}  // namespace nse2d
