// Spatial operators of the reference (ops.cu) as gfx950 kernels: velocities, mapped Laplacian,
// 9-point mapped Arakawa Jacobian (and its split/test variants), Thom wall closure, RHS,
// manufactured forcing and the SSPRK(3,3) combinations. One thread per grid point, 64x4 blocks
// so that each wavefront covers 64 consecutive x (one 512-byte row segment per field load).
#include "nse2d_common.hpp"
#include "nse2d_ops.hpp"
#include "../../include/nse2d_hip.h"

#include <atomic>

namespace nse2d {

namespace {
std::atomic<int> g_wall_type{0};
}
int wall_type() { return g_wall_type.load(); }

constexpr int TX = 64, TY = 4;
inline dim3 grid2d(int Nx, int Ny) { return dim3(ceil_div(Nx, TX), ceil_div(Ny, TY)); }

// ---- ops.cu:47-63
__global__ void __launch_bounds__(256) uv_from_psi_mapped_kernel(
    const double* __restrict__ psi, double* __restrict__ u, double* __restrict__ v, int Nx, int Ny,
    double inv_dx, double inv_deta, const double* __restrict__ a_node) {
    const int ix = blockIdx.x * TX + threadIdx.x, iy = blockIdx.y * TY + threadIdx.y;
    if (ix >= Nx || iy >= Ny) return;
    const int xm = wrap_x(ix - 1, Nx), xp = wrap_x(ix + 1, Nx);
    const int ym = (iy == 0 ? 0 : iy - 1), yp = (iy == Ny - 1 ? Ny - 1 : iy + 1);
    const double dpsideta = (psi[idx2(yp, ix, Nx)] - psi[idx2(ym, ix, Nx)]) * 0.5 * inv_deta;
    const double dpsidx = (psi[idx2(iy, xp, Nx)] - psi[idx2(iy, xm, Nx)]) * 0.5 * inv_dx;
    u[idx2(iy, ix, Nx)] = a_node[iy] * dpsideta;
    v[idx2(iy, ix, Nx)] = -dpsidx;
}

// ---- ops.cu:72-93
__global__ void __launch_bounds__(256) omega_from_psi_interior_mapped_kernel(
    double* __restrict__ w, const double* __restrict__ psi, int Nx, int Ny, double inv_dx2,
    double inv_deta2, const double* __restrict__ a_node, const double* __restrict__ a_edge) {
    const int ix = blockIdx.x * TX + threadIdx.x, iy = blockIdx.y * TY + threadIdx.y;
    if (ix >= Nx || iy >= Ny || iy == 0 || iy == Ny - 1) return;
    const int xm = wrap_x(ix - 1, Nx), xp = wrap_x(ix + 1, Nx);
    const size_t c = idx2(iy, ix, Nx);
    const double lapx = psi[idx2(iy, xp, Nx)] - 2.0 * psi[c] + psi[idx2(iy, xm, Nx)];
    const double lapy = a_node[iy] * (a_edge[iy] * (psi[idx2(iy + 1, ix, Nx)] - psi[c]) -
                                      a_edge[iy - 1] * (psi[c] - psi[idx2(iy - 1, ix, Nx)]));
    w[c] = -(inv_dx2 * lapx + inv_deta2 * lapy);
}

// ---- ops.cu:102-158 (second-order one-sided Thom closure on the stretched grid)
__global__ void __launch_bounds__(256) thom_wall_kernel(double* __restrict__ w,
                                                        const double* __restrict__ psi, int Nx,
                                                        int Ny, const double* __restrict__ dy_edge,
                                                        double psi_bot, double psi_top,
                                                        int free_slip) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nx) return;
    double wb, wt;
    thom_values(psi, Nx, Ny, dy_edge, psi_bot, psi_top, free_slip, i, wb, wt);
    w[i] = wb;
    w[idx2(Ny - 1, i, Nx)] = wt;
}

// ---- ops.cu:283-326
__global__ void __launch_bounds__(256) arakawa_J_mapped_kernel(
    const double* __restrict__ psi, const double* __restrict__ omg, double* __restrict__ J,
    int Nx, int Ny, double inv_dx, double inv_deta, const double* __restrict__ a_node) {
    const int ix = blockIdx.x * TX + threadIdx.x, iy = blockIdx.y * TY + threadIdx.y;
    if (ix >= Nx || iy >= Ny) return;
    const size_t c = idx2(iy, ix, Nx);
    if (iy == 0 || iy == Ny - 1) { J[c] = 0.0; return; }
    const Stencil9 P = Stencil9::load(psi, Nx, ix, iy), W = Stencil9::load(omg, Nx, ix, iy);
    double J1, J2, J3;
    arakawa_terms(P, W, a_node[iy], 0.25 * inv_dx * inv_deta, J1, J2, J3);
    J[c] = -((J1 + J2 + J3) * (1.0 / 3.0));
}

// ---- ops.cu:415-467
__global__ void __launch_bounds__(256) arakawa_J_mapped_split_kernel(
    const double* __restrict__ psi, const double* __restrict__ omg, double* __restrict__ J1o,
    double* __restrict__ J2o, double* __restrict__ J3o, double* __restrict__ Jadv, int Nx,
    int Ny, double inv_dx, double inv_deta, const double* __restrict__ a_node) {
    const int ix = blockIdx.x * TX + threadIdx.x, iy = blockIdx.y * TY + threadIdx.y;
    if (ix >= Nx || iy >= Ny) return;
    const size_t c = idx2(iy, ix, Nx);
    if (iy == 0 || iy == Ny - 1) {
        if (J1o) J1o[c] = 0.0;
        if (J2o) J2o[c] = 0.0;
        if (J3o) J3o[c] = 0.0;
        if (Jadv) Jadv[c] = 0.0;
        return;
    }
    const Stencil9 P = Stencil9::load(psi, Nx, ix, iy), W = Stencil9::load(omg, Nx, ix, iy);
    double J1, J2, J3;
    arakawa_terms(P, W, a_node[iy], 0.25 * inv_dx * inv_deta, J1, J2, J3);
    if (J1o) J1o[c] = J1;
    if (J2o) J2o[c] = J2;
    if (J3o) J3o[c] = J3;
    if (Jadv) Jadv[c] = -((J1 + J2 + J3) * (1.0 / 3.0));
}

// ---- ops.cu:339-371
__global__ void __launch_bounds__(256) adv_from_uv_kernel(
    const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ omg,
    double* __restrict__ Jadv, int Nx, int Ny, double inv_dx, double inv_deta,
    const double* __restrict__ a_node) {
    const int ix = blockIdx.x * TX + threadIdx.x, iy = blockIdx.y * TY + threadIdx.y;
    if (ix >= Nx || iy >= Ny) return;
    const size_t c = idx2(iy, ix, Nx);
    if (iy == 0 || iy == Ny - 1) { Jadv[c] = 0.0; return; }
    const int xm = wrap_x(ix - 1, Nx), xp = wrap_x(ix + 1, Nx);
    const double dx_omg = (omg[idx2(iy, xp, Nx)] - omg[idx2(iy, xm, Nx)]) * (0.5 * inv_dx);
    const double dy_omg =
        a_node[iy] * (omg[idx2(iy + 1, ix, Nx)] - omg[idx2(iy - 1, ix, Nx)]) * (0.5 * inv_deta);
    Jadv[c] = u[c] * dx_omg + v[c] * dy_omg;
}

// ---- ops.cu:372-409
__global__ void __launch_bounds__(256) adv_flux_kernel(
    const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ omg,
    double* __restrict__ Jadv, int Nx, int Ny, double inv_dx, double inv_deta,
    const double* __restrict__ a_node) {
    const int ix = blockIdx.x * TX + threadIdx.x, iy = blockIdx.y * TY + threadIdx.y;
    if (ix >= Nx || iy >= Ny) return;
    const size_t c = idx2(iy, ix, Nx);
    if (iy == 0 || iy == Ny - 1) { Jadv[c] = 0.0; return; }
    const int xm = wrap_x(ix - 1, Nx), xp = wrap_x(ix + 1, Nx);
    const size_t cxm = idx2(iy, xm, Nx), cxp = idx2(iy, xp, Nx);
    const size_t cym = idx2(iy - 1, ix, Nx), cyp = idx2(iy + 1, ix, Nx);
    const double u_ip = (u[cxp] + u[c]) * 0.5, u_im = (u[c] + u[cxm]) * 0.5;
    const double w_ip = (omg[cxp] + omg[c]) * 0.5, w_im = (omg[c] + omg[cxm]) * 0.5;
    const double dFdx = (u_ip * w_ip - u_im * w_im) * inv_dx;
    const double v_jp = (v[cyp] + v[c]) * 0.5, v_jm = (v[c] + v[cym]) * 0.5;
    const double w_jp = (omg[cyp] + omg[c]) * 0.5, w_jm = (omg[c] + omg[cym]) * 0.5;
    const double dFdy = a_node[iy] * ((v_jp * w_jp - v_jm * w_jm) * (0.5 * inv_deta * 2.0));
    Jadv[c] = dFdx + dFdy;
}

// ---- ops.cu:471-499
__global__ void __launch_bounds__(256) laplacian_fd_mapped_kernel(
    const double* __restrict__ q, double* __restrict__ out, int Nx, int Ny, double inv_dx2,
    double inv_deta2, const double* __restrict__ a_node, const double* __restrict__ a_edge) {
    const int ix = blockIdx.x * TX + threadIdx.x, iy = blockIdx.y * TY + threadIdx.y;
    if (ix >= Nx || iy >= Ny) return;
    const size_t c = idx2(iy, ix, Nx);
    if (iy == 0 || iy == Ny - 1) { out[c] = 0.0; return; }
    out[c] = mapped_laplacian(q[idx2(iy, wrap_x(ix - 1, Nx), Nx)], q[c],
                              q[idx2(iy, wrap_x(ix + 1, Nx), Nx)], q[idx2(iy - 1, ix, Nx)],
                              q[idx2(iy + 1, ix, Nx)], a_node[iy], a_edge[iy], a_edge[iy - 1],
                              inv_dx2, inv_deta2);
}

// ---- ops.cu:515-534 (forcing evaluated per point like the reference)
__global__ void __launch_bounds__(256) rhs_mapped_kernel(
    const double* __restrict__ J, const double* __restrict__ lap, const double* __restrict__ omg,
    double* __restrict__ rhs, int Nx, int Ny, double nu, double F0, int nforce, double h,
    double lin_drag, const double* __restrict__ y_node) {
    const int ix = blockIdx.x * TX + threadIdx.x, iy = blockIdx.y * TY + threadIdx.y;
    if (ix >= Nx || iy >= Ny) return;
    const size_t c = idx2(iy, ix, Nx);
    const double f = kolmogorov_f_omega(F0, nforce, h, y_node[iy]);
    rhs[c] = -J[c] + nu * lap[c] + f - lin_drag * omg[c];
}

// ---- ops.cu:559-626
__global__ void __launch_bounds__(256) mms_add_forcing_kernel(double* __restrict__ rhs, int Nx,
                                                              int Ny, MmsParams m,
                                                              const double* __restrict__ y_node) {
    const int ix = blockIdx.x * TX + threadIdx.x, iy = blockIdx.y * TY + threadIdx.y;
    if (ix >= Nx || iy >= Ny) return;
    rhs[idx2(iy, ix, Nx)] += mms_f_omega(m, ix, y_node[iy]);
}

// ---- ops.cu:644-655
__global__ void __launch_bounds__(256) axpy_kernel(double* y, const double* x, const double* r,
                                                   double dt, size_t N) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) y[i] = x[i] + dt * r[i];
}
__global__ void __launch_bounds__(256) ssprk2_kernel(double* w2, const double* w, const double* w1,
                                                     const double* rhs, double dt, size_t N) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) w2[i] = 0.75 * w[i] + 0.25 * (w1[i] + dt * rhs[i]);
}
// wnp1 may alias w (in-place final stage, main.cu:530): each thread reads then writes its own
// element only.
__global__ void __launch_bounds__(256) ssprk3_kernel(double* wnp1, const double* w,
                                                     const double* w2, const double* rhs,
                                                     double dt, size_t N) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) wnp1[i] = (1.0 / 3.0) * w[i] + (2.0 / 3.0) * (w2[i] + dt * rhs[i]);
}

void launch_thom(double* w, const double* psi, int Nx, int Ny, const double* dy_edge,
                 double psi_bot, double psi_top, int free_slip, hipStream_t s) {
    thom_wall_kernel<<<ceil_div(Nx, 256), 256, 0, s>>>(w, psi, Nx, Ny, dy_edge, psi_bot, psi_top,
                                                       free_slip);
    NSE2D_LAUNCH_CHECK();
}

}  // namespace nse2d

using namespace nse2d;

extern "C" {

void nse2d_set_wall_type(int free_slip) { g_wall_type.store(free_slip ? 1 : 0); }
int nse2d_get_wall_type(void) { return g_wall_type.load(); }

void launch_uv_from_psi_mapped(const double* psi, double* u, double* v, int Nx, int Ny,
                               double inv_dx, double inv_deta, const double* d_a_node,
                               hipStream_t s) {
    if (Nx <= 0 || Ny <= 0) return;
    uv_from_psi_mapped_kernel<<<grid2d(Nx, Ny), dim3(TX, TY), 0, s>>>(psi, u, v, Nx, Ny, inv_dx,
                                                                       inv_deta, d_a_node);
    NSE2D_LAUNCH_CHECK();
}

void launch_omega_from_psi_interior_mapped(double* w, const double* psi, int Nx, int Ny,
                                           double inv_dx2, double inv_deta2,
                                           const double* d_a_node, const double* d_a_edge,
                                           hipStream_t s) {
    if (Nx <= 0 || Ny <= 2) return;
    omega_from_psi_interior_mapped_kernel<<<grid2d(Nx, Ny), dim3(TX, TY), 0, s>>>(
        w, psi, Nx, Ny, inv_dx2, inv_deta2, d_a_node, d_a_edge);
    NSE2D_LAUNCH_CHECK();
}

void launch_thom_wall_vorticity_mapped(double* w, const double* psi, int Nx, int Ny,
                                       const double* d_dy_edge, double psi_bot, double psi_top,
                                       hipStream_t s) {
    if (Nx <= 0 || Ny < 3) return;
    launch_thom(w, psi, Nx, Ny, d_dy_edge, psi_bot, psi_top, wall_type(), s);
}

void launch_arakawa_J_mapped(const double* psi, const double* omg, double* J, int Nx, int Ny,
                             double inv_dx, double inv_deta, const double* d_a_node,
                             hipStream_t s) {
    if (Nx <= 0 || Ny <= 0) return;
    arakawa_J_mapped_kernel<<<grid2d(Nx, Ny), dim3(TX, TY), 0, s>>>(psi, omg, J, Nx, Ny, inv_dx,
                                                                     inv_deta, d_a_node);
    NSE2D_LAUNCH_CHECK();
}

void launch_laplacian_fd_mapped(const double* d_q, double* d_out, int Nx, int Ny,
                                double inv_dx2, double inv_deta2, const double* d_a_node,
                                const double* d_a_edge, hipStream_t s) {
    if (Nx <= 0 || Ny <= 0) return;
    laplacian_fd_mapped_kernel<<<grid2d(Nx, Ny), dim3(TX, TY), 0, s>>>(
        d_q, d_out, Nx, Ny, inv_dx2, inv_deta2, d_a_node, d_a_edge);
    NSE2D_LAUNCH_CHECK();
}

void launch_rhs_mapped(const double* J, const double* lap, const double* omg, double* rhs,
                       int Nx, int Ny, double nu, double F0, int nforce, double h,
                       double lin_drag, const double* d_y_node, hipStream_t s) {
    if (Nx <= 0 || Ny <= 0) return;
    rhs_mapped_kernel<<<grid2d(Nx, Ny), dim3(TX, TY), 0, s>>>(J, lap, omg, rhs, Nx, Ny, nu, F0,
                                                               nforce, h, lin_drag, d_y_node);
    NSE2D_LAUNCH_CHECK();
}

void launch_mms_add_forcing(double* d_rhs, int Nx, int Ny, double Lx, double h, double t,
                            double nu, double r, double A0, double eps, double Om, int kx_mode,
                            const double* d_y_node, hipStream_t s) {
    if (Nx <= 0 || Ny <= 0) return;
    const MmsParams m = make_mms_params(Nx, Lx, h, t, nu, r, A0, eps, Om, kx_mode);
    mms_add_forcing_kernel<<<grid2d(Nx, Ny), dim3(TX, TY), 0, s>>>(d_rhs, Nx, Ny, m, d_y_node);
    NSE2D_LAUNCH_CHECK();
}

void launch_axpy(double* y, const double* x, const double* r, double dt, size_t N,
                 hipStream_t s) {
    if (N == 0) return;
    axpy_kernel<<<ceil_div(N, 256), 256, 0, s>>>(y, x, r, dt, N);
    NSE2D_LAUNCH_CHECK();
}
void launch_ssprk2(double* w2, const double* w, const double* w1, const double* rhs, double dt,
                   size_t N, hipStream_t s) {
    if (N == 0) return;
    ssprk2_kernel<<<ceil_div(N, 256), 256, 0, s>>>(w2, w, w1, rhs, dt, N);
    NSE2D_LAUNCH_CHECK();
}
void launch_ssprk3(double* wnp1, const double* w, const double* w2, const double* rhs,
                   double dt, size_t N, hipStream_t s) {
    if (N == 0) return;
    ssprk3_kernel<<<ceil_div(N, 256), 256, 0, s>>>(wnp1, w, w2, rhs, dt, N);
    NSE2D_LAUNCH_CHECK();
}

// ---- test_api.cu:8-92
void testapi_uv_from_psi_mapped(const double* d_psi, double* d_u, double* d_v, int Nx, int Ny,
                                double inv_dx, double inv_deta, const double* d_a_node,
                                hipStream_t s) {
    launch_uv_from_psi_mapped(d_psi, d_u, d_v, Nx, Ny, inv_dx, inv_deta, d_a_node, s);
}
void testapi_omega_from_psi_interior_mapped(double* d_w, const double* d_psi, int Nx, int Ny,
                                            double inv_dx2, double inv_deta2,
                                            const double* d_a_node, const double* d_a_edge,
                                            hipStream_t s) {
    launch_omega_from_psi_interior_mapped(d_w, d_psi, Nx, Ny, inv_dx2, inv_deta2, d_a_node,
                                          d_a_edge, s);
}
void testapi_thom_wall_vorticity_mapped(double* d_w, const double* d_psi, int Nx, int Ny,
                                        const double* d_dy_edge, double psi_bot,
                                        double psi_top, hipStream_t s) {
    launch_thom_wall_vorticity_mapped(d_w, d_psi, Nx, Ny, d_dy_edge, psi_bot, psi_top, s);
}
void testapi_arakawa_J_mapped(const double* d_psi, const double* d_w, double* d_J, int Nx,
                              int Ny, double inv_dx, double inv_deta, const double* d_a_node,
                              hipStream_t s) {
    launch_arakawa_J_mapped(d_psi, d_w, d_J, Nx, Ny, inv_dx, inv_deta, d_a_node, s);
}
void testapi_compute_adv_from_uv_mapped(const double* d_u, const double* d_v,
                                        const double* d_omg, double* d_Jref, int Nx, int Ny,
                                        double inv_dx, double inv_deta, const double* d_a_node,
                                        hipStream_t s) {
    if (Nx <= 0 || Ny <= 0) return;
    adv_from_uv_kernel<<<grid2d(Nx, Ny), dim3(TX, TY), 0, s>>>(d_u, d_v, d_omg, d_Jref, Nx, Ny,
                                                                inv_dx, inv_deta, d_a_node);
    NSE2D_LAUNCH_CHECK();
}
void testapi_compute_adv_flux_mapped(const double* d_u, const double* d_v, const double* d_omg,
                                     double* d_Jref, int Nx, int Ny, double inv_dx,
                                     double inv_deta, const double* d_a_node, hipStream_t s) {
    if (Nx <= 0 || Ny <= 0) return;
    adv_flux_kernel<<<grid2d(Nx, Ny), dim3(TX, TY), 0, s>>>(d_u, d_v, d_omg, d_Jref, Nx, Ny,
                                                             inv_dx, inv_deta, d_a_node);
    NSE2D_LAUNCH_CHECK();
}
void testapi_arakawa_J_mapped_split(const double* d_psi, const double* d_omg, double* d_J1,
                                    double* d_J2, double* d_J3, double* d_Jadv, int Nx, int Ny,
                                    double inv_dx, double inv_deta, const double* d_a_node,
                                    hipStream_t s) {
    if (Nx <= 0 || Ny <= 0) return;
    arakawa_J_mapped_split_kernel<<<grid2d(Nx, Ny), dim3(TX, TY), 0, s>>>(
        d_psi, d_omg, d_J1, d_J2, d_J3, d_Jadv, Nx, Ny, inv_dx, inv_deta, d_a_node);
    NSE2D_LAUNCH_CHECK();
}

}  // extern "C"
