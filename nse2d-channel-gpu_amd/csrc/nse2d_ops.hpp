// Device-side building blocks shared by the per-operator kernels (ops.hip) and the fused
// SSPRK stage kernel (stage.hip). Each helper restates the arithmetic of one reference kernel
// in the reference's operation order.
#pragma once
#include "nse2d_common.hpp"

#include <cmath>

namespace nse2d {

// The 3x3 neighbourhood of (ix, iy) with periodic x; rows iy-1..iy+1 must exist.
struct Stencil9 {
    double c, xp, xm, yp, ym, pp, pm, mp, mm;  // pp=(x+1,y+1) pm=(x+1,y-1) mp=(x-1,y+1) mm=(x-1,y-1)
    __device__ static __forceinline__ Stencil9 load(const double* __restrict__ f, int Nx, int ix, int iy) {
        const int xm_ = wrap_x(ix - 1, Nx), xp_ = wrap_x(ix + 1, Nx);
        const double* r0 = f + (size_t)(iy - 1) * Nx;
        const double* r1 = f + (size_t)iy * Nx;
        const double* r2 = f + (size_t)(iy + 1) * Nx;
        Stencil9 s;
        s.c = r1[ix]; s.xp = r1[xp_]; s.xm = r1[xm_];
        s.yp = r2[ix]; s.ym = r0[ix];
        s.pp = r2[xp_]; s.pm = r0[xp_]; s.mp = r2[xm_]; s.mm = r0[xm_];
        return s;
    }
};

// ops.cu:305-322: the three Arakawa terms with the row metric a on every eta difference.
__device__ __forceinline__ void arakawa_terms(const Stencil9& P, const Stencil9& W, double a,
                                     double inv4dxde, double& J1, double& J2, double& J3) {
    J1 = ((P.xp - P.xm) * (a * (W.yp - W.ym)) - (a * (P.yp - P.ym)) * (W.xp - W.xm)) * inv4dxde;
    J2 = (P.xp * (a * (W.pp - W.pm)) - P.xm * (a * (W.mp - W.mm)) -
          a * (P.yp * (W.pp - W.mp) - P.ym * (W.pm - W.mm))) * inv4dxde;
    J3 = (-W.xp * (a * (P.pp - P.pm)) + W.xm * (a * (P.mp - P.mm)) +
          a * (W.yp * (P.pp - P.mp) - W.ym * (P.pm - P.mm))) * inv4dxde;
}

// ops.cu:488-498: delta_xx + a_j [a_{j+1/2}(q_{j+1}-q_j) - a_{j-1/2}(q_j-q_{j-1})] / deta^2
__device__ __forceinline__ double mapped_laplacian(double qxm, double qc, double qxp, double qym,
                                          double qyp, double a, double ap, double am,
                                          double inv_dx2, double inv_deta2) {
    const double dxx = (qxp - 2 * qc + qxm) * inv_dx2;
    const double dyy = a * (ap * (qyp - qc) - am * (qc - qym)) * inv_deta2;
    return dxx + dyy;
}

// ops.cu:526-531: Kolmogorov vorticity forcing f(y) = -F0 (n pi / 2h) cos(n pi (y+h) / 2h)
__host__ __device__ inline double kolmogorov_f_omega(double F0, int nforce, double h, double y) {
    if (F0 == 0.0) return 0.0;
    const double arg = (double)nforce * kPi * ((y + h) / (2.0 * h));
    return -F0 * ((double)nforce * kPi / (2.0 * h)) * cos(arg);
}

// ops.cu:565-574: constants of the manufactured forcing at time t.
struct MmsParams {
    double dx, k, c, h, A, Aprime, nu, r;
};
inline MmsParams make_mms_params(int Nx, double Lx, double h, double t, double nu, double r,
                                 double A0, double eps, double Om, int kx_mode) {
    MmsParams m;
    m.dx = Lx / (double)Nx;
    m.k = 2.0 * kPi * (double)kx_mode / Lx;
    m.c = kPi / (2.0 * h);
    m.h = h;
    m.A = A0 * (1.0 + eps * std::sin(Om * t));
    m.Aprime = A0 * (eps * Om * std::cos(Om * t));
    m.nu = nu;
    m.r = r;
    return m;
}
// ops.cu:583-624: f_omega(x_i, y) for psi_ex = A(t) sin(kx) sin^2(theta)
__device__ __forceinline__ double mms_f_omega(const MmsParams& m, int i, double y) {
    const double x = (double)i * m.dx;
    const double theta = m.c * (y + m.h);
    const double two_th = 2.0 * theta;
    const double s_th = sin(theta);
    const double s = s_th * s_th;
    const double cos2 = cos(two_th);
    const double sin2 = sin(two_th);
    const double c2 = m.c * m.c;
    const double k = m.k;
    const double g = (k * k) * s - 2.0 * c2 * cos2;
    const double gpp = 2.0 * c2 * cos2 * ((k * k) + 4.0 * c2);
    const double sk = sin(k * x);
    const double s2k = sin(2.0 * k * x);
    const double term_lin = m.Aprime * g + (m.nu * m.A) * ((k * k) * g - gpp) + (m.r * m.A) * g;
    const double c3 = c2 * m.c;
    const double term_adv = -(m.A * m.A) * k * c3 * s2k * sin2;
    return sk * term_lin + term_adv;
}

// ops.cu:113-151: wall vorticity at column i from psi rows 1,2 and Ny-2,Ny-3.
// A = 2(1/h1^2 + 1/(h1 h2)), B = -2 h1 / (h2 (h1+h2)^2); omega_wall = -(A dpsi1 + B dpsi2).
__device__ __forceinline__ void thom_values(const double* __restrict__ psi, int Nx, int Ny,
                                   const double* __restrict__ dy_edge, double psi_bot,
                                   double psi_top, int free_slip, int i, double& wb, double& wt) {
    if (free_slip) { wb = 0.0; wt = 0.0; return; }
    {
        const double dy1 = dy_edge[0], dy2 = dy_edge[1];
        const double A = 2.0 * (1.0 / (dy1 * dy1) + 1.0 / (dy1 * dy2));
        const double B = (-2.0) * dy1 / (dy2 * (dy1 + dy2) * (dy1 + dy2));
        const double d2 = A * (psi[idx2(1, i, Nx)] - psi_bot) + B * (psi[idx2(2, i, Nx)] - psi_bot);
        wb = -d2;
    }
    {
        const double dy1 = dy_edge[Ny - 2], dy2 = dy_edge[Ny - 3];
        const double A = 2.0 * (1.0 / (dy1 * dy1) + 1.0 / (dy1 * dy2));
        const double B = (-2.0) * dy1 / (dy2 * (dy1 + dy2) * (dy1 + dy2));
        const double d2 = A * (psi[idx2(Ny - 2, i, Nx)] - psi_top) +
                          B * (psi[idx2(Ny - 3, i, Nx)] - psi_top);
        wt = -d2;
    }
}

// diag.cu:702-816: maxima of |u|, |v| and |v| a over interior row j (rows 0 and Ny-1 give 0), by one
// workgroup (blockDim.x a multiple of 64); partial[a * Ny + j] receives maximum a. `sm` needs 48 doubles.
__device__ __forceinline__ void row_max_uv_body(const double* __restrict__ psi, int Nx, int Ny, double inv_dx,
                                       double inv_deta, const double* __restrict__ a_node, int j,
                                       double* __restrict__ partial, double* sm) {
    double m[3] = {0, 0, 0};
    if (j > 0 && j < Ny - 1) {
        const double a = a_node[j];
        for (int i = threadIdx.x; i < Nx; i += blockDim.x) {
            const int ip = (i + 1 < Nx) ? (i + 1) : 0, im = (i > 0) ? (i - 1) : (Nx - 1);
            const double dpsideta = (psi[idx2(j + 1, i, Nx)] - psi[idx2(j - 1, i, Nx)]) * (0.5 * inv_deta);
            const double dpsidx = (psi[idx2(j, ip, Nx)] - psi[idx2(j, im, Nx)]) * (0.5 * inv_dx);
            const double au = fabs(a * dpsideta), av = fabs(-dpsidx);
            m[0] = fmax(m[0], au);
            m[1] = fmax(m[1], av);
            m[2] = fmax(m[2], av * a);
        }
    }
    block_max<3>(m, sm);
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; ++a) partial[(size_t)a * Ny + j] = m[a];
}


void launch_thom(double* w, const double* psi, int Nx, int Ny, const double* dy_edge,
                 double psi_bot, double psi_top, int free_slip, hipStream_t s);

}  // namespace nse2d
