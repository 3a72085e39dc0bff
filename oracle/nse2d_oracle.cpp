// nse2d oracle: FP64 OpenMP C++ restatement of the reference solver's per-step path.
// TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); see nse2d_oracle.hpp.
//
// Reductions are summed per row and then over rows in index order, so results do not depend on
// the OpenMP thread count.
#include "nse2d_oracle.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using Real = double;
const double PI = 3.14159265358979323846264338327950288;

inline size_t ID(int j, int i, int Nx) { return (size_t)j * (size_t)Nx + (size_t)i; }
inline int wrapx(int i, int Nx) { return i < 0 ? i + Nx : (i >= Nx ? i - Nx : i); }

// Sum of per-row partials in row order: deterministic for any thread count.
template <class F>
double row_sum(int j0, int j1, F f) {
    std::vector<double> part((size_t)std::max(0, j1 - j0), 0.0);
#pragma omp parallel for schedule(static)
    for (int j = j0; j < j1; ++j) part[(size_t)(j - j0)] = f(j);
    double s = 0.0;
    for (double p : part) s += p;
    return s;
}

}  // namespace

// =============================================================================================
// grid.cu:7-94
// =============================================================================================
namespace {
// edges, spacings, trapezoid weights, grid.cu:83-93
void edges_and_weights(int Ny, const double* y, const double* a_node, double* a_edge, double* dy_edge,
                       double* w_node) {
    for (int j = 0; j < Ny - 1; ++j) a_edge[j] = 0.5 * (a_node[j] + a_node[j + 1]);
    for (int j = 0; j < Ny - 1; ++j) dy_edge[j] = y[j + 1] - y[j];
    for (int j = 0; j < Ny; ++j) w_node[j] = 0.0;
    w_node[0] = 0.5 * dy_edge[0];
    w_node[Ny - 1] = 0.5 * dy_edge[Ny - 2];
    for (int j = 1; j < Ny - 1; ++j) w_node[j] = 0.5 * (dy_edge[j - 1] + dy_edge[j]);
}
}  // namespace

extern "C" int orc_build_y_metrics(int Ny, double h, int stretch, double beta, double* y,
                                   double* a_node, double* a_edge, double* dy_edge,
                                   double* w_node) {
    if (Ny < 2) return -1;
    if (Ny <= 2 || stretch == 0) {  // uniform branch, grid.cu:10-27
        const double dy = (2.0 * h) / (double)(Ny - 1);
        for (int j = 0; j < Ny; ++j) y[j] = -h + j * dy;
        const double deta = 2.0 / (double)(Ny - 1);
        const double y_eta = dy / deta;
        const double a = 1.0 / y_eta;
        for (int j = 0; j < Ny; ++j) a_node[j] = a;
        for (int j = 0; j < Ny - 1; ++j) { a_edge[j] = a; dy_edge[j] = dy; }
        for (int j = 0; j < Ny; ++j) w_node[j] = dy;
        w_node[0] *= 0.5; w_node[Ny - 1] *= 0.5;
        return 0;
    }
    if (stretch != 1) return -2;  // tabulated grids: orc_table_y_metrics
    // tanh branch, grid.cu:33-45
    const double b = (beta > 0 ? beta : 2.5);
    for (int j = 0; j < Ny; ++j) {
        const double e = -1.0 + 2.0 * (double)j / (double)(Ny - 1);
        y[j] = h * std::tanh(b * e) / std::tanh(b);
    }
    for (int j = 0; j < Ny; ++j) {
        const double e = -1.0 + 2.0 * (double)j / (double)(Ny - 1);
        const double sech = 1.0 / std::cosh(b * e);
        const double y_eta = h * b * (sech * sech) / std::tanh(b);
        a_node[j] = 1.0 / y_eta;
    }
    edges_and_weights(Ny, y, a_node, a_edge, dy_edge, w_node);
    return 0;
}

// Tabulated branch, grid.cu:46-78: y given (the reference reads dataset /y from an HDF5 file;
// the caller passes the Ny values), a_node = 1/y_eta with y_eta from centred differences in eta
// and one-sided ones at the two walls.
extern "C" int orc_table_y_metrics(int Ny, const double* y_in, double* y, double* a_node, double* a_edge,
                                   double* dy_edge, double* w_node) {
    if (Ny < 3) return -1;
    for (int j = 0; j < Ny; ++j) y[j] = y_in[j];
    const double deta = 2.0 / (double)(Ny - 1);
    for (int j = 0; j < Ny; ++j) {
        double yeta;
        if (j == 0) yeta = (y[1] - y[0]) / deta;
        else if (j == Ny - 1) yeta = (y[Ny - 1] - y[Ny - 2]) / deta;
        else yeta = (y[j + 1] - y[j - 1]) / (2.0 * deta);
        a_node[j] = 1.0 / yeta;
    }
    edges_and_weights(Ny, y, a_node, a_edge, dy_edge, w_node);
    return 0;
}

// =============================================================================================
// ops.cu
// =============================================================================================
extern "C" void orc_uv_from_psi_mapped(const double* psi, double* u, double* v, int Nx, int Ny,
                                       double inv_dx, double inv_deta, const double* a_node) {
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < Ny; ++iy) {
        const int ym = (iy == 0 ? 0 : iy - 1), yp = (iy == Ny - 1 ? Ny - 1 : iy + 1);
        for (int ix = 0; ix < Nx; ++ix) {
            const int xm = wrapx(ix - 1, Nx), xp = wrapx(ix + 1, Nx);
            const double dpsideta = (psi[ID(yp, ix, Nx)] - psi[ID(ym, ix, Nx)]) * 0.5 * inv_deta;
            const double dpsidx = (psi[ID(iy, xp, Nx)] - psi[ID(iy, xm, Nx)]) * 0.5 * inv_dx;
            u[ID(iy, ix, Nx)] = a_node[iy] * dpsideta;
            v[ID(iy, ix, Nx)] = -dpsidx;
        }
    }
}

extern "C" void orc_omega_from_psi_interior_mapped(double* w, const double* psi, int Nx, int Ny,
                                                   double inv_dx2, double inv_deta2,
                                                   const double* a_node, const double* a_edge) {
#pragma omp parallel for schedule(static)
    for (int iy = 1; iy < Ny - 1; ++iy) {
        for (int ix = 0; ix < Nx; ++ix) {
            const int xm = wrapx(ix - 1, Nx), xp = wrapx(ix + 1, Nx);
            const size_t idx = ID(iy, ix, Nx);
            const double lapx = psi[ID(iy, xp, Nx)] - 2.0 * psi[idx] + psi[ID(iy, xm, Nx)];
            const double aN = a_node[iy], ap = a_edge[iy], am = a_edge[iy - 1];
            const double lapy = aN * (ap * (psi[ID(iy + 1, ix, Nx)] - psi[idx]) -
                                      am * (psi[idx] - psi[ID(iy - 1, ix, Nx)]));
            w[idx] = -(inv_dx2 * lapx + inv_deta2 * lapy);
        }
    }
}

extern "C" void orc_thom_wall_vorticity_mapped(double* w, const double* psi, int Nx, int Ny,
                                               const double* dy_edge, double psi_bot,
                                               double psi_top, int free_slip) {
    for (int i = 0; i < Nx; ++i) {
        if (free_slip) {  // ops.cu:108-111
            w[i] = 0;
            w[ID(Ny - 1, i, Nx)] = 0;
            continue;
        }
        {  // bottom wall, ops.cu:115-132
            const double dy1 = dy_edge[0], dy2 = dy_edge[1];
            const double A = 2.0 * (1.0 / (dy1 * dy1) + 1.0 / (dy1 * dy2));
            const double B = (-2.0) * dy1 / (dy2 * (dy1 + dy2) * (dy1 + dy2));
            const double d2psi = A * (psi[ID(1, i, Nx)] - psi_bot) + B * (psi[ID(2, i, Nx)] - psi_bot);
            w[i] = -d2psi;
        }
        {  // top wall, ops.cu:135-151
            const double dy1 = dy_edge[Ny - 2], dy2 = dy_edge[Ny - 3];
            const double A = 2.0 * (1.0 / (dy1 * dy1) + 1.0 / (dy1 * dy2));
            const double B = (-2.0) * dy1 / (dy2 * (dy1 + dy2) * (dy1 + dy2));
            const double d2psi = A * (psi[ID(Ny - 2, i, Nx)] - psi_top) +
                                 B * (psi[ID(Ny - 3, i, Nx)] - psi_top);
            w[ID(Ny - 1, i, Nx)] = -d2psi;
        }
    }
}

namespace {
// 9-point mapped Arakawa terms at an interior point, ops.cu:305-322.
inline void arakawa_terms(const double* psi, const double* omg, int Nx, int ix, int iy,
                          double inv_dx, double inv_deta, double a, double& J1, double& J2,
                          double& J3) {
    const int xm = wrapx(ix - 1, Nx), xp = wrapx(ix + 1, Nx), ym = iy - 1, yp = iy + 1;
    const size_t cxp = ID(iy, xp, Nx), cxm = ID(iy, xm, Nx);
    const size_t cyp = ID(yp, ix, Nx), cym = ID(ym, ix, Nx);
    const size_t ip_jp = ID(yp, xp, Nx), ip_jm = ID(ym, xp, Nx);
    const size_t im_jp = ID(yp, xm, Nx), im_jm = ID(ym, xm, Nx);
    const double inv4dxde = 0.25 * inv_dx * inv_deta;
    J1 = ((psi[cxp] - psi[cxm]) * (a * (omg[cyp] - omg[cym])) -
          (a * (psi[cyp] - psi[cym])) * (omg[cxp] - omg[cxm])) * inv4dxde;
    J2 = (psi[cxp] * (a * (omg[ip_jp] - omg[ip_jm])) - psi[cxm] * (a * (omg[im_jp] - omg[im_jm])) -
          a * (psi[cyp] * (omg[ip_jp] - omg[im_jp]) - psi[cym] * (omg[ip_jm] - omg[im_jm]))) *
         inv4dxde;
    J3 = (-omg[cxp] * (a * (psi[ip_jp] - psi[ip_jm])) + omg[cxm] * (a * (psi[im_jp] - psi[im_jm])) +
          a * (omg[cyp] * (psi[ip_jp] - psi[im_jp]) - omg[cym] * (psi[ip_jm] - psi[im_jm]))) *
         inv4dxde;
}
}  // namespace

extern "C" void orc_arakawa_J_mapped(const double* psi, const double* w, double* J, int Nx,
                                     int Ny, double inv_dx, double inv_deta,
                                     const double* a_node) {
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < Ny; ++iy) {
        for (int ix = 0; ix < Nx; ++ix) {
            if (iy == 0 || iy == Ny - 1) { J[ID(iy, ix, Nx)] = 0.0; continue; }
            double J1, J2, J3;
            arakawa_terms(psi, w, Nx, ix, iy, inv_dx, inv_deta, a_node[iy], J1, J2, J3);
            const double Jbr = (J1 + J2 + J3) * (1.0 / 3.0);
            J[ID(iy, ix, Nx)] = -Jbr;
        }
    }
}

extern "C" void orc_arakawa_J_mapped_split(const double* psi, const double* w, double* J1o,
                                           double* J2o, double* J3o, double* Jadv, int Nx,
                                           int Ny, double inv_dx, double inv_deta,
                                           const double* a_node) {
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < Ny; ++iy) {
        for (int ix = 0; ix < Nx; ++ix) {
            const size_t c = ID(iy, ix, Nx);
            if (iy == 0 || iy == Ny - 1) {
                if (J1o) J1o[c] = 0; if (J2o) J2o[c] = 0; if (J3o) J3o[c] = 0; if (Jadv) Jadv[c] = 0;
                continue;
            }
            double J1, J2, J3;
            arakawa_terms(psi, w, Nx, ix, iy, inv_dx, inv_deta, a_node[iy], J1, J2, J3);
            if (J1o) J1o[c] = J1;
            if (J2o) J2o[c] = J2;
            if (J3o) J3o[c] = J3;
            const double Jbr = (J1 + J2 + J3) * (1.0 / 3.0);
            if (Jadv) Jadv[c] = -Jbr;
        }
    }
}

extern "C" void orc_compute_adv_from_uv_mapped(const double* u, const double* v, const double* w,
                                               double* Jadv, int Nx, int Ny, double inv_dx,
                                               double inv_deta, const double* a_node) {
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < Ny; ++iy) {
        for (int ix = 0; ix < Nx; ++ix) {
            const size_t c = ID(iy, ix, Nx);
            if (iy == 0 || iy == Ny - 1) { Jadv[c] = 0; continue; }
            const int xm = wrapx(ix - 1, Nx), xp = wrapx(ix + 1, Nx);
            const double inv2dx = 0.5 * inv_dx, inv2de = 0.5 * inv_deta;
            const double a = a_node[iy];
            const double dx_omg = (w[ID(iy, xp, Nx)] - w[ID(iy, xm, Nx)]) * inv2dx;
            const double dy_omg = a * (w[ID(iy + 1, ix, Nx)] - w[ID(iy - 1, ix, Nx)]) * inv2de;
            Jadv[c] = u[c] * dx_omg + v[c] * dy_omg;
        }
    }
}

extern "C" void orc_compute_adv_flux_mapped(const double* u, const double* v, const double* w,
                                            double* Jadv, int Nx, int Ny, double inv_dx,
                                            double inv_deta, const double* a_node) {
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < Ny; ++iy) {
        for (int ix = 0; ix < Nx; ++ix) {
            const size_t c = ID(iy, ix, Nx);
            if (iy == 0 || iy == Ny - 1) { Jadv[c] = 0; continue; }
            const int xm = wrapx(ix - 1, Nx), xp = wrapx(ix + 1, Nx);
            const size_t cxm = ID(iy, xm, Nx), cxp = ID(iy, xp, Nx);
            const size_t cym = ID(iy - 1, ix, Nx), cyp = ID(iy + 1, ix, Nx);
            const double u_ip = (u[cxp] + u[c]) * 0.5, u_im = (u[c] + u[cxm]) * 0.5;
            const double w_ip = (w[cxp] + w[c]) * 0.5, w_im = (w[c] + w[cxm]) * 0.5;
            const double dFdx = (u_ip * w_ip - u_im * w_im) * inv_dx;
            const double v_jp = (v[cyp] + v[c]) * 0.5, v_jm = (v[c] + v[cym]) * 0.5;
            const double w_jp = (w[cyp] + w[c]) * 0.5, w_jm = (w[c] + w[cym]) * 0.5;
            const double dFdy = a_node[iy] * ((v_jp * w_jp - v_jm * w_jm) * (0.5 * inv_deta * 2.0));
            Jadv[c] = dFdx + dFdy;
        }
    }
}

extern "C" void orc_laplacian_fd_mapped(const double* q, double* out, int Nx, int Ny,
                                        double inv_dx2, double inv_deta2, const double* a_node,
                                        const double* a_edge) {
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < Ny; ++iy) {
        for (int ix = 0; ix < Nx; ++ix) {
            const size_t id = ID(iy, ix, Nx);
            if (iy == 0 || iy == Ny - 1) { out[id] = 0.0; continue; }
            const int xm = wrapx(ix - 1, Nx), xp = wrapx(ix + 1, Nx);
            const double dxx = (q[ID(iy, xp, Nx)] - 2 * q[id] + q[ID(iy, xm, Nx)]) * inv_dx2;
            const double a = a_node[iy], ap = a_edge[iy], am = a_edge[iy - 1];
            const double dyy = a * (ap * (q[ID(iy + 1, ix, Nx)] - q[id]) -
                                    am * (q[id] - q[ID(iy - 1, ix, Nx)])) * inv_deta2;
            out[id] = dxx + dyy;
        }
    }
}

extern "C" void orc_rhs_mapped(const double* J, const double* lap, const double* w, double* rhs,
                               int Nx, int Ny, double nu, double F0, int nforce, double h,
                               double lin_drag, const double* y_node) {
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < Ny; ++iy) {
        double f = 0.0;
        if (F0 != 0.0) {
            const double y = y_node[iy];
            const double arg = (double)nforce * PI * ((y + h) / (2.0 * h));
            f = -F0 * ((double)nforce * PI / (2.0 * h)) * std::cos(arg);
        }
        for (int ix = 0; ix < Nx; ++ix) {
            const size_t idx = ID(iy, ix, Nx);
            rhs[idx] = -J[idx] + nu * lap[idx] + f - lin_drag * w[idx];
        }
    }
}

extern "C" void orc_mms_add_forcing(double* rhs, int Nx, int Ny, double Lx, double h, double t,
                                    double nu, double r, double A0, double eps, double Om,
                                    int kx_mode, const double* y_node) {
    const double dx = Lx / (double)Nx;
    const double k = 2.0 * PI * (double)kx_mode / Lx;
    const double c = PI / (2.0 * h);
    const double A = A0 * (1.0 + eps * std::sin(Om * t));
    const double Aprime = A0 * (eps * Om * std::cos(Om * t));
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Ny; ++j) {
        for (int i = 0; i < Nx; ++i) {
            const double y = y_node[j];
            const double x = (double)i * dx;
            const double theta = c * (y + h);
            const double two_th = 2.0 * theta;
            const double s_th = std::sin(theta);
            const double s = s_th * s_th;
            const double cos2 = std::cos(two_th);
            const double sin2 = std::sin(two_th);
            const double c2 = c * c;
            const double g = (k * k) * s - 2.0 * c2 * cos2;
            const double gpp = 2.0 * c2 * cos2 * ((k * k) + 4.0 * c2);
            const double sk = std::sin(k * x);
            const double s2k = std::sin(2.0 * k * x);
            const double term_lin = Aprime * g + (nu * A) * ((k * k) * g - gpp) + (r * A) * g;
            const double c3 = c2 * c;
            const double term_adv = -(A * A) * k * c3 * s2k * sin2;
            rhs[ID(j, i, Nx)] += sk * term_lin + term_adv;
        }
    }
}

extern "C" void orc_axpy(double* y, const double* x, const double* r, double dt, size_t N) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)N; ++i) y[i] = x[i] + dt * r[i];
}
extern "C" void orc_ssprk2(double* w2, const double* w, const double* w1, const double* rhs,
                           double dt, size_t N) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)N; ++i) w2[i] = 0.75 * w[i] + 0.25 * (w1[i] + dt * rhs[i]);
}
extern "C" void orc_ssprk3(double* wnp1, const double* w, const double* w2, const double* rhs,
                           double dt, size_t N) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)N; ++i)
        wnp1[i] = (1.0 / 3.0) * w[i] + (2.0 / 3.0) * (w2[i] + dt * rhs[i]);
}

// =============================================================================================
// Poisson: poisson.cu + poisson.cuh:83-118, with the project's own FFT in place of cuFFT.
// =============================================================================================
namespace {
// Evaluation order of the solve (bit flags): 1 = power-of-two FFTs by decimation in frequency
// instead of in time, 2 = y elimination from the top wall down instead of from the bottom up. A
// measurement tool with no reference counterpart: two FP64 evaluations of the same solver that
// round differently show how closely any two FP64 implementations can agree.
std::atomic<int> g_eval_order{0};

struct FFTPlan {
    int N = 0;
    bool pow2 = false;
    std::vector<std::complex<double>> tw;  // exp(-2 pi i k / N), k < N
    std::vector<int> rev;
    explicit FFTPlan(int n) : N(n) {
        pow2 = n > 0 && (n & (n - 1)) == 0;
        tw.resize(n);
        for (int k = 0; k < n; ++k) {
            const double ang = -2.0 * PI * (double)k / (double)n;
            tw[k] = std::complex<double>(std::cos(ang), std::sin(ang));
        }
        if (pow2) {
            rev.resize(n);
            int lg = 0;
            while ((1 << lg) < n) ++lg;
            for (int i = 0; i < n; ++i) {
                int r = 0;
                for (int b = 0; b < lg; ++b) r |= ((i >> b) & 1) << (lg - 1 - b);
                rev[i] = r;
            }
        }
    }
    // In place on one row of N complex values.
    void run(std::complex<double>* a, bool inverse, std::vector<std::complex<double>>& scratch) const {
        if (!pow2) {
            scratch.assign(a, a + N);
            for (int k = 0; k < N; ++k) {
                std::complex<double> s(0, 0);
                for (int n = 0; n < N; ++n) {
                    const long long m = ((long long)k * n) % N;
                    const std::complex<double> w = inverse ? std::conj(tw[(size_t)m]) : tw[(size_t)m];
                    s += scratch[(size_t)n] * w;
                }
                a[k] = s;
            }
            return;
        }
        if (g_eval_order.load(std::memory_order_relaxed) & 1) {
            // decimation in frequency (Gentleman-Sande): butterflies from the full length down,
            // the bit-reversal permutation last. The same transform as the decimation-in-time
            // loop below, with every value rounded in another order.
            for (int len = N; len >= 2; len >>= 1) {
                const int half = len >> 1, step = N / len;
                for (int s = 0; s < N; s += len) {
                    for (int k = 0; k < half; ++k) {
                        const std::complex<double> w = inverse ? std::conj(tw[(size_t)k * step]) : tw[(size_t)k * step];
                        const std::complex<double> u = a[s + k], v = a[s + k + half];
                        a[s + k] = u + v;
                        a[s + k + half] = (u - v) * w;
                    }
                }
            }
            for (int i = 0; i < N; ++i)
                if (i < rev[i]) std::swap(a[i], a[rev[i]]);
            return;
        }
        for (int i = 0; i < N; ++i)
            if (i < rev[i]) std::swap(a[i], a[rev[i]]);
        for (int len = 2; len <= N; len <<= 1) {
            const int half = len >> 1, step = N / len;
            for (int s = 0; s < N; s += len) {
                for (int k = 0; k < half; ++k) {
                    const std::complex<double> w = inverse ? std::conj(tw[(size_t)k * step]) : tw[(size_t)k * step];
                    const std::complex<double> u = a[s + k];
                    const std::complex<double> v = a[s + k + half] * w;
                    a[s + k] = u + v;
                    a[s + k + half] = u - v;
                }
            }
        }
    }
};
}  // namespace

extern "C" void orc_set_eval_order(int flags) { g_eval_order.store(flags & 3); }
extern "C" int orc_get_eval_order(void) { return g_eval_order.load(); }

extern "C" void orc_fft_z2z(double* data, int N, int howmany, int inverse) {
    FFTPlan plan(N);
#pragma omp parallel
    {
        std::vector<std::complex<double>> scratch;
#pragma omp for schedule(static)
        for (int r = 0; r < howmany; ++r)
            plan.run(reinterpret_cast<std::complex<double>*>(data) + (size_t)r * N, inverse != 0, scratch);
    }
}

extern "C" void orc_dft_naive(const double* in, double* out, int N, int inverse) {
    const double sgn = inverse ? 1.0 : -1.0;
    for (int k = 0; k < N; ++k) {
        long double sr = 0, si = 0;
        for (int n = 0; n < N; ++n) {
            const long long m = ((long long)k * n) % N;
            const long double ang = sgn * 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)N;
            const long double c = std::cos(ang), s = std::sin(ang);
            sr += in[2 * n] * c - in[2 * n + 1] * s;
            si += in[2 * n] * s + in[2 * n + 1] * c;
        }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
}

extern "C" void orc_fill_kx2(double* kx2, int Nx, double Lx) {
    for (int k = 0; k < Nx; ++k) {
        const int kphys_i = (k <= Nx / 2 ? k : k - Nx);
        const double kphys = (2.0 * PI * (double)kphys_i) / Lx;
        kx2[k] = kphys * kphys;
    }
}

extern "C" void orc_precompute_cp_mapped(double* cp, const double* a_node, const double* a_edge,
                                         const double* kx2, int Nx, int Ny, double inv_deta2) {
    const int Ni = Ny - 2;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Nx; ++k) {
        int j = 1;
        double an = a_node[j], ae_p = a_edge[j], ae_m = a_edge[j - 1];
        double a_sub = an * ae_m * inv_deta2, c_sup = an * ae_p * inv_deta2;
        double b = -(a_sub + c_sup) - kx2[k];
        double cp_prev = c_sup / b;
        cp[(size_t)k * Ni + 0] = cp_prev;
        for (int r = 1; r < Ni; ++r) {
            j = r + 1;
            an = a_node[j]; ae_p = a_edge[j]; ae_m = a_edge[j - 1];
            a_sub = an * ae_m * inv_deta2;
            c_sup = an * ae_p * inv_deta2;
            b = -(a_sub + c_sup) - kx2[k];
            const double den = b - a_sub * cp_prev;
            const double cpi = c_sup / den;
            cp[(size_t)k * Ni + r] = cpi;
            cp_prev = cpi;
        }
    }
}

extern "C" void orc_poisson_tridiag_mapped(double* hat, const double* cp, const double* a_node,
                                           const double* a_edge, const double* kx2, int Nx,
                                           int Ny, double inv_deta2, double psi_bot,
                                           double psi_top) {
    typedef std::complex<double> C;
    C* H = reinterpret_cast<C*>(hat);
    const int Ni = Ny - 2;
    if (g_eval_order.load() & 2) {
        // The same systems eliminated from the top interior row down (x_r = e_r - q_r x_{r-1},
        // q_r = a_r / den_r, e_r = (f_r - c_r e_{r+1}) / den_r, den_r = b_r - c_r q_{r+1}), then
        // substituted upward: exact in real arithmetic, rounded in another order.
#pragma omp parallel
        {
            std::vector<double> q(Ni);
            std::vector<C> e(Ni);
#pragma omp for schedule(static)
            for (int k = 0; k < Nx; ++k) {
                C* ph = H + k;
                double qn = 0.0;
                C en(0.0, 0.0);
                for (int r = Ni - 1; r >= 0; --r) {
                    const int j = r + 1;
                    const double a_sub = a_node[j] * a_edge[j - 1] * inv_deta2;
                    const double c_sup = a_node[j] * a_edge[j] * inv_deta2;
                    const double b = -(a_sub + c_sup) - kx2[k];
                    C f = -ph[(size_t)Nx * j];
                    if (k == 0 && r == 0) f = C(f.real() - a_sub * psi_bot * (double)Nx, f.imag());
                    if (k == 0 && r == Ni - 1) f = C(f.real() - c_sup * psi_top * (double)Nx, f.imag());
                    const double den = b - c_sup * qn;
                    q[r] = a_sub / den;
                    e[r] = C((f.real() - c_sup * en.real()) / den, (f.imag() - c_sup * en.imag()) / den);
                    qn = q[r];
                    en = e[r];
                }
                C xp = e[0];
                ph[(size_t)Nx] = xp;
                for (int r = 1; r < Ni; ++r) {
                    xp = C(e[r].real() - q[r] * xp.real(), e[r].imag() - q[r] * xp.imag());
                    ph[(size_t)Nx * (r + 1)] = xp;
                }
                ph[0] = (k == 0 ? C((double)Nx * psi_bot, 0.0) : C(0.0, 0.0));
                ph[(size_t)Nx * (Ny - 1)] = (k == 0 ? C((double)Nx * psi_top, 0.0) : C(0.0, 0.0));
            }
        }
        return;
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Nx; ++k) {
        C* ph = H + k;  // stride Nx
        int j = 1;
        double an = a_node[j], ae_p = a_edge[j], ae_m = a_edge[j - 1];
        double a_sub = an * ae_m * inv_deta2, c_sup = an * ae_p * inv_deta2;
        double b = -(a_sub + c_sup) - kx2[k];
        C dp_prev(0.0, 0.0);
        C rhs0 = ph[(size_t)Nx * j];
        rhs0 = C(-rhs0.real(), -rhs0.imag());
        if (k == 0) rhs0 = C(rhs0.real() - a_sub * psi_bot * (double)Nx, rhs0.imag());
        const double den0 = b;
        C dp0((rhs0.real() - a_sub * dp_prev.real()) / den0, (rhs0.imag() - a_sub * dp_prev.imag()) / den0);
        ph[(size_t)Nx * j] = dp0;
        dp_prev = dp0;
        for (int r = 1; r < Ni; ++r) {
            j = r + 1;
            an = a_node[j]; ae_p = a_edge[j]; ae_m = a_edge[j - 1];
            a_sub = an * ae_m * inv_deta2;
            c_sup = an * ae_p * inv_deta2;
            C rhs = ph[(size_t)Nx * j];
            rhs = C(-rhs.real(), -rhs.imag());
            if (k == 0 && r == Ni - 1) rhs = C(rhs.real() - c_sup * psi_top * (double)Nx, rhs.imag());
            const double cpj = cp[(size_t)k * Ni + r];
            const double den = c_sup / cpj;  // poisson.cu:117, equals b - a_sub*cp_prev
            C dpj((rhs.real() - a_sub * dp_prev.real()) / den, (rhs.imag() - a_sub * dp_prev.imag()) / den);
            ph[(size_t)Nx * j] = dpj;
            dp_prev = dpj;
        }
        C xnext = ph[(size_t)Nx * Ni];
        for (int r = Ni - 2; r >= 0; --r) {
            j = r + 1;
            const double cpj = cp[(size_t)k * Ni + r];
            const C dpj = ph[(size_t)Nx * j];
            const C xj(dpj.real() - cpj * xnext.real(), dpj.imag() - cpj * xnext.imag());
            ph[(size_t)Nx * j] = xj;
            xnext = xj;
        }
        ph[0] = (k == 0 ? C((double)Nx * psi_bot, 0.0) : C(0.0, 0.0));
        ph[(size_t)Nx * (Ny - 1)] = (k == 0 ? C((double)Nx * psi_top, 0.0) : C(0.0, 0.0));
    }
}

namespace {
// Poisson2D_Z2Z with enable_mapped (poisson.cuh:18-80): owns kx2, cp and the spectral buffer.
struct OPoisson {
    int Nx = 0, Ny = 0;
    std::vector<double> kx2, cp, hat;
    const double *a_node = nullptr, *a_edge = nullptr;
    double inv_deta2 = 0;
    FFTPlan plan;
    bool extended = false;  // solve in 80-bit arithmetic instead (measurement runs, --poisson_ld)
    double Lx_ = 0, inv_deta_ = 0;
    OPoisson(int nx, int ny, double Lx, const double* an, const double* ae, double inv_deta)
        : Nx(nx), Ny(ny), plan(nx), Lx_(Lx), inv_deta_(inv_deta) {
        kx2.resize(Nx);
        orc_fill_kx2(kx2.data(), Nx, Lx);
        cp.resize((size_t)Nx * (size_t)std::max(0, Ny - 2));
        a_node = an; a_edge = ae; inv_deta2 = inv_deta * inv_deta;
        orc_precompute_cp_mapped(cp.data(), a_node, a_edge, kx2.data(), Nx, Ny, inv_deta2);
        hat.resize((size_t)2 * Nx * Ny);
    }
    void solve(const double* omega, double* psi, double psi_bot, double psi_top) {
        if (extended) {
            orc_poisson_solve_ld(omega, psi, Nx, Ny, Lx_, a_node, a_edge, inv_deta_, psi_bot, psi_top);
            return;
        }
        typedef std::complex<double> C;
        C* H = reinterpret_cast<C*>(hat.data());
#pragma omp parallel
        {
            std::vector<C> scratch;
#pragma omp for schedule(static)
            for (int j = 0; j < Ny; ++j) {  // real_to_complex + forward FFT (poisson.cu:8-11, poisson.cuh:96)
                C* row = H + (size_t)j * Nx;
                for (int i = 0; i < Nx; ++i) row[i] = C(omega[ID(j, i, Nx)], 0.0);
                plan.run(row, false, scratch);
            }
        }
        orc_poisson_tridiag_mapped(hat.data(), cp.data(), a_node, a_edge, kx2.data(), Nx, Ny,
                                   inv_deta2, psi_bot, psi_top);
        const double s = (double)(1.0 / (double)Nx);
#pragma omp parallel
        {
            std::vector<C> scratch;
#pragma omp for schedule(static)
            for (int j = 0; j < Ny; ++j) {  // inverse FFT + complex_to_real_scaled (poisson.cu:12-15)
                C* row = H + (size_t)j * Nx;
                plan.run(row, true, scratch);
                for (int i = 0; i < Nx; ++i) psi[ID(j, i, Nx)] = s * row[i].real();
            }
        }
    }
};
}  // namespace

extern "C" void orc_poisson_solve(const double* omega, double* psi, int Nx, int Ny, double Lx,
                                  const double* a_node, const double* a_edge, double inv_deta,
                                  double psi_bot, double psi_top) {
    OPoisson P(Nx, Ny, Lx, a_node, a_edge, inv_deta);
    P.solve(omega, psi, psi_bot, psi_top);
}

// Extended-precision (80-bit long double) solve of the same discrete operator: coefficients as
// the reference forms them in double (a_sub, c_sup, kx^2), FFTs and elimination in long double.
// A measurement tool, not a restatement: it shows how far any FP64 evaluation of the solve, the
// reference's included, is from the exact solution of the operator it defines.
namespace {
typedef std::complex<long double> CL;
void fft_ld(CL* a, int N, bool inverse) {
    for (int i = 1, j = 0; i < N; ++i) {  // bit reversal
        int bit = N >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    const long double pi = 3.14159265358979323846264338327950288L;
    for (int len = 2; len <= N; len <<= 1) {
        const int half = len >> 1;
        for (int k = 0; k < half; ++k) {
            const long double ang = (inverse ? 2.0L : -2.0L) * pi * (long double)k / (long double)len;
            const CL w(std::cos(ang), std::sin(ang));
            for (int s = 0; s < N; s += len) {
                const CL u = a[s + k], v = a[s + k + half] * w;
                a[s + k] = u + v;
                a[s + k + half] = u - v;
            }
        }
    }
}
}  // namespace

extern "C" void orc_poisson_solve_ld(const double* omega, double* psi, int Nx, int Ny, double Lx,
                                     const double* a_node, const double* a_edge, double inv_deta,
                                     double psi_bot, double psi_top) {
    if ((Nx & (Nx - 1)) != 0) return;
    const int Ni = Ny - 2;
    const double inv_deta2 = inv_deta * inv_deta;
    std::vector<double> kx2(Nx);
    orc_fill_kx2(kx2.data(), Nx, Lx);
    std::vector<CL> H((size_t)Nx * Ny);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Ny; ++j) {
        CL* row = H.data() + (size_t)j * Nx;
        for (int i = 0; i < Nx; ++i) row[i] = CL(omega[ID(j, i, Nx)], 0.0L);
        fft_ld(row, Nx, false);
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Nx; ++k) {
        std::vector<long double> cp(Ni);
        std::vector<CL> dp(Ni);
        long double cp_prev = 0.0L;
        CL dp_prev(0.0L, 0.0L);
        for (int r = 0; r < Ni; ++r) {
            const int j = r + 1;
            const double a_sub = a_node[j] * a_edge[j - 1] * inv_deta2, c_sup = a_node[j] * a_edge[j] * inv_deta2;
            const long double b = -((long double)a_sub + (long double)c_sup) - (long double)kx2[k];
            CL rhs = -H[(size_t)j * Nx + k];
            if (k == 0 && r == 0) rhs -= CL((long double)a_sub * psi_bot * (long double)Nx, 0.0L);
            if (k == 0 && r == Ni - 1) rhs -= CL((long double)c_sup * psi_top * (long double)Nx, 0.0L);
            const long double den = b - (long double)a_sub * cp_prev;
            cp[r] = (long double)c_sup / den;
            dp[r] = (rhs - (long double)a_sub * dp_prev) / den;
            cp_prev = cp[r];
            dp_prev = dp[r];
        }
        CL x = dp[Ni - 1];
        H[(size_t)Ni * Nx + k] = x;
        for (int r = Ni - 2; r >= 0; --r) {
            x = dp[r] - cp[r] * x;
            H[(size_t)(r + 1) * Nx + k] = x;
        }
        H[k] = (k == 0 ? CL((long double)Nx * psi_bot, 0.0L) : CL(0.0L, 0.0L));
        H[(size_t)(Ny - 1) * Nx + k] = (k == 0 ? CL((long double)Nx * psi_top, 0.0L) : CL(0.0L, 0.0L));
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Ny; ++j) {
        CL* row = H.data() + (size_t)j * Nx;
        fft_ld(row, Nx, true);
        for (int i = 0; i < Nx; ++i) psi[ID(j, i, Nx)] = (double)(row[i].real() / (long double)Nx);
    }
}

// Spectra1D::accumulate (diag.cuh:165-174) with spectra1d_rowReduce_kernel (diag.cu:876-893):
// D2Z of every row of u and v (the real rows as complex rows, k = 0..Nx/2 kept), then
// acc[k] += norm1d * wx * sum_j |X_j(k)|^2, norm1d = 1/(Nx^2 Ny), wx = 2 for 0 < k < Nx/2.
// acc holds E_u in [0, Nx/2] and E_v in [Nx/2+1, Nx+1].
extern "C" void orc_spectra1d_accumulate(const double* u, const double* v, int Nx, int Ny, double* acc) {
    typedef std::complex<double> C;
    const int nk = Nx / 2 + 1;
    const double norm1d = 1.0 / ((double)Nx * (double)Nx * (double)Ny);
    FFTPlan plan(Nx);
    for (int c = 0; c < 2; ++c) {
        const double* f = c == 0 ? u : v;
        std::vector<double> sums((size_t)Ny * nk);
#pragma omp parallel
        {
            std::vector<C> row(Nx), scratch;
#pragma omp for schedule(static)
            for (int j = 0; j < Ny; ++j) {
                for (int i = 0; i < Nx; ++i) row[i] = C(f[ID(j, i, Nx)], 0.0);
                plan.run(row.data(), false, scratch);
                for (int k = 0; k < nk; ++k) sums[(size_t)j * nk + k] = row[k].real() * row[k].real() + row[k].imag() * row[k].imag();
            }
        }
        for (int k = 0; k < nk; ++k) {
            const double wx = (k > 0 && k < nk - 1) ? 2.0 : 1.0;
            double sum = 0.0;
            for (int j = 0; j < Ny; ++j) sum += sums[(size_t)j * nk + k];
            acc[(size_t)c * nk + k] += norm1d * wx * sum;
        }
    }
}

// =============================================================================================
// diag.cu reductions
// =============================================================================================
extern "C" void orc_enstrophy_diss_mapped(const double* omg, int Nx, int Ny, double inv_dx,
                                          double inv_deta, const double* a_node,
                                          const double* w_node, double* out2) {
    std::vector<double> pw(Ny, 0.0), pg(Ny, 0.0);
#pragma omp parallel for schedule(static)
    for (int iy = 1; iy < Ny - 1; ++iy) {
        double lw2 = 0, lg2 = 0;
        for (int ix = 0; ix < Nx; ++ix) {
            const int xm = (ix == 0 ? Nx - 1 : ix - 1), xp = (ix == Nx - 1 ? 0 : ix + 1);
            const double w = omg[ID(iy, ix, Nx)];
            const double wx = (omg[ID(iy, xp, Nx)] - omg[ID(iy, xm, Nx)]) * 0.5 * inv_dx;
            const double wy = a_node[iy] * (omg[ID(iy + 1, ix, Nx)] - omg[ID(iy - 1, ix, Nx)]) * 0.5 * inv_deta;
            const double wt = w_node[iy];
            lw2 += wt * (w * w);
            lg2 += wt * (wx * wx + wy * wy);
        }
        pw[iy] = lw2; pg[iy] = lg2;
    }
    double a = 0, b = 0;
    for (int j = 0; j < Ny; ++j) { a += pw[j]; b += pg[j]; }
    out2[0] = a; out2[1] = b;
}

extern "C" void orc_enstrophy_diss_mapped_full(const double* omg, int Nx, int Ny, double inv_dx,
                                               double inv_deta, const double* a_node,
                                               const double* w_node, double* out2) {
    std::vector<double> pw(Ny, 0.0), pg(Ny, 0.0);
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < Ny; ++iy) {
        double lw2 = 0, lg2 = 0;
        for (int ix = 0; ix < Nx; ++ix) {
            const int xm = (ix == 0 ? Nx - 1 : ix - 1), xp = (ix == Nx - 1 ? 0 : ix + 1);
            const double w = omg[ID(iy, ix, Nx)];
            const double wx = (omg[ID(iy, xp, Nx)] - omg[ID(iy, xm, Nx)]) * 0.5 * inv_dx;
            double weta;
            if (iy == 0) {
                const double w0 = omg[ID(0, ix, Nx)], w1 = omg[ID(1, ix, Nx)], w2 = omg[ID(2, ix, Nx)];
                weta = (-3.0 * w0 + 4.0 * w1 - 1.0 * w2) * (0.5 * inv_deta);
            } else if (iy == Ny - 1) {
                const double wN = omg[ID(Ny - 1, ix, Nx)], wNm1 = omg[ID(Ny - 2, ix, Nx)],
                             wNm2 = omg[ID(Ny - 3, ix, Nx)];
                weta = (3.0 * wN - 4.0 * wNm1 + 1.0 * wNm2) * (0.5 * inv_deta);
            } else {
                weta = (omg[ID(iy + 1, ix, Nx)] - omg[ID(iy - 1, ix, Nx)]) * 0.5 * inv_deta;
            }
            const double wy = a_node[iy] * weta;
            const double wt = w_node[iy];
            lw2 += wt * (w * w);
            lg2 += wt * (wx * wx + wy * wy);
        }
        pw[iy] = lw2; pg[iy] = lg2;
    }
    double a = 0, b = 0;
    for (int j = 0; j < Ny; ++j) { a += pw[j]; b += pg[j]; }
    out2[0] = a; out2[1] = b;
}

extern "C" void orc_enstrophy_diss_balance_mapped(const double* w, int Nx, int Ny, double inv_dx2,
                                                  double inv_deta2, const double* a_node,
                                                  const double* a_edge, const double* w_node,
                                                  double* out2) {
    std::vector<double> lap((size_t)Nx * Ny);
    orc_laplacian_fd_mapped(w, lap.data(), Nx, Ny, inv_dx2, inv_deta2, a_node, a_edge);
    std::vector<double> pw(Ny, 0.0), pe(Ny, 0.0);
#pragma omp parallel for schedule(static)
    for (int j = 1; j < Ny - 1; ++j) {
        double a = 0, b = 0;
        for (int i = 0; i < Nx; ++i) {
            const double wv = w[ID(j, i, Nx)], l = lap[ID(j, i, Nx)];
            a += w_node[j] * (wv * wv);
            b += w_node[j] * (wv * l);
        }
        pw[j] = a; pe[j] = b;
    }
    double a = 0, b = 0;
    for (int j = 0; j < Ny; ++j) { a += pw[j]; b += pe[j]; }
    out2[0] = a; out2[1] = b;
}

extern "C" void orc_enstrophy_sbp_mapped(const double* omg, int Nx, int Ny, double inv_dx,
                                         double inv_deta, const double* a_edge,
                                         const double* wt_node, double* out2) {
    std::vector<double> pe(Ny, 0.0), pp(Ny, 0.0);
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < Ny; ++iy) {
        double le = 0, lp = 0;
        for (int ix = 0; ix < Nx; ++ix) {
            const int xp = (ix == Nx - 1 ? 0 : ix + 1);
            const double w0 = omg[ID(iy, ix, Nx)];
            if (iy >= 1 && iy <= Ny - 2) {
                const double dxw = (omg[ID(iy, xp, Nx)] - w0) * inv_dx;
                le += wt_node[iy] * dxw * dxw;
            }
            if (iy <= Ny - 2) {
                const double dw = omg[ID(iy + 1, ix, Nx)] - w0;
                le += a_edge[iy] * dw * dw * inv_deta;
            }
            if (iy == 0) lp -= w0 * a_edge[0] * (omg[ID(1, ix, Nx)] - w0) * inv_deta;
            if (iy == Ny - 1) lp += w0 * a_edge[Ny - 2] * (w0 - omg[ID(Ny - 2, ix, Nx)]) * inv_deta;
        }
        pe[iy] = le; pp[iy] = lp;
    }
    double a = 0, b = 0;
    for (int j = 0; j < Ny; ++j) { a += pe[j]; b += pp[j]; }
    out2[0] = a; out2[1] = b;
}

extern "C" double orc_kinetic_energy_mapped(const double* psi, int Nx, int Ny, double inv_dx,
                                            double inv_deta, const double* a_node,
                                            const double* w_node) {
    return row_sum(1, Ny - 1, [&](int iy) {
        double s = 0;
        for (int ix = 0; ix < Nx; ++ix) {
            const int xm = (ix == 0 ? Nx - 1 : ix - 1), xp = (ix == Nx - 1 ? 0 : ix + 1);
            const double u = (psi[ID(iy + 1, ix, Nx)] - psi[ID(iy - 1, ix, Nx)]) * 0.5 * inv_deta * a_node[iy];
            const double v = -(psi[ID(iy, xp, Nx)] - psi[ID(iy, xm, Nx)]) * 0.5 * inv_dx;
            s += 0.5 * (u * u + v * v) * w_node[iy];
        }
        return s;
    });
}

extern "C" double orc_energy_dissipation_mapped(const double* psi, const double* omega, int Nx,
                                                int Ny, double inv_dx, double inv_deta,
                                                const double* a_node, const double* a_edge,
                                                const double* w_node) {
    const double inv_dx2 = inv_dx * inv_dx, inv_deta2 = inv_deta * inv_deta;
    return row_sum(0, Ny, [&](int j) {
        double loc = 0;
        for (int i = 0; i < Nx; ++i) {
            if (j == 0 || j == Ny - 1) {  // wall rows: (u_y)^2 = omega^2, diag.cu:279-283
                const double uy = -omega[ID(j, i, Nx)];
                loc += (uy * uy) * w_node[j];
                continue;
            }
            const int im = (i == 0 ? Nx - 1 : i - 1), ip = (i == Nx - 1 ? 0 : i + 1);
            const double a = a_node[j], ap = a_edge[j], am = a_edge[j - 1];
            const double psi_x_jp = (psi[ID(j + 1, ip, Nx)] - psi[ID(j + 1, im, Nx)]) * (0.5 * inv_dx);
            const double psi_x_jm = (psi[ID(j - 1, ip, Nx)] - psi[ID(j - 1, im, Nx)]) * (0.5 * inv_dx);
            const double ux = a * (psi_x_jp - psi_x_jm) * (0.5 * inv_deta);
            const double vy = -ux;
            const double vx = -(psi[ID(j, ip, Nx)] - 2.0 * psi[ID(j, i, Nx)] + psi[ID(j, im, Nx)]) * inv_dx2;
            const double uy = a * (ap * (psi[ID(j + 1, i, Nx)] - psi[ID(j, i, Nx)]) -
                                   am * (psi[ID(j, i, Nx)] - psi[ID(j - 1, i, Nx)])) * inv_deta2;
            const double g2 = ux * ux + uy * uy + vx * vx + vy * vy;
            loc += g2 * w_node[j];
        }
        return loc;
    });
}

extern "C" double orc_power_input_mapped(const double* psi, int Nx, int Ny, double inv_deta,
                                         double h, double F0, int nforce, const double* a_node,
                                         const double* y_node, const double* w_node) {
    return row_sum(0, Ny, [&](int iy) {
        double loc = 0;
        const int ym = (iy == 0 ? 0 : iy - 1), yp = (iy == Ny - 1 ? Ny - 1 : iy + 1);
        for (int ix = 0; ix < Nx; ++ix) {
            const double u = (psi[ID(yp, ix, Nx)] - psi[ID(ym, ix, Nx)]) * 0.5 * inv_deta * a_node[iy];
            const double y = y_node[iy];
            const double Fx = F0 * std::sin((double)nforce * PI * ((y + h) / (2.0 * h)));
            loc += u * Fx * w_node[iy];
        }
        return loc;
    });
}

extern "C" double orc_enstrophy_input_mapped_full(const double* omega, int Nx, int Ny, double h,
                                                  double F0, int nforce, const double* y_node,
                                                  const double* w_node) {
    return row_sum(0, Ny, [&](int iy) {
        double loc = 0;
        for (int ix = 0; ix < Nx; ++ix) {
            const double w = omega[ID(iy, ix, Nx)];
            const double y = y_node[iy];
            const double arg = (double)nforce * PI * ((y + h) / (2.0 * h));
            const double f_omega = -F0 * ((double)nforce * PI / (2.0 * h)) * std::cos(arg);
            loc += w * f_omega * w_node[iy];
        }
        return loc;
    });
}

extern "C" double orc_enstrophy_input_mapped(const double* omega, int Nx, int Ny, double h,
                                             double F0, int nforce, const double* y_node,
                                             const double* w_node) {
    return row_sum(1, Ny - 1, [&](int iy) {
        double loc = 0;
        for (int ix = 0; ix < Nx; ++ix) {
            const double w = omega[ID(iy, ix, Nx)];
            const double y = y_node[iy];
            const double arg = (double)nforce * PI * ((y + h) / (2.0 * h));
            const double f_omega = -F0 * ((double)nforce * PI / (2.0 * h)) * std::cos(arg);
            loc += w * f_omega * w_node[iy];
        }
        return loc;
    });
}

extern "C" void orc_row_moments_from_psi_mapped(const double* psi, int Nx, int Ny, double inv_dx,
                                                double inv_deta, const double* a_node,
                                                double* su, double* sv, double* su2, double* sv2,
                                                double* suv) {
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < Ny; ++iy) {
        double lu = 0, lv = 0, lu2 = 0, lv2 = 0, luv = 0;
        const int ym = (iy == 0 ? 0 : iy - 1), yp = (iy == Ny - 1 ? Ny - 1 : iy + 1);
        for (int ix = 0; ix < Nx; ++ix) {
            const int xm = (ix == 0 ? Nx - 1 : ix - 1), xp = (ix == Nx - 1 ? 0 : ix + 1);
            const double u = (psi[ID(yp, ix, Nx)] - psi[ID(ym, ix, Nx)]) * 0.5 * inv_deta * a_node[iy];
            const double v = -(psi[ID(iy, xp, Nx)] - psi[ID(iy, xm, Nx)]) * 0.5 * inv_dx;
            lu += u; lv += v; lu2 += u * u; lv2 += v * v; luv += u * v;
        }
        su[iy] = lu; sv[iy] = lv; su2[iy] = lu2; sv2[iy] = lv2; suv[iy] = luv;
    }
}

extern "C" void orc_mean_wall_omega(const double* omega, int Nx, int Ny, double* out2) {
    double lb = 0, lt = 0;
    for (int ix = 0; ix < Nx; ++ix) { lb += omega[ix]; lt += omega[ID(Ny - 1, ix, Nx)]; }
    out2[0] = lb / (double)Nx;
    out2[1] = lt / (double)Nx;
}

extern "C" void orc_mean_wall_omega_omegay_mapped(const double* omega, int Nx, int Ny,
                                                  double inv_deta, const double* a_node,
                                                  double* out2) {
    double lb = 0, lt = 0;
    for (int ix = 0; ix < Nx; ++ix) {
        const double w0 = omega[ID(0, ix, Nx)], w1 = omega[ID(1, ix, Nx)], w2 = omega[ID(2, ix, Nx)];
        const double weta_bot = (-3.0 * w0 + 4.0 * w1 - 1.0 * w2) * (0.5 * inv_deta);
        lb += w0 * (a_node[0] * weta_bot);
        const double wN = omega[ID(Ny - 1, ix, Nx)], wNm1 = omega[ID(Ny - 2, ix, Nx)],
                     wNm2 = omega[ID(Ny - 3, ix, Nx)];
        const double weta_top = (3.0 * wN - 4.0 * wNm1 + 1.0 * wNm2) * (0.5 * inv_deta);
        lt += wN * (a_node[Ny - 1] * weta_top);
    }
    out2[0] = lb / (double)Nx;
    out2[1] = lt / (double)Nx;
}

extern "C" double orc_omega_rhs_balance_mapped(const double* omega, const double* rhs, int Nx,
                                               int Ny, const double* w_node) {
    return row_sum(0, Ny, [&](int j) {
        double s = 0;
        for (int i = 0; i < Nx; ++i) s += w_node[j] * omega[ID(j, i, Nx)] * rhs[ID(j, i, Nx)];
        return s;
    });
}

extern "C" void orc_max_uv_from_psi_mapped(const double* psi, int Nx, int Ny, double inv_dx,
                                           double inv_deta, const double* a_node, double* out3) {
    if (Nx < 2 || Ny < 3) { out3[0] = out3[1] = out3[2] = 0; return; }  // diag.cu:826-832
    std::vector<double> pu(Ny, 0.0), pv(Ny, 0.0), pm(Ny, 0.0);
#pragma omp parallel for schedule(static)
    for (int j = 1; j < Ny - 1; ++j) {
        double umax = 0, vmax = 0, vmaxm = 0;
        for (int i = 0; i < Nx; ++i) {
            const int ip = (i + 1 < Nx) ? (i + 1) : 0, im = (i > 0) ? (i - 1) : (Nx - 1);
            const double dpsideta = (psi[ID(j + 1, i, Nx)] - psi[ID(j - 1, i, Nx)]) * (0.5 * inv_deta);
            const double dpsidx = (psi[ID(j, ip, Nx)] - psi[ID(j, im, Nx)]) * (0.5 * inv_dx);
            const double a = a_node[j];
            const double au = std::fabs(a * dpsideta), av = std::fabs(-dpsidx), avm = av * a;
            if (au > umax) umax = au;
            if (av > vmax) vmax = av;
            if (avm > vmaxm) vmaxm = avm;
        }
        pu[j] = umax; pv[j] = vmax; pm[j] = vmaxm;
    }
    double u = 0, v = 0, m = 0;
    for (int j = 0; j < Ny; ++j) { u = std::max(u, pu[j]); v = std::max(v, pv[j]); m = std::max(m, pm[j]); }
    out3[0] = u; out3[1] = v; out3[2] = m;
}

// =============================================================================================
// Host side: parameters, parse_args (init.cu:7-96), initial conditions (init.cu:99-378),
// scheduling helpers (io.cu:162-178), writers (io.cu:262-312) and the time loop (main.cu).
// =============================================================================================
namespace {

struct OGrid {  // common.cuh:42-57
    int Nx{1024};
    int Ny{1025};
    double Lx{12.566370612};
    double h{1.0};
    std::vector<double> y, a_node, a_edge, dy_edge, w_node;
    double dx() const { return Lx / (double)Nx; }
    double dy() const { return (2.0 * h) / (double)(Ny - 1); }
    double Ly() const { return 2.0 * h; }
};

struct OParams {  // common.cuh:58-114
    double Ub{1.0};
    std::string stretch{"none"}, ytable;
    double beta{0.0};
    double Reb{10000.0};
    double nu_override{-1};
    double nu(double h) const { return (nu_override >= 0 ? nu_override : (2.0 * h * Ub) / Reb); }
    double dt_init{1e-3}, dt_max{1e-2}, cfl{0.5}, cvisc{0.5};
    bool adapt{true};
    double t_end{200.0};
    double F0{0.0};
    int nforce{1};
    double lin_drag{0.0};
    std::string init{"load"};
    double amp{0.5}, alpha{1.2};
    int my{1};
    double phase{0.0}, rand_amp{0.0};
    unsigned long long seed{123456789ULL};
    std::string load_path{""};
    bool load_use_psi{false};
    std::string snap_fields{"all"};
    int rand_nx{48}, rand_mymax{8};
    double rand_k0{0.0}, rand_sigma{0.6}, rand_amp_abs{0.0};
    bool mms{false};
    double mms_A0{1e-3}, mms_eps{0.1}, mms_Om{1.0};
    int mms_kx{1};
    bool closure_diag{false};
    bool resume{false};
    // options of this project (no reference counterpart)
    bool free_slip{false};
    long long max_steps{-1};
    int band_kmax{32};
    bool poisson_ld{false};  // measurement: extended-precision Poisson solves
};

struct ORun {
    OGrid G;
    OParams P;
    int statsEvery = 50, timeEvery = 50, spectraEvery = 0, profilesWriteEvery = 0;
    bool profile = false;
    double snap_save_dt = 0.1, diag_save_dt = 0.2;
    std::string snap_fmt = "h5", outdir = "out";
};

void parse_args(int argc, char** argv, ORun& R) {
    OGrid& G = R.G;
    OParams& P = R.P;
    for (int i = 0; i < argc; i++) {
        auto eq = [&](const char* a, const char* b) { return std::strcmp(a, b) == 0; };
        const bool nx = i + 1 < argc;
        if (eq(argv[i], "--Nx") && nx) G.Nx = std::atoi(argv[++i]);
        else if (eq(argv[i], "--Ny") && nx) G.Ny = std::atoi(argv[++i]);
        else if (eq(argv[i], "--Lx") && nx) G.Lx = std::atof(argv[++i]);
        else if (eq(argv[i], "--h") && nx) G.h = std::atof(argv[++i]);
        else if (eq(argv[i], "--Ub") && nx) P.Ub = std::atof(argv[++i]);
        else if (eq(argv[i], "--Reb") && nx) P.Reb = std::atof(argv[++i]);
        else if (eq(argv[i], "--dt") && nx) P.dt_init = std::atof(argv[++i]);
        else if (eq(argv[i], "--dtmax") && nx) P.dt_max = std::atof(argv[++i]);
        else if (eq(argv[i], "--cfl") && nx) P.cfl = std::atof(argv[++i]);
        else if (eq(argv[i], "--cvisc") && nx) P.cvisc = std::atof(argv[++i]);
        else if (eq(argv[i], "--no-adapt")) P.adapt = false;
        else if (eq(argv[i], "--tend") && nx) P.t_end = std::atof(argv[++i]);
        else if (eq(argv[i], "--F0") && nx) P.F0 = std::atof(argv[++i]);
        else if (eq(argv[i], "--nforce") && nx) P.nforce = std::atoi(argv[++i]);
        else if (eq(argv[i], "--drag") && nx) P.lin_drag = std::atof(argv[++i]);
        else if (eq(argv[i], "--nu") && nx) P.nu_override = std::atof(argv[++i]);
        else if (eq(argv[i], "--stretch") && nx) P.stretch = argv[++i];
        else if (eq(argv[i], "--beta") && nx) P.beta = std::atof(argv[++i]);
        else if (eq(argv[i], "--ytable") && nx) P.ytable = argv[++i];
        else if (eq(argv[i], "--integrator") && nx) ++i;
        else if (eq(argv[i], "--init") && nx) P.init = argv[++i];
        else if (eq(argv[i], "--amp") && nx) P.amp = std::atof(argv[++i]);
        else if (eq(argv[i], "--alpha") && nx) P.alpha = std::atof(argv[++i]);
        else if (eq(argv[i], "--my") && nx) P.my = std::atoi(argv[++i]);
        else if (eq(argv[i], "--phase") && nx) P.phase = std::atof(argv[++i]);
        else if (eq(argv[i], "--rand_amp") && nx) P.rand_amp = std::atof(argv[++i]);
        else if (eq(argv[i], "--seed") && nx) P.seed = (unsigned long long)std::stoull(argv[++i]);
        else if (eq(argv[i], "--load_path") && nx) P.load_path = argv[++i];
        else if (eq(argv[i], "--load_use_psi")) P.load_use_psi = true;
        else if (eq(argv[i], "--rand_nx") && nx) P.rand_nx = std::atoi(argv[++i]);
        else if (eq(argv[i], "--rand_mymax") && nx) P.rand_mymax = std::atoi(argv[++i]);
        else if (eq(argv[i], "--rand_k0") && nx) P.rand_k0 = std::atof(argv[++i]);
        else if (eq(argv[i], "--rand_sigma") && nx) P.rand_sigma = std::atof(argv[++i]);
        else if (eq(argv[i], "--rand_amp_abs") && nx) P.rand_amp_abs = std::atof(argv[++i]);
        else if (eq(argv[i], "--mms")) {
            P.mms = true;
            if (i + 1 < argc && argv[i + 1][0] != '-') P.mms = (std::atoi(argv[++i]) != 0);
        }
        else if (eq(argv[i], "--mmsA0") && nx) P.mms_A0 = std::atof(argv[++i]);
        else if (eq(argv[i], "--mmsEps") && nx) P.mms_eps = std::atof(argv[++i]);
        else if (eq(argv[i], "--mmsOm") && nx) P.mms_Om = std::atof(argv[++i]);
        else if (eq(argv[i], "--mmsKx") && nx) P.mms_kx = std::atoi(argv[++i]);
        else if (eq(argv[i], "--statsEvery") && nx) R.statsEvery = std::atoi(argv[++i]);
        else if (eq(argv[i], "--timeEvery") && nx) R.timeEvery = std::atoi(argv[++i]);
        else if (eq(argv[i], "--profile")) R.profile = true;
        else if (eq(argv[i], "--closure_diag")) P.closure_diag = true;
        else if (eq(argv[i], "--resume")) P.resume = true;
        else if (eq(argv[i], "--profilesWriteEvery") && nx) R.profilesWriteEvery = std::atoi(argv[++i]);
        else if (eq(argv[i], "--snap_save") && nx) R.snap_save_dt = std::atof(argv[++i]);
        else if (eq(argv[i], "--snap_fmt") && nx) R.snap_fmt = argv[++i];
        else if (eq(argv[i], "--snap_fields") && nx) P.snap_fields = argv[++i];
        else if (eq(argv[i], "--diag_save") && nx) R.diag_save_dt = std::atof(argv[++i]);
        else if (eq(argv[i], "--diagsave") && nx) R.diag_save_dt = std::atof(argv[++i]);
        else if (eq(argv[i], "--spectraEvery") && nx) R.spectraEvery = std::atoi(argv[++i]);
        else if (eq(argv[i], "--outdir") && nx) R.outdir = argv[++i];
        // options of this project
        else if (eq(argv[i], "--walls") && nx) P.free_slip = (std::strcmp(argv[++i], "freeslip") == 0);
        else if (eq(argv[i], "--max_steps") && nx) P.max_steps = std::atoll(argv[++i]);
        else if (eq(argv[i], "--band_kmax") && nx) P.band_kmax = std::atoi(argv[++i]);
        else if (eq(argv[i], "--poisson_ld")) P.poisson_ld = true;
    }
    if (P.init == "laminar" && P.rand_amp > 0.0 && P.amp <= 0.0) { P.init = "rand"; P.amp = P.rand_amp; }
    if (P.init == "mms") P.mms = true;
}

bool build_grid(ORun& R) {
    OGrid& G = R.G;
    const int Ny = G.Ny;
    G.y.assign(Ny, 0); G.a_node.assign(Ny, 0); G.a_edge.assign(Ny - 1, 0);
    G.dy_edge.assign(Ny - 1, 0); G.w_node.assign(Ny, 0);
    int st;
    if (R.P.stretch == "none" || R.P.stretch == "uniform") st = 0;
    else if (R.P.stretch == "tanh") st = 1;
    else if (R.P.stretch == "table" && Ny > 2) {
        // the Ny values of y as whitespace-separated text (in place of the reference's HDF5 /y)
        std::ifstream f(R.P.ytable);
        std::vector<double> yt((size_t)Ny);
        for (int j = 0; j < Ny; ++j)
            if (!(f >> yt[(size_t)j])) { std::fprintf(stderr, "oracle: --ytable needs %d values\n", Ny); return false; }
        return orc_table_y_metrics(Ny, yt.data(), G.y.data(), G.a_node.data(), G.a_edge.data(), G.dy_edge.data(),
                                   G.w_node.data()) == 0;
    }
    else { std::fprintf(stderr, "oracle: unsupported --stretch %s\n", R.P.stretch.c_str()); return false; }
    return orc_build_y_metrics(Ny, G.h, st, R.P.beta, G.y.data(), G.a_node.data(), G.a_edge.data(),
                               G.dy_edge.data(), G.w_node.data()) == 0;
}

// ---- snapshots: raw little-endian float64 fields in one .bin file plus a JSON header.
// (The reference writes HDF5, io.cu:203-259, which this image has no library for.)
std::string snap_base(const std::string& outdir, double t) {
    std::ostringstream fn;
    fn << outdir << "/snaps/snap_t" << std::fixed << std::setprecision(6) << t;
    return fn.str();
}
void ensure_dir(const std::string& p) {
    std::string cur;
    std::stringstream ss(p);
    std::string part;
    if (!p.empty() && p[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        ::mkdir(cur.c_str(), 0755);
    }
}
void write_snapshot(const ORun& R, double t, const std::vector<std::pair<std::string, const double*>>& fields) {
    const OGrid& G = R.G;
    ensure_dir(R.outdir + "/snaps");
    const std::string base = snap_base(R.outdir, t);
    std::ofstream b(base + ".bin", std::ios::binary);
    const size_t N = (size_t)G.Nx * G.Ny;
    for (auto& f : fields) b.write(reinterpret_cast<const char*>(f.second), (std::streamsize)(N * sizeof(double)));
    b.write(reinterpret_cast<const char*>(G.y.data()), (std::streamsize)(G.Ny * sizeof(double)));
    std::ofstream j(base + ".json");
    j << std::setprecision(17);
    j << "{\"t\": " << t << ", \"Nx\": " << G.Nx << ", \"Ny\": " << G.Ny << ", \"Lx\": " << G.Lx
      << ", \"Ly\": " << G.Ly() << ", \"h\": " << G.h << ", \"dx\": " << G.dx() << ", \"dy\": " << G.dy()
      << ", \"nu\": " << R.P.nu(G.h) << ", \"beta\": " << R.P.beta
      << ", \"stretch_code\": " << (R.P.stretch == "tanh" ? 1 : 0) << ", \"fields\": [";
    for (size_t k = 0; k < fields.size(); ++k) j << (k ? ", " : "") << "\"" << fields[k].first << "\"";
    j << (fields.empty() ? "" : ", ") << "\"y\"], \"dtype\": \"<f8\"}\n";
}
// The time "t" of a snapshot header (path with or without the .json/.bin suffix).
bool read_snapshot_time(std::string base, double* t) {
    for (const char* ext : {".json", ".bin"})
        if (base.size() > std::strlen(ext) && base.substr(base.size() - std::strlen(ext)) == ext)
            base = base.substr(0, base.size() - std::strlen(ext));
    std::ifstream j(base + ".json");
    if (!j) return false;
    std::string s((std::istreambuf_iterator<char>(j)), std::istreambuf_iterator<char>());
    const size_t p = s.find("\"t\"");
    if (p == std::string::npos) return false;
    *t = std::atof(s.c_str() + s.find(':', p) + 1);
    return true;
}
bool read_snapshot_field(const std::string& base, const std::string& name, int Nx, int Ny, double* out) {
    std::ifstream j(base + ".json");
    if (!j) return false;
    std::string s((std::istreambuf_iterator<char>(j)), std::istreambuf_iterator<char>());
    const size_t fpos = s.find("\"fields\"");
    if (fpos == std::string::npos) return false;
    size_t k = 0, pos = s.find('[', fpos);
    const size_t end = s.find(']', pos);
    int index = -1;
    while (pos < end) {
        const size_t q0 = s.find('"', pos);
        if (q0 == std::string::npos || q0 > end) break;
        const size_t q1 = s.find('"', q0 + 1);
        if (s.substr(q0 + 1, q1 - q0 - 1) == name) index = (int)k;
        ++k;
        pos = q1 + 1;
    }
    if (index < 0) return false;
    std::ifstream b(base + ".bin", std::ios::binary);
    const size_t N = (size_t)Nx * Ny;
    b.seekg((std::streamoff)((size_t)index * N * sizeof(double)));
    b.read(reinterpret_cast<char*>(out), (std::streamsize)(N * sizeof(double)));
    return (bool)b;
}

// ---- portable seeded normals for the benchmark-configuration ICs (no reference counterpart).
// splitmix64 -> uniform in (0,1) -> Box-Muller pairs.
struct Normal {
    unsigned long long s;
    explicit Normal(unsigned long long seed) : s(seed) {}
    unsigned long long next() {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uni() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    void pair(double& a, double& b) {
        const double u1 = uni(), u2 = uni();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double th = 2.0 * PI * u2;
        a = r * std::cos(th);
        b = r * std::sin(th);
    }
};

bool init_omega_with_IC(const ORun& R, std::vector<double>& h_w) {
    const OGrid& G = R.G;
    const OParams& P = R.P;
    const int Nx = G.Nx, Ny = G.Ny;
    const double Lx = G.Lx, h = G.h, dy = G.dy();
    const double Ub = P.Ub;
    h_w.assign((size_t)Nx * Ny, 0.0);
    std::vector<double> y(Ny);
    if (!G.y.empty()) y = G.y;
    else for (int j = 0; j < Ny; ++j) y[j] = -h + j * dy;
    if (P.init == "mms") {  // init.cu:118-137
        const double k = 2.0 * PI * (double)P.mms_kx / Lx;
        const double c = PI / (2.0 * h);
        const double A0 = P.mms_A0;
        for (int j = 0; j < Ny; ++j) {
            const double theta = c * (y[j] + h);
            double s = std::sin(theta); s *= s;
            const double cos2 = std::cos(2.0 * theta);
            const double g = (k * k) * s - 2.0 * (c * c) * cos2;
            for (int i = 0; i < Nx; ++i) {
                const double x = (double)i * (Lx / (double)Nx);
                h_w[ID(j, i, Nx)] = A0 * std::sin(k * x) * g;
            }
        }
        return true;
    }
    if (P.init == "diffusion") {  // init.cu:151-171
        const int my0 = (P.my > 0 ? P.my : 1);
        const int n = 2 * my0;
        const double mu = (double)n * PI / (2.0 * h);
        const double Au = (P.amp != 0.0) ? P.amp : 1.0;
        for (int j = 0; j < Ny; ++j) {
            const double omega_j = -Au * mu * std::cos(mu * (y[j] + h));
            for (int i = 0; i < Nx; ++i) h_w[ID(j, i, Nx)] = omega_j;
        }
        return true;
    }
    if (P.init == "diffusion2d") {  // init.cu:173-204
        const int mx = 1, my = (P.my > 0 ? P.my : 1);
        const double alpha = (P.alpha > 0.0) ? P.alpha : (2.0 * PI * (double)mx / Lx);
        const double beta = (double)my * PI / (2.0 * h);
        const double lambda = alpha * alpha + beta * beta;
        const double Apsi = (P.amp != 0.0) ? P.amp : 1.0;
        for (int j = 0; j < Ny; ++j) {
            const double sj = std::sin(beta * (y[j] + h));
            for (int i = 0; i < Nx; ++i) {
                const double x = (double)i * (Lx / (double)Nx);
                h_w[ID(j, i, Nx)] = lambda * (Apsi * std::sin(alpha * x) * sj);
            }
        }
        return true;
    }
    unsigned long long seed = P.seed;  // init.cu:207-212
    auto rnd01 = [&]() {
        seed = 6364136223846793005ULL * seed + 1ULL;
        return ((seed >> 11) * 1.1102230246251565e-16);
    };
    if (P.init == "load") {  // init.cu:215-279, own snapshot format; psi path: interior -(dxx+dyy)
        if (P.load_path.empty()) { std::fprintf(stderr, "oracle: --init load needs --load_path\n"); return false; }
        std::string base = P.load_path;
        if (base.size() > 5 && base.substr(base.size() - 5) == ".json") base = base.substr(0, base.size() - 5);
        if (base.size() > 4 && base.substr(base.size() - 4) == ".bin") base = base.substr(0, base.size() - 4);
        if (!P.load_use_psi && read_snapshot_field(base, "omega", Nx, Ny, h_w.data())) return true;
        std::vector<double> psi((size_t)Nx * Ny);
        if (!read_snapshot_field(base, "psi", Nx, Ny, psi.data())) return false;
        const double dx = G.dx();
        const double inv_dx2 = 1.0 / (dx * dx), inv_dy2 = 1.0 / (dy * dy);
        for (int j = 1; j < Ny - 1; ++j)
            for (int i = 0; i < Nx; ++i) {
                const double pc = psi[ID(j, i, Nx)];
                const double pxx = (psi[ID(j, wrapx(i + 1, Nx), Nx)] - 2 * pc + psi[ID(j, wrapx(i - 1, Nx), Nx)]) * inv_dx2;
                const double pyy = (psi[ID(j + 1, i, Nx)] - 2 * pc + psi[ID(j - 1, i, Nx)]) * inv_dy2;
                h_w[ID(j, i, Nx)] = -(pxx + pyy);
            }
        const double psi_bot = 0.0, psi_top = 2.0 * h * Ub;
        for (int i = 0; i < Nx; ++i) {
            h_w[ID(0, i, Nx)] = -2.0 * (psi[ID(1, i, Nx)] - psi_bot) * inv_dy2;
            h_w[ID(Ny - 1, i, Nx)] = -2.0 * (psi[ID(Ny - 2, i, Nx)] - psi_top) * inv_dy2;
        }
        return true;
    }
    // laminar base, init.cu:284-288
    for (int j = 0; j < Ny; ++j) {
        const double wj = 3.0 * Ub * y[j] / (h * h);
        for (int i = 0; i < Nx; ++i) h_w[ID(j, i, Nx)] = wj;
    }
    if (P.init == "mix") {  // init.cu:290-296
        const double scale = std::max(1e-12, P.amp) * (Ub / h);
        const size_t N = (size_t)Nx * Ny;
        for (size_t idx = 0; idx < N; ++idx) h_w[idx] += scale * (rnd01() - 0.5);
    }
    if (P.init == "ts" || P.init == "mix") {  // init.cu:299-312
        const double alpha = (P.alpha > 0 ? P.alpha : (2.0 * PI / Lx));
        const double mu = PI * (double)P.my / (2.0 * h);
        const double Apsi = (P.amp * Ub) / std::max(1e-12, mu);
        for (int j = 0; j < Ny; ++j) {
            const double sy = std::sin(mu * (y[j] + h));
            for (int i = 0; i < Nx; ++i) {
                const double x = (double)i * (Lx / (double)Nx);
                const double psi_p = Apsi * sy * std::cos(alpha * x + P.phase);
                h_w[ID(j, i, Nx)] += (alpha * alpha + mu * mu) * psi_p;
            }
        }
    }
    if (P.init == "rand") {  // init.cu:314-362
        const double dx = G.dx();
        const double alpha_min = 2.0 * PI / Lx;
        const double alpha_nyq = PI / dx;
        const double alpha_max = 0.35 * alpha_nyq;
        const int nxm = std::max(1, P.rand_nx);
        const int mymx = std::max(1, P.rand_mymax);
        const double k0 = (P.rand_k0 > 0 ? P.rand_k0 : (6.0 * 2.0 * PI / Lx));
        const double sig = std::max(0.05, P.rand_sigma);
        std::mt19937_64 rng(P.seed);
        std::uniform_real_distribution<double> U01(0.0, 1.0);
        std::normal_distribution<double> N0(0.0, 1.0);
        auto sample_alpha = [&]() {
            double a = std::exp(std::log(k0) + sig * N0(rng));
            if (a < alpha_min) a = alpha_min;
            if (a > alpha_max) a = alpha_max;
            return a;
        };
        const double A0 = std::max(1e-12, (P.rand_amp_abs > 0 ? P.rand_amp_abs : P.amp)) / std::sqrt((double)nxm);
        std::vector<std::vector<double>> sY(mymx + 1, std::vector<double>(Ny, 0.0));
        for (int my = 1; my <= mymx; ++my) {
            const double mu = PI * (double)my / (2.0 * h);
            for (int j = 0; j < Ny; ++j) sY[my][j] = std::sin(mu * (y[j] + h));
        }
        for (int m = 0; m < nxm; ++m) {
            const double alpha = sample_alpha();
            const int my = 1 + (int)std::floor(U01(rng) * mymx);
            const double phi = 2.0 * PI * U01(rng);
            const double mu = PI * (double)my / (2.0 * h);
            const double k = std::sqrt(alpha * alpha + mu * mu);
            const double wamp = A0 * std::exp(-0.5 * std::pow(std::log(k / k0) / sig, 2));
            for (int j = 0; j < Ny; ++j) {
                const double sj = sY[my][j];
                for (int i = 0; i < Nx; ++i) {
                    const double x = ((double)i * (Lx / (double)Nx));
                    h_w[ID(j, i, Nx)] += wamp * std::cos(alpha * x + phi) * sj;
                }
            }
        }
    }
    // Benchmark-configuration ICs (no reference counterpart): laminar base plus seeded noise
    // tapered by (1-(y/h)^2)^2.
    //   lamnoise: i.i.d. normal noise of standard deviation amp at every point (row-major order).
    //   lamband:  per row, sum over k=1..K of (a cos(k 2pi x/Lx) + b sin(k 2pi x/Lx)) / sqrt(K)
    //             with (a,b) one normal pair per (j,k), rows then k ascending; scaled by amp.
    if (P.init == "lamnoise") {
        Normal g(P.seed);
        const size_t N = (size_t)Nx * Ny;
        double a = 0, b = 0;
        for (size_t idx = 0; idx < N; ++idx) {
            if ((idx & 1) == 0) g.pair(a, b);
            const double z = (idx & 1) ? b : a;
            const int j = (int)(idx / (size_t)Nx);
            const double yy = y[j] / h;
            const double taper = (1.0 - yy * yy) * (1.0 - yy * yy);
            h_w[idx] += P.amp * taper * z;
        }
    }
    if (P.init == "lamband") {
        Normal g(P.seed);
        const int K = std::max(1, P.band_kmax);
        std::vector<double> ct(Nx), st(Nx);
        for (int m = 0; m < Nx; ++m) {
            const double ang = 2.0 * PI * (double)m / (double)Nx;
            ct[m] = std::cos(ang); st[m] = std::sin(ang);
        }
        std::vector<double> ca((size_t)Ny * K), cb((size_t)Ny * K);
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < K; ++k) g.pair(ca[(size_t)j * K + k], cb[(size_t)j * K + k]);
        const double norm = 1.0 / std::sqrt((double)K);
#pragma omp parallel for schedule(static)
        for (int j = 0; j < Ny; ++j) {
            const double yy = y[j] / h;
            const double taper = (1.0 - yy * yy) * (1.0 - yy * yy);
            for (int i = 0; i < Nx; ++i) {
                double s = 0.0;
                for (int k = 1; k <= K; ++k) {
                    const int m = (int)(((long long)k * i) % Nx);
                    s += ca[(size_t)j * K + k - 1] * ct[m] + cb[(size_t)j * K + k - 1] * st[m];
                }
                h_w[ID(j, i, Nx)] += P.amp * taper * (s * norm);
            }
        }
    }
    return true;
}

// ---- io.cu:162-178
double time_tol(double t) { return 64.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, std::fabs(t)); }
double sched_tol(double t, double dt) {
    const double tol_abs = time_tol(t);
    const double tol_dt = 1e-6 * std::max(1e-300, std::fabs(dt));
    return std::max(tol_abs, tol_dt);
}
double next_after(double t0, double dt) {  // io.cu:173-178
    if (dt <= 0.0) return 1e300;
    const double k = std::floor((t0 + time_tol(t0)) / dt);
    const double t1 = (k + 1.0) * dt;
    return (t1 <= t0 + time_tol(t0)) ? t1 + dt : t1;
}

// ---- io.cu:262-312
void write_profiles_csv(const std::string& path, const OGrid& G, const OParams& P, double utau_mean,
                        double T_accum, const std::vector<double>& Hu, const std::vector<double>& Hv,
                        const std::vector<double>& Hu2, const std::vector<double>& Hv2,
                        const std::vector<double>& Huv) {
    std::ofstream f(path, std::ios::out);
    f << "j,y,yplus,U,V,uu,vv,uv,Uplus,tau_total\n";
    const bool mapped_out = (P.stretch != "none" && P.beta > 0.0);
    for (int j = 0; j < G.Ny; ++j) {
        const double y = (mapped_out ? G.y[j] : (-G.h + (double)j * G.dy()));
        const double ybot = y + G.h;
        const double Ubar = (T_accum > 0 ? Hu[j] / T_accum : 0.0);
        const double Vbar = (T_accum > 0 ? Hv[j] / T_accum : 0.0);
        const double U2b = (T_accum > 0 ? Hu2[j] / T_accum : 0.0);
        const double V2b = (T_accum > 0 ? Hv2[j] / T_accum : 0.0);
        const double UVb = (T_accum > 0 ? Huv[j] / T_accum : 0.0);
        const double uu = std::fmax(0.0, U2b - Ubar * Ubar);
        const double vv = std::fmax(0.0, V2b - Vbar * Vbar);
        const double uv = UVb - Ubar * Vbar;
        const double Ujp = (T_accum > 0 ? Hu[std::min(j + 1, G.Ny - 1)] / T_accum : 0.0);
        const double Ujm = (T_accum > 0 ? Hu[std::max(j - 1, 0)] / T_accum : 0.0);
        double dUdy;
        if (mapped_out) {
            const double inv_deta = (G.Ny > 1) ? (0.5 * (G.Ny - 1)) : 0.0;
            const double a = G.a_node[j];
            if (j == 0) dUdy = a * (Ujp - Ubar) * inv_deta;
            else if (j == G.Ny - 1) dUdy = a * (Ubar - Ujm) * inv_deta;
            else dUdy = a * (Ujp - Ujm) * 0.5 * inv_deta;
        } else {
            if (j == 0) dUdy = (Ujp - Ubar) / G.dy();
            else if (j == G.Ny - 1) dUdy = (Ubar - Ujm) / G.dy();
            else dUdy = (Ujp - Ujm) * 0.5 / G.dy();
        }
        const double tau_total = P.nu(G.h) * dUdy - uv;
        const double yplus = (utau_mean > 0 ? utau_mean * ybot / P.nu(G.h) : 0.0);
        const double Uplus = (utau_mean > 0 ? Ubar / utau_mean : 0.0);
        f.setf(std::ios::scientific);
        f.precision(8);
        f << j << "," << y << "," << yplus << "," << Ubar << "," << Vbar << "," << uu << "," << vv << ","
          << uv << "," << Uplus << "," << tau_total << "\n";
    }
}

// ---- the run state and the stage sequence of main.cu:482-538
struct OState {
    ORun R;
    double nu = 0, dx = 0, dy = 0, inv_dx = 0, inv_dx2 = 0, deta = 0, inv_deta = 0, inv_deta2 = 0;
    double inv_dymin2 = 0, invA_full = 0, meanFx_A = 0, psi_bot = 0, psi_top = 0;
    size_t N = 0;
    std::vector<double> w, w1, w2, psi, rhs, J, lap, wt_node;
    OPoisson* poisson = nullptr;
    double Jwall_cum[3] = {0, 0, 0};
    ~OState() { delete poisson; }

    bool setup(int argc, char** argv) {
        parse_args(argc, argv, R);
        OGrid& G = R.G;
        OParams& P = R.P;
        if (G.Nx < 2 || G.Ny < 4) { std::fprintf(stderr, "oracle: grid too small\n"); return false; }
        if (!build_grid(R)) return false;
        nu = P.nu(G.h);
        dx = G.dx(); dy = G.dy();
        inv_dx = 1.0 / dx; inv_dx2 = inv_dx * inv_dx;
        deta = (G.Ny > 1) ? 2.0 / (double)(G.Ny - 1) : 1.0;
        inv_deta = 1.0 / deta; inv_deta2 = inv_deta * inv_deta;
        double dy_min = std::fabs(G.dy_edge[0]);  // main.cu:56-62
        for (int j = 1; j < G.Ny - 1; ++j) dy_min = std::min(dy_min, std::fabs(G.dy_edge[j]));
        inv_dymin2 = 1.0 / (dy_min * dy_min);
        double Ly_full = 0.0;  // main.cu:64-71
        for (int j = 0; j < G.Ny; ++j) Ly_full += G.w_node[j];
        invA_full = 1.0 / ((double)G.Nx * Ly_full);
        meanFx_A = 0.0;  // main.cu:78-87
        if (P.F0 != 0.0) {
            for (int j = 0; j < G.Ny; ++j) {
                const double arg = (double)P.nforce * PI * ((G.y[j] + G.h) / (2.0 * G.h));
                meanFx_A += P.F0 * std::sin(arg) * G.w_node[j];
            }
            meanFx_A /= Ly_full;
        }
        N = (size_t)G.Nx * G.Ny;
        w.assign(N, 0); w1.assign(N, 0); w2.assign(N, 0); psi.assign(N, 0);
        rhs.assign(N, 0); J.assign(N, 0); lap.assign(N, 0);
        wt_node.assign(G.Ny, 0.0);  // main.cu:252-254
        for (int j = 1; j < G.Ny - 1; ++j) wt_node[j] = (2.0 / (double)(G.Ny - 1)) / G.a_node[j];
        if (!init_omega_with_IC(R, w)) return false;
        poisson = new OPoisson(G.Nx, G.Ny, G.Lx, G.a_node.data(), G.a_edge.data(), inv_deta);
        poisson->extended = P.poisson_ld;
        psi_bot = 0.0; psi_top = 2.0 * G.h * P.Ub;  // main.cu:211-213
        poisson->solve(w.data(), psi.data(), psi_bot, psi_top);
        thom(w);
        return true;
    }
    void thom(std::vector<double>& f) {
        orc_thom_wall_vorticity_mapped(f.data(), psi.data(), R.G.Nx, R.G.Ny, R.G.dy_edge.data(),
                                       psi_bot, psi_top, R.P.free_slip ? 1 : 0);
    }
    void build_rhs(const std::vector<double>& ws, double t_mms) {
        const OGrid& G = R.G;
        const OParams& P = R.P;
        orc_arakawa_J_mapped(psi.data(), ws.data(), J.data(), G.Nx, G.Ny, inv_dx, inv_deta, G.a_node.data());
        orc_laplacian_fd_mapped(ws.data(), lap.data(), G.Nx, G.Ny, inv_dx2, inv_deta2, G.a_node.data(), G.a_edge.data());
        orc_rhs_mapped(J.data(), lap.data(), ws.data(), rhs.data(), G.Nx, G.Ny, nu, P.F0, P.nforce, G.h,
                       P.lin_drag, G.y.data());
        if (P.mms)
            orc_mms_add_forcing(rhs.data(), G.Nx, G.Ny, G.Lx, G.h, t_mms, nu, P.lin_drag, P.mms_A0,
                                P.mms_eps, P.mms_Om, P.mms_kx, G.y.data());
    }
    double wall_row_enstrophy(const std::vector<double>& f) const {  // main.cu:284-289
        const OGrid& G = R.G;
        const double wq_b = G.w_node[0], wq_t = G.w_node[G.Ny - 1];
        double s = 0.0;
        for (int i = 0; i < G.Nx; ++i) {
            const double b = f[(size_t)i], t = f[ID(G.Ny - 1, i, G.Nx)];
            s += wq_b * b * b + wq_t * t * t;
        }
        return 0.5 * s * invA_full;
    }
    double wall_row_omega_rhs(const std::vector<double>& f, const std::vector<double>& r) const {  // main.cu:290-296
        const OGrid& G = R.G;
        const double wq_b = G.w_node[0], wq_t = G.w_node[G.Ny - 1];
        double s = 0.0;
        for (int i = 0; i < G.Nx; ++i)
            s += wq_b * f[(size_t)i] * r[(size_t)i] + wq_t * f[ID(G.Ny - 1, i, G.Nx)] * r[ID(G.Ny - 1, i, G.Nx)];
        return s * invA_full;
    }
    void close_stage(std::vector<double>& f, int s) {
        if (R.P.closure_diag) {
            const double Jb = wall_row_enstrophy(f);
            thom(f);
            Jwall_cum[s] += wall_row_enstrophy(f) - Jb;
        } else {
            thom(f);
        }
    }
    // One SSPRK(3,3) step, main.cu:482-538.
    void step(double t, double dt) {
        const OGrid& G = R.G;
        build_rhs(w, t);
        orc_axpy(w1.data(), w.data(), rhs.data(), dt, N);
        poisson->solve(w1.data(), psi.data(), psi_bot, psi_top);
        close_stage(w1, 0);
        build_rhs(w1, t + dt);
        orc_ssprk2(w2.data(), w.data(), w1.data(), rhs.data(), dt, N);
        poisson->solve(w2.data(), psi.data(), psi_bot, psi_top);
        close_stage(w2, 1);
        build_rhs(w2, t + 0.5 * dt);
        orc_ssprk3(w.data(), w.data(), w2.data(), rhs.data(), dt, N);
        poisson->solve(w.data(), psi.data(), psi_bot, psi_top);
        close_stage(w, 2);
        (void)G;
    }
    void maxuv(double* out3) const {
        orc_max_uv_from_psi_mapped(psi.data(), R.G.Nx, R.G.Ny, inv_dx, inv_deta, R.G.a_node.data(), out3);
    }
    double adapt_dt() const {  // main.cu:440-458
        const OParams& P = R.P;
        double m[3];
        maxuv(m);
        const double umax = m[0], vmaxm = m[2];
        const double dt_x = (umax > 0 ? P.cfl * (dx / umax) : 1e30);
        const double dt_y = (vmaxm > 0 ? P.cfl * (deta / vmaxm) : 1e30);
        const double dt_adv = std::fmin(dt_x, dt_y);
        const double dt_vis = (nu > 0) ? P.cvisc * 0.5 / (nu * (inv_dx2 + inv_dymin2)) : 1e30;
        return std::fmin(P.dt_max, std::fmin(dt_adv, dt_vis));
    }
};

void write_diag_header(const std::string& path, bool closure) {
    std::ofstream f(path, std::ios::out);
    f << "step,t,dt,umax,vmax,K,Omega,dOmega_rhs,eta,BOmega_wall,Pin,epsilon,POmega,tauw_bot,tauw_top,"
         "dpdx,utau_avg,Cf,Re_tau";
    if (closure)
        f << ",dOmega_rhs_same,dOmega_rhs_same_int,Omega_int,eta_int,Jwall_s1_cum,Jwall_s2_cum,Jwall_s3_cum,"
             "Omega_int_w,dOmega_same_int_w,A_w,P_int_w,eta_h,Phi_h";
    f << "\n";
}

int run_main_impl(int argc, char** argv) {
    OState S;
    if (!S.setup(argc, argv)) return 1;
    ORun& R = S.R;
    OGrid& G = R.G;
    OParams& P = R.P;
    const double nu = S.nu;
    ensure_dir(R.outdir);
    const std::string diag_path = R.outdir + "/diagnostics.csv";
    const std::string prof_path = R.outdir + "/profiles_timeavg.csv";
    {
        std::ofstream pl(R.outdir + "/progress.log", std::ios::out);
        pl << "oracle run Nx=" << G.Nx << " Ny=" << G.Ny << "\n";
    }
    // main.cu:160-165: the header only into an empty file (append-safe on resume)
    {
        std::ifstream chk(diag_path, std::ios::binary | std::ios::ate);
        if (!chk.good() || chk.tellg() <= 0) write_diag_header(diag_path, P.closure_diag);
    }
    { std::ofstream f(prof_path, std::ios::out); f << "j,y,yplus,U,V,uu,vv,uv,Uplus,tau_total\n"; }
    const size_t Ny = (size_t)G.Ny;
    // main.cu:129-147: a resumed run starts at the snapshot's time (its header's "t" in place of the
    // HDF5 attribute); step and dt are not restored (the step counter starts at 0 and dt at dt_init)
    double t_start = 0.0;
    if (P.resume && !read_snapshot_time(P.load_path, &t_start)) {
        std::fprintf(stderr, "oracle: --resume needs a readable --load_path snapshot\n");
        return 1;
    }
    auto snap_fields = [&](std::vector<double>& u, std::vector<double>& v) {
        std::vector<std::pair<std::string, const double*>> fs;
        fs.push_back({"omega", S.w.data()});
        if (P.snap_fields == "all" || P.snap_fields == "omega_psi") fs.push_back({"psi", S.psi.data()});
        if (P.snap_fields == "all") {
            orc_uv_from_psi_mapped(S.psi.data(), u.data(), v.data(), G.Nx, G.Ny, S.inv_dx, S.inv_deta, G.a_node.data());
            fs.push_back({"u", u.data()});
            fs.push_back({"v", v.data()});
        }
        return fs;
    };
    std::vector<double> ubuf(S.N), vbuf(S.N);
    const bool use_snap = (R.snap_save_dt > 0), use_diag = (R.diag_save_dt > 0);
    double next_snap_t = use_snap ? (P.resume ? next_after(t_start, R.snap_save_dt) : 0.0) : 1e300;  // main.cu:313
    double next_diag_t = use_diag ? (P.resume ? next_after(t_start, R.diag_save_dt) : 0.0) : 1e300;  // main.cu:314
    double tauw_abs_dt_sum = 0.0, T_diag = 0.0, last_diag_t = t_start;

    if (use_snap && !P.resume) {  // main.cu:321-324 (t = 0 outputs only for new runs)
        write_snapshot(R, 0.0, snap_fields(ubuf, vbuf));
        next_snap_t = R.snap_save_dt;
    }
    auto diag_row = [&](long long stepno, double t, double dt_col, bool t0_row, double sum_omega_rhs_mixed) {
        double m[3];
        S.maxuv(m);
        const double umax = m[0], vmax = m[1];
        const double sumK = orc_kinetic_energy_mapped(S.psi.data(), G.Nx, G.Ny, S.inv_dx, S.inv_deta,
                                                      G.a_node.data(), G.w_node.data());
        double ew[2];
        orc_enstrophy_diss_mapped_full(S.w.data(), G.Nx, G.Ny, S.inv_dx, S.inv_deta, G.a_node.data(),
                                       G.w_node.data(), ew);
        double wwy[2];
        orc_mean_wall_omega_omegay_mapped(S.w.data(), G.Nx, G.Ny, S.inv_deta, G.a_node.data(), wwy);
        double sum_omega_rhs = sum_omega_rhs_mixed;
        const double qnan = std::numeric_limits<double>::quiet_NaN();
        double cols[13];
        for (double& c : cols) c = qnan;
        if (!t0_row && P.closure_diag) {  // main.cu:591-611
            S.build_rhs(S.w, t);
            sum_omega_rhs = orc_omega_rhs_balance_mapped(S.w.data(), S.rhs.data(), G.Nx, G.Ny, G.w_node.data());
            const double wall_omega_rhs_same = S.wall_row_omega_rhs(S.w, S.rhs);
            double eint[2];
            orc_enstrophy_diss_mapped(S.w.data(), G.Nx, G.Ny, S.inv_dx, S.inv_deta, G.a_node.data(), G.w_node.data(), eint);
            const double t2_D = orc_omega_rhs_balance_mapped(S.w.data(), S.rhs.data(), G.Nx, G.Ny, S.wt_node.data());
            const double t2_J = orc_omega_rhs_balance_mapped(S.w.data(), S.J.data(), G.Nx, G.Ny, S.wt_node.data());
            double t2w[2];
            orc_enstrophy_diss_mapped(S.w.data(), G.Nx, G.Ny, S.inv_dx, S.inv_deta, G.a_node.data(), S.wt_node.data(), t2w);
            const double t2_P = orc_enstrophy_input_mapped_full(S.w.data(), G.Nx, G.Ny, G.h, P.F0, P.nforce,
                                                                G.y.data(), S.wt_node.data());
            double sbp[2];
            orc_enstrophy_sbp_mapped(S.w.data(), G.Nx, G.Ny, S.inv_dx, S.inv_deta, G.a_edge.data(), S.wt_node.data(), sbp);
            cols[0] = sum_omega_rhs * S.invA_full;
            cols[1] = cols[0] - wall_omega_rhs_same;
            cols[2] = 0.5 * eint[0] * S.invA_full;
            cols[3] = nu * eint[1] * S.invA_full;
            cols[4] = S.Jwall_cum[0]; cols[5] = S.Jwall_cum[1]; cols[6] = S.Jwall_cum[2];
            cols[7] = 0.5 * t2w[0] * S.invA_full;
            cols[8] = t2_D * S.invA_full;
            cols[9] = -t2_J * S.invA_full;
            cols[10] = t2_P * S.invA_full;
            cols[11] = nu * sbp[0] * S.invA_full;
            cols[12] = nu * sbp[1] * S.invA_full;
        }
        const double Pin = orc_power_input_mapped(S.psi.data(), G.Nx, G.Ny, S.inv_deta, G.h, P.F0, P.nforce,
                                                  G.a_node.data(), G.y.data(), G.w_node.data());
        const double sum_eps = orc_energy_dissipation_mapped(S.psi.data(), S.w.data(), G.Nx, G.Ny, S.inv_dx,
                                                             S.inv_deta, G.a_node.data(), G.a_edge.data(), G.w_node.data());
        const double p_omega_sum = orc_enstrophy_input_mapped_full(S.w.data(), G.Nx, G.Ny, G.h, P.F0, P.nforce,
                                                                   G.y.data(), G.w_node.data());
        const double K = sumK * S.invA_full;
        const double Omega = 0.5 * ew[0] * S.invA_full;
        const double eta = nu * ew[1] * S.invA_full;
        const double BOmega_wall = (nu > 0.0) ? (nu * (wwy[1] - wwy[0]) / (2.0 * G.h)) : 0.0;
        const double dOmega_rhs = sum_omega_rhs_mixed * S.invA_full;
        const double Pin1 = Pin * S.invA_full;
        const double epsilon = nu * sum_eps * S.invA_full;
        const double P_Omega = p_omega_sum * S.invA_full;
        double mw[2];
        orc_mean_wall_omega(S.w.data(), G.Nx, G.Ny, mw);
        const double tauw_bot = -nu * mw[0], tauw_top = -nu * mw[1];
        const double utau = std::sqrt(0.5 * (std::fabs(tauw_bot) + std::fabs(tauw_top)));
        const double Cf = (P.Ub != 0.0) ? 2.0 * (0.5 * (std::fabs(tauw_bot) + std::fabs(tauw_top))) / (P.Ub * P.Ub) : qnan;
        const double Re_tau = (nu > 0.0) ? utau * G.h / nu : 0.0;
        const double dpdx = (tauw_top - tauw_bot) / (2.0 * G.h) + S.meanFx_A - P.lin_drag * P.Ub;
        if (!t0_row) {
            double dtw = next_diag_t - last_diag_t;
            if (dtw < 0) dtw = 0;
            last_diag_t = next_diag_t;
            if (dtw > 0) { tauw_abs_dt_sum += 0.5 * (std::fabs(tauw_bot) + std::fabs(tauw_top)) * dtw; T_diag += dtw; }
        }
        std::ofstream f(diag_path, std::ios::app);
        f.setf(std::ios::scientific);
        f.precision(8);
        f << stepno << "," << t << "," << dt_col << "," << umax << "," << vmax << "," << K << "," << Omega << ","
          << dOmega_rhs << "," << eta << "," << BOmega_wall << "," << Pin1 << "," << epsilon << "," << P_Omega
          << "," << tauw_bot << "," << tauw_top << "," << dpdx << "," << utau << "," << Cf << "," << Re_tau;
        if (P.closure_diag) {
            if (t0_row)
                f << "," << dOmega_rhs << ",nan,nan,nan," << 0.0 << "," << 0.0 << "," << 0.0 << ",nan,nan,nan,nan,nan,nan";
            else
                for (double c : cols) f << "," << c;
        }
        f << "\n";
    };
    if (use_diag && !P.resume) {  // t = 0 row, main.cu:325-421
        S.build_rhs(S.w, t_start);
        const double mixed = orc_omega_rhs_balance_mapped(S.w.data(), S.rhs.data(), G.Nx, G.Ny, G.w_node.data());
        diag_row(0, 0.0, P.dt_init, true, mixed);
        next_diag_t = R.diag_save_dt;
    }
    std::vector<double> Hu(Ny, 0.0), Hv(Ny, 0.0), Hu2(Ny, 0.0), Hv2(Ny, 0.0), Huv(Ny, 0.0);
    std::vector<double> su(Ny), sv(Ny), su2(Ny), sv2(Ny), suv(Ny);
    // x-spectra (main.cu:265-268): velocity fields and the accumulators
    std::vector<double> su_f, sv_f, spec_acc;
    long long spec_samples = 0;
    if (R.spectraEvery > 0) {
        su_f.assign(S.N, 0.0);
        sv_f.assign(S.N, 0.0);
        spec_acc.assign(2 * (size_t)(G.Nx / 2 + 1), 0.0);
    }
    double T_accum = 0.0, last_stats_t = t_start;
    double t = t_start, dt = P.dt_init;
    long long step = 0;
    while (t < P.t_end - 0.5 * dt) {  // main.cu:434
        if (P.max_steps >= 0 && step >= P.max_steps) break;
        if (P.adapt) dt = S.adapt_dt();
        if (P.adapt && (use_snap || use_diag)) {  // main.cu:467-480
            double t_next = 1e300;
            if (use_snap) t_next = std::min(t_next, next_snap_t);
            if (use_diag) t_next = std::min(t_next, next_diag_t);
            if (t_next < 1e299) {
                const double rem = t_next - t;
                const double tol = std::max(time_tol(t_next), 1e-6 * dt);
                if (rem > tol && rem < 0.5 * dt) dt = rem;
            }
        }
        S.step(t, dt);
        t += dt;
        step++;
        if (use_snap && std::fabs(t - next_snap_t) <= sched_tol(t, dt)) t = next_snap_t;  // main.cu:544-551
        if (use_diag && std::fabs(t - next_diag_t) <= sched_tol(t, dt)) t = next_diag_t;
        if (use_snap) {
            while (t + sched_tol(t, dt) >= next_snap_t && next_snap_t <= P.t_end + sched_tol(P.t_end, dt)) {
                write_snapshot(R, next_snap_t, snap_fields(ubuf, vbuf));
                next_snap_t += R.snap_save_dt;
            }
        }
        if (use_diag && (t + time_tol(t) >= next_diag_t)) {  // main.cu:566-692
            const double mixed = orc_omega_rhs_balance_mapped(S.w.data(), S.rhs.data(), G.Nx, G.Ny, G.w_node.data());
            diag_row(step, t, dt, false, mixed);
            next_diag_t += R.diag_save_dt;
        }
        if (step % R.statsEvery == 0) {  // main.cu:695-733
            const double dtw = t - last_stats_t;
            if (dtw > 0) {
                orc_row_moments_from_psi_mapped(S.psi.data(), G.Nx, G.Ny, S.inv_dx, S.inv_deta, G.a_node.data(),
                                                su.data(), sv.data(), su2.data(), sv2.data(), suv.data());
                for (size_t j = 0; j < Ny; j++) {
                    Hu[j] += su[j] / (double)G.Nx * dtw; Hv[j] += sv[j] / (double)G.Nx * dtw;
                    Hu2[j] += su2[j] / (double)G.Nx * dtw; Hv2[j] += sv2[j] / (double)G.Nx * dtw;
                    Huv[j] += suv[j] / (double)G.Nx * dtw;
                }
                T_accum += dtw;
                last_stats_t = t;
                if (R.profilesWriteEvery > 0 && (step % R.profilesWriteEvery) == 0) {
                    double ul = 0.0;
                    if (T_diag > 0) ul = std::sqrt(tauw_abs_dt_sum / T_diag);
                    else {
                        double mw[2];
                        orc_mean_wall_omega(S.w.data(), G.Nx, G.Ny, mw);
                        ul = std::sqrt(0.5 * (std::fabs(-nu * mw[0]) + std::fabs(-nu * mw[1])));
                    }
                    write_profiles_csv(prof_path, G, P, ul, T_accum, Hu, Hv, Hu2, Hv2, Huv);
                }
            }
        }
        if (R.spectraEvery > 0 && (step % R.spectraEvery == 0)) {  // main.cu:735-744
            orc_uv_from_psi_mapped(S.psi.data(), su_f.data(), sv_f.data(), G.Nx, G.Ny, S.inv_dx, S.inv_deta,
                                   G.a_node.data());
            orc_spectra1d_accumulate(su_f.data(), sv_f.data(), G.Nx, G.Ny, spec_acc.data());
            ++spec_samples;
        }
    }
    double utau_mean = 0.0;  // main.cu:757-774
    if (T_diag > 0) utau_mean = std::sqrt(tauw_abs_dt_sum / T_diag);
    else {
        double mw[2];
        orc_mean_wall_omega(S.w.data(), G.Nx, G.Ny, mw);
        utau_mean = std::sqrt(0.5 * (std::fabs(-nu * mw[0]) + std::fabs(-nu * mw[1])));
    }
    write_profiles_csv(prof_path, G, P, utau_mean, T_accum, Hu, Hv, Hu2, Hv2, Huv);
    if (R.spectraEvery > 0) {  // main.cu:776-780, diag.cuh:175-190
        const int nk = G.Nx / 2 + 1;
        std::ofstream f(R.outdir + "/spectra_kx.csv");
        f.setf(std::ios::scientific);
        f.precision(10);
        f << "kx,E_u,E_v,E_tot\n";
        for (int k = 0; k < nk; ++k) {
            double eu = spec_acc[k], ev = spec_acc[nk + k];
            if (spec_samples > 0) { eu /= spec_samples; ev /= spec_samples; }
            const double kx = (2.0 * PI * (double)k) / G.Lx;  // diag.cu:870-875
            f << kx << "," << eu << "," << ev << "," << (eu + ev) << "\n";
        }
    }
    return 0;
}

std::vector<char*> argv_with_prog(int argc, char** argv) {
    std::vector<char*> v;
    for (int i = 0; i < argc; ++i) v.push_back(argv[i]);
    return v;
}

}  // namespace

extern "C" int orc_init_omega(int argc, char** argv, double* w_out) {
    ORun R;
    parse_args(argc, argv, R);
    if (!build_grid(R)) return 1;
    std::vector<double> w;
    if (!init_omega_with_IC(R, w)) return 2;
    std::memcpy(w_out, w.data(), w.size() * sizeof(double));
    return 0;
}

extern "C" int orc_run_main(int argc, char** argv) {
    try {
        auto v = argv_with_prog(argc, argv);
        return run_main_impl((int)v.size(), v.data());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "oracle: %s\n", e.what());
        return 3;
    }
}

extern "C" double orc_time_steps(int argc, char** argv, int warm, int steps) {
    OState S;
    if (!S.setup(argc, argv)) return -1.0;
    double t = 0.0;
    double dt = S.R.P.dt_init;
    for (int s = 0; s < warm; ++s) {
        if (S.R.P.adapt) dt = S.adapt_dt();
        S.step(t, dt);
        t += dt;
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; ++s) {
        if (S.R.P.adapt) dt = S.adapt_dt();
        S.step(t, dt);
        t += dt;
    }
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

extern "C" int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
