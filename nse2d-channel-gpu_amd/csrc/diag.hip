// Run-time diagnostics of the reference (diag.cu): budgets, wall quantities, row moments and the
// CFL maxima. Each reduction is two launches: one workgroup per grid row sums its row with
// wavefront shuffles and writes a per-row partial, then one workgroup adds the partials in row
// order. The result is deterministic (the reference's one-atomic-per-block sums are not) and no
// output memset is needed.
#include "nse2d_common.hpp"
#include "nse2d_ops.hpp"
#include "../../include/nse2d_hip.h"

#include <mutex>
#include <string>

// Row partials of launch_max_uv_from_psi_mapped_noalloc: a fixed array that each GPU holds from
// the library's load until the process ends (3 x NSE2D_NOALLOC_MAX_ROWS doubles, 393 KB). Host
// code never uses this name as an address: the device address comes from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[3 * NSE2D_NOALLOC_MAX_ROWS];

namespace nse2d {

void set_last_error(const std::string& msg);

namespace {

constexpr int RB = 256;  // threads per row workgroup

template <int NACC>
struct Outs {
    double* p[NACC];
};

template <int NACC, class F>
__global__ void __launch_bounds__(RB) row_sum_kernel(F f, int Nx, int Ny, double* __restrict__ partial) {
    __shared__ double sm[NACC * 16];
    const int iy = blockIdx.x;
    double acc[NACC];
#pragma unroll
    for (int a = 0; a < NACC; ++a) acc[a] = 0.0;
    if (f.active(iy))
        for (int ix = threadIdx.x; ix < Nx; ix += RB) f(ix, iy, acc);
    block_sum<NACC>(acc, sm);
    if (threadIdx.x == 0)
#pragma unroll
        for (int a = 0; a < NACC; ++a) partial[(size_t)a * Ny + iy] = acc[a];
}

template <int NACC, bool MAX>
__global__ void __launch_bounds__(RB) final_kernel(const double* __restrict__ partial, int n,
                                                   Outs<NACC> out, double scale) {
    __shared__ double sm[NACC * 16];
    double acc[NACC];
#pragma unroll
    for (int a = 0; a < NACC; ++a) {
        double s = 0.0;
        for (int i = threadIdx.x; i < n; i += RB) {
            const double v = partial[(size_t)a * n + i];
            s = MAX ? fmax(s, v) : s + v;
        }
        acc[a] = s;
    }
    if (MAX) block_max<NACC>(acc, sm);
    else block_sum<NACC>(acc, sm);
    if (threadIdx.x == 0)
#pragma unroll
        for (int a = 0; a < NACC; ++a) *out.p[a] = acc[a] * scale;
}

template <int NACC, class F>
void reduce_rows(const F& f, int Nx, int Ny, Outs<NACC> out, hipStream_t s, double scale = 1.0) {
    double* partial = scratch_doubles((size_t)NACC * Ny);
    row_sum_kernel<NACC, F><<<Ny, RB, 0, s>>>(f, Nx, Ny, partial);
    NSE2D_LAUNCH_CHECK();
    final_kernel<NACC, false><<<1, RB, 0, s>>>(partial, Ny, out, scale);
    NSE2D_LAUNCH_CHECK();
}

// ---- diag.cu:6-38 (interior rows; also used with the operator weights wt = deta / a_j)
struct EnstDiss {
    const double* w; int Nx, Ny; double inv_dx, inv_deta; const double *a_node, *w_node;
    __device__ bool active(int iy) const { return iy > 0 && iy < Ny - 1; }
    __device__ void operator()(int ix, int iy, double (&acc)[2]) const {
        const int xm = (ix == 0 ? Nx - 1 : ix - 1), xp = (ix == Nx - 1 ? 0 : ix + 1);
        const double wc = w[idx2(iy, ix, Nx)];
        const double wx = (w[idx2(iy, xp, Nx)] - w[idx2(iy, xm, Nx)]) * 0.5 * inv_dx;
        const double wy = a_node[iy] * (w[idx2(iy + 1, ix, Nx)] - w[idx2(iy - 1, ix, Nx)]) * 0.5 * inv_deta;
        const double wt = w_node[iy];
        acc[0] += wt * (wc * wc);
        acc[1] += wt * (wx * wx + wy * wy);
    }
};

// ---- diag.cu:43-101 (all rows; second-order one-sided eta derivative on the wall rows)
struct EnstDissFull {
    const double* w; int Nx, Ny; double inv_dx, inv_deta; const double *a_node, *w_node;
    __device__ bool active(int) const { return true; }
    __device__ void operator()(int ix, int iy, double (&acc)[2]) const {
        const int xm = (ix == 0 ? Nx - 1 : ix - 1), xp = (ix == Nx - 1 ? 0 : ix + 1);
        const double wc = w[idx2(iy, ix, Nx)];
        const double wx = (w[idx2(iy, xp, Nx)] - w[idx2(iy, xm, Nx)]) * 0.5 * inv_dx;
        double weta;
        if (iy == 0)
            weta = (-3.0 * w[idx2(0, ix, Nx)] + 4.0 * w[idx2(1, ix, Nx)] - 1.0 * w[idx2(2, ix, Nx)]) * (0.5 * inv_deta);
        else if (iy == Ny - 1)
            weta = (3.0 * w[idx2(Ny - 1, ix, Nx)] - 4.0 * w[idx2(Ny - 2, ix, Nx)] + 1.0 * w[idx2(Ny - 3, ix, Nx)]) *
                   (0.5 * inv_deta);
        else
            weta = (w[idx2(iy + 1, ix, Nx)] - w[idx2(iy - 1, ix, Nx)]) * 0.5 * inv_deta;
        const double wy = a_node[iy] * weta;
        const double wt = w_node[iy];
        acc[0] += wt * (wc * wc);
        acc[1] += wt * (wx * wx + wy * wy);
    }
};

// ---- diag.cu:103-156 (interior rows; lap supplied)
struct EnstBalance {
    const double *w, *lap; int Nx, Ny; const double* w_node;
    __device__ bool active(int iy) const { return iy > 0 && iy < Ny - 1; }
    __device__ void operator()(int ix, int iy, double (&acc)[2]) const {
        const double wc = w[idx2(iy, ix, Nx)], l = lap[idx2(iy, ix, Nx)], wt = w_node[iy];
        acc[0] += wt * (wc * wc);
        acc[1] += wt * (wc * l);
    }
};

// ---- diag.cu:158-199 (summation-by-parts remainders)
struct EnstSbp {
    const double* w; int Nx, Ny; double inv_dx, inv_deta; const double *a_edge, *wt_node;
    __device__ bool active(int) const { return true; }
    __device__ void operator()(int ix, int iy, double (&acc)[2]) const {
        const int xp = (ix == Nx - 1 ? 0 : ix + 1);
        const double w0 = w[idx2(iy, ix, Nx)];
        if (iy >= 1 && iy <= Ny - 2) {
            const double dxw = (w[idx2(iy, xp, Nx)] - w0) * inv_dx;
            acc[0] += wt_node[iy] * dxw * dxw;
        }
        if (iy <= Ny - 2) {
            const double dw = w[idx2(iy + 1, ix, Nx)] - w0;
            acc[0] += a_edge[iy] * dw * dw * inv_deta;
        }
        if (iy == 0) acc[1] -= w0 * a_edge[0] * (w[idx2(1, ix, Nx)] - w0) * inv_deta;
        if (iy == Ny - 1) acc[1] += w0 * a_edge[Ny - 2] * (w0 - w[idx2(Ny - 2, ix, Nx)]) * inv_deta;
    }
};

// ---- diag.cu:201-231
struct KineticEnergy {
    const double* psi; int Nx, Ny; double inv_dx, inv_deta; const double *a_node, *w_node;
    __device__ bool active(int iy) const { return iy > 0 && iy < Ny - 1; }
    __device__ void operator()(int ix, int iy, double (&acc)[1]) const {
        const int xm = (ix == 0 ? Nx - 1 : ix - 1), xp = (ix == Nx - 1 ? 0 : ix + 1);
        const double u = (psi[idx2(iy + 1, ix, Nx)] - psi[idx2(iy - 1, ix, Nx)]) * 0.5 * inv_deta * a_node[iy];
        const double v = -(psi[idx2(iy, xp, Nx)] - psi[idx2(iy, xm, Nx)]) * 0.5 * inv_dx;
        acc[0] += 0.5 * (u * u + v * v) * w_node[iy];
    }
};

// ---- diag.cu:250-348
struct EnergyDiss {
    const double *psi, *omega; int Nx, Ny; double inv_dx, inv_deta; const double *a_node, *a_edge, *w_node;
    __device__ bool active(int) const { return true; }
    __device__ void operator()(int i, int j, double (&acc)[1]) const {
        if (j == 0 || j == Ny - 1) {
            const double uy = -omega[idx2(j, i, Nx)];
            acc[0] += (uy * uy) * w_node[j];
            return;
        }
        const double inv_dx2 = inv_dx * inv_dx, inv_deta2 = inv_deta * inv_deta;
        const int im = (i == 0 ? Nx - 1 : i - 1), ip = (i == Nx - 1 ? 0 : i + 1);
        const double a = a_node[j], ap = a_edge[j], am = a_edge[j - 1];
        const double psi_x_jp = (psi[idx2(j + 1, ip, Nx)] - psi[idx2(j + 1, im, Nx)]) * (0.5 * inv_dx);
        const double psi_x_jm = (psi[idx2(j - 1, ip, Nx)] - psi[idx2(j - 1, im, Nx)]) * (0.5 * inv_dx);
        const double ux = a * (psi_x_jp - psi_x_jm) * (0.5 * inv_deta);
        const double vy = -ux;
        const double pc = psi[idx2(j, i, Nx)];
        const double vx = -(psi[idx2(j, ip, Nx)] - 2.0 * pc + psi[idx2(j, im, Nx)]) * inv_dx2;
        const double uy = a * (ap * (psi[idx2(j + 1, i, Nx)] - pc) - am * (pc - psi[idx2(j - 1, i, Nx)])) * inv_deta2;
        acc[0] += (ux * ux + uy * uy + vx * vx + vy * vy) * w_node[j];
    }
};

// ---- diag.cu:350-376 (clamped y neighbours on the wall rows)
struct PowerInput {
    const double* psi; int Nx, Ny; double inv_deta, h, F0; int nforce;
    const double *a_node, *y_node, *w_node;
    __device__ bool active(int) const { return true; }
    __device__ void operator()(int ix, int iy, double (&acc)[1]) const {
        const int ym = (iy == 0 ? 0 : iy - 1), yp = (iy == Ny - 1 ? Ny - 1 : iy + 1);
        const double u = (psi[idx2(yp, ix, Nx)] - psi[idx2(ym, ix, Nx)]) * 0.5 * inv_deta * a_node[iy];
        const double y = y_node[iy];
        const double Fx = F0 * sin((double)nforce * kPi * ((y + h) / (2.0 * h)));
        acc[0] += u * Fx * w_node[iy];
    }
};

// ---- diag.cu:378-467 (full: all rows; otherwise interior rows)
struct EnstInput {
    const double* w; int Nx, Ny; double h, F0; int nforce; const double *y_node, *w_node; bool full;
    __device__ bool active(int iy) const { return full || (iy > 0 && iy < Ny - 1); }
    __device__ void operator()(int ix, int iy, double (&acc)[1]) const {
        const double y = y_node[iy];
        const double arg = (double)nforce * kPi * ((y + h) / (2.0 * h));
        const double f_omega = -F0 * ((double)nforce * kPi / (2.0 * h)) * cos(arg);
        acc[0] += w[idx2(iy, ix, Nx)] * f_omega * w_node[iy];
    }
};

// ---- diag.cu:628-691 (all rows; 64-bit indexing, unlike the reference's int)
struct OmegaRhs {
    const double *w, *rhs; int Nx, Ny; const double* w_node;
    __device__ bool active(int) const { return true; }
    __device__ void operator()(int ix, int iy, double (&acc)[1]) const {
        const size_t c = idx2(iy, ix, Nx);
        acc[0] += w_node[iy] * w[c] * rhs[c];
    }
};

// The five row-moment sums of one point (diag.cu:468-499: u with clamped y neighbours, centred v),
// with the multiply-adds written out and no other contraction, so that every kernel using this
// rounds it alike.
__device__ __forceinline__ void moments_point(const double* __restrict__ psi, int Nx, int ix, int iy, int ym, int yp,
                                              double inv_dx, double inv_deta, double a, double (&acc)[5]) {
    // no contraction of u = d * a into the sum acc[0] + u: whether the compiler would fuse it
    // depends on the other uses of u in the calling kernel
#pragma clang fp contract(off)
    const int xm = (ix == 0 ? Nx - 1 : ix - 1), xp = (ix == Nx - 1 ? 0 : ix + 1);
    const double u = (psi[idx2(yp, ix, Nx)] - psi[idx2(ym, ix, Nx)]) * 0.5 * inv_deta * a;
    const double v = -(psi[idx2(iy, xp, Nx)] - psi[idx2(iy, xm, Nx)]) * 0.5 * inv_dx;
    acc[0] += u;
    acc[1] += v;
    acc[2] = __builtin_fma(u, u, acc[2]);
    acc[3] = __builtin_fma(v, v, acc[3]);
    acc[4] = __builtin_fma(u, v, acc[4]);
}

// ---- diag.cu:468-499: one workgroup per row, five sums, clamped wall rows
__global__ void __launch_bounds__(RB) row_moments_kernel(const double* __restrict__ psi, int Nx, int Ny,
                                                         double inv_dx, double inv_deta,
                                                         const double* __restrict__ a_node, double* su,
                                                         double* sv, double* su2, double* sv2, double* suv) {
    __shared__ double sm[5 * 16];
    const int iy = blockIdx.x;
    double acc[5] = {0, 0, 0, 0, 0};
    const int ym = (iy == 0 ? 0 : iy - 1), yp = (iy == Ny - 1 ? Ny - 1 : iy + 1);
    const double a = a_node[iy];
    for (int ix = threadIdx.x; ix < Nx; ix += RB) moments_point(psi, Nx, ix, iy, ym, yp, inv_dx, inv_deta, a, acc);
    block_sum<5>(acc, sm);
    if (threadIdx.x == 0) { su[iy] = acc[0]; sv[iy] = acc[1]; su2[iy] = acc[2]; sv2[iy] = acc[3]; suv[iy] = acc[4]; }
}

// omega * a * omega_eta on a wall row (diag.cu:516-583): second-order one-sided eta derivative
// into the channel (rows 0, 1, 2 at the bottom, Ny-1, Ny-2, Ny-3 at the top). The multiply-adds
// are written out, so that every kernel using this rounds it alike.
__device__ __forceinline__ double wall_omega_omegay(double w0, double w1, double w2, double a, double half_inv_deta,
                                                    bool bottom) {
#pragma clang fp contract(off)
    const double t = bottom ? __builtin_fma(-3.0, w0, 4.0 * w1) - w2 : __builtin_fma(3.0, w0, -(4.0 * w1)) + w2;
    return w0 * (a * (t * half_inv_deta));
}

// ---- diag.cu:500-512 and 516-583: x-means over the two wall rows (one workgroup)
template <bool WITH_GRAD>
__global__ void __launch_bounds__(RB) wall_means_kernel(const double* __restrict__ w, int Nx, int Ny,
                                                        double inv_deta, const double* __restrict__ a_node,
                                                        double* mean_bot, double* mean_top) {
    __shared__ double sm[2 * 16];
    double acc[2] = {0, 0};
    for (int ix = threadIdx.x; ix < Nx; ix += RB) {
        const double w0 = w[idx2(0, ix, Nx)], wN = w[idx2(Ny - 1, ix, Nx)];
        if (WITH_GRAD) {
            acc[0] += wall_omega_omegay(w0, w[idx2(1, ix, Nx)], w[idx2(2, ix, Nx)], a_node[0], 0.5 * inv_deta, true);
            acc[1] += wall_omega_omegay(wN, w[idx2(Ny - 2, ix, Nx)], w[idx2(Ny - 3, ix, Nx)], a_node[Ny - 1],
                                        0.5 * inv_deta, false);
        } else {
            acc[0] += w0;
            acc[1] += wN;
        }
    }
    block_sum<2>(acc, sm);
    if (threadIdx.x == 0) {
        *mean_bot = acc[0] / (double)Nx;
        *mean_top = acc[1] / (double)Nx;
    }
}

__global__ void __launch_bounds__(RB) row_max_uv_kernel(const double* __restrict__ psi, int Nx, int Ny,
                                                        double inv_dx, double inv_deta,
                                                        const double* __restrict__ a_node,
                                                        double* __restrict__ partial) {
    __shared__ double sm[3 * 16];
    const int j = xcd_tile(blockIdx.x, gridDim.x);  // neighbouring rows share an L2
    row_max_uv_body(psi, Nx, Ny, inv_dx, inv_deta, a_node, j, partial, sm);
}

__global__ void zero3_kernel(double* a, double* b, double* c) {
    if (threadIdx.x == 0) { *a = 0.0; *b = 0.0; *c = 0.0; }
}

// ---- One diagnostics row in one pass (main.cu:565-692 without --closure_diag): workgroup iy reads
// rows iy-1..iy+1 of psi and omega (the neighbour rows hit in L2) and row iy of rhs, and forms the
// row partials of every reduction of the row with the per-point bodies of the launchers above,
// each accumulator in the same order as its own launcher's row kernel; the wall rows' workgroups
// also form the wall means. The final kernel then adds (or, for the CFL maxima, takes the maximum
// of) the row partials in row order, as the launchers' final kernels do: every output equals the
// separate launchers' bit for bit. Outputs: NSE2D_DIAG_* of include/nse2d_hip.h.
struct DiagRowArgs {
    const double *psi, *omega, *rhs;
    int Nx, Ny;
    double inv_dx, inv_deta, h, F0;
    int nforce;
    const double *a_node, *a_edge, *y_node, *w_node;
    double *out, *moments;  // moments: [5][Ny] row sums of u, v, u^2, v^2, uv, or nullptr
};

constexpr int kDiagSums = 7;  // K, W2, GW2, omega*rhs, Pin, eps, P_Omega

__global__ void __launch_bounds__(RB) diag_row_kernel(DiagRowArgs A, double* __restrict__ partial) {
    __shared__ double sm_s[kDiagSums * 16], sm_m[3 * 16], sm_w[5 * 16];
    const int iy = blockIdx.x, Nx = A.Nx, Ny = A.Ny;
    const KineticEnergy ke{A.psi, Nx, Ny, A.inv_dx, A.inv_deta, A.a_node, A.w_node};
    const EnstDissFull ed{A.omega, Nx, Ny, A.inv_dx, A.inv_deta, A.a_node, A.w_node};
    const OmegaRhs orh{A.omega, A.rhs, Nx, Ny, A.w_node};
    const PowerInput pin{A.psi, Nx, Ny, A.inv_deta, A.h, A.F0, A.nforce, A.a_node, A.y_node, A.w_node};
    const EnergyDiss eps{A.psi, A.omega, Nx, Ny, A.inv_dx, A.inv_deta, A.a_node, A.a_edge, A.w_node};
    const EnstInput pom{A.omega, Nx, Ny, A.h, A.F0, A.nforce, A.y_node, A.w_node, true};
    const bool interior = iy > 0 && iy < Ny - 1;
    double k1[1] = {0.0}, e2[2] = {0.0, 0.0}, o1[1] = {0.0}, p1[1] = {0.0}, s1[1] = {0.0}, q1[1] = {0.0};
    double m[3] = {0.0, 0.0, 0.0};
    double mo[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const int ym = (iy == 0 ? 0 : iy - 1), yp = (iy == Ny - 1 ? Ny - 1 : iy + 1);
    const double a = A.a_node[iy];
    for (int ix = threadIdx.x; ix < Nx; ix += RB) {
        if (interior) {
            ke(ix, iy, k1);
            // row_max_uv_body (diag.cu:702-866)
            const int ip = (ix + 1 < Nx) ? (ix + 1) : 0, im = (ix > 0) ? (ix - 1) : (Nx - 1);
            const double dpsideta = (A.psi[idx2(iy + 1, ix, Nx)] - A.psi[idx2(iy - 1, ix, Nx)]) * (0.5 * A.inv_deta);
            const double dpsidx = (A.psi[idx2(iy, ip, Nx)] - A.psi[idx2(iy, im, Nx)]) * (0.5 * A.inv_dx);
            const double au = fabs(a * dpsideta), av = fabs(-dpsidx);
            m[0] = fmax(m[0], au);
            m[1] = fmax(m[1], av);
            m[2] = fmax(m[2], av * a);
        }
        ed(ix, iy, e2);
        orh(ix, iy, o1);
        pin(ix, iy, p1);
        eps(ix, iy, s1);
        pom(ix, iy, q1);
        if (A.moments) moments_point(A.psi, Nx, ix, iy, ym, yp, A.inv_dx, A.inv_deta, a, mo);  // diag.cu:468-499
    }
    double sums[kDiagSums] = {k1[0], e2[0], e2[1], o1[0], p1[0], s1[0], q1[0]};
    block_sum<kDiagSums>(sums, sm_s);
    block_max<3>(m, sm_m);
    if (A.moments) block_sum<5>(mo, sm_w);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kDiagSums; ++k) partial[(size_t)k * Ny + iy] = sums[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) partial[(size_t)(kDiagSums + k) * Ny + iy] = m[k];
        if (A.moments)
#pragma unroll
            for (int k = 0; k < 5; ++k) A.moments[(size_t)k * Ny + iy] = mo[k];
    }
    if (iy != 0 && iy != Ny - 1) return;  // uniform over the workgroup
    // wall means (diag.cu:500-512, 516-583), one wall per workgroup
    __syncthreads();  // sm_w is reused
    const bool bot = (iy == 0);
    double wa[2] = {0.0, 0.0};
    for (int ix = threadIdx.x; ix < Nx; ix += RB) {
        const double w0 = A.omega[idx2(bot ? 0 : Ny - 1, ix, Nx)];
        const double w1 = A.omega[idx2(bot ? 1 : Ny - 2, ix, Nx)], w2 = A.omega[idx2(bot ? 2 : Ny - 3, ix, Nx)];
        wa[0] += wall_omega_omegay(w0, w1, w2, A.a_node[bot ? 0 : Ny - 1], 0.5 * A.inv_deta, bot);
        wa[1] += w0;
    }
    block_sum<2>(wa, sm_w);
    if (threadIdx.x == 0) {
        A.out[bot ? NSE2D_DIAG_WALL_OMEGA_OMEGAY_BOT : NSE2D_DIAG_WALL_OMEGA_OMEGAY_TOP] = wa[0] / (double)Nx;
        A.out[bot ? NSE2D_DIAG_WALL_OMEGA_BOT : NSE2D_DIAG_WALL_OMEGA_TOP] = wa[1] / (double)Nx;
    }
}

// The row partials of diag_row_kernel in row order: sums, then the CFL maxima.
__global__ void __launch_bounds__(RB) diag_final_kernel(const double* __restrict__ partial, int n,
                                                        double* __restrict__ out) {
    __shared__ double sm_s[kDiagSums * 16], sm_m[3 * 16];
    double s[kDiagSums], m[3];
#pragma unroll
    for (int a = 0; a < kDiagSums + 3; ++a) {
        if (a < kDiagSums) s[a] = 0.0;
        else m[a - kDiagSums] = 0.0;
    }
    // the partials of this thread (rows tid, tid + RB, ...) added in row order, as final_kernel
    // does, with eight loads issued before their additions
    constexpr int B = 8;
    for (int i0 = threadIdx.x; i0 < n; i0 += B * RB) {
#pragma unroll
        for (int a = 0; a < kDiagSums + 3; ++a) {
            double v[B];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int i = i0 + b * RB;
                v[b] = (i < n) ? partial[(size_t)a * n + i] : 0.0;
            }
#pragma unroll
            for (int b = 0; b < B; ++b) {
                if (i0 + b * RB >= n) break;
                if (a < kDiagSums) s[a] += v[b];
                else m[a - kDiagSums] = fmax(m[a - kDiagSums], v[b]);
            }
        }
    }
    block_sum<kDiagSums>(s, sm_s);
    block_max<3>(m, sm_m);
    if (threadIdx.x == 0) {
        out[NSE2D_DIAG_K] = s[0];
        out[NSE2D_DIAG_OMEGA_W2] = s[1];
        out[NSE2D_DIAG_OMEGA_GRAD2] = s[2];
        out[NSE2D_DIAG_OMEGA_RHS] = s[3];
        out[NSE2D_DIAG_PIN] = s[4];
        out[NSE2D_DIAG_EPS] = s[5];
        out[NSE2D_DIAG_POMEGA] = s[6];
        out[NSE2D_DIAG_UMAX] = m[0];
        out[NSE2D_DIAG_VMAX] = m[1];
        out[NSE2D_DIAG_VMAXM] = m[2];
    }
}

// The two launches of the CFL maxima with the row partials in `partial` (3 x Ny doubles; every
// row writes its three entries, so nothing needs clearing between calls).
void launch_max_uv_into(const double* psi, int Nx, int Ny, double inv_dx, double inv_deta, const double* a_node,
                        double* umax, double* vmax, double* vmaxm, double* partial, hipStream_t s) {
    row_max_uv_kernel<<<Ny, RB, 0, s>>>(psi, Nx, Ny, inv_dx, inv_deta, a_node, partial);
    NSE2D_LAUNCH_CHECK();
    final_kernel<3, true><<<1, RB, 0, s>>>(partial, Ny, Outs<3>{{umax, vmax, vmaxm}}, 1.0);
    NSE2D_LAUNCH_CHECK();
}

void launch_zero3(double* umax, double* vmax, double* vmaxm, hipStream_t s) {  // diag.cu:826-832
    zero3_kernel<<<1, 64, 0, s>>>(umax, vmax, vmaxm);
    NSE2D_LAUNCH_CHECK();
}

// Device address of `scratch` on the current device, looked up once per device with that device
// current (each GPU holds its own copy at its own address). nullptr with the error set on failure.
constexpr int kNoallocDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kNoallocDevices] = {};

double* noalloc_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoallocDevices) {
        set_last_error("launch_max_uv_from_psi_mapped_noalloc: no usable current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !p) {
            set_last_error(std::string("launch_max_uv_from_psi_mapped_noalloc: hipGetSymbolAddress failed: ") +
                           hipGetErrorString(rc));
            return nullptr;
        }
        g_noalloc_ptr[dev] = static_cast<double*>(p);
    }
    return g_noalloc_ptr[dev];
}

}  // namespace

void launch_max_uv(const double* psi, int Nx, int Ny, double inv_dx, double inv_deta, const double* a_node,
                   double* umax, double* vmax, double* vmaxm, hipStream_t s) {
    if (Nx < 2 || Ny < 3) {
        launch_zero3(umax, vmax, vmaxm, s);
        return;
    }
    launch_max_uv_into(psi, Nx, Ny, inv_dx, inv_deta, a_node, umax, vmax, vmaxm, scratch_doubles((size_t)3 * Ny), s);
}

}  // namespace nse2d

using namespace nse2d;

extern "C" {

void launch_enstrophy_diss_mapped(const double* omg, int Nx, int Ny, double inv_dx, double inv_deta,
                                  const double* d_a_node, const double* d_w_node, double* d_sum_w2,
                                  double* d_sum_gw2, hipStream_t s) {
    reduce_rows<2>(EnstDiss{omg, Nx, Ny, inv_dx, inv_deta, d_a_node, d_w_node}, Nx, Ny,
                   Outs<2>{{d_sum_w2, d_sum_gw2}}, s);
}

void launch_enstrophy_diss_mapped_full(const double* omg, int Nx, int Ny, double inv_dx, double inv_deta,
                                       const double* d_a_node, const double* d_w_node,
                                       double* d_sum_w2_full, double* d_sum_gw2_full, hipStream_t s) {
    reduce_rows<2>(EnstDissFull{omg, Nx, Ny, inv_dx, inv_deta, d_a_node, d_w_node}, Nx, Ny,
                   Outs<2>{{d_sum_w2_full, d_sum_gw2_full}}, s);
}

void launch_enstrophy_sbp_mapped(const double* omg, int Nx, int Ny, double inv_dx, double inv_deta,
                                 const double* d_a_edge, const double* d_wt_node, double* d_sum_eta_h,
                                 double* d_sum_phi_h, hipStream_t s) {
    reduce_rows<2>(EnstSbp{omg, Nx, Ny, inv_dx, inv_deta, d_a_edge, d_wt_node}, Nx, Ny,
                   Outs<2>{{d_sum_eta_h, d_sum_phi_h}}, s);
}

void launch_kinetic_energy_mapped(const double* psi, int Nx, int Ny, double inv_dx, double inv_deta,
                                  const double* d_a_node, const double* d_w_node, double* d_sum,
                                  hipStream_t s) {
    reduce_rows<1>(KineticEnergy{psi, Nx, Ny, inv_dx, inv_deta, d_a_node, d_w_node}, Nx, Ny,
                   Outs<1>{{d_sum}}, s);
}

void launch_energy_dissipation_mapped(const double* d_psi, const double* d_omega, int Nx, int Ny,
                                      double inv_dx, double inv_deta, const double* d_a_node,
                                      const double* d_a_edge, const double* d_w_node, double* d_sum_eps,
                                      hipStream_t s) {
    reduce_rows<1>(EnergyDiss{d_psi, d_omega, Nx, Ny, inv_dx, inv_deta, d_a_node, d_a_edge, d_w_node}, Nx,
                   Ny, Outs<1>{{d_sum_eps}}, s);
}

void launch_power_input_mapped(const double* psi, int Nx, int Ny, double inv_deta, double h, double F0,
                               int nforce, const double* d_a_node, const double* d_y_node,
                               const double* d_w_node, double* d_sum, hipStream_t s) {
    reduce_rows<1>(PowerInput{psi, Nx, Ny, inv_deta, h, F0, nforce, d_a_node, d_y_node, d_w_node}, Nx, Ny,
                   Outs<1>{{d_sum}}, s);
}

void launch_enstrophy_input_mapped_full(const double* omega, int Nx, int Ny, double h, double F0,
                                        int nforce, const double* d_y_node, const double* d_w_node,
                                        double* d_sum, hipStream_t s) {
    reduce_rows<1>(EnstInput{omega, Nx, Ny, h, F0, nforce, d_y_node, d_w_node, true}, Nx, Ny,
                   Outs<1>{{d_sum}}, s);
}

void launch_enstrophy_input_mapped(const double* omega, int Nx, int Ny, double h, double F0, int nforce,
                                   const double* d_y_node, const double* d_w_node, double* d_sum,
                                   hipStream_t s) {
    reduce_rows<1>(EnstInput{omega, Nx, Ny, h, F0, nforce, d_y_node, d_w_node, false}, Nx, Ny,
                   Outs<1>{{d_sum}}, s);
}

void launch_row_moments_from_psi_mapped(const double* psi, int Nx, int Ny, double inv_dx, double inv_deta,
                                        const double* d_a_node, double* su, double* sv, double* su2,
                                        double* sv2, double* suv, hipStream_t s) {
    row_moments_kernel<<<Ny, RB, 0, s>>>(psi, Nx, Ny, inv_dx, inv_deta, d_a_node, su, sv, su2, sv2, suv);
    NSE2D_LAUNCH_CHECK();
}

void launch_mean_wall_omega(const double* omega, int Nx, int Ny, double* d_mean_bot, double* d_mean_top,
                            hipStream_t s) {
    wall_means_kernel<false><<<1, RB, 0, s>>>(omega, Nx, Ny, 0.0, nullptr, d_mean_bot, d_mean_top);
    NSE2D_LAUNCH_CHECK();
}

void launch_mean_wall_omega_omegay_mapped(const double* omega, int Nx, int Ny, double inv_deta,
                                          const double* d_a_node, double* d_mean_bot, double* d_mean_top,
                                          hipStream_t s) {
    wall_means_kernel<true><<<1, RB, 0, s>>>(omega, Nx, Ny, inv_deta, d_a_node, d_mean_bot, d_mean_top);
    NSE2D_LAUNCH_CHECK();
}

void launch_enstrophy_diss_balance_mapped(const double* d_omega, int Nx, int Ny, double inv_dx2,
                                          double inv_deta2, const double* d_a_node, const double* d_a_edge,
                                          const double* d_w_node, double* d_sum_w2, double* d_sum_eta,
                                          hipStream_t s) {
    // The Laplacian lives after the row partials in the library scratch (diag.cu:599-612 uses a
    // function-static buffer for it).
    const size_t N = (size_t)Nx * Ny;
    double* lap = scratch_doubles(N + (size_t)2 * Ny + 64) + (size_t)2 * Ny + 64;
    launch_laplacian_fd_mapped(d_omega, lap, Nx, Ny, inv_dx2, inv_deta2, d_a_node, d_a_edge, s);
    double* partial = scratch_doubles(N + (size_t)2 * Ny + 64);
    row_sum_kernel<2, EnstBalance><<<Ny, RB, 0, s>>>(EnstBalance{d_omega, lap, Nx, Ny, d_w_node}, Nx, Ny, partial);
    NSE2D_LAUNCH_CHECK();
    final_kernel<2, false><<<1, RB, 0, s>>>(partial, Ny, Outs<2>{{d_sum_w2, d_sum_eta}}, 1.0);
    NSE2D_LAUNCH_CHECK();
}

void launch_omega_rhs_balance_mapped(const double* d_omega, const double* d_rhs, int Nx, int Ny,
                                     const double* d_w_node, double* d_sum_omega_rhs, hipStream_t stream) {
    reduce_rows<1>(OmegaRhs{d_omega, d_rhs, Nx, Ny, d_w_node}, Nx, Ny, Outs<1>{{d_sum_omega_rhs}}, stream);
}

void launch_max_uv_from_psi_mapped(const double* d_psi, int Nx, int Ny, double inv_dx, double inv_deta,
                                   const double* d_a_node, double* d_umax, double* d_vmax, double* d_vmaxm,
                                   hipStream_t stream) {
    launch_max_uv(d_psi, Nx, Ny, inv_dx, inv_deta, d_a_node, d_umax, d_vmax, d_vmaxm, stream);
}

int nse2d_diag_row(const double* d_psi, const double* d_omega, const double* d_rhs, int Nx, int Ny, double inv_dx,
                   double inv_deta, double h, double F0, int nforce, const double* d_a_node, const double* d_a_edge,
                   const double* d_y_node, const double* d_w_node, double* d_out, double* d_moments,
                   hipStream_t stream) {
    if (Nx < 2 || Ny < 4 || !d_psi || !d_omega || !d_rhs || !d_out) {
        set_last_error("nse2d_diag_row: needs Nx >= 2, Ny >= 4 and the psi, omega, rhs and output pointers");
        return -1;
    }
    double* partial = scratch_doubles((size_t)(kDiagSums + 3) * Ny);
    const DiagRowArgs A{d_psi, d_omega, d_rhs, Nx, Ny, inv_dx, inv_deta, h, F0, nforce,
                        d_a_node, d_a_edge, d_y_node, d_w_node, d_out, d_moments};
    diag_row_kernel<<<Ny, RB, 0, stream>>>(A, partial);
    NSE2D_LAUNCH_CHECK();
    diag_final_kernel<<<1, RB, 0, stream>>>(partial, Ny, d_out);
    NSE2D_LAUNCH_CHECK();
    return 0;
}

int launch_max_uv_from_psi_mapped_noalloc(const double* d_psi, int Nx, int Ny, double inv_dx, double inv_deta,
                                          const double* d_a_node, double* d_umax, double* d_vmax,
                                          double* d_vmaxm, hipStream_t stream) {
    if (Ny > NSE2D_NOALLOC_MAX_ROWS) {
        set_last_error("launch_max_uv_from_psi_mapped_noalloc: Ny = " + std::to_string(Ny) + " exceeds the " +
                       std::to_string(NSE2D_NOALLOC_MAX_ROWS) + " rows of its fixed scratch");
        return -1;
    }
    if (Nx < 2 || Ny < 3) {
        launch_zero3(d_umax, d_vmax, d_vmaxm, stream);
        return 0;
    }
    double* partial = noalloc_scratch();
    if (!partial) return -1;
    launch_max_uv_into(d_psi, Nx, Ny, inv_dx, inv_deta, d_a_node, d_umax, d_vmax, d_vmaxm, partial, stream);
    return 0;
}

}  // extern "C"
