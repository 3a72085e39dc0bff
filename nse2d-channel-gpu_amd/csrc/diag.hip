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

// ---- diag.cu:468-499: one workgroup per row, five sums, clamped wall rows
__global__ void __launch_bounds__(RB) row_moments_kernel(const double* __restrict__ psi, int Nx, int Ny,
                                                         double inv_dx, double inv_deta,
                                                         const double* __restrict__ a_node, double* su,
                                                         double* sv, double* su2, double* sv2, double* suv) {
    __shared__ double sm[5 * 16];
    const int iy = blockIdx.x;
    double acc[5] = {0, 0, 0, 0, 0};
    const int ym = (iy == 0 ? 0 : iy - 1), yp = (iy == Ny - 1 ? Ny - 1 : iy + 1);
    for (int ix = threadIdx.x; ix < Nx; ix += RB) {
        const int xm = (ix == 0 ? Nx - 1 : ix - 1), xp = (ix == Nx - 1 ? 0 : ix + 1);
        const double u = (psi[idx2(yp, ix, Nx)] - psi[idx2(ym, ix, Nx)]) * 0.5 * inv_deta * a_node[iy];
        const double v = -(psi[idx2(iy, xp, Nx)] - psi[idx2(iy, xm, Nx)]) * 0.5 * inv_dx;
        acc[0] += u; acc[1] += v; acc[2] += u * u; acc[3] += v * v; acc[4] += u * v;
    }
    block_sum<5>(acc, sm);
    if (threadIdx.x == 0) { su[iy] = acc[0]; sv[iy] = acc[1]; su2[iy] = acc[2]; sv2[iy] = acc[3]; suv[iy] = acc[4]; }
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
            const double weta_bot = (-3.0 * w0 + 4.0 * w[idx2(1, ix, Nx)] - 1.0 * w[idx2(2, ix, Nx)]) * (0.5 * inv_deta);
            const double weta_top = (3.0 * wN - 4.0 * w[idx2(Ny - 2, ix, Nx)] + 1.0 * w[idx2(Ny - 3, ix, Nx)]) * (0.5 * inv_deta);
            acc[0] += w0 * (a_node[0] * weta_bot);
            acc[1] += wN * (a_node[Ny - 1] * weta_top);
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

// Several fields (ensemble members) in one launch: workgroup b (after the XCD renumbering) takes
// row b - row_begin of the member whose row range holds it.
__global__ void __launch_bounds__(RB) row_max_uv_group_kernel(const MaxUVJob* __restrict__ jobs, int n) {
    __shared__ double sm[3 * 16];
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    int m = 0;
    while (m + 1 < n && jobs[m + 1].row_begin <= b) ++m;
    const MaxUVJob& J = jobs[m];
    row_max_uv_body(J.psi, J.Nx, J.Ny, J.inv_dx, J.inv_deta, J.a_node, b - J.row_begin, J.partial, sm);
}

// One workgroup per member: the maxima over its rows, in row order.
__global__ void __launch_bounds__(RB) final_max_group_kernel(const MaxUVJob* __restrict__ jobs) {
    __shared__ double sm[3 * 16];
    const MaxUVJob& J = jobs[blockIdx.x];
    double acc[3];
    for (int a = 0; a < 3; ++a) {
        double v = 0.0;
        for (int i = threadIdx.x; i < J.Ny; i += RB) v = fmax(v, J.partial[(size_t)a * J.Ny + i]);
        acc[a] = v;
    }
    block_max<3>(acc, sm);
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; ++a) J.out3[a] = acc[a];
}

__global__ void zero3_kernel(double* a, double* b, double* c) {
    if (threadIdx.x == 0) { *a = 0.0; *b = 0.0; *c = 0.0; }
}

}  // namespace

void launch_max_uv_group(const MaxUVJob* d_jobs, int n, int total_rows, hipStream_t s) {
    row_max_uv_group_kernel<<<total_rows, RB, 0, s>>>(d_jobs, n);
    NSE2D_LAUNCH_CHECK();
    final_max_group_kernel<<<n, RB, 0, s>>>(d_jobs);
    NSE2D_LAUNCH_CHECK();
}

namespace {

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
