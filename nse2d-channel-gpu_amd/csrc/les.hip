// Learned subfilter closure of a coarse channel (LES), trained on filtered DNS of this solver.
// No reference counterpart: the reference runs DNS only. Everything FP64 stays FP64 (the filter,
// the coarse Poisson solve, the Thom closure, the RHS and the stepping); only the CNN runs in
// bf16/fp32 (closure.hip).
//
// Filter F (fine Nx_f x Ny_f -> coarse Nx_c x Ny_c, Nx_f = rx Nx_c, Ny_f - 1 = ry (Ny_c - 1)):
//   1. box filter of width ry fine cells in the mapped coordinate eta, by the trapezoid rule on
//      the fine nodes, centred on the coarse nodes (fine row ry J); the wall rows are kept;
//   2. sharp spectral cut in x, |k_x| < Nx_c / 2 (the coarse Nyquist mode is removed), sampled at
//      every rx-th fine node: one forward row transform at Nx_f, truncation of the half spectrum
//      scaled by Nx_c / Nx_f, one inverse row transform at Nx_c.
// The coarse nodes coincide with every ry-th fine node of the same tanh map, so F acts on nodal
// values only.
//
// Closure target of a fine state (omega_f, psi_f): with omega_c = F(omega_f) whose wall rows are
// replaced by the Thom closure of psi_c = solve_c(omega_c) (the state an LES holds),
//   target = F(R_f(omega_f, psi_f)) - R_c(omega_c, psi_c)    on rows 1..Ny_c-2,
// where R = -J + nu lap + f(y) - r omega is the solver's own RHS (ops.cu:515-544) on each grid.
// Adding the exact target to the coarse RHS reproduces the filtered fine RHS; it is dominated by
// the subfilter vorticity-flux divergence F(J_f) - J_c, and also carries the commutation error of
// the viscous term.
//
// CNN inputs: (omega, psi, u, v) of the coarse state, u and v as ops.cu:47-63 (clamped wall
// differences), each normalised by its mean and standard deviation over the training set; the
// CNN predicts the normalised target.
#include "nse2d_cnn.hpp"
#include "nse2d_common.hpp"
#include "nse2d_poisson.hpp"
#include "../../include/nse2d_hip.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <string>
#include <vector>

namespace nse2d {

void set_last_error(const std::string& msg);

namespace {

constexpr int kFeat = 4;  // omega, psi, u, v

struct Norm {
    float mean[kFeat + 1];     // inputs, then target
    float inv_std[kFeat + 1];
};

// Step 1 of F: coarse row J from fine rows ry J - ry/2 .. ry J + ry/2 (trapezoid weights 1/2 at
// the ends), divided by ry; rows 0 and Ny_c - 1 are the fine wall rows.
__global__ void ybox_kernel(const double* __restrict__ f, double* __restrict__ out, int Nx, int Nyc, int ry) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y;
    if (x >= Nx) return;
    const int jc = ry * J;
    double v;
    if (J == 0 || J == Nyc - 1 || ry == 1) {
        v = f[idx2(jc, x, Nx)];
    } else {
        const int h = ry / 2;
        double sum = 0.5 * f[idx2(jc - h, x, Nx)];
        for (int m = -h + 1; m < h; ++m) sum += f[idx2(jc + m, x, Nx)];
        sum += 0.5 * f[idx2(jc + h, x, Nx)];
        v = sum / (double)ry;
    }
    out[idx2(J, x, Nx)] = v;
}

// Step 2 of F: half spectra [rows][Mf] -> [rows][Mc] (Mc <= Mf), modes 0..Mc-1 scaled by `scale`.
// Slot 0 packs (X_0, X_M) on both sides; the coarse X_Mc is cut, so its slot part is zero.
__global__ void truncate_kernel(const double2* __restrict__ hf, double2* __restrict__ hc, int Mf, int Mc,
                                double scale) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y;
    if (k >= Mc) return;
    double2 z = hf[(size_t)J * Mf + k];
    if (k == 0) z.y = 0.0;
    hc[(size_t)J * Mc + k] = make_double2(z.x * scale, z.y * scale);
}

// CNN inputs of a coarse state, [Ny][Nx][4] fp32, each channel as (value - mean) * inv_std;
// u = a_j (psi_{j+1} - psi_{j-1}) / (2 d_eta), v = -(psi_{i+1} - psi_{i-1}) / (2 dx) with the
// wall rows' differences clamped to the row itself (ops.cu:47-63).
// PACKED: out is the network's packed bf16 input ([Ny][Nx][32]); the inputs go to channels 0..3
// as bf16 of their fp32 values, as the network's own packing rounds them.
template <bool PACKED, typename OutT>
__global__ void features_kernel(const double* __restrict__ w, const double* __restrict__ psi, int Nx, int Ny,
                                double inv_dx, double inv_deta, const double* __restrict__ a_node,
                                OutT* __restrict__ out, Norm n) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y;
    if (ix >= Nx) return;
    const int xm = wrap_x(ix - 1, Nx), xp = wrap_x(ix + 1, Nx);
    const int ym = (iy == 0 ? 0 : iy - 1), yp = (iy == Ny - 1 ? Ny - 1 : iy + 1);
    const size_t c = idx2(iy, ix, Nx);
    const double u = a_node[iy] * ((psi[idx2(yp, ix, Nx)] - psi[idx2(ym, ix, Nx)]) * 0.5 * inv_deta);
    const double v = -((psi[idx2(iy, xp, Nx)] - psi[idx2(iy, xm, Nx)]) * 0.5 * inv_dx);
    const double val[kFeat] = {w[c], psi[c], u, v};
    float4 o;
    o.x = (float)((val[0] - n.mean[0]) * n.inv_std[0]);
    o.y = (float)((val[1] - n.mean[1]) * n.inv_std[1]);
    o.z = (float)((val[2] - n.mean[2]) * n.inv_std[2]);
    o.w = (float)((val[3] - n.mean[3]) * n.inv_std[3]);
    if constexpr (PACKED) {
        typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
        const bf16x4 p = {(__bf16)o.x, (__bf16)o.y, (__bf16)o.z, (__bf16)o.w};
        *reinterpret_cast<bf16x4*>(out + c * 32) = p;
    } else {
        reinterpret_cast<float4*>(out)[c] = o;
    }
}

// target = rf - rc on the interior rows, 0 on the wall rows; fp64 copy and/or fp32 copy.
__global__ void target_kernel(const double* __restrict__ rf, const double* __restrict__ rc, int Nx, int Ny,
                              double* __restrict__ t64, float* __restrict__ t32) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y;
    if (ix >= Nx) return;
    const size_t c = idx2(iy, ix, Nx);
    const double t = (iy > 0 && iy < Ny - 1) ? rf[c] - rc[c] : 0.0;
    if (t64) t64[c] = t;
    if (t32) t32[c] = (float)t;
}

// Normalised mini-batch: pairs idx[0..B) of the raw set into Xb [B][npix][4], Yb [B][npix].
__global__ void gather_kernel(const float* __restrict__ X, const float* __restrict__ Y, const int* __restrict__ idx,
                              size_t npix, Norm n, float* __restrict__ Xb, float* __restrict__ Yb) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (p >= npix) return;
    const size_t src = (size_t)idx[b] * npix + p, dst = (size_t)b * npix + p;
    const float4 x = reinterpret_cast<const float4*>(X)[src];
    float4 o;
    o.x = (x.x - n.mean[0]) * n.inv_std[0];
    o.y = (x.y - n.mean[1]) * n.inv_std[1];
    o.z = (x.z - n.mean[2]) * n.inv_std[2];
    o.w = (x.w - n.mean[3]) * n.inv_std[3];
    reinterpret_cast<float4*>(Xb)[dst] = o;
    Yb[dst] = (Y[src] - n.mean[4]) * n.inv_std[4];
}

// Per-block partial sums over pairs [0, n): for the 4 inputs (all rows) and the target (rows
// 1..Ny-2): sum and sum of squares, 10 doubles per block.
__global__ void __launch_bounds__(256) moments_kernel(const float* __restrict__ X, const float* __restrict__ Y,
                                                      size_t total, int Nx, int Ny, double* __restrict__ partial) {
    __shared__ double sm[10 * 16];
    double acc[10];
#pragma unroll
    for (int a = 0; a < 10; ++a) acc[a] = 0.0;
    const size_t npix = (size_t)Nx * Ny;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const float4 x = reinterpret_cast<const float4*>(X)[e];
        const double v[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int c = 0; c < 4; ++c) { acc[2 * c] += v[c]; acc[2 * c + 1] += v[c] * v[c]; }
        const int iy = (int)((e % npix) / Nx);
        if (iy > 0 && iy < Ny - 1) { const double t = Y[e]; acc[8] += t; acc[9] += t * t; }
    }
    block_sum<10>(acc, sm);
    if (threadIdx.x == 0)
        for (int a = 0; a < 10; ++a) partial[(size_t)blockIdx.x * 10 + a] = acc[a];
}

// Sum over pairs [first, first + count) of ((target - mean) * inv_std)^2 and of the number of
// interior points, the loss of a zero closure in normalised units (per-block partials).
__global__ void __launch_bounds__(256) zero_loss_kernel(const float* __restrict__ Y, size_t begin, size_t end, int Nx,
                                                        int Ny, float mean, float inv_std, double* __restrict__ partial) {
    __shared__ double sm[2 * 16];
    double acc[2] = {0.0, 0.0};
    const size_t npix = (size_t)Nx * Ny;
    for (size_t e = begin + (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < end; e += (size_t)gridDim.x * blockDim.x) {
        const int iy = (int)((e % npix) / Nx);
        if (iy > 0 && iy < Ny - 1) {
            const double z = ((double)Y[e] - mean) * inv_std;
            acc[0] += z * z;
            acc[1] += 1.0;
        }
    }
    block_sum<2>(acc, sm);
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = acc[0]; partial[2 * blockIdx.x + 1] = acc[1]; }
}

struct Grid {
    nse2d_channel_desc d;
    double inv_dx = 0, inv_deta = 0, inv_dx2 = 0, inv_deta2 = 0;
    size_t N = 0;
    void set(const nse2d_channel_desc& desc) {
        d = desc;
        inv_dx = (double)d.Nx / d.Lx;
        inv_deta = (double)(d.Ny - 1) / 2.0;
        inv_dx2 = inv_dx * inv_dx;
        inv_deta2 = inv_deta * inv_deta;
        N = (size_t)d.Nx * d.Ny;
    }
};

}  // namespace

struct LesImpl {
    hipStream_t s = nullptr;
    Grid f, c;
    int rx = 1, ry = 1;
    PoissonImpl* Pf = nullptr;  // row transform at Nx_f over Ny_c rows (filter)
    PoissonImpl* Pc = nullptr;  // coarse mapped solver; its row transform also finishes the filter
    double *tmp = nullptr;                       // [Ny_c][Nx_f]
    double *fJ = nullptr, *flap = nullptr, *frhs = nullptr;
    double *cw = nullptr, *cpsi = nullptr, *cJ = nullptr, *clap = nullptr, *crhs = nullptr, *cfr = nullptr;
    double *partial = nullptr;                   // reduction partials (host-summed)
    int max_pairs = 0, npairs = 0, max_batch = 0;
    float *X = nullptr, *Y = nullptr;            // raw training set [max_pairs][Ny_c][Nx_c]([4])
    float *Xb = nullptr, *Yb = nullptr;          // normalised batch
    int* d_idx = nullptr;
    Norm norm{};
    double mean[kFeat + 1] = {0}, stdv[kFeat + 1] = {1, 1, 1, 1, 1};
    nse2d_cnn_t cnn = nullptr;
    int attach_mode = 1;   // 1: closure of every stage state; 2: of the step's first stage, held
    const double* attached_w = nullptr;  // omega^n of the attached channel (its first stage state)
    int constraints = 0;  // NSE2D_CLOSURE_* flags applied to every closure evaluation
};

namespace {

constexpr int kRedBlocks = 512;

void set_norm(LesImpl* L) {
    for (int k = 0; k <= kFeat; ++k) {
        L->norm.mean[k] = (float)L->mean[k];
        L->norm.inv_std[k] = (float)(1.0 / (L->stdv[k] > 0 ? L->stdv[k] : 1.0));
    }
}

void filter(LesImpl* L, const double* fine, double* coarse) {
    const int Nxf = L->f.d.Nx, Nxc = L->c.d.Nx, Nyc = L->c.d.Ny;
    ybox_kernel<<<dim3(ceil_div(Nxf, 256), Nyc), 256, 0, L->s>>>(fine, L->tmp, Nxf, Nyc, L->ry);
    NSE2D_LAUNCH_CHECK();
    poisson_forward(L->Pf, L->tmp, L->s);
    const int Mf = Nxf / 2, Mc = Nxc / 2;
    truncate_kernel<<<dim3(ceil_div(Mc, 256), Nyc), 256, 0, L->s>>>(
        static_cast<const double2*>(poisson_hat(L->Pf)), static_cast<double2*>(poisson_hat(L->Pc)), Mf, Mc,
        (double)Nxc / (double)Nxf);
    NSE2D_LAUNCH_CHECK();
    poisson_inverse(L->Pc, coarse, L->s);
}

// R = -J(psi, w) + nu lap(w) + f(y) - r w on grid g (ops.cu:283-334, 471-544)
void rhs(LesImpl* L, const Grid& g, const double* w, const double* psi, double* J, double* lap, double* out) {
    const nse2d_channel_desc& d = g.d;
    launch_arakawa_J_mapped(psi, w, J, d.Nx, d.Ny, g.inv_dx, g.inv_deta, d.d_a_node, L->s);
    launch_laplacian_fd_mapped(w, lap, d.Nx, d.Ny, g.inv_dx2, g.inv_deta2, d.d_a_node, d.d_a_edge, L->s);
    launch_rhs_mapped(J, lap, w, out, d.Nx, d.Ny, d.nu, d.F0, d.nforce, d.h, d.lin_drag, d.d_y_node, L->s);
}

void features(LesImpl* L, const double* w, const double* psi, float* out, const Norm& n) {
    const Grid& g = L->c;
    features_kernel<false><<<dim3(ceil_div(g.d.Nx, 256), g.d.Ny), 256, 0, L->s>>>(
        w, psi, g.d.Nx, g.d.Ny, g.inv_dx, g.inv_deta, g.d.d_a_node, out, n);
    NSE2D_LAUNCH_CHECK();
}

Norm raw_norm() {
    Norm n;
    for (int k = 0; k <= kFeat; ++k) { n.mean[k] = 0.0f; n.inv_std[k] = 1.0f; }
    return n;
}

// The closure term of the coarse state (w, psi) into dst, on L's stream: the inputs go straight
// into the network's packed input, the scaled term straight into dst.
// Mean-free closure: the x-average of C over each row is subtracted, so the term leaves the
// mean vorticity profile (and with it the mean velocity and the wall shear) to the resolved
// dynamics. One workgroup of 256 per row; the row sum is taken in double.
__global__ void __launch_bounds__(256) mean_free_rows_kernel(float* __restrict__ C, int Nx) {
    __shared__ double part[256];
    float* row = C + (size_t)blockIdx.x * Nx;
    double acc = 0.0;
    for (int i = threadIdx.x; i < Nx; i += 256) acc += (double)row[i];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    const float mean = (float)(part[0] / Nx);
    for (int i = threadIdx.x; i < Nx; i += 256) row[i] -= mean;
}

// Backscatter removal: where omega * C > 0 the closure term would add enstrophy (d(omega^2/2)/dt
// gains omega * C); those points get C = 0, so the closure only removes enstrophy.
__global__ void dissipative_clip_kernel(float* __restrict__ C, const double* __restrict__ w, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (double)C[i] * w[i] > 0.0) C[i] = 0.0f;
}

void closure_eval(LesImpl* L, const double* w, const double* psi, float* dst) {
    const Grid& g = L->c;
    features_kernel<true><<<dim3(ceil_div(g.d.Nx, 256), g.d.Ny), 256, 0, L->s>>>(
        w, psi, g.d.Nx, g.d.Ny, g.inv_dx, g.inv_deta, g.d.d_a_node, cnn_packed_input(L->cnn), L->norm);
    NSE2D_LAUNCH_CHECK();
    cnn_infer_packed_closure(L->cnn, (float)L->stdv[kFeat], (float)L->mean[kFeat], dst);
    if (L->constraints & NSE2D_CLOSURE_MEAN_FREE) {
        mean_free_rows_kernel<<<g.d.Ny, 256, 0, L->s>>>(dst, g.d.Nx);
        NSE2D_LAUNCH_CHECK();
    }
    if (L->constraints & NSE2D_CLOSURE_DISSIPATIVE) {
        const size_t n = (size_t)g.d.Nx * g.d.Ny;
        dissipative_clip_kernel<<<(unsigned)((n + 255) / 256), 256, 0, L->s>>>(dst, w, n);
        NSE2D_LAUNCH_CHECK();
    }
}

// The closure hook of a coarse channel (nse2d_closure_fn). The channel runs on L's stream
// (checked by nse2d_les_attach). Mode 2: the channel calls the hook before each of the three
// stages of a step with that stage's state; only the first stage's state (the channel's own
// omega^n field) is evaluated, and the channel's buffer keeps that term for the other two.
void closure_hook(void* ctx, const double* w, const double* psi, float* dst, hipStream_t s) {
    LesImpl* L = static_cast<LesImpl*>(ctx);
    (void)s;
    if (L->attach_mode == 2 && w != L->attached_w) return;
    closure_eval(L, w, psi, dst);
}

std::vector<double> read_partials(LesImpl* L, size_t n) {
    std::vector<double> h(n);
    NSE2D_HIP_CHECK(hipMemcpyAsync(h.data(), L->partial, n * sizeof(double), hipMemcpyDeviceToHost, L->s));
    NSE2D_HIP_CHECK(hipStreamSynchronize(L->s));
    return h;
}

void les_destroy(LesImpl* L) {
    if (!L) return;
    (void)hipStreamSynchronize(L->s);
    for (void* p : {(void*)L->tmp, (void*)L->fJ, (void*)L->flap, (void*)L->frhs, (void*)L->cw, (void*)L->cpsi,
                    (void*)L->cJ, (void*)L->clap, (void*)L->crhs, (void*)L->cfr, (void*)L->partial, (void*)L->X,
                    (void*)L->Y, (void*)L->Xb, (void*)L->Yb, (void*)L->d_idx})
        (void)hipFree(p);
    if (L->Pf) poisson_destroy(L->Pf);
    if (L->Pc) poisson_destroy(L->Pc);
    if (L->cnn) nse2d_cnn_destroy(L->cnn);
    delete L;
}

LesImpl* les_create(const nse2d_channel_desc* fine, const nse2d_channel_desc* coarse, int max_pairs, int max_batch,
                    unsigned long long seed, hipStream_t s) {
    if (!fine || !coarse) { set_last_error("nse2d_les_create: null descriptor"); return nullptr; }
    const nse2d_channel_desc &F = *fine, &C = *coarse;
    for (const nse2d_channel_desc* d : {fine, coarse})
        if (!d->d_a_node || !d->d_a_edge || !d->d_y_node || !d->d_dy_edge) {
            set_last_error("nse2d_les_create: metric arrays must be set on both grids");
            return nullptr;
        }
    if (C.Nx < 64 || C.Nx % 64 != 0 || F.Nx % C.Nx != 0 || C.Ny < 3 || (F.Ny - 1) % (C.Ny - 1) != 0) {
        set_last_error("nse2d_les_create: needs Nx_c a multiple of 64 dividing Nx_f and Ny_c - 1 dividing Ny_f - 1");
        return nullptr;
    }
    const int ry = (F.Ny - 1) / (C.Ny - 1);
    if (ry != 1 && ry % 2 != 0) { set_last_error("nse2d_les_create: the y ratio must be 1 or even"); return nullptr; }
    if (std::fabs(F.Lx - C.Lx) > 1e-12 * F.Lx || std::fabs(F.h - C.h) > 1e-12 * F.h) {
        set_last_error("nse2d_les_create: the grids must share Lx and h");
        return nullptr;
    }
    if (max_pairs < 1 || max_batch < 1) { set_last_error("nse2d_les_create: max_pairs and max_batch >= 1"); return nullptr; }
    auto* L = new LesImpl;
    L->s = s;
    L->f.set(F);
    L->c.set(C);
    L->rx = F.Nx / C.Nx;
    L->ry = ry;
    L->max_pairs = max_pairs;
    L->max_batch = max_batch;
    L->Pf = poisson_create(F.Nx, C.Ny, F.Lx, s);
    L->Pc = poisson_create(C.Nx, C.Ny, C.Lx, s);
    if (!L->Pf || !L->Pc) { les_destroy(L); return nullptr; }
    poisson_enable_mapped(L->Pc, C.d_a_node, C.d_a_edge, L->c.inv_deta, s);
    NSE2D_HIP_CHECK(hipMalloc(&L->tmp, sizeof(double) * (size_t)C.Ny * F.Nx));
    for (double** p : {&L->fJ, &L->flap, &L->frhs}) NSE2D_HIP_CHECK(hipMalloc(p, sizeof(double) * L->f.N));
    for (double** p : {&L->cw, &L->cpsi, &L->cJ, &L->clap, &L->crhs, &L->cfr})
        NSE2D_HIP_CHECK(hipMalloc(p, sizeof(double) * L->c.N));
    NSE2D_HIP_CHECK(hipMalloc(&L->partial, sizeof(double) * 10 * kRedBlocks));
    NSE2D_HIP_CHECK(hipMalloc(&L->X, sizeof(float) * kFeat * L->c.N * max_pairs));
    NSE2D_HIP_CHECK(hipMalloc(&L->Y, sizeof(float) * L->c.N * max_pairs));
    NSE2D_HIP_CHECK(hipMalloc(&L->Xb, sizeof(float) * kFeat * L->c.N * max_batch));
    NSE2D_HIP_CHECK(hipMalloc(&L->Yb, sizeof(float) * L->c.N * max_batch));
    NSE2D_HIP_CHECK(hipMalloc(&L->d_idx, sizeof(int) * max_batch));
    set_norm(L);
    L->cnn = nse2d_cnn_create(C.Ny, C.Nx, max_batch, seed, s);
    if (!L->cnn) { les_destroy(L); return nullptr; }
    NSE2D_HIP_CHECK(hipStreamSynchronize(s));
    return L;
}

int les_pair(LesImpl* L, const double* wf, const double* psif, double* wc_out, double* psic_out, double* t_out,
             int append) {
    if (append && L->npairs >= L->max_pairs) { set_last_error("nse2d_les_pair: the training set is full"); return -1; }
    const nse2d_channel_desc& C = L->c.d;
    rhs(L, L->f, wf, psif, L->fJ, L->flap, L->frhs);
    filter(L, L->frhs, L->cfr);
    filter(L, wf, L->cw);
    poisson_solve(L->Pc, L->cw, L->cpsi, C.psi_bot, C.psi_top, L->s);
    launch_thom_wall_vorticity_mapped(L->cw, L->cpsi, C.Nx, C.Ny, C.d_dy_edge, C.psi_bot, C.psi_top, L->s);
    rhs(L, L->c, L->cw, L->cpsi, L->cJ, L->clap, L->crhs);
    float* t32 = append ? L->Y + (size_t)L->npairs * L->c.N : nullptr;
    target_kernel<<<dim3(ceil_div(C.Nx, 256), C.Ny), 256, 0, L->s>>>(L->cfr, L->crhs, C.Nx, C.Ny, t_out, t32);
    NSE2D_LAUNCH_CHECK();
    if (append) {
        features(L, L->cw, L->cpsi, L->X + (size_t)L->npairs * kFeat * L->c.N, raw_norm());
        L->npairs++;
    }
    if (wc_out) NSE2D_HIP_CHECK(hipMemcpyAsync(wc_out, L->cw, sizeof(double) * L->c.N, hipMemcpyDeviceToDevice, L->s));
    if (psic_out)
        NSE2D_HIP_CHECK(hipMemcpyAsync(psic_out, L->cpsi, sizeof(double) * L->c.N, hipMemcpyDeviceToDevice, L->s));
    return 0;
}

int les_fit_scales(LesImpl* L, int n) {
    if (n < 1 || n > L->npairs) { set_last_error("nse2d_les_fit_scales: n outside [1, pairs]"); return -1; }
    const size_t total = (size_t)n * L->c.N;
    moments_kernel<<<kRedBlocks, 256, 0, L->s>>>(L->X, L->Y, total, L->c.d.Nx, L->c.d.Ny, L->partial);
    NSE2D_LAUNCH_CHECK();
    const std::vector<double> p = read_partials(L, 10 * (size_t)kRedBlocks);
    double acc[10] = {0};
    for (int b = 0; b < kRedBlocks; ++b)
        for (int a = 0; a < 10; ++a) acc[a] += p[(size_t)b * 10 + a];
    const double cnt_all = (double)total, cnt_int = (double)n * (double)L->c.d.Nx * (L->c.d.Ny - 2);
    for (int k = 0; k <= kFeat; ++k) {
        const double cnt = k < kFeat ? cnt_all : cnt_int;
        const double m = acc[2 * k] / cnt;
        const double var = std::max(acc[2 * k + 1] / cnt - m * m, 0.0);
        L->mean[k] = m;
        L->stdv[k] = var > 0 ? std::sqrt(var) : 1.0;
    }
    set_norm(L);
    return 0;
}

void gather(LesImpl* L, const int* idx_host, int B) {
    NSE2D_HIP_CHECK(hipMemcpyAsync(L->d_idx, idx_host, sizeof(int) * B, hipMemcpyHostToDevice, L->s));
    gather_kernel<<<dim3(ceil_div(L->c.N, 256), B), 256, 0, L->s>>>(L->X, L->Y, L->d_idx, L->c.N, L->norm, L->Xb,
                                                                     L->Yb);
    NSE2D_LAUNCH_CHECK();
}

double les_train(LesImpl* L, int n_train, int steps, int batch, float lr0, unsigned long long seed, float* losses) {
    if (n_train < 1 || n_train > L->npairs || batch < 1 || batch > L->max_batch || batch > n_train || steps < 1) {
        set_last_error("nse2d_les_train: needs 1 <= batch <= min(max_batch, n_train), n_train <= pairs, steps >= 1");
        return -1.0;
    }
    std::mt19937_64 rng(seed);
    std::vector<int> order(n_train);
    std::iota(order.begin(), order.end(), 0);
    size_t pos = (size_t)n_train;  // forces a shuffle first
    std::vector<int> idx(batch);
    double tail = 0.0;
    int tail_n = 0;
    for (int st = 0; st < steps; ++st) {
        for (int b = 0; b < batch; ++b) {
            if (pos >= order.size()) { std::shuffle(order.begin(), order.end(), rng); pos = 0; }
            idx[b] = order[pos++];
        }
        gather(L, idx.data(), batch);
        // cosine decay from lr0 to lr0 / 20
        const double lr = lr0 * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(kPi * (double)st / (double)steps)));
        const double loss = nse2d_cnn_train_step(L->cnn, L->Xb, L->Yb, batch, (float)lr);
        if (losses) losses[st] = (float)loss;
        if (st >= steps - 10) { tail += loss; tail_n++; }
    }
    return tail / std::max(tail_n, 1);
}

int les_eval(LesImpl* L, int first, int count, double* loss_cnn, double* loss_zero) {
    if (first < 0 || count < 1 || first + count > L->npairs) {
        set_last_error("nse2d_les_eval: pair range outside the training set");
        return -1;
    }
    double sum = 0.0;
    int done = 0;
    std::vector<int> idx(L->max_batch);
    while (done < count) {
        const int B = std::min(L->max_batch, count - done);
        for (int b = 0; b < B; ++b) idx[b] = first + done + b;
        gather(L, idx.data(), B);
        sum += nse2d_cnn_eval_loss(L->cnn, L->Xb, L->Yb, B) * B;
        done += B;
    }
    if (loss_cnn) *loss_cnn = sum / count;
    if (loss_zero) {
        const size_t N = L->c.N;
        zero_loss_kernel<<<kRedBlocks, 256, 0, L->s>>>(L->Y, (size_t)first * N, (size_t)(first + count) * N, L->c.d.Nx,
                                                       L->c.d.Ny, L->norm.mean[kFeat], L->norm.inv_std[kFeat],
                                                       L->partial);
        NSE2D_LAUNCH_CHECK();
        const std::vector<double> p = read_partials(L, 2 * (size_t)kRedBlocks);
        double a = 0.0, c = 0.0;
        for (int b = 0; b < kRedBlocks; ++b) { a += p[2 * b]; c += p[2 * b + 1]; }
        *loss_zero = c > 0 ? a / c : 0.0;
    }
    return 0;
}

}  // namespace
}  // namespace nse2d

struct nse2d_les {
    nse2d::LesImpl* impl;
};

using namespace nse2d;

extern "C" {

nse2d_les_t nse2d_les_create(const nse2d_channel_desc* fine, const nse2d_channel_desc* coarse, int max_pairs,
                             int max_batch, unsigned long long seed, hipStream_t s) {
    LesImpl* L = les_create(fine, coarse, max_pairs, max_batch, seed, s);
    return L ? new nse2d_les{L} : nullptr;
}
void nse2d_les_destroy(nse2d_les_t L) {
    if (!L) return;
    les_destroy(L->impl);
    delete L;
}
nse2d_cnn_t nse2d_les_cnn(nse2d_les_t L) { return L->impl->cnn; }
int nse2d_les_filter(nse2d_les_t L, const double* d_fine, double* d_coarse) {
    filter(L->impl, d_fine, d_coarse);
    return 0;
}
int nse2d_les_pair(nse2d_les_t L, const double* d_omega_f, const double* d_psi_f, double* d_omega_c, double* d_psi_c,
                   double* d_target, int append) {
    return les_pair(L->impl, d_omega_f, d_psi_f, d_omega_c, d_psi_c, d_target, append);
}
int nse2d_les_num_pairs(nse2d_les_t L) { return L->impl->npairs; }
void nse2d_les_clear_pairs(nse2d_les_t L) { L->impl->npairs = 0; }
int nse2d_les_fit_scales(nse2d_les_t L, int n) { return les_fit_scales(L->impl, n); }
void nse2d_les_get_scales(nse2d_les_t L, double* mean5, double* std5) {
    for (int k = 0; k <= kFeat; ++k) {
        if (mean5) mean5[k] = L->impl->mean[k];
        if (std5) std5[k] = L->impl->stdv[k];
    }
}
int nse2d_les_set_scales(nse2d_les_t L, const double* mean5, const double* std5) {
    for (int k = 0; k <= kFeat; ++k)
        if (!(std5[k] > 0)) { set_last_error("nse2d_les_set_scales: standard deviations must be positive"); return -1; }
    for (int k = 0; k <= kFeat; ++k) { L->impl->mean[k] = mean5[k]; L->impl->stdv[k] = std5[k]; }
    set_norm(L->impl);
    return 0;
}
double nse2d_les_train(nse2d_les_t L, int n_train, int steps, int batch, float lr0, unsigned long long seed,
                       float* losses) {
    return les_train(L->impl, n_train, steps, batch, lr0, seed, losses);
}
int nse2d_les_eval(nse2d_les_t L, int first, int count, double* loss_cnn, double* loss_zero) {
    return les_eval(L->impl, first, count, loss_cnn, loss_zero);
}
int nse2d_les_closure(nse2d_les_t L, const double* d_omega_c, const double* d_psi_c, float* d_out) {
    closure_eval(L->impl, d_omega_c, d_psi_c, d_out);
    return 0;
}
int nse2d_les_set_constraints(nse2d_les_t L, int flags) {
    if (flags & ~(NSE2D_CLOSURE_DISSIPATIVE | NSE2D_CLOSURE_MEAN_FREE)) {
        set_last_error("nse2d_les_set_constraints: unknown flag");
        return -1;
    }
    L->impl->constraints = flags;
    return 0;
}
int nse2d_les_attach(nse2d_les_t L, nse2d_channel_t coarse_ch, int on) {
    if (on < 0 || on > 2) { set_last_error("nse2d_les_attach: mode must be 0, 1 or 2"); return -1; }
    if (!coarse_ch) { set_last_error("nse2d_les_attach: null channel"); return -1; }
    if (on && nse2d_channel_stream(coarse_ch) != L->impl->s) {
        // the stage kernel reads the closure buffer that the hook's inference writes: one stream
        // orders them
        set_last_error("nse2d_les_attach: the channel must run on the stream the LES was created with");
        return -1;
    }
    L->impl->attach_mode = on;
    L->impl->attached_w = nse2d_channel_field(coarse_ch, NSE2D_FIELD_OMEGA);
    return nse2d_channel_set_closure(coarse_ch, on ? closure_hook : nullptr, L->impl);
}

}  // extern "C"
