// Poisson solver for the streamfunction: -lap_h psi = omega with spectral x (FFT) and the
// mapped second-order y operator, psi = psi_bot / psi_top on the walls.
// Replaces poisson.cu + poisson.cuh (cuFFT Z2Z + one-thread-per-kx Thomas).
//
// Spectral layout: for each row j the Nx real values are transformed as one complex FFT of
// M = Nx/2 points (even samples in the real part, odd samples in the imaginary part), then
// split into the half spectrum X[0..M]. X[0] and X[M] are real for real data, so they share
// one complex slot: hat[j][0] = (X[0], X[M]). A row of the spectrum is therefore exactly M
// complex values, [Ny][M] row-major, kx fastest: the Thomas sweep runs one lane per kx and
// every lane load of a row is coalesced. Lane 0 solves two real systems (kx = 0 with the wall
// values, kx = Nyquist without).
#include "nse2d_common.hpp"
#include "nse2d_poisson.hpp"
#include "../../include/nse2d_hip.h"

#include <cmath>
#include <string>
#include <vector>

namespace nse2d {

void set_last_error(const std::string& msg);

namespace {

struct cplx { double x, y; };
__device__ inline cplx cadd(cplx a, cplx b) { return {a.x + b.x, a.y + b.y}; }
__device__ inline cplx csub(cplx a, cplx b) { return {a.x - b.x, a.y - b.y}; }
__device__ inline cplx cmul(cplx a, cplx b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline cplx cconj(cplx a) { return {a.x, -a.y}; }
__device__ inline cplx cscale(cplx a, double s) { return {a.x * s, a.y * s}; }

// In-LDS Stockham FFT of M points (M a power of two), radix 4 with one leading radix-2 pass
// when log2(M) is odd. Each pass: every thread loads its R inputs, barrier, then stores its R
// outputs (every element is read and written exactly once per pass, so one buffer suffices).
// tw[q] = exp(-2 pi i q / (2M)); the M-point twiddle exp(-2 pi i p / M) is tw[2p].
// IT4 = work items per thread in a radix-4 pass (blockDim.x * IT4 >= M/4).
template <bool INV, int IT4>
__device__ void lds_fft(cplx* __restrict__ buf, int M, int log2M, const cplx* __restrict__ tw) {
    const int T = blockDim.x, tid = threadIdx.x;
    int Ns = 1;
    if (log2M & 1) {  // radix-2 pass with Ns = 1 (no twiddles)
        const int Q = M >> 1;
        cplx v0[2 * IT4], v1[2 * IT4];
#pragma unroll
        for (int n = 0; n < 2 * IT4; ++n) {
            const int j = tid + n * T;
            if (j < Q) { v0[n] = buf[j]; v1[n] = buf[j + Q]; }
        }
        __syncthreads();
#pragma unroll
        for (int n = 0; n < 2 * IT4; ++n) {
            const int j = tid + n * T;
            if (j < Q) {
                buf[2 * j] = cadd(v0[n], v1[n]);
                buf[2 * j + 1] = csub(v0[n], v1[n]);
            }
        }
        __syncthreads();
        Ns = 2;
    }
    const int Q = M >> 2;
    for (; Ns < M; Ns <<= 2) {
        cplx v[IT4][4];
#pragma unroll
        for (int n = 0; n < IT4; ++n) {
            const int j = tid + n * T;
            if (j < Q) {
                v[n][0] = buf[j]; v[n][1] = buf[j + Q]; v[n][2] = buf[j + 2 * Q]; v[n][3] = buf[j + 3 * Q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int n = 0; n < IT4; ++n) {
            const int j = tid + n * T;
            if (j >= Q) continue;
            const int k = j & (Ns - 1);
            // twiddle exp(-+2 pi i r k / (4 Ns)) = tw[r * step], step = 2 k M / (4 Ns)
            const int step = (M / (4 * Ns)) * 2 * k;
            cplx a0 = v[n][0], a1 = v[n][1], a2 = v[n][2], a3 = v[n][3];
            if (k) {
                cplx w1 = tw[step], w2 = tw[2 * step], w3 = tw[3 * step];
                if (INV) { w1 = cconj(w1); w2 = cconj(w2); w3 = cconj(w3); }
                a1 = cmul(a1, w1); a2 = cmul(a2, w2); a3 = cmul(a3, w3);
            }
            const cplx t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3);
            const cplx d = csub(a1, a3);
            const cplx t3 = INV ? cplx{-d.y, d.x} : cplx{d.y, -d.x};  // (+-i) * (a1 - a3)
            const int o = (j - k) * 4 + k;
            buf[o] = cadd(t0, t2);
            buf[o + Ns] = cadd(t1, t3);
            buf[o + 2 * Ns] = csub(t0, t2);
            buf[o + 3 * Ns] = csub(t1, t3);
        }
        __syncthreads();
    }
}

// Forward real FFT of each row of omega ([Ny][Nx]) into the packed half spectrum ([Ny][M]).
template <int IT4>
__global__ void __launch_bounds__(1024) fft_rows_forward_kernel(const double* __restrict__ omega,
                                                               cplx* __restrict__ hat, int Nx,
                                                               int log2M,
                                                               const cplx* __restrict__ tw) {
    extern __shared__ cplx lds[];
    const int M = Nx >> 1, j = blockIdx.x, T = blockDim.x;
    const cplx* src = reinterpret_cast<const cplx*>(omega + (size_t)j * Nx);
    for (int n = threadIdx.x; n < M; n += T) lds[n] = src[n];
    __syncthreads();
    lds_fft<false, IT4>(lds, M, log2M, tw);
    cplx* dst = hat + (size_t)j * M;
    for (int k = threadIdx.x; k <= (M >> 1); k += T) {
        const cplx Zk = lds[k], Zm = lds[(M - k) & (M - 1)];
        if (k == 0) {
            dst[0] = {Zk.x + Zk.y, Zk.x - Zk.y};  // (X[0], X[M])
            continue;
        }
        const cplx E = cscale(cadd(Zk, cconj(Zm)), 0.5);
        const cplx D = csub(Zk, cconj(Zm));
        const cplx O = {0.5 * D.y, -0.5 * D.x};  // -i/2 * D
        const cplx WO = cmul(tw[k], O);
        dst[k] = cadd(E, WO);
        if (k != M - k) dst[M - k] = cconj(csub(E, WO));
    }
}

// Inverse: packed half spectrum row -> real row, scaled by `scale` (1/Nx).
template <int IT4>
__global__ void __launch_bounds__(1024) fft_rows_inverse_kernel(const cplx* __restrict__ hat,
                                                               double* __restrict__ psi, int Nx,
                                                               int log2M,
                                                               const cplx* __restrict__ tw,
                                                               double scale) {
    extern __shared__ cplx lds[];
    const int M = Nx >> 1, j = blockIdx.x, T = blockDim.x;
    const cplx* src = hat + (size_t)j * M;
    for (int k = threadIdx.x; k <= (M >> 1); k += T) {
        if (k == 0) {
            const cplx p = src[0];  // (X[0], X[M])
            lds[0] = {p.x + p.y, p.x - p.y};
            continue;
        }
        const cplx Xk = src[k], Xm = src[M - k];
        const cplx A = cadd(Xk, cconj(Xm)), B = csub(Xk, cconj(Xm));
        const cplx iWB = cmul(cplx{0.0, 1.0}, cmul(cconj(tw[k]), B));
        lds[k] = cadd(A, iWB);
        if (k != M - k) lds[M - k] = cconj(csub(A, iWB));
    }
    __syncthreads();
    lds_fft<true, IT4>(lds, M, log2M, tw);
    cplx* dst = reinterpret_cast<cplx*>(psi + (size_t)j * Nx);
    for (int n = threadIdx.x; n < M; n += T) dst[n] = cscale(lds[n], scale);
}

// poisson.cu:38-68 restated per column k = 0..M (M+1 columns, column M = Nyquist), written
// transposed: cp[r][k] (row r = j-1 for j = 1..Ny-2).
__global__ void precompute_cp_kernel(double* __restrict__ cp, const double* __restrict__ a_sub,
                                     const double* __restrict__ c_sup,
                                     const double* __restrict__ kx2, int M, int Ni) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > M) return;
    const int ld = M + 1;
    double b = -(a_sub[0] + c_sup[0]) - kx2[k];
    double cp_prev = c_sup[0] / b;
    cp[k] = cp_prev;
    for (int r = 1; r < Ni; ++r) {
        b = -(a_sub[r] + c_sup[r]) - kx2[k];
        const double den = b - a_sub[r] * cp_prev;
        const double cpi = c_sup[r] / den;
        cp[(size_t)r * ld + k] = cpi;
        cp_prev = cpi;
    }
}

// Row coefficients of the mapped operator, poisson.cu:83-88,103-107.
__global__ void row_coeff_kernel(double* __restrict__ a_sub, double* __restrict__ c_sup,
                                 const double* __restrict__ a_node,
                                 const double* __restrict__ a_edge, int Ny, double inv_deta2) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= Ny - 2) return;
    const int j = r + 1;
    a_sub[r] = a_node[j] * a_edge[j - 1] * inv_deta2;
    c_sup[r] = a_node[j] * a_edge[j] * inv_deta2;
}

// poisson.cu:70-133, one lane per packed column k, rows j = 1..Ny-2, in place on hat.
// The right-hand side is -omega_hat; the wall values enter the kx = 0 system only, scaled by
// Nx (unnormalised transform). Forward sweep divides by den = c_sup / cp like the reference.
__global__ void __launch_bounds__(64) thomas_kernel(cplx* __restrict__ hat,
                                                    const double* __restrict__ cp,
                                                    const double* __restrict__ a_sub,
                                                    const double* __restrict__ c_sup,
                                                    const double* __restrict__ kx2, int M, int Ny,
                                                    double bc_bot, double bc_top) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    const int Ni = Ny - 2, ld = M + 1;
    // real part uses column kr, imaginary part column ki (different only for lane 0)
    const int kr = k, ki = (k == 0 ? M : k);
    cplx* col = hat + k;
    cplx dp_prev;
    {
        const double br = -(a_sub[0] + c_sup[0]) - kx2[kr];
        const double bi = -(a_sub[0] + c_sup[0]) - kx2[ki];
        cplx rhs = col[(size_t)M * 1];
        rhs.x = -rhs.x; rhs.y = -rhs.y;
        if (k == 0) rhs.x -= a_sub[0] * bc_bot;
        dp_prev = {(rhs.x - a_sub[0] * 0.0) / br, (rhs.y - a_sub[0] * 0.0) / bi};
        col[(size_t)M * 1] = dp_prev;
    }
    for (int r = 1; r < Ni; ++r) {
        const int j = r + 1;
        const double as = a_sub[r], cs = c_sup[r];
        cplx rhs = col[(size_t)M * j];
        rhs.x = -rhs.x; rhs.y = -rhs.y;
        if (k == 0 && r == Ni - 1) rhs.x -= cs * bc_top;
        const double den_r = cs / cp[(size_t)r * ld + kr];
        const double den_i = cs / cp[(size_t)r * ld + ki];
        dp_prev = {(rhs.x - as * dp_prev.x) / den_r, (rhs.y - as * dp_prev.y) / den_i};
        col[(size_t)M * j] = dp_prev;
    }
    cplx xnext = col[(size_t)M * Ni];
    for (int r = Ni - 2; r >= 0; --r) {
        const int j = r + 1;
        const cplx dp = col[(size_t)M * j];
        const cplx xj = {dp.x - cp[(size_t)r * ld + kr] * xnext.x, dp.y - cp[(size_t)r * ld + ki] * xnext.y};
        col[(size_t)M * j] = xj;
        xnext = xj;
    }
    col[0] = (k == 0 ? cplx{bc_bot, 0.0} : cplx{0.0, 0.0});
    col[(size_t)M * (Ny - 1)] = (k == 0 ? cplx{bc_top, 0.0} : cplx{0.0, 0.0});
}

int ilog2(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

}  // namespace

struct PoissonImpl {
    int Nx = 0, Ny = 0, M = 0, log2M = 0, threads = 256, it4 = 1;
    double Lx = 0, inv_deta2 = 0;
    bool mapped = false, profile = false;
    float acc_fft_ms = 0.f, acc_tridiag_ms = 0.f;
    cplx* d_tw = nullptr;   // [Nx] exp(-2 pi i q / Nx)
    double* d_kx2 = nullptr;  // [M+1]
    double* d_cp = nullptr;   // [Ny-2][M+1]
    double* d_asub = nullptr;  // [Ny-2]
    double* d_csup = nullptr;  // [Ny-2]
    cplx* d_hat = nullptr;    // [Ny][M]
    size_t lds_bytes = 0;
};

PoissonImpl* poisson_create(int Nx, int Ny, double Lx, hipStream_t s) {
    if (Nx < 4 || Nx > 16384 || (Nx & (Nx - 1)) != 0) {
        set_last_error("poisson2d_create: Nx must be a power of two in [4, 16384], got " + std::to_string(Nx));
        return nullptr;
    }
    if (Ny < 4) {
        set_last_error("poisson2d_create: Ny must be >= 4, got " + std::to_string(Ny));
        return nullptr;
    }
    auto* p = new PoissonImpl;
    p->Nx = Nx; p->Ny = Ny; p->Lx = Lx;
    p->M = Nx / 2;
    p->log2M = ilog2(p->M);
    p->lds_bytes = (size_t)p->M * sizeof(cplx);
    int dev = 0, max_lds = 0;
    NSE2D_HIP_CHECK(hipGetDevice(&dev));
    NSE2D_HIP_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (p->lds_bytes > (size_t)max_lds) {
        set_last_error("poisson2d_create: row FFT needs " + std::to_string(p->lds_bytes) +
                       " bytes of LDS, device allows " + std::to_string(max_lds));
        delete p;
        return nullptr;
    }
    // threads per row FFT: one radix-4 item per thread up to M = 4096, two beyond
    p->threads = std::min(1024, std::max(64, p->M / 4));
    p->it4 = (p->M / 4 + p->threads - 1) / p->threads;
    if (p->it4 < 1) p->it4 = 1;
    if (p->lds_bytes > 65536) {
        const void* fns[4] = {(const void*)fft_rows_forward_kernel<1>, (const void*)fft_rows_forward_kernel<2>,
                              (const void*)fft_rows_inverse_kernel<1>, (const void*)fft_rows_inverse_kernel<2>};
        for (const void* f : fns)
            NSE2D_HIP_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_bytes));
    }
    // twiddles exp(-2 pi i q / Nx), computed in long double
    std::vector<cplx> tw(Nx);
    for (int q = 0; q < Nx; ++q) {
        const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)q / (long double)Nx;
        tw[q] = {(double)std::cos(ang), (double)std::sin(ang)};
    }
    // kx^2 for k = 0..M (poisson.cu:16-22: k <= Nx/2 maps to +k, so the Nyquist is +pi Nx/Lx)
    std::vector<double> kx2(p->M + 1);
    for (int k = 0; k <= p->M; ++k) {
        const double kphys = (2.0 * kPi * (double)k) / Lx;
        kx2[k] = kphys * kphys;
    }
    NSE2D_HIP_CHECK(hipMalloc(&p->d_tw, sizeof(cplx) * Nx));
    NSE2D_HIP_CHECK(hipMalloc(&p->d_kx2, sizeof(double) * (p->M + 1)));
    NSE2D_HIP_CHECK(hipMalloc(&p->d_cp, sizeof(double) * (size_t)(Ny - 2) * (p->M + 1)));
    NSE2D_HIP_CHECK(hipMalloc(&p->d_asub, sizeof(double) * (Ny - 2)));
    NSE2D_HIP_CHECK(hipMalloc(&p->d_csup, sizeof(double) * (Ny - 2)));
    NSE2D_HIP_CHECK(hipMalloc(&p->d_hat, sizeof(cplx) * (size_t)Ny * p->M));
    NSE2D_HIP_CHECK(hipMemcpyAsync(p->d_tw, tw.data(), sizeof(cplx) * Nx, hipMemcpyHostToDevice, s));
    NSE2D_HIP_CHECK(hipMemcpyAsync(p->d_kx2, kx2.data(), sizeof(double) * (p->M + 1), hipMemcpyHostToDevice, s));
    NSE2D_HIP_CHECK(hipStreamSynchronize(s));
    return p;
}

int poisson_enable_mapped(PoissonImpl* p, const double* a_node, const double* a_edge,
                          double inv_deta, hipStream_t s) {
    if (!p) { set_last_error("poisson2d_enable_mapped: null plan"); return -1; }
    p->inv_deta2 = inv_deta * inv_deta;
    const int Ni = p->Ny - 2;
    row_coeff_kernel<<<ceil_div(Ni, 256), 256, 0, s>>>(p->d_asub, p->d_csup, a_node, a_edge, p->Ny, p->inv_deta2);
    NSE2D_LAUNCH_CHECK();
    precompute_cp_kernel<<<ceil_div(p->M + 1, 64), 64, 0, s>>>(p->d_cp, p->d_asub, p->d_csup, p->d_kx2, p->M, Ni);
    NSE2D_LAUNCH_CHECK();
    p->mapped = true;
    return 0;
}

void poisson_forward(PoissonImpl* p, const double* omega, hipStream_t s) {
    if (p->it4 == 1)
        fft_rows_forward_kernel<1><<<p->Ny, p->threads, p->lds_bytes, s>>>(omega, p->d_hat, p->Nx, p->log2M, p->d_tw);
    else
        fft_rows_forward_kernel<2><<<p->Ny, p->threads, p->lds_bytes, s>>>(omega, p->d_hat, p->Nx, p->log2M, p->d_tw);
    NSE2D_LAUNCH_CHECK();
}
void poisson_tridiag(PoissonImpl* p, double psi_bot, double psi_top, hipStream_t s) {
    thomas_kernel<<<ceil_div(p->M, 64), 64, 0, s>>>(p->d_hat, p->d_cp, p->d_asub, p->d_csup, p->d_kx2, p->M, p->Ny,
                                                   (double)p->Nx * psi_bot, (double)p->Nx * psi_top);
    NSE2D_LAUNCH_CHECK();
}
void poisson_inverse(PoissonImpl* p, double* psi, hipStream_t s) {
    const double scale = (double)(1.0 / (double)p->Nx);
    if (p->it4 == 1)
        fft_rows_inverse_kernel<1><<<p->Ny, p->threads, p->lds_bytes, s>>>(p->d_hat, psi, p->Nx, p->log2M, p->d_tw, scale);
    else
        fft_rows_inverse_kernel<2><<<p->Ny, p->threads, p->lds_bytes, s>>>(p->d_hat, psi, p->Nx, p->log2M, p->d_tw, scale);
    NSE2D_LAUNCH_CHECK();
}

void poisson_solve(PoissonImpl* p, const double* omega, double* psi, double psi_bot, double psi_top,
                   hipStream_t s) {
    if (!p->mapped) {
        std::fprintf(stderr, "poisson2d_solve: enable_mapped() was not called\n");
        std::abort();
    }
    if (!p->profile) {
        poisson_forward(p, omega, s);
        poisson_tridiag(p, psi_bot, psi_top, s);
        poisson_inverse(p, psi, s);
        return;
    }
    hipEvent_t e[6];
    for (auto& ev : e) NSE2D_HIP_CHECK(hipEventCreate(&ev));
    NSE2D_HIP_CHECK(hipEventRecord(e[0], s));
    poisson_forward(p, omega, s);
    NSE2D_HIP_CHECK(hipEventRecord(e[1], s));
    poisson_tridiag(p, psi_bot, psi_top, s);
    NSE2D_HIP_CHECK(hipEventRecord(e[2], s));
    poisson_inverse(p, psi, s);
    NSE2D_HIP_CHECK(hipEventRecord(e[3], s));
    NSE2D_HIP_CHECK(hipEventSynchronize(e[3]));
    float a = 0, b = 0, c = 0;
    NSE2D_HIP_CHECK(hipEventElapsedTime(&a, e[0], e[1]));
    NSE2D_HIP_CHECK(hipEventElapsedTime(&b, e[1], e[2]));
    NSE2D_HIP_CHECK(hipEventElapsedTime(&c, e[2], e[3]));
    p->acc_fft_ms += a + c;
    p->acc_tridiag_ms += b;
    for (auto& ev : e) NSE2D_HIP_CHECK(hipEventDestroy(ev));
}

void poisson_destroy(PoissonImpl* p) {
    if (!p) return;
    (void)hipFree(p->d_tw); (void)hipFree(p->d_kx2); (void)hipFree(p->d_cp);
    (void)hipFree(p->d_asub); (void)hipFree(p->d_csup); (void)hipFree(p->d_hat);
    delete p;
}

}  // namespace nse2d

struct nse2d_poisson {
    nse2d::PoissonImpl* impl;
};

extern "C" {

nse2d_poisson_t poisson2d_create(int Nx, int Ny, double Lx, hipStream_t s) {
    nse2d::PoissonImpl* impl = nse2d::poisson_create(Nx, Ny, Lx, s);
    if (!impl) return nullptr;
    return new nse2d_poisson{impl};
}
int poisson2d_enable_mapped(nse2d_poisson_t p, const double* d_a_node, const double* d_a_edge,
                            double inv_deta, hipStream_t s) {
    if (!p) { nse2d::set_last_error("poisson2d_enable_mapped: null plan"); return -1; }
    return nse2d::poisson_enable_mapped(p->impl, d_a_node, d_a_edge, inv_deta, s);
}
void poisson2d_solve(nse2d_poisson_t p, const double* d_omega, double* d_psi, double dy,
                     double psi_bot, double psi_top, hipStream_t s) {
    (void)dy;
    nse2d::poisson_solve(p->impl, d_omega, d_psi, psi_bot, psi_top, s);
}
void poisson2d_set_profile(nse2d_poisson_t p, int on) { p->impl->profile = (on != 0); }
void poisson2d_profile_reset(nse2d_poisson_t p) { p->impl->acc_fft_ms = p->impl->acc_tridiag_ms = 0.f; }
void poisson2d_profile_get(nse2d_poisson_t p, float* fft_ms, float* tridiag_ms) {
    *fft_ms = p->impl->acc_fft_ms;
    *tridiag_ms = p->impl->acc_tridiag_ms;
}
void poisson2d_destroy(nse2d_poisson_t p) {
    if (!p) return;
    nse2d::poisson_destroy(p->impl);
    delete p;
}

}  // extern "C"
