// Poisson solver for the streamfunction: -lap_h psi = omega with spectral x (FFT) and the
// mapped second-order y operator, psi = psi_bot / psi_top on the walls.
// Replaces poisson.cu + poisson.cuh (cuFFT Z2Z + one-thread-per-kx Thomas).
//
// Row FFTs: fft_rows.hpp (register radix-16/8/4 Stockham passes) for Nx/2 in [256, 8192]; the
// LDS radix-4 kernels below for the smaller rows.
//
// Spectral layout: for each row j the Nx real values are transformed as one complex FFT of
// M = Nx/2 points (even samples in the real part, odd samples in the imaginary part), then
// split into the half spectrum X[0..M]. X[0] and X[M] are real for real data, so they share one
// complex slot: hat[j][0] = (X[0], X[M]). A spectrum row is exactly M complex values, [Ny][M]
// row-major with kx fastest, so that the y-direction solve reads and writes whole row segments;
// slot 0 carries two real systems (kx = 0 with the wall values, kx = Nyquist without).
//
// y-direction solve (chunked Thomas): a workgroup owns KC kx columns and splits the Ny-2
// interior rows into NCH chunks of RC rows, one lane per (column, chunk). Each lane runs the
// forward elimination of its chunk with zero carry-in, keeping the local values in registers
// and the chunk's affine transfer maps; an LDS scan composes the maps into the true carries;
// the lanes then finish the back substitution and write the solution. HBM traffic is one read
// and one write of the spectrum; the elimination coefficients are recomputed from the values
// at the chunk boundaries (precomputed once per plan).
#include "nse2d_common.hpp"
#include "nse2d_poisson.hpp"
#include "fft_rows.hpp"
#include "../../include/nse2d_hip.h"

#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

namespace nse2d {

void set_last_error(const std::string& msg);

namespace {

using fft::cplx;
using fft::cadd;
using fft::csub;
using fft::cmul;
using fft::cconj;
using fft::cscale;

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
                                                               cplx* __restrict__ hat, int Nx, int log2M,
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
                                                               double* __restrict__ psi, int Nx, int log2M,
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

// Row coefficients of the mapped operator, poisson.cu:83-88,103-107:
// a_sub = a_j a_{j-1/2} / deta^2, c_sup = a_j a_{j+1/2} / deta^2 for interior row r = j-1.
__global__ void row_coeff_kernel(double* __restrict__ a_sub, double* __restrict__ c_sup,
                                 const double* __restrict__ a_node,
                                 const double* __restrict__ a_edge, int Ny, double inv_deta2) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= Ny - 2) return;
    const int j = r + 1;
    a_sub[r] = a_node[j] * a_edge[j - 1] * inv_deta2;
    c_sup[r] = a_node[j] * a_edge[j] * inv_deta2;
}

// 1/x for a pivot of the y-direction systems (|x| >= a + c > 0, far inside the normal range):
// the hardware reciprocal refined by two Newton steps, within an ulp of the quotient. The IEEE
// division adds scaling and fix-up steps for denormals and overflow, which a pivot never is.
// Every kernel that forms a pivot uses this, so their recurrences agree bit for bit.
__device__ __forceinline__ double pivot_recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// Elimination recurrence of poisson.cu:38-68 (cp_r = c_r / (b_r - a_r cp_{r-1}), b_r =
// -(a_r + c_r) - kx^2), one thread per kx (k = 0..M, k = M the Nyquist), storing the value that
// enters each chunk: cpb[c][k] = cp_{c*RC - 1} for c >= 1. Same operations as the solve kernel,
// so the chunked solve reproduces the serial recurrence exactly.
__global__ void chunk_boundary_cp_kernel(double* __restrict__ cpb, const double* __restrict__ a_sub,
                                         const double* __restrict__ c_sup, const double* __restrict__ kx2,
                                         int ncols, int Ni, int RC, int NCH) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncols) return;
    double cp_prev = 0.0;
    for (int c = 0; c < NCH; ++c) {
        if (c > 0) cpb[(size_t)c * ncols + k] = cp_prev;
        for (int i = 0; i < RC; ++i) {
            const int r = c * RC + i;
            if (r >= Ni) break;
            const double a = a_sub[r], cc = c_sup[r];
            const double b = -(a + cc) - kx2[k];
            const double gi = pivot_recip(b - a * cp_prev);
            cp_prev = cc * gi;
        }
    }
}

// The pivots g_r of the single system with kx^2 = kk[0] (the Nyquist system), serially, with the
// operations of the solve kernel's sweep.
__global__ void column_pivot_kernel(double* __restrict__ g, const double* __restrict__ a_sub,
                                    const double* __restrict__ c_sup, const double* __restrict__ kk, int Ni) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double cp_prev = 0.0;
    for (int r = 0; r < Ni; ++r) {
        const double a = a_sub[r], cc = c_sup[r];
        const double gi = pivot_recip((-(a + cc) - kk[0]) - a * cp_prev);
        g[r] = gi;
        cp_prev = cc * gi;
    }
}

// Chunked Thomas solve (see the header comment), in place on hat. For interior row r (grid row
// r+1) and spectral slot k: a_r x_{r-1} + b_r x_r + c_r x_{r+1} = -omega_hat, with the wall values
// entering the kx = 0 system (scaled by Nx, unnormalised transform; poisson.cu:93-114) and
// psi_hat = Nx psi_wall at kx = 0, 0 otherwise on the wall rows (poisson.cu:131-132).
//   forward: dp_r = (f_r - a_r dp_{r-1}) g_r, g_r = 1 / (b_r - a_r cp_{r-1}), cp_r = c_r g_r
//   backward: x_r = dp_r - cp_r x_{r+1}
// Per chunk [s, e]: dp_r = dpL_r + P_r D with P_r = prod_{s..r} (-a_i g_i), D = dp_{s-1};
// x_s = S + RP D + QS X with X = x_{e+1}, S = sum_r W_r dpL_r, RP = sum_r W_r P_r,
// W_r = prod_{s..r-1} (-cp_i), QS = prod_{s..e} (-cp_i).
// Real and imaginary parts share the coefficient chain of their column, except in slot 0 (real:
// kx = 0, imaginary: Nyquist), whose imaginary part takes the Nyquist pivots g from a table.
template <int RC, int KC, int T>
struct ThomasShared {
    // The row coefficients a_r and c_r of every row, and the Nyquist system's pivots, staged in
    // LDS when they fit (up to 2048 rows): the serial loops then read them at LDS latency instead
    // of one dependent L2 round trip per row (measured 38 -> 25 us per solve at 4096x2049).
    static constexpr bool COEF = (T / KC) * RC <= 2048;
    static constexpr int NAGG = 4 * (T / 64) * KC;  // the carry scans' wavefront totals
    // Row r sits at r + r / RC: the chunks of one wavefront read their rows at the same time,
    // and without the padding word per chunk those reads (RC doubles apart) fall in one bank.
    static_assert((RC & (RC - 1)) == 0, "chunk heights are powers of two");
    static constexpr int LOG2RC = fft::ilog2c(RC);
    static constexpr int NCOEF = COEF ? (T / KC) * (RC + 1) : 1;
    __device__ static int pad(int r) { return r + (r >> LOG2RC); }
    // The pivots g of each lane's first GS rows, kept from the forward sweep for the second pass
    // (which otherwise recomputes the serial pivot recurrence: 4.2 of 27 us per solve at
    // 4096x2049). As many rows as fit beside the rest in the CU's 160 KB; chunks of 8 rows or
    // fewer keep to 80 KB, so that two workgroups fit a CU where their registers allow it.
    static constexpr int LDS_BUDGET = (RC >= 16 ? 160 : 80) * 1024;
    static constexpr int FREE_ROWS = (LDS_BUDGET - 8 * (NAGG + 3 * NCOEF)) / (8 * T);
    static constexpr int GS = FREE_ROWS < RC ? (FREE_ROWS > 0 ? FREE_ROWS : 0) : RC;
    double agg[NAGG];
    double ca[NCOEF], cc[NCOEF], gq[NCOEF];
    double gs[GS > 0 ? GS : 1][T];
};

// The solve of column group `grp` (columns grp*KC .. grp*KC+KC-1) of one spectrum. Each lane
// carries two coefficient chains, x for the real part and y for the imaginary part, which are the
// same chain (same operations, same values) except in slot 0, the first column of group 0: there
// the real part is the kx = 0 system, with the wall values, and the imaginary part the Nyquist
// system, whose pivots g are read from the table gny (computed once per plan by the same
// recurrence). Every workgroup runs the same code. (Measured at 4096x2049: slot 0 in a code path
// of its own, in group 0 or in a 257th workgroup, ended 5 to 9 us after the other groups. The
// Nyquist system solved by group 0 after its columns took 9 us, and a second run of the same
// code right after it 5 us (its instructions then cached); a 257th workgroup moreover waits for a
// CU to free, one 512-lane workgroup filling a CU's registers.)
//
// Issue order: the small coefficient loads are staged in LDS first, before any column load is
// issued (under the flood of every workgroup's column loads they return only after most of the
// columns, measured 8 of 24 us). The lane's RC column values are then loads issued LA rows ahead
// of the forward sweep, unconditionally (a row past the last interior row re-reads that row and
// is skipped by the sweeps): loads complete in issue order, so each elimination row waits only
// for its own value, and the sweep runs while the column arrives instead of after the last load
// has been issued (issue stalls while the memory queues are full; at 4096x2049 the issue of a
// whole column took 6.6 us). Measured at 4096x2049: 28.6 -> 24.8 us per solve.
template <int RC, int KC, int T>
__device__ __forceinline__ void thomas_group(ThomasShared<RC, KC, T>& shm, int grp, cplx* __restrict__ hat, int M,
                                             int Ni, const double* __restrict__ a_sub,
                                             const double* __restrict__ c_sup, const double* __restrict__ kx2,
                                             const double* __restrict__ cpb, const double* __restrict__ gny,
                                             double bc_bot, double bc_top) {
    constexpr int NCH = T / KC;
    constexpr int NR = NCH * RC;  // rows covered by the chunks (>= Ni)
    constexpr bool COEF = ThomasShared<RC, KC, T>::COEF;
    constexpr int GS = ThomasShared<RC, KC, T>::GS;
    const int t = threadIdx.x, kc = t % KC, c = t / KC;
    const int k = grp * KC + kc;
    const bool colok = k < M;
    const bool split = k == 0;  // slot 0: kx = 0 (real part, with the wall values), Nyquist (imaginary)
    const int r0 = c * RC;
    const int nr = colok ? max(0, min(RC, Ni - r0)) : 0;
    const int kl = colok ? k : M - 1;
    // coefficients: row coefficients of the rows this lane stages
    constexpr int NS = (NR + T - 1) / T;
    double sa[COEF ? NS : 1], sc[COEF ? NS : 1], sg[COEF ? NS : 1];
    if (COEF) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const int r = min(t + q * T, Ni - 1);
            sa[q] = a_sub[r];
            sc[q] = c_sup[r];
            sg[q] = gny[r];
        }
    }
    const double kkx = kx2[kl];
    const double cpx0 = cpb[(size_t)c * (M + 1) + kl];
    if (COEF) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const int r = t + q * T;
            // rows Ni..NR-1 receive row Ni-1's values and are never read
            if (NR % T == 0 || r < NR) {
                const int rp = ThomasShared<RC, KC, T>::pad(r);
                shm.ca[rp] = sa[q];
                shm.cc[rp] = sc[q];
                shm.gq[rp] = sg[q];
            }
        }
        __syncthreads();  // uniform over the workgroup
    }
    constexpr int LA = RC < 4 ? RC : 4;  // column loads in flight ahead of the sweep
    // the lane's column as 32-bit byte offsets from the spectrum's base (a scalar base address
    // and one offset register per access instead of a 64-bit address per row)
    const char* hb = reinterpret_cast<const char*>(hat);
    auto col = [&](int row) {
        return *reinterpret_cast<const cplx*>(hb + ((unsigned)row * (unsigned)M + (unsigned)kl) * 16u);
    };
    cplx f[RC];
#pragma unroll
    for (int i = 0; i < LA; ++i) f[i] = col(min(r0 + i + 1, Ni));
    // the coefficients of row r0 + i of the lane's chunk (i < nr), in LDS at pad(r0) + i, so that
    // the row is an immediate offset; A1, C1, G1 also take rows past the last interior row (the
    // Q pass reads a row ahead): LDS holds copies of that row there, global memory is clamped
    using SH = ThomasShared<RC, KC, T>;
    const int pb = SH::pad(r0);
    auto A = [&](int i) { return COEF ? shm.ca[pb + i] : a_sub[r0 + i]; };
    auto C = [&](int i) { return COEF ? shm.cc[pb + i] : c_sup[r0 + i]; };
    auto G = [&](int i) { return COEF ? shm.gq[pb + i] : gny[r0 + i]; };
    auto A1 = [&](int i) { return COEF ? shm.ca[pb + i] : a_sub[min(r0 + i, Ni - 1)]; };
    auto C1 = [&](int i) { return COEF ? shm.cc[pb + i] : c_sup[min(r0 + i, Ni - 1)]; };
    auto G1 = [&](int i) { return COEF ? shm.gq[pb + i] : gny[min(r0 + i, Ni - 1)]; };
    const double cpx_in = (c > 0 && nr > 0) ? cpx0 : 0.0;
    double cpx = cpx_in;
    cplx dpl = {0.0, 0.0}, S = {0.0, 0.0};
    double Wx = 1.0, Px = 1.0, RPx = 0.0, Wy = 1.0, Py = 1.0, RPy = 0.0;
#pragma unroll
    for (int i = 0; i < RC; ++i) {
        if (i + LA < RC) f[i + LA] = col(min(r0 + i + LA + 1, Ni));
        // keeps the scheduler from hoisting the later rows' loads above this row's elimination
        __builtin_amdgcn_sched_barrier(0);
        // (a branch per row, as in the sweeps below: unguarded, the unrolled sweep holds more
        // registers and spills at RC = 32)
        if (i >= nr) continue;
        const double a = A(i), cc = C(i);
        const double rx = -f[i].x, ry = -f[i].y;
        const double gx = pivot_recip((-(a + cc) - kkx) - a * cpx);
        if (i < GS) shm.gs[i][t] = gx;
        double gy = gx;
        if (split) gy = G(i);
        cpx = cc * gx;
        dpl = {(rx - a * dpl.x) * gx, (ry - a * dpl.y) * gy};
        f[i] = dpl;
        Px *= -a * gx;
        Py *= -a * gy;
        S = {S.x + Wx * dpl.x, S.y + Wy * dpl.y};
        RPx += Wx * Px;
        RPy += Wy * Py;
        Wx *= -cpx;
        Wy *= -(cc * gy);
    }
    // Every column load has returned (the sweep has used each row). Said explicitly: rows past the
    // chunk's end keep their loaded values in f, so without this the compiler finds a load
    // possibly pending on every f[i] it reads later, and before each back-substitution row it
    // waited for the previous rows' stores as well (vector memory counts complete in order).
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    // Chunk carries. Forward: D_c = dp at the end of chunk c-1, the composition over chunks
    // 0..c-1 of the maps D -> E + P D (E = the chunk's local dp at its end), applied to the
    // carry into chunk 0. Backward: X_c = x at the start of chunk c+1, the composition over
    // chunks c+1..NCH-1 of X -> h + Q X (h = S + RP D, Q = W), applied to the carry out of the
    // last chunk. The wall values are these two end carries: with dp_{-1} = x_{-1} = psi_bot_hat
    // and x_Ni = psi_top_hat the first and last rows' recurrences take the wall terms
    // -a_0 psi_bot_hat and -c_{Ni-1} psi_top_hat of their right-hand sides; both are zero except in
    // the kx = 0 system. Chunk c of column kc sits in lane t = c*KC + kc, so the CPW = 64/KC
    // chunks of one wavefront are composed with lane shuffles (no barrier); the wavefronts'
    // totals then meet in LDS (one barrier per direction).
    constexpr int CPW = 64 / KC, NW = T / 64;
    const int lane = t & 63, w = t >> 6;
    const bool first_in_wave = lane < KC, last_in_wave = lane >= 64 - KC;
    double* agg = shm.agg;  // [4][NW][KC] wavefront totals (the scans' scratch)
    {
        double ex = dpl.x, ey = dpl.y, px = Px, py = Py;
#pragma unroll
        for (int d = 1; d < CPW; d <<= 1) {
            const double ux = __shfl_up(ex, d * KC, 64), uy = __shfl_up(ey, d * KC, 64);
            const double upx = __shfl_up(px, d * KC, 64), upy = __shfl_up(py, d * KC, 64);
            if (lane >= d * KC) {
                ex += px * ux;
                ey += py * uy;
                px *= upx;
                py *= upy;
            }
        }
        if (last_in_wave) {
            agg[(0 * NW + w) * KC + kc] = ex;
            agg[(1 * NW + w) * KC + kc] = ey;
            agg[(2 * NW + w) * KC + kc] = px;
            agg[(3 * NW + w) * KC + kc] = py;
        }
        // exclusive within the wavefront: the inclusive value of the previous chunk
        double dx = __shfl_up(ex, KC, 64), dy = __shfl_up(ey, KC, 64);
        double qx = __shfl_up(px, KC, 64), qy = __shfl_up(py, KC, 64);
        if (first_in_wave) { dx = 0.0; dy = 0.0; qx = 1.0; qy = 1.0; }
        __syncthreads();
        // carry into this wavefront: the earlier wavefronts' totals, composed in order on the
        // carry into chunk 0. Each total is read one step ahead of its use (reading it in the
        // step, every step waited a full LDS latency, up to NW - 1 of them in series).
        double cx = split ? bc_bot : 0.0, cy = 0.0;
        double tn[4];  // wavefront v's totals: E (real, imaginary), P (real, imaginary)
#pragma unroll
        for (int q = 0; q < 4; ++q) tn[q] = agg[(q * NW + 0) * KC + kc];
#pragma unroll 1
        for (int v = 0; v < w; ++v) {
            const double tc[4] = {tn[0], tn[1], tn[2], tn[3]};
#pragma unroll
            for (int q = 0; q < 4; ++q) tn[q] = agg[(q * NW + v + 1) * KC + kc];  // v + 1 <= w < NW
            cx = tc[0] + tc[2] * cx;
            cy = tc[1] + tc[3] * cy;
        }
        dpl = {dx + qx * cx, dy + qy * cy};  // now D of this chunk
    }
    const cplx D = dpl;
    __syncthreads();
    {
        double hx = S.x + RPx * D.x, hy = S.y + RPy * D.y, qx = Wx, qy = Wy;
#pragma unroll
        for (int d = 1; d < CPW; d <<= 1) {
            const double ux = __shfl_down(hx, d * KC, 64), uy = __shfl_down(hy, d * KC, 64);
            const double uqx = __shfl_down(qx, d * KC, 64), uqy = __shfl_down(qy, d * KC, 64);
            if (lane + d * KC < 64) {
                hx += qx * ux;
                hy += qy * uy;
                qx *= uqx;
                qy *= uqy;
            }
        }
        if (first_in_wave) {
            agg[(0 * NW + w) * KC + kc] = hx;
            agg[(1 * NW + w) * KC + kc] = hy;
            agg[(2 * NW + w) * KC + kc] = qx;
            agg[(3 * NW + w) * KC + kc] = qy;
        }
        double xx = __shfl_down(hx, KC, 64), xy = __shfl_down(hy, KC, 64);
        double rx = __shfl_down(qx, KC, 64), ry = __shfl_down(qy, KC, 64);
        if (last_in_wave) { xx = 0.0; xy = 0.0; rx = 1.0; ry = 1.0; }
        __syncthreads();
        double cx = split ? bc_top : 0.0, cy = 0.0;
        double tn[4];  // as in the forward composition, read one step ahead
#pragma unroll
        for (int q = 0; q < 4; ++q) tn[q] = agg[(q * NW + NW - 1) * KC + kc];
#pragma unroll 1
        for (int v = NW - 1; v > w; --v) {
            const double tc[4] = {tn[0], tn[1], tn[2], tn[3]};
#pragma unroll
            for (int q = 0; q < 4; ++q) tn[q] = agg[(q * NW + v - 1) * KC + kc];  // v - 1 >= w >= 0
            cx = tc[0] + tc[2] * cx;
            cy = tc[1] + tc[3] * cy;
        }
        S = {xx + rx * cx, xy + ry * cy};  // now X of this chunk
    }
    const cplx X = S;
    // true dp, then back substitution from x_{e+1} = X. The pivots g_r depend on the
    // coefficients only: the first GS rows' come back from LDS, the rest are recomputed (the same
    // operations as in the forward sweep, so the same values), and they are kept as
    // cp_r = c_r g_r, the back substitution's factor (the y chain's differs only in slot 0 and
    // is recomputed there from the table).
    double cp[RC];
    double Qx = 1.0, Qy = 1.0;
    cpx = cpx_in;
    // each row's coefficients and stored pivot are read one row ahead, outside the row's branch
    // (inside it, every row waited a full LDS latency for its own reads)
    double an = A1(0), cn = C1(0), qn = G1(0), gn = GS > 0 ? shm.gs[0][t] : 0.0;
#pragma unroll
    for (int i = 0; i < RC; ++i) {
        const double a = an, cc = cn, gq = qn, gsv = gn;
        if (i + 1 < RC) {
            an = A1(i + 1);
            cn = C1(i + 1);
            qn = G1(i + 1);
            if (i + 1 < GS) gn = shm.gs[i + 1][t];
        }
        if (i < nr) {
            const double gx = i < GS ? gsv : pivot_recip((-(a + cc) - kkx) - a * cpx);
            double gy = gx;
            if (split) gy = gq;
            cpx = cc * gx;
            cp[i] = cpx;
            Qx *= -a * gx;
            Qy *= -a * gy;
            f[i] = {f[i].x + Qx * D.x, f[i].y + Qy * D.y};
        }
    }
    cplx xn = X;
    // the rows' store addresses from a base the compiler cannot tie to the loads' row indices
    // (it kept those live from the first loads to here, two values in scratch)
    int rbase = r0 + 1;
    asm volatile("" : "+v"(rbase));
    // 32-bit byte offsets from the spectrum's base, as for the loads (a 64-bit address per row
    // took a 64-bit multiply-add per row)
    char* const hw = reinterpret_cast<char*>(hat);
    const unsigned off0 = ((unsigned)rbase * (unsigned)M + (unsigned)k) * 16u, rowb = (unsigned)M * 16u;
#pragma unroll
    for (int i = RC - 1; i >= 0; --i) {
        if (i < nr) {
            double cpy = cp[i];
            if (split) cpy = C(i) * G(i);
            xn = {f[i].x - cp[i] * xn.x, f[i].y - cpy * xn.y};
            *reinterpret_cast<cplx*>(hw + (off0 + (unsigned)i * rowb)) = xn;  // row r0 + i + 1
        }
    }
    if (c == 0 && colok) {
        hat[k] = (split ? cplx{bc_bot, 0.0} : cplx{0.0, 0.0});
        hat[(size_t)(Ni + 1) * M + k] = (split ? cplx{bc_top, 0.0} : cplx{0.0, 0.0});
    }
}

// The solve of one spectrum: workgroup b owns the KC columns of group xcd_tile(b).
template <int RC, int KC, int T>
__global__ void __launch_bounds__(T) thomas_chunked_kernel(cplx* __restrict__ hat, int M, int Ni,
                                                           const double* __restrict__ a_sub,
                                                           const double* __restrict__ c_sup,
                                                           const double* __restrict__ kx2,
                                                           const double* __restrict__ cpb,
                                                           const double* __restrict__ gny, double bc_bot,
                                                           double bc_top) {
    __shared__ ThomasShared<RC, KC, T> S;
    // column groups contiguous per XCD: groups narrower than a 128-byte line share lines in one L2
    thomas_group<RC, KC, T>(S, xcd_tile(blockIdx.x, gridDim.x), hat, M, Ni, a_sub, c_sup, kx2, cpb, gny, bc_bot,
                            bc_top);
}

// A group of n equally sized spectra sharing one plan: workgroup b solves column group b % groups
// of spectrum b / groups (b after the XCD mapping).
template <int RC, int KC, int T>
__global__ void __launch_bounds__(T) thomas_chunked_group_kernel(cplx* const* __restrict__ hats, int groups,
                                                                 int M, int Ni, const double* __restrict__ a_sub,
                                                                 const double* __restrict__ c_sup,
                                                                 const double* __restrict__ kx2,
                                                                 const double* __restrict__ cpb,
                                                                 const double* __restrict__ gny, double bc_bot,
                                                                 double bc_top) {
    __shared__ ThomasShared<RC, KC, T> S;
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int m = b / groups;
    thomas_group<RC, KC, T>(S, b - m * groups, hats[m], M, Ni, a_sub, c_sup, kx2, cpb, gny, bc_bot, bc_top);
}

int ilog2(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

}  // namespace

struct PoissonImpl {
    int Nx = 0, Ny = 0, M = 0, log2M = 0, threads = 256, it4 = 1;
    int rc = 1, kc = 8, nch = 64;  // chunked Thomas layout
    double Lx = 0, inv_deta2 = 0;
    bool mapped = false, profile = false;
    float acc_fft_ms = 0.f, acc_tridiag_ms = 0.f;
    int fft_radix = 0;         // register row FFT radix (0: the LDS radix-4 kernels)
    cplx* d_tw = nullptr;      // [Nx] exp(-2 pi i q / Nx)
    cplx* d_twp = nullptr;     // per-pass twiddles of the register row FFT
    double* d_kx2 = nullptr;   // [M+1]
    double* d_cpb = nullptr;   // [NCH][M+1] elimination coefficient entering each chunk
    double* d_gny = nullptr;   // [Ny-2] pivots g of the Nyquist system (slot 0, imaginary part)
    double* d_asub = nullptr;  // [Ny-2]
    double* d_csup = nullptr;  // [Ny-2]
    cplx* d_hat = nullptr;     // [Ny][M]
    size_t lds_bytes = 0;
};

namespace {
// Register row FFT plans (M = Nx/2, radix R): instantiated pairs and the default radix per M.
// NSE2D_FFT_RADIX selects another instantiated radix (used to tune on the device).
bool fft_instantiated(int M, int R) {
    switch (M) {
        case 256: return R == 4;
        case 512: return R == 4 || R == 8;
        case 1024: return R == 8 || R == 16;
        case 2048: return R == 8 || R == 16;
        case 4096: case 8192: return R == 16;
        default: return false;
    }
}
int choose_fft_radix(int M, int want) {
    if (want >= 0 && (want == 0 || fft_instantiated(M, want))) return want;
    if (const char* e = std::getenv("NSE2D_FFT_RADIX")) {
        const int r = std::atoi(e);
        if (r == 0 || fft_instantiated(M, r)) return r;
    }
    const int def = M <= 256 ? 4 : (M <= 2048 ? 8 : 16);
    return fft_instantiated(M, def) ? def : 0;
}
// Twiddles of passes p >= 1 of the register FFT: [k][q-1] = exp(-2 pi i q k / (Rp Ns)).
std::vector<cplx> fft_pass_twiddles(int M, int R) {
    int logm = 0, logr = 0;
    while ((1 << logm) < M) ++logm;
    while ((1 << logr) < R) ++logr;
    const int nfull = logm / logr, rem = logm % logr, npass = nfull + (rem ? 1 : 0);
    std::vector<cplx> t;
    int ns = R;
    for (int p = 1; p < npass; ++p) {
        const int rp = p < nfull ? R : (1 << rem);
        for (int k = 0; k < ns; ++k)
            for (int q = 1; q < rp; ++q) {
                const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)(q * k) /
                                        (long double)(rp * ns);
                t.push_back({(double)std::cos(ang), (double)std::sin(ang)});
            }
        ns *= rp;
    }
    return t;
}
template <int M, int R>
void launch_rows_forward(const double* omega, cplx* hat, int rows, const cplx* twp, const cplx* tw, hipStream_t s,
                         int row0 = 0) {
    static_assert(fft::Plan<M, R>::T >= 64, "row FFT needs at least one wavefront");
    fft::rows_forward<M, R><<<rows, M / R, 0, s>>>(omega, hat, twp, tw, row0);
}
template <int M, int R>
void launch_rows_inverse(const cplx* hat, double* psi, int rows, const cplx* twp, const cplx* tw, double scale,
                         const fft::WallClosure& wc, hipStream_t s) {
    fft::rows_inverse<M, R><<<rows, M / R, 0, s>>>(hat, psi, twp, tw, scale, wc);
}
#define NSE2D_FFT_PLANS(X) X(256, 4) X(512, 4) X(512, 8) X(1024, 8) X(1024, 16) X(2048, 8) X(2048, 16) \
    X(4096, 16) X(8192, 16)

constexpr int kThomasThreads = 512;
// Chunk geometry of the y solve: KC columns per workgroup, NCH = 512/KC chunks of RC rows
// (RC a power of two, NCH * RC >= Ny-2). Instantiated: KC = 8 with RC <= 32, KC = 4 with
// RC 16/32, KC = 2 with RC 8/16/32. The default takes the widest KC; NSE2D_THOMAS_KC selects
// another column width (used to tune on the device).
bool thomas_instantiated(int kc, int rc) {
    if (kc == 8) return rc >= 1 && rc <= 32;
    if (kc == 4) return rc == 16 || rc == 32;
    if (kc == 2) return rc == 8 || rc == 16 || rc == 32;
    return false;
}
void choose_thomas_layout(PoissonImpl* p, int want) {
    const int Ni = p->Ny - 2;
    int pref = 8;
    if (const char* e = std::getenv("NSE2D_THOMAS_KC")) pref = std::atoi(e);
    if (want > 0) pref = want;

    for (int kc : {pref, 8, 4, 2}) {
        if (kc != 8 && kc != 4 && kc != 2) continue;
        const int nch = kThomasThreads / kc;
        int rc = 1;
        while (rc * nch < Ni) rc <<= 1;
        while (rc <= 32 && !thomas_instantiated(kc, rc)) rc <<= 1;
        if (rc <= 32) { p->kc = kc; p->nch = nch; p->rc = rc; return; }
    }
    p->kc = 0;
}
}  // namespace

PoissonImpl* poisson_create(int Nx, int Ny, double Lx, hipStream_t s) {
    return poisson_create_with(Nx, Ny, Lx, -1, -1, s);
}

PoissonImpl* poisson_create_with(int Nx, int Ny, double Lx, int fft_radix, int thomas_kc, hipStream_t s) {
    if (Nx < 4 || Nx > 16384 || (Nx & (Nx - 1)) != 0) {
        set_last_error("poisson2d_create: Nx must be a power of two in [4, 16384], got " + std::to_string(Nx));
        return nullptr;
    }
    if (Ny < 4 || Ny > 16386) {
        set_last_error("poisson2d_create: Ny must be in [4, 16386], got " + std::to_string(Ny));
        return nullptr;
    }
    auto* p = new PoissonImpl;
    p->Nx = Nx; p->Ny = Ny; p->Lx = Lx;
    p->M = Nx / 2;
    choose_thomas_layout(p, thomas_kc);
    if (p->kc == 0) {
        set_last_error("poisson2d_create: Ny = " + std::to_string(Ny) + " exceeds the y-solve layouts");
        delete p;
        return nullptr;
    }
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
    p->fft_radix = choose_fft_radix(p->M, fft_radix);
    if (p->fft_radix) {
        const std::vector<cplx> twp = fft_pass_twiddles(p->M, p->fft_radix);
        NSE2D_HIP_CHECK(hipMalloc(&p->d_twp, sizeof(cplx) * std::max<size_t>(1, twp.size())));
        if (!twp.empty())
            NSE2D_HIP_CHECK(hipMemcpyAsync(p->d_twp, twp.data(), sizeof(cplx) * twp.size(), hipMemcpyHostToDevice, s));
    }
    NSE2D_HIP_CHECK(hipMalloc(&p->d_tw, sizeof(cplx) * Nx));
    NSE2D_HIP_CHECK(hipMalloc(&p->d_kx2, sizeof(double) * (p->M + 1)));
    NSE2D_HIP_CHECK(hipMalloc(&p->d_cpb, sizeof(double) * (size_t)p->nch * (p->M + 1)));
    NSE2D_HIP_CHECK(hipMalloc(&p->d_gny, sizeof(double) * (size_t)(Ny - 2)));
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
    chunk_boundary_cp_kernel<<<ceil_div(p->M + 1, 64), 64, 0, s>>>(p->d_cpb, p->d_asub, p->d_csup, p->d_kx2,
                                                                  p->M + 1, Ni, p->rc, p->nch);
    NSE2D_LAUNCH_CHECK();
    column_pivot_kernel<<<1, 64, 0, s>>>(p->d_gny, p->d_asub, p->d_csup, p->d_kx2 + p->M, Ni);
    NSE2D_LAUNCH_CHECK();
    p->mapped = true;
    return 0;
}

void poisson_forward(PoissonImpl* p, const double* omega, hipStream_t s) {
    if (p->fft_radix) {
        switch (p->M * 100 + p->fft_radix) {
#define NSE2D_FWD_CASE(M, R) \
    case M * 100 + R: launch_rows_forward<M, R>(omega, p->d_hat, p->Ny, p->d_twp, p->d_tw, s); break;
            NSE2D_FFT_PLANS(NSE2D_FWD_CASE)
#undef NSE2D_FWD_CASE
            default: std::abort();
        }
        NSE2D_LAUNCH_CHECK();
        return;
    }
    if (p->it4 == 1)
        fft_rows_forward_kernel<1><<<p->Ny, p->threads, p->lds_bytes, s>>>(omega, p->d_hat, p->Nx, p->log2M, p->d_tw);
    else
        fft_rows_forward_kernel<2><<<p->Ny, p->threads, p->lds_bytes, s>>>(omega, p->d_hat, p->Nx, p->log2M, p->d_tw);
    NSE2D_LAUNCH_CHECK();
}
void poisson_forward_interior(PoissonImpl* p, const double* omega, hipStream_t s) {
    if (!p->fft_radix) {
        poisson_forward(p, omega, s);
        return;
    }
    const int nrows = p->Ny - 2;
    switch (p->M * 100 + p->fft_radix) {
#define NSE2D_FWDI_CASE(M, R)                                                                            \
    case M * 100 + R:                                                                                    \
        fft::rows_forward2<M, R><<<ceil_div(nrows, 2), M / R, 0, s>>>(omega, p->d_hat, p->d_twp, p->d_tw, 1, \
                                                                      nrows);                            \
        break;
        NSE2D_FFT_PLANS(NSE2D_FWDI_CASE)
#undef NSE2D_FWDI_CASE
        default: std::abort();
    }
    NSE2D_LAUNCH_CHECK();
}
void poisson_tridiag(PoissonImpl* p, double psi_bot, double psi_top, hipStream_t s) {
    const int Ni = p->Ny - 2;
    const double bb = (double)p->Nx * psi_bot, bt = (double)p->Nx * psi_top;
    const unsigned grid = ceil_div(p->M, p->kc);
#define NSE2D_THOMAS(RC, KC)                                                                              \
    thomas_chunked_kernel<RC, KC, kThomasThreads><<<grid, kThomasThreads, 0, s>>>(                        \
        p->d_hat, p->M, Ni, p->d_asub, p->d_csup, p->d_kx2, p->d_cpb, p->d_gny, bb, bt)
    switch (p->kc * 100 + p->rc) {
        case 801: NSE2D_THOMAS(1, 8); break;
        case 802: NSE2D_THOMAS(2, 8); break;
        case 804: NSE2D_THOMAS(4, 8); break;
        case 808: NSE2D_THOMAS(8, 8); break;
        case 816: NSE2D_THOMAS(16, 8); break;
        case 832: NSE2D_THOMAS(32, 8); break;
        case 416: NSE2D_THOMAS(16, 4); break;
        case 432: NSE2D_THOMAS(32, 4); break;
        case 208: NSE2D_THOMAS(8, 2); break;
        case 216: NSE2D_THOMAS(16, 2); break;
        default: NSE2D_THOMAS(32, 2); break;
    }
#undef NSE2D_THOMAS
    NSE2D_LAUNCH_CHECK();
}
bool poisson_fuses_closure(const PoissonImpl* p) { return p->fft_radix != 0; }
bool poisson_row_fft(const PoissonImpl* p, int* M, int* R, const void** twp, const void** tw) {
    if (!p->fft_radix) return false;
    *M = p->M;
    *R = p->fft_radix;
    *twp = p->d_twp;
    *tw = p->d_tw;
    return true;
}

bool poisson_inverse_closure(PoissonImpl* p, double* psi, double* omega, const double* dy_edge, double psi_bot,
                             double psi_top, int free_slip, hipStream_t s) {
    if (!p->fft_radix) return false;
    const double scale = (double)(1.0 / (double)p->Nx);
    const fft::WallClosure wc{omega, dy_edge, psi_bot, psi_top, p->Ny, free_slip, 1};
    switch (p->M * 100 + p->fft_radix) {
#define NSE2D_INVC_CASE(M, R) \
    case M * 100 + R: launch_rows_inverse<M, R>(p->d_hat, psi, p->Ny - 2, p->d_twp, p->d_tw, scale, wc, s); break;
        NSE2D_FFT_PLANS(NSE2D_INVC_CASE)
#undef NSE2D_INVC_CASE
        default: std::abort();
    }
    NSE2D_LAUNCH_CHECK();
    return true;
}

void poisson_inverse(PoissonImpl* p, double* psi, hipStream_t s) {
    const double scale = (double)(1.0 / (double)p->Nx);
    if (p->fft_radix) {
        const fft::WallClosure none{nullptr, nullptr, 0.0, 0.0, p->Ny, 0};
        switch (p->M * 100 + p->fft_radix) {
#define NSE2D_INV_CASE(M, R) \
    case M * 100 + R: launch_rows_inverse<M, R>(p->d_hat, psi, p->Ny, p->d_twp, p->d_tw, scale, none, s); break;
            NSE2D_FFT_PLANS(NSE2D_INV_CASE)
#undef NSE2D_INV_CASE
            default: std::abort();
        }
        NSE2D_LAUNCH_CHECK();
        return;
    }
    if (p->it4 == 1)
        fft_rows_inverse_kernel<1><<<p->Ny, p->threads, p->lds_bytes, s>>>(p->d_hat, psi, p->Nx, p->log2M, p->d_tw, scale);
    else
        fft_rows_inverse_kernel<2><<<p->Ny, p->threads, p->lds_bytes, s>>>(p->d_hat, psi, p->Nx, p->log2M, p->d_tw, scale);
    NSE2D_LAUNCH_CHECK();
}

// ---- group launches: n equally sized fields that share plan p's coefficients, each with its own
// spectrum buffer (poisson_hat_bytes). Pointer tables are device arrays of n pointers.
size_t poisson_hat_bytes(const PoissonImpl* p) { return sizeof(cplx) * (size_t)p->Ny * p->M; }
void* poisson_hat(PoissonImpl* p) { return p->d_hat; }

template <int M, int R>
void launch_rows_forward_group(const double* const* om, cplx* const* hats, int n, int rows, const cplx* twp,
                               const cplx* tw, hipStream_t s) {
    fft::rows_forward_group<M, R><<<n * rows, M / R, 0, s>>>(om, hats, rows, twp, tw);
}
template <int M, int R>
void launch_rows_inverse_group(const cplx* const* hats, double* const* psis, double* const* om, int n, int rows,
                               const cplx* twp, const cplx* tw, double scale, const fft::WallClosure& wc,
                               hipStream_t s) {
    fft::rows_inverse_group<M, R><<<n * rows, M / R, 0, s>>>(hats, psis, om, rows, twp, tw, scale, wc);
}

int poisson_group_supported(const PoissonImpl* p) { return p->fft_radix != 0; }
int poisson_fft_radix(const PoissonImpl* p) { return p->fft_radix; }
int poisson_thomas_kc(const PoissonImpl* p) { return p->kc; }

void poisson_forward_group(PoissonImpl* p, const double* const* d_omegas, void* const* d_hats, int n,
                           hipStream_t s) {
    cplx* const* hats = reinterpret_cast<cplx* const*>(d_hats);
    switch (p->M * 100 + p->fft_radix) {
#define NSE2D_FWDG_CASE(M, R) \
    case M * 100 + R: launch_rows_forward_group<M, R>(d_omegas, hats, n, p->Ny, p->d_twp, p->d_tw, s); break;
        NSE2D_FFT_PLANS(NSE2D_FWDG_CASE)
#undef NSE2D_FWDG_CASE
        default: std::abort();
    }
    NSE2D_LAUNCH_CHECK();
}

void poisson_tridiag_group(PoissonImpl* p, void* const* d_hats, int n, double psi_bot, double psi_top,
                           hipStream_t s) {
    cplx* const* hats = reinterpret_cast<cplx* const*>(d_hats);
    const int Ni = p->Ny - 2;
    const double bb = (double)p->Nx * psi_bot, bt = (double)p->Nx * psi_top;
    const int groups = (int)ceil_div(p->M, p->kc);
#define NSE2D_THOMAS_G(RC, KC)                                                                           \
    thomas_chunked_group_kernel<RC, KC, kThomasThreads><<<n * groups, kThomasThreads, 0, s>>>(           \
        hats, groups, p->M, Ni, p->d_asub, p->d_csup, p->d_kx2, p->d_cpb, p->d_gny, bb, bt)
    switch (p->kc * 100 + p->rc) {
        case 801: NSE2D_THOMAS_G(1, 8); break;
        case 802: NSE2D_THOMAS_G(2, 8); break;
        case 804: NSE2D_THOMAS_G(4, 8); break;
        case 808: NSE2D_THOMAS_G(8, 8); break;
        case 816: NSE2D_THOMAS_G(16, 8); break;
        case 832: NSE2D_THOMAS_G(32, 8); break;
        case 416: NSE2D_THOMAS_G(16, 4); break;
        case 432: NSE2D_THOMAS_G(32, 4); break;
        case 208: NSE2D_THOMAS_G(8, 2); break;
        case 216: NSE2D_THOMAS_G(16, 2); break;
        default: NSE2D_THOMAS_G(32, 2); break;
    }
#undef NSE2D_THOMAS_G
    NSE2D_LAUNCH_CHECK();
}

void poisson_inverse_group(PoissonImpl* p, void* const* d_hats, double* const* d_psis, double* const* d_omegas,
                           int n, const double* dy_edge, double psi_bot, double psi_top, int free_slip,
                           hipStream_t s) {
    const cplx* const* hats = reinterpret_cast<const cplx* const*>(d_hats);
    const double scale = (double)(1.0 / (double)p->Nx);
    const fft::WallClosure wc{nullptr, dy_edge, psi_bot, psi_top, p->Ny, free_slip};
    switch (p->M * 100 + p->fft_radix) {
#define NSE2D_INVG_CASE(M, R)                                                                             \
    case M * 100 + R:                                                                                     \
        launch_rows_inverse_group<M, R>(hats, d_psis, d_omegas, n, p->Ny, p->d_twp, p->d_tw, scale, wc, s); \
        break;
        NSE2D_FFT_PLANS(NSE2D_INVG_CASE)
#undef NSE2D_INVG_CASE
        default: std::abort();
    }
    NSE2D_LAUNCH_CHECK();
}

void poisson_solve(PoissonImpl* p, const double* omega, double* psi, double psi_bot, double psi_top,
                   hipStream_t s) {
    if (!p->mapped) {
        std::fprintf(stderr, "poisson2d_solve: enable_mapped() was not called\n");
        std::abort();
    }
    // the solve reads the interior rows' spectrum only: the same forward kernel as the stepper's
    if (!p->profile) {
        poisson_forward_interior(p, omega, s);
        poisson_tridiag(p, psi_bot, psi_top, s);
        poisson_inverse(p, psi, s);
        return;
    }
    hipEvent_t e[6];
    for (auto& ev : e) NSE2D_HIP_CHECK(hipEventCreate(&ev));
    NSE2D_HIP_CHECK(hipEventRecord(e[0], s));
    poisson_forward_interior(p, omega, s);
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

// ---- x-spectra of u and v (diag.cuh:123-191, diag.cu:869-898): every accumulate transforms
// each row of u and v (half spectrum, the plan's row FFT) and adds, per k = 0..Nx/2,
//   E(k) += wx(k) / (Nx^2 Ny) * sum_j |X_j(k)|^2,   wx = 2 for 0 < k < Nx/2, 1 otherwise.
namespace {
__global__ void spectra_rows_kernel(const cplx* __restrict__ hat, int Ny, int M, double norm1d,
                                    double* __restrict__ acc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > M) return;
    const double wx = (k > 0 && k < M) ? 2.0 : 1.0;
    double sum = 0.0;
    for (int j = 0; j < Ny; ++j) {
        const cplx* row = hat + (size_t)j * M;
        if (k == 0) { const double r = row[0].x; sum += r * r; }          // X_0 (real)
        else if (k == M) { const double r = row[0].y; sum += r * r; }     // X_M (real)
        else { const cplx z = row[k]; sum += (z.x * z.x + z.y * z.y); }
    }
    acc[k] += norm1d * wx * sum;
}
}  // namespace

struct Spectra1DImpl {
    PoissonImpl* plan = nullptr;  // only its row FFT and spectrum buffer are used
    int Nx = 0, Ny = 0, nk = 0, samples = 0;
    double Lx = 0, norm1d = 0;
    double* d_acc = nullptr;      // [2][nk]: E_u, E_v sums
};

Spectra1DImpl* spectra1d_create(int Nx, int Ny, double Lx, hipStream_t s) {
    PoissonImpl* plan = poisson_create(Nx, Ny, Lx, s);
    if (!plan) return nullptr;
    auto* sp = new Spectra1DImpl;
    sp->plan = plan;
    sp->Nx = Nx; sp->Ny = Ny; sp->Lx = Lx; sp->nk = Nx / 2 + 1;
    sp->norm1d = 1.0 / ((double)Nx * (double)Nx * (double)Ny);
    NSE2D_HIP_CHECK(hipMalloc(&sp->d_acc, sizeof(double) * 2 * sp->nk));
    NSE2D_HIP_CHECK(hipMemsetAsync(sp->d_acc, 0, sizeof(double) * 2 * sp->nk, s));
    return sp;
}

void spectra1d_accumulate(Spectra1DImpl* sp, const double* d_u, const double* d_v, hipStream_t s) {
    const int M = sp->Nx / 2;
    for (int c = 0; c < 2; ++c) {
        poisson_forward(sp->plan, c == 0 ? d_u : d_v, s);
        spectra_rows_kernel<<<ceil_div(M + 1, 256), 256, 0, s>>>(sp->plan->d_hat, sp->Ny, M, sp->norm1d,
                                                                 sp->d_acc + (size_t)c * sp->nk);
        NSE2D_LAUNCH_CHECK();
    }
    sp->samples++;
}

int spectra1d_read(Spectra1DImpl* sp, double* kx, double* Eu, double* Ev, hipStream_t s) {
    std::vector<double> acc(2 * (size_t)sp->nk);
    NSE2D_HIP_CHECK(hipMemcpyAsync(acc.data(), sp->d_acc, sizeof(double) * acc.size(), hipMemcpyDeviceToHost, s));
    NSE2D_HIP_CHECK(hipStreamSynchronize(s));
    for (int k = 0; k < sp->nk; ++k) {
        if (kx) kx[k] = (2.0 * kPi * (double)k) / sp->Lx;  // diag.cu:870-875
        double eu = acc[k], ev = acc[sp->nk + k];
        if (sp->samples > 0) { eu /= sp->samples; ev /= sp->samples; }
        if (Eu) Eu[k] = eu;
        if (Ev) Ev[k] = ev;
    }
    return sp->samples;
}

void spectra1d_destroy(Spectra1DImpl* sp) {
    if (!sp) return;
    poisson_destroy(sp->plan);
    (void)hipFree(sp->d_acc);
    delete sp;
}

void poisson_destroy(PoissonImpl* p) {
    if (!p) return;
    (void)hipFree(p->d_tw); (void)hipFree(p->d_twp); (void)hipFree(p->d_kx2); (void)hipFree(p->d_cpb);
    (void)hipFree(p->d_gny);
    (void)hipFree(p->d_asub); (void)hipFree(p->d_csup); (void)hipFree(p->d_hat);
    delete p;
}

}  // namespace nse2d

struct nse2d_poisson {
    nse2d::PoissonImpl* impl;
};
struct nse2d_spectra1d {
    nse2d::Spectra1DImpl* impl;
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

nse2d_spectra1d_t nse2d_spectra1d_create(int Nx, int Ny, double Lx, hipStream_t s) {
    nse2d::Spectra1DImpl* impl = nse2d::spectra1d_create(Nx, Ny, Lx, s);
    if (!impl) return nullptr;
    return new nse2d_spectra1d{impl};
}
void nse2d_spectra1d_accumulate(nse2d_spectra1d_t sp, const double* d_u, const double* d_v, hipStream_t s) {
    nse2d::spectra1d_accumulate(sp->impl, d_u, d_v, s);
}
int nse2d_spectra1d_read(nse2d_spectra1d_t sp, double* kx, double* Eu, double* Ev, hipStream_t s) {
    return nse2d::spectra1d_read(sp->impl, kx, Eu, Ev, s);
}
void nse2d_spectra1d_destroy(nse2d_spectra1d_t sp) {
    if (!sp) return;
    nse2d::spectra1d_destroy(sp->impl);
    delete sp;
}

}  // extern "C"
