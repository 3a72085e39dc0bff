// Row FFTs of the Poisson solve for M = Nx/2 in [256, 8192]: one workgroup per row, T = M/R
// threads, each holding R complex values in registers (R = 4, 8 or 16).
//
// Stockham autosort passes p = 0..NPASS-1 of radix Rp (R, the last pass 2^(log2 M mod log2 R)
// when R does not divide the exponent), sub-transform size Ns = prod of the earlier radices:
//   item j (0 <= j < M/Rp), k = j mod Ns:  v_q = x[j + q M/Rp] * W_{Rp Ns}^{q k}  (q < Rp)
//   y[(j - k) Rp + k + q Ns] = DFT_Rp(v)_q
// Data move through LDS between passes; the forward transform's first pass reads the row from
// HBM straight into registers and the inverse's last pass writes the row to HBM from registers,
// so a row costs one HBM read and one HBM write and NPASS-1 LDS exchanges (two for M = 2048).
// LDS index p is stored at p + p/16: the stride-Rp stores of the first pass and the (Ns >= 16)
// blocked stores of the later passes then fall on distinct banks.
//
// Real rows: the Nx real values are the M complex values z_n = (x_2n, x_2n+1); the half spectrum
// X[0..M] follows from Z by the usual split (forward) and is rebuilt before the inverse:
//   X_k = (Z_k + conj Z_{M-k})/2 - i/2 W^k (Z_k - conj Z_{M-k}),  W = exp(-2 pi i / Nx)
//   Z_k = (X_k + conj X_{M-k}) + i conj(W^k) (X_k - conj X_{M-k})   (twice the true Z)
// with X_0 and X_M (both real) packed into slot 0 as (X_0, X_M).
#pragma once
#include "nse2d_common.hpp"

#include <utility>

namespace nse2d {
namespace fft {

struct cplx { double x, y; };
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.x - b.x, a.y - b.y}; }
// Complex products with the fused multiply-adds written out: left to the compiler, a*b - c*d may
// be fused on either product depending on the surrounding code, so two kernels that share these
// functions (one row or two rows per workgroup, single channel or ensemble) could differ in the
// last bit.
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
    return {__builtin_fma(a.x, b.x, -(a.y * b.y)), __builtin_fma(a.x, b.y, a.y * b.x)};
}
__device__ __forceinline__ cplx cmulc(cplx a, cplx b) {  // a conj(b)
    return {__builtin_fma(a.x, b.x, a.y * b.y), __builtin_fma(a.y, b.x, -(a.x * b.y))};
}
__device__ __forceinline__ cplx cconj(cplx a) { return {a.x, -a.y}; }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return {a.x * s, a.y * s}; }

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }

template <int M, int R>
struct Plan {
    static_assert((M & (M - 1)) == 0 && (R == 4 || R == 8 || R == 16) && M >= R * 64, "unsupported row FFT plan");
    static constexpr int T = M / R;
    static constexpr int LOGM = ilog2c(M), LOGR = ilog2c(R);
    static constexpr int NFULL = LOGM / LOGR, REM = LOGM % LOGR;
    static constexpr int NPASS = NFULL + (REM ? 1 : 0);
    static constexpr int LDS = M + (M >> 4);
    static constexpr int radix(int p) { return p < NFULL ? R : (1 << REM); }
    static constexpr int ns(int p) {
        int n = 1;
        for (int i = 0; i < p; ++i) n *= radix(i);
        return n;
    }
    // offset of pass p's twiddle block in the plan table: entries [k][q-1], k < Ns, 1 <= q < Rp
    static constexpr int twoff(int p) {
        int o = 0;
        for (int i = 1; i < p; ++i) o += ns(i) * (radix(i) - 1);
        return o;
    }
    static constexpr int TWSIZE = twoff(NPASS);
};

__device__ __forceinline__ int lds_pad(int p) { return p + (p >> 4); }

// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ cplx mul_mi(cplx a) { return INV ? cplx{-a.y, a.x} : cplx{a.y, -a.x}; }

// a * exp(-+2 pi i e / N) for N in {8, 16}; e is a constant after unrolling, so the table
// reads fold into immediates and the exact cases (multiples of pi/2) skip the multiply.
template <int N, bool INV>
__device__ __forceinline__ cplx twiddle_const(cplx a, int e) {
    // cos and sin of m pi / 8, m = 0..15
    constexpr double C1 = 0.92387953251128675613, S1 = 0.38268343236508977173, R2 = 0.70710678118654752440;
    constexpr double cs[16] = {1, C1, R2, S1, 0, -S1, -R2, -C1, -1, -C1, -R2, -S1, 0, S1, R2, C1};
    constexpr double sn[16] = {0, S1, R2, C1, 1, C1, R2, S1, 0, -S1, -R2, -C1, -1, -C1, -R2, -S1};
    const int m = (e * (16 / N)) & 15;
    if (m == 0) return a;
    if (m == 8) return {-a.x, -a.y};
    if (m == 4) return mul_mi<INV>(a);
    if (m == 12) return mul_mi<!INV>(a);
    const double c = cs[m], s = sn[m];
    // forward: a (c - i s); inverse: a (c + i s)
    return INV ? cplx{__builtin_fma(a.x, c, -(a.y * s)), __builtin_fma(a.y, c, a.x * s)}
               : cplx{__builtin_fma(a.x, c, a.y * s), __builtin_fma(a.y, c, -(a.x * s))};
}

// In-register DFT of N = 1, 2, 4, 8 or 16 values x[0], x[S], .., x[(N-1)S], natural order in and
// out. N = N1 * 4: DFT-4 over n2 of x[n1 + N1 n2], twiddle W_N^{n1 k2}, DFT-N1 over n1, and the
// result X[4 k1 + k2] (found at n1 = k1, slot k2) is moved to natural order.
template <int N, bool INV, int S = 1>
__device__ __forceinline__ void dft(cplx* x) {
    if constexpr (N == 2) {
        const cplx a = x[0], b = x[S];
        x[0] = cadd(a, b);
        x[S] = csub(a, b);
    } else if constexpr (N == 4) {
        const cplx a0 = x[0], a1 = x[S], a2 = x[2 * S], a3 = x[3 * S];
        const cplx t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_mi<INV>(csub(a1, a3));
        x[0] = cadd(t0, t2);
        x[S] = cadd(t1, t3);
        x[2 * S] = csub(t0, t2);
        x[3 * S] = csub(t1, t3);
    } else if constexpr (N == 8 || N == 16) {
        constexpr int N1 = N / 4;
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) dft<4, INV, N1 * S>(x + n1 * S);
#pragma unroll
        for (int n1 = 1; n1 < N1; ++n1)
#pragma unroll
            for (int k2 = 1; k2 < 4; ++k2) x[(n1 + N1 * k2) * S] = twiddle_const<N, INV>(x[(n1 + N1 * k2) * S], n1 * k2);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) dft<N1, INV, S>(x + N1 * k2 * S);
        cplx t[N];
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1)
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) t[4 * k1 + k2] = x[(k1 + N1 * k2) * S];
#pragma unroll
        for (int n = 0; n < N; ++n) x[n * S] = t[n];
    }
}

// Pass p >= 1 (or p = 0 of the inverse) from LDS; writes LDS, or, for the inverse's last pass,
// leaves the scaled outputs in v: v[i Rp + q] is output element tid + i T + q Ns (store_row).
template <int M, int R, bool INV, int p>
__device__ __forceinline__ void lds_pass(cplx (&v)[R], cplx* lds, const cplx* __restrict__ twp, double scale) {
    using P = Plan<M, R>;
    constexpr int T = P::T, Rp = P::radix(p), Ns = P::ns(p), NI = R / Rp, MR = M / Rp;
    constexpr bool LAST = (p == P::NPASS - 1);
    const int tid = threadIdx.x;
    // the pass's twiddles first: they do not depend on the data, so their memory latency runs
    // under the LDS reads and the barrier instead of after them
    cplx wv[NI][Rp > 1 ? Rp - 1 : 1];
    if constexpr (Ns > 1) {
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const cplx* w = twp + P::twoff(p) + ((tid + i * T) & (Ns - 1)) * (Rp - 1);
#pragma unroll
            for (int q = 1; q < Rp; ++q) wv[i][q - 1] = w[q - 1];
        }
    }
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int q = 0; q < Rp; ++q) v[i * Rp + q] = lds[lds_pad(tid + i * T + q * MR)];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int j = tid + i * T, k = j & (Ns - 1);
        if constexpr (Ns > 1) {
#pragma unroll
            for (int q = 1; q < Rp; ++q)
                v[i * Rp + q] = INV ? cmulc(v[i * Rp + q], wv[i][q - 1]) : cmul(v[i * Rp + q], wv[i][q - 1]);
        }
        dft<Rp, INV>(v + i * Rp);
        const int base = (j - k) * Rp + k;
        if (LAST && INV) {
            // last pass: j < Ns, so base = j = tid + i T; the outputs stay in registers
#pragma unroll
            for (int q = 0; q < Rp; ++q) v[i * Rp + q] = cscale(v[i * Rp + q], scale);
        } else {
#pragma unroll
            for (int q = 0; q < Rp; ++q) lds[lds_pad(base + q * Ns)] = v[i * Rp + q];
        }
    }
    if (!(LAST && INV)) __syncthreads();
}

template <int M, int R, bool INV, int... Ps>
__device__ __forceinline__ void lds_passes(cplx (&v)[R], cplx* lds, const cplx* twp, double scale,
                                  std::integer_sequence<int, Ps...>) {
    (lds_pass<M, R, INV, Ps + (INV ? 0 : 1)>(v, lds, twp, scale), ...);
}

// Output element of v[i Rp + q] after the inverse's last pass: tid + i T + q Ns. The stores of
// one (i, q) are contiguous across the workgroup.
template <int M, int R>
struct LastPass {
    using P = Plan<M, R>;
    static constexpr int Rp = P::radix(P::NPASS - 1), Ns = P::ns(P::NPASS - 1), NI = R / Rp;
    __device__ static int pos(int i, int q) { return (int)threadIdx.x + i * P::T + q * Ns; }
};

// The R complex values z_{tid + q T} of one real row (the first pass's inputs).
template <int M, int R>
__device__ __forceinline__ void load_row(const double* __restrict__ src_row, cplx (&v)[R]) {
    const cplx* src = reinterpret_cast<const cplx*>(src_row);
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = src[(int)threadIdx.x + q * Plan<M, R>::T];
}

// One forward row from its loaded inputs: -> packed half spectrum (M complex). Uses all of lds.
template <int M, int R>
__device__ __forceinline__ void forward_row_from(cplx (&v)[R], cplx* __restrict__ dst, cplx* lds,
                                                 const cplx* __restrict__ twp, const cplx* __restrict__ tw) {
    using P = Plan<M, R>;
    constexpr int T = P::T;
    const int tid = threadIdx.x;
    dft<R, false>(v);  // pass 0: Ns = 1, no twiddles
#pragma unroll
    for (int q = 0; q < R; ++q) lds[lds_pad(tid * R + q)] = v[q];
    __syncthreads();
    lds_passes<M, R, false>(v, lds, twp, 1.0, std::make_integer_sequence<int, P::NPASS - 1>{});
    // the split factors of this thread's slots k = tid + j T, loaded together (clamped to
    // k <= M/2): one memory latency instead of one per slot
    constexpr int J = (M / 2 + 1 + T - 1) / T;
    cplx Wk[J];
#pragma unroll
    for (int j = 0; j < J; ++j) Wk[j] = tw[min(tid + j * T, M / 2)];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int k = tid + j * T;
        if (k > M / 2) break;  // only the last slot is partial
        const cplx Zk = lds[lds_pad(k)], Zm = lds[lds_pad((M - k) & (M - 1))];
        if (k == 0) {
            dst[0] = {Zk.x + Zk.y, Zk.x - Zk.y};  // (X[0], X[M])
            continue;
        }
        const cplx E = cscale(cadd(Zk, cconj(Zm)), 0.5);
        const cplx D = csub(Zk, cconj(Zm));
        const cplx O = {0.5 * D.y, -0.5 * D.x};  // -i/2 * D
        const cplx WO = cmul(Wk[j], O);
        dst[k] = cadd(E, WO);
        if (k != M - k) dst[M - k] = cconj(csub(E, WO));
    }
}

// One forward row: Nx = 2M reals -> packed half spectrum (M complex). Uses all of lds.
template <int M, int R>
__device__ __forceinline__ void forward_row(const double* __restrict__ src_row, cplx* __restrict__ dst,
                                   cplx* lds, const cplx* __restrict__ twp, const cplx* __restrict__ tw) {
    cplx v[R];
    load_row<M, R>(src_row, v);
    forward_row_from<M, R>(v, dst, lds, twp, tw);
}

// Forward: real rows of omega ([rows][2M]) -> packed half spectra ([rows][M]).
template <int M, int R>
__global__ void __launch_bounds__(M / R) rows_forward(const double* __restrict__ omega, cplx* __restrict__ hat,
                                                      const cplx* __restrict__ twp, const cplx* __restrict__ tw,
                                                      int row0) {
    __shared__ cplx lds[Plan<M, R>::LDS];
    const size_t row = (size_t)blockIdx.x + row0;
    forward_row<M, R>(omega + row * 2 * M, hat + row * M, lds, twp, tw);
}

// Two rows per workgroup, rows row0 + b and row0 + b + gridDim.x (when < row0 + nrows): the
// second row's loads are issued before the first row's LDS passes, so that they stream while the
// first row is transformed and the chip does not alternate between loading and computing. They
// are unconditional (the first row again when there is no second): under a branch, the compiler
// waited for them before the first row's transform, taking the wait that the path without them
// needs (loads complete in order). The scheduling barriers keep the first row's loads, the second
// row's loads and the first row's arithmetic in that order.
template <int M, int R>
__global__ void __launch_bounds__(M / R) rows_forward2(const double* __restrict__ omega, cplx* __restrict__ hat,
                                                       const cplx* __restrict__ twp, const cplx* __restrict__ tw,
                                                       int row0, int nrows) {
    __shared__ cplx lds[Plan<M, R>::LDS];
    const int ra = blockIdx.x, rb = blockIdx.x + gridDim.x;
    const bool two = rb < nrows;
    cplx va[R], vb[R];
    load_row<M, R>(omega + (size_t)(row0 + ra) * 2 * M, va);
    __builtin_amdgcn_sched_barrier(0);
    load_row<M, R>(omega + (size_t)(row0 + (two ? rb : ra)) * 2 * M, vb);
    __builtin_amdgcn_sched_barrier(0);
    forward_row_from<M, R>(va, hat + (size_t)(row0 + ra) * M, lds, twp, tw);
    if (two) {
        __syncthreads();  // lds is reused
        forward_row_from<M, R>(vb, hat + (size_t)(row0 + rb) * M, lds, twp, tw);
    }
}

// The same for a group of equally sized fields (ensemble members): workgroup b transforms row
// b % rows of field b / rows.
template <int M, int R>
__global__ void __launch_bounds__(M / R) rows_forward_group(const double* const* __restrict__ omegas,
                                                            cplx* const* __restrict__ hats, int rows,
                                                            const cplx* __restrict__ twp,
                                                            const cplx* __restrict__ tw) {
    __shared__ cplx lds[Plan<M, R>::LDS];
    const int m = blockIdx.x / rows;
    const size_t row = blockIdx.x - m * rows;
    forward_row<M, R>(omegas[m] + row * 2 * M, hats[m] + row * M, lds, twp, tw);
}

// One inverse row: packed half spectrum -> the row's M complex values (x_2n, x_2n+1), scaled by
// `scale`, left in v (element LastPass::pos(i, q) in v[i Rp + q]). Uses all of lds.
template <int M, int R>
__device__ __forceinline__ void inverse_row(const cplx* __restrict__ src, cplx (&v)[R], cplx* lds,
                                   const cplx* __restrict__ twp, const cplx* __restrict__ tw, double scale) {
    using P = Plan<M, R>;
    constexpr int T = P::T;
    constexpr int J = (M / 2 + 1 + T - 1) / T;  // slots k = tid + j T of the half spectrum
    const int tid = threadIdx.x;
    // all loads first (clamped to a valid k, k = 0 pairing with itself), so that a row costs one
    // memory latency instead of one per slot
    cplx Xk[J], Xm[J], Wk[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int k = min(tid + j * T, M / 2);
        Xk[j] = src[k];
        Xm[j] = src[k == 0 ? 0 : M - k];
        Wk[j] = tw[k];
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int k = tid + j * T;
        if (k > M / 2) break;  // only the last slot is partial
        if (k == 0) {
            const cplx p = Xk[0];  // (X[0], X[M])
            lds[0] = {p.x + p.y, p.x - p.y};
            continue;
        }
        const cplx A = cadd(Xk[j], cconj(Xm[j])), B = csub(Xk[j], cconj(Xm[j]));
        const cplx WB = cmulc(B, Wk[j]);          // conj(W^k) B
        const cplx iWB = {-WB.y, WB.x};
        lds[lds_pad(k)] = cadd(A, iWB);
        if (k != M - k) lds[lds_pad(M - k)] = cconj(csub(A, iWB));
    }
    __syncthreads();
    lds_passes<M, R, true>(v, lds, twp, scale, std::make_integer_sequence<int, P::NPASS>{});
}

// Optional Thom closure of the wall rows (ops.cu:113-151), fused into the inverse transform:
// the workgroup of row 1 also transforms row 2 (kept in registers, not stored) and writes
// omega's bottom row; the workgroup of row Ny-2 does the same with row Ny-3 for the top row.
struct WallClosure {
    double* omega;           // nullptr: no closure
    const double* dy_edge;
    double psi_bot, psi_top;
    int Ny, free_slip;
    // nonzero: the launch covers the interior rows only (Ny - 2 workgroups) and the workgroups of
    // rows 1 and Ny-2 also write psi's wall rows, psi_bot and psi_top. The transform of a wall row
    // of the solved spectrum, (Nx psi_wall, 0, ..., 0), is that constant exactly (Nx a power of
    // two), so the result is the same as transforming the wall rows.
    int fill_walls;
};

// Inverse of workgroup b of nrow (one field): packed half spectra -> real rows, scaled by `scale`
// (1/Nx), and the wall closure. Workgroup b -> row: 0 -> Ny-2, b in [1, Ny-2] -> b-1, Ny-1 ->
// Ny-1, so that the two closure workgroups (rows 1 and Ny-2, each transforming two rows) are
// dispatched first.
template <int M, int R>
__device__ __forceinline__ void inverse_block(const cplx* __restrict__ hat, double* __restrict__ psi, cplx* lds,
                                     const cplx* __restrict__ twp, const cplx* __restrict__ tw, double scale,
                                     const WallClosure& wc, int b, int nrow) {
    using L = LastPass<M, R>;
    // interior launch: b = 0 -> Ny-2, b in [1, Ny-3] -> b
    const int row = wc.fill_walls ? (b == 0 ? wc.Ny - 2 : b) : ((b == 0) ? nrow - 2 : (b < nrow - 1 ? b - 1 : b));
    cplx v[R];
    inverse_row<M, R>(hat + (size_t)row * M, v, lds, twp, tw, scale);
    cplx* dst = reinterpret_cast<cplx*>(psi + (size_t)row * 2 * M);
#pragma unroll
    for (int i = 0; i < L::NI; ++i)
#pragma unroll
        for (int q = 0; q < L::Rp; ++q) dst[L::pos(i, q)] = v[i * L::Rp + q];
    if (wc.fill_walls && (row == 1 || row == wc.Ny - 2)) {
        const bool bot = (row == 1);
        const double pw = bot ? wc.psi_bot : wc.psi_top;
        cplx* wall = reinterpret_cast<cplx*>(psi + (size_t)(bot ? 0 : wc.Ny - 1) * 2 * M);
#pragma unroll
        for (int i = 0; i < L::NI; ++i)
#pragma unroll
            for (int q = 0; q < L::Rp; ++q) wall[L::pos(i, q)] = {pw, pw};
    }
    if (!wc.omega || (row != 1 && row != wc.Ny - 2)) return;  // uniform over the workgroup
    const bool bottom = (row == 1);
    cplx* wrow = reinterpret_cast<cplx*>(wc.omega + (size_t)(bottom ? 0 : wc.Ny - 1) * 2 * M);
    if (wc.free_slip) {  // ops.cu:108-111
#pragma unroll
        for (int i = 0; i < L::NI; ++i)
#pragma unroll
            for (int q = 0; q < L::Rp; ++q) wrow[L::pos(i, q)] = {0.0, 0.0};
        return;
    }
    __syncthreads();  // lds is reused
    cplx v2[R];
    // Twiddle pointers made opaque to the compiler for the second transform: otherwise it keeps
    // the first transform's twiddle loads live in registers for reuse here (about 64 more VGPRs
    // in every workgroup, one wave per SIMD fewer), although only two workgroups take this path.
    const cplx* twp2 = twp;
    const cplx* tw2 = tw;
    asm volatile("" : "+s"(twp2), "+s"(tw2));
    inverse_row<M, R>(hat + (size_t)(bottom ? 2 : wc.Ny - 3) * M, v2, lds, twp2, tw2, scale);
    const double dy1 = wc.dy_edge[bottom ? 0 : wc.Ny - 2], dy2 = wc.dy_edge[bottom ? 1 : wc.Ny - 3];
    const double A = 2.0 * (1.0 / (dy1 * dy1) + 1.0 / (dy1 * dy2));
    const double B = (-2.0) * dy1 / (dy2 * (dy1 + dy2) * (dy1 + dy2));
    const double pw = bottom ? wc.psi_bot : wc.psi_top;
#pragma unroll
    for (int i = 0; i < L::NI; ++i)
#pragma unroll
        for (int q = 0; q < L::Rp; ++q) {
            const cplx a = v[i * L::Rp + q], c = v2[i * L::Rp + q];
            wrow[L::pos(i, q)] = {-(A * (a.x - pw) + B * (c.x - pw)), -(A * (a.y - pw) + B * (c.y - pw))};
        }
}

template <int M, int R>
__global__ void __launch_bounds__(M / R) rows_inverse(const cplx* __restrict__ hat, double* __restrict__ psi,
                                                      const cplx* __restrict__ twp, const cplx* __restrict__ tw,
                                                      double scale, WallClosure wc) {
    __shared__ cplx lds[Plan<M, R>::LDS];
    inverse_block<M, R>(hat, psi, lds, twp, tw, scale, wc, blockIdx.x, gridDim.x);
}

// The same for a group of equally sized fields: workgroup b handles block b % rows of field
// b / rows; omegas (may be null) gives each field's vorticity for the closure.
template <int M, int R>
__global__ void __launch_bounds__(M / R) rows_inverse_group(const cplx* const* __restrict__ hats,
                                                            double* const* __restrict__ psis,
                                                            double* const* __restrict__ omegas, int rows,
                                                            const cplx* __restrict__ twp,
                                                            const cplx* __restrict__ tw, double scale,
                                                            WallClosure wc) {
    __shared__ cplx lds[Plan<M, R>::LDS];
    const int m = blockIdx.x / rows;
    wc.omega = omegas ? omegas[m] : nullptr;
    inverse_block<M, R>(hats[m], psis[m], lds, twp, tw, scale, wc, blockIdx.x - m * rows, rows);
}

}  // namespace fft
}  // namespace nse2d
