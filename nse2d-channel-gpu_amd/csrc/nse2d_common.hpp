// Shared device/host helpers of the nse2d HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

// HIP runtime failures abort the process, as the reference's CUDA_CHECK does (common.cuh:116-119).
#define NSE2D_HIP_CHECK(e)                                                                      \
    do {                                                                                        \
        hipError_t err__ = (e);                                                                 \
        if (err__ != hipSuccess) {                                                              \
            std::fprintf(stderr, "HIP error %s:%d: %s\n", __FILE__, __LINE__,                   \
                         hipGetErrorString(err__));                                             \
            std::abort();                                                                       \
        }                                                                                       \
    } while (0)

#define NSE2D_LAUNCH_CHECK() NSE2D_HIP_CHECK(hipGetLastError())

namespace nse2d {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr int kWave = 64;

__host__ __device__ inline size_t idx2(int j, int i, int Nx) {
    return (size_t)j * (size_t)Nx + (size_t)i;
}
__host__ __device__ inline int wrap_x(int i, int Nx) { return i < 0 ? i + Nx : (i >= Nx ? i - Nx : i); }

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// MI355X dispatches the workgroups of a 1-D grid round-robin over its 8 XCDs (workgroup b runs on
// XCD b % 8), and each XCD has its own L2. xcd_tile(b, nb) renumbers the nb workgroups so that
// each XCD receives one contiguous range of logical tiles, in order: tiles that share halo rows
// or cache lines then meet in the same L2 instead of being fetched from HBM once per XCD.
// The map is a bijection of [0, nb) for any nb.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int x = b % kXcds, q = nb / kXcds, r = nb % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

// Sum over the 64 lanes of a wavefront; every lane gets the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Whole-wavefront lane shifts by one (GFX9 DPP wave_shr:1 / wave_shl:1, all 64 lanes active):
// wave_shr1 gives lane i the value of lane i-1 and lane 0 its own `edge`; wave_shl1 gives lane i
// the value of lane i+1 and lane 63 its own `edge`.
__device__ __forceinline__ double dpp_shift_(double v, double edge, int ctrl_is_shr) {
    const unsigned long long vb = __double_as_longlong(v), eb = __double_as_longlong(edge);
    int lo, hi;
    if (ctrl_is_shr) {
        lo = __builtin_amdgcn_update_dpp((int)(unsigned)eb, (int)(unsigned)vb, 0x138, 0xf, 0xf, false);
        hi = __builtin_amdgcn_update_dpp((int)(unsigned)(eb >> 32), (int)(unsigned)(vb >> 32), 0x138, 0xf, 0xf, false);
    } else {
        lo = __builtin_amdgcn_update_dpp((int)(unsigned)eb, (int)(unsigned)vb, 0x130, 0xf, 0xf, false);
        hi = __builtin_amdgcn_update_dpp((int)(unsigned)(eb >> 32), (int)(unsigned)(vb >> 32), 0x130, 0xf, 0xf, false);
    }
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double wave_shr1(double v, double edge) { return dpp_shift_(v, edge, 1); }
__device__ __forceinline__ double wave_shl1(double v, double edge) { return dpp_shift_(v, edge, 0); }

// Block-wide sum of NACC accumulators; result valid in thread 0. blockDim.x must be a multiple
// of 64 and at most 1024. `sm` needs NACC*16 doubles.
template <int NACC>
__device__ __forceinline__ void block_sum(double (&v)[NACC], double* sm) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int a = 0; a < NACC; ++a) v[a] = wave_sum(v[a]);
    if (lane == 0)
#pragma unroll
        for (int a = 0; a < NACC; ++a) sm[a * 16 + wid] = v[a];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < NACC; ++a) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += sm[a * 16 + w];
            v[a] = s;
        }
    }
}

template <int NACC>
__device__ __forceinline__ void block_max(double (&v)[NACC], double* sm) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int a = 0; a < NACC; ++a) v[a] = wave_max(v[a]);
    if (lane == 0)
#pragma unroll
        for (int a = 0; a < NACC; ++a) sm[a * 16 + wid] = v[a];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < NACC; ++a) {
            double s = sm[a * 16];
            for (int w = 1; w < nw; ++w) s = fmax(s, sm[a * 16 + w]);
            v[a] = s;
        }
    }
}

// Library-owned device scratch (per device, grown on demand, never shrunk). Used by the
// two-pass deterministic reductions. Not safe for concurrent use from several host threads
// or streams on one device, like the reference's function-static scratch (diag.cu:843-855).
double* scratch_doubles(size_t n);

// Runtime wall-type selection: 0 = no-slip Thom closure, 1 = free-slip (omega_wall = 0).
// The reference selects this at compile time (ops.cu:108, FREE_SLIP_WALLS).
int wall_type();

}  // namespace nse2d
