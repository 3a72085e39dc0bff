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

// Sum over the 64 lanes of a wavefront; every lane gets the total.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Block-wide sum of NACC accumulators; result valid in thread 0. blockDim.x must be a multiple
// of 64 and at most 1024. `sm` needs NACC*16 doubles.
template <int NACC>
__device__ inline void block_sum(double (&v)[NACC], double* sm) {
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
__device__ inline void block_max(double (&v)[NACC], double* sm) {
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
