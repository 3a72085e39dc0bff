// Learned subfilter closure for the coarse channel (no reference counterpart): a CNN of six 3x3
// convolutions (4 -> 64 -> 64 -> 64 -> 64 -> 64 -> 1 channels, GELU after the first five, about
// 150k parameters) on [batch][H][W] images with x periodic and zero padding beyond the walls in y.
// Activations and weights are bf16, accumulation fp32 (MFMA 32x32x16 bf16 on gfx950); the master
// weights and the Adam moments are fp32.
//
// Layouts: activations NHWC [b][y][x][c] bf16 with c fastest; the 4 input channels are padded to
// 32 (zeros). Weights of a 64-output layer: W[t][co][ci] (t = 3*(dy+1) + (dx+1)), and a transposed
// copy WT[t][ci][co] for the input-gradient pass. The last layer (64 -> 1) is a plain fp32 dot
// product per pixel.
//
// Convolution as implicit GEMM: a workgroup of 4 wavefronts owns 64 pixels of one image row and
// all 64 output channels; wavefront w computes the 32x32 block (pixels 32(w&1).., channels
// 32(w>>1)..) with MFMA 32x32x16: lane l supplies A[l%32][8(l/32)..+7] (8 consecutive input
// channels of one pixel, one 16-byte load) and B[8(l/32)..+7][l%32] (8 consecutive input
// channels of one weight row, one 16-byte load); result element (8(r/4) + 4(l/32) + r%4, l%32)
// lands in accumulator r of lane l. The weight gradient is the same product with the pixel
// dimension as K: tiles of 64 pixels are staged transposed ([channel][pixel]) in LDS, with the
// three x-shifted copies of the input rows, so that each lane reads 8 consecutive pixels.
#include "nse2d_common.hpp"
#include "nse2d_cnn.hpp"
#include "../../include/nse2d_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace nse2d {

void set_last_error(const std::string& msg);

namespace cnn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kC = 64;      // hidden channels
constexpr int kCin0 = 32;   // padded input channels of the first layer
constexpr int kInputs = 4;  // real input channels
constexpr int kLayers = 6;

// GELU, tanh form (Hendrycks and Gimpel): 0.5 z (1 + tanh(u)), u = s (z + 0.044715 z^3),
// s = sqrt(2/pi), evaluated as z sigma(2u) with sigma(2u) = 1 / (1 + 2^(k (z + 0.044715 z^3))),
// k = -2 s log2(e) (0.5 (1 + tanh u) = sigma(2u)): the native base-2 exponential and reciprocal
// (v_exp_f32, v_rcp_f32, 1 ulp each), about 1e-7 relative, far below the bf16 resolution of the
// values it feeds, and fewer instructions than the tanh form (the epilogue of every convolution
// applies it to each output). Saturates correctly: the exponential -> inf gives 0, -> 0 gives z.
constexpr float kGeluS = 0.7978845608028654f, kGeluC = 0.044715f;
constexpr float kGeluK = -2.0f * 0.7978845608028654f * 1.4426950408889634f;
__device__ inline float gelu_sigma(float z) {
    const float p = __builtin_fmaf(kGeluC * z * z, z, z);
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(kGeluK * p));
}
__device__ inline float gelu(float z) { return z * gelu_sigma(z); }
// d/dz z sigma(2u) = sigma + 2 z sigma (1 - sigma) u',  u' = s (1 + 3 * 0.044715 z^2)
__device__ inline float gelu_grad(float z) {
    const float sg = gelu_sigma(z);
    const float du2 = __builtin_fmaf(2.0f * kGeluS * 3.0f * kGeluC, z * z, 2.0f * kGeluS);
    return __builtin_fmaf(z * sg * (1.0f - sg), du2, sg);
}

__device__ inline bf16x8 load8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

enum { MODE_FWD_GELU = 0, MODE_FWD_LINEAR = 1, MODE_BWD_DATA = 2 };

// out[b][y][x][co] for co < 64 over the 3x3 neighbourhood of `in` (CIN channels).
//   MODE_FWD_GELU:  z = conv + bias; zout = bf16(z); aout = bf16(gelu(z))
//   MODE_BWD_DATA:  taps mirrored, weights WT; g = conv; aout = bf16(g * gelu'(zprev))
template <int CIN, int MODE>
__global__ void __launch_bounds__(256) conv_mfma_kernel(const bf16* __restrict__ in, const bf16* __restrict__ w,
                                                        const float* __restrict__ bias, bf16* __restrict__ zout,
                                                        bf16* __restrict__ aout, const bf16* __restrict__ zprev,
                                                        int H, int W) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nseg = W >> 6;
    const int seg = blockIdx.x % nseg, y = (blockIdx.x / nseg) % H, b = blockIdx.x / (nseg * H);
    const int ph = wave & 1, ch = wave >> 1;
    const int i = lane & 31, kh = lane >> 5;
    const int x = (seg << 6) + (ph << 5) + i;  // this lane's A-row pixel
    const int co = (ch << 5) + i;              // this lane's B-column channel
    floatx16 acc = {0};
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        int dy = t / 3 - 1, dx = t % 3 - 1;
        if (MODE == MODE_BWD_DATA) { dy = -dy; dx = -dx; }
        const int yy = y + dy;
        if (yy < 0 || yy >= H) continue;  // zero padding beyond the walls (uniform over the workgroup)
        int xx = x + dx;
        xx = xx < 0 ? xx + W : (xx >= W ? xx - W : xx);
        const bf16* arow = in + (((size_t)b * H + yy) * W + xx) * CIN + kh * 8;
        const bf16* brow = w + ((size_t)t * kC + co) * CIN + kh * 8;
#pragma unroll
        for (int kb = 0; kb < CIN / 16; ++kb)
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(load8(arow + kb * 16), load8(brow + kb * 16), acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int px = (seg << 6) + (ph << 5) + ((r >> 2) << 3) + (kh << 2) + (r & 3);
        const int oc = (ch << 5) + i;
        const size_t o = (((size_t)b * H + y) * W + px) * kC + oc;
        if (MODE == MODE_BWD_DATA) {
            aout[o] = (bf16)(acc[r] * gelu_grad((float)zprev[o]));
        } else {
            const float z = acc[r] + bias[oc];
            const bf16 zb = (bf16)z;
            if (zout) zout[o] = zb;  // null for inference
            if (MODE == MODE_FWD_GELU) aout[o] = (bf16)gelu((float)zb);
        }
    }
}

// The same convolution with its operands staged in LDS (the kernel used by the network): a
// workgroup of 4 wavefronts owns 4 image rows x 64 pixels x all 64 output channels; wavefront w
// computes row y0 + w as a 2 x 2 block of 32x32 MFMA tiles (pixels 32 pb.., channels 32 cb..),
// so each pair of A and B fragments feeds four MFMAs. LDS holds the input halo tile (rows y0-1 ..
// y0+4, pixels x0-1 .. x0+64, all CIN channels; rows beyond the walls are zero) and the weights of
// one tap, double-buffered: the next tap's weights are fetched from global memory while the
// current tap computes. Rows of both LDS arrays are padded by 8 channels (16 bytes) so that the
// 16-byte fragment reads of consecutive lanes fall on different banks. zout may be null (no
// pre-activations kept: inference).
constexpr int kTileRows = 4;
template <int CIN>
struct ConvLds {
    static constexpr int CP = CIN + 8;  // padded channel stride (bf16 elements)
    bf16 in[kTileRows + 2][66][CP];
    bf16 w[2][kC][CP];
};

template <int CIN, int MODE>
__global__ void __launch_bounds__(256) conv_lds_kernel(const bf16* __restrict__ in, const bf16* __restrict__ w,
                                                       const float* __restrict__ bias, bf16* __restrict__ zout,
                                                       bf16* __restrict__ aout, const bf16* __restrict__ zprev,
                                                       int H, int W) {
    using S = ConvLds<CIN>;
    constexpr int C8 = CIN / 8;
    __shared__ alignas(16) S sh;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nseg = W >> 6, nyt = (H + kTileRows - 1) / kTileRows;
    const int seg = blockIdx.x % nseg, yt = (blockIdx.x / nseg) % nyt, b = blockIdx.x / (nseg * nyt);
    const int x0 = seg << 6, y0 = yt * kTileRows;
    const int i = lane & 31, kh = lane >> 5;
    // weights of tap t into registers (16 bytes per chunk, 64 x CIN / 8 chunks)
    constexpr int WCH = kC * C8 / 256;  // chunks per thread (1 or 2)
    bf16x8 wreg[WCH];
    auto fetch_w = [&](int t) {
#pragma unroll
        for (int q = 0; q < WCH; ++q) {
            const int e = tid + q * 256, n = e / C8, c8 = (e % C8) * 8;
            wreg[q] = load8(w + ((size_t)t * kC + n) * CIN + c8);
        }
    };
    auto store_w = [&](int buf) {
#pragma unroll
        for (int q = 0; q < WCH; ++q) {
            const int e = tid + q * 256, n = e / C8, c8 = (e % C8) * 8;
            *reinterpret_cast<bf16x8*>(&sh.w[buf][n][c8]) = wreg[q];
        }
    };
    fetch_w(0);
    // input halo tile
    for (int e = tid; e < (kTileRows + 2) * 66 * C8; e += 256) {
        const int r = e / (66 * C8), rem = e % (66 * C8), q = rem / C8, c8 = (rem % C8) * 8;
        const int yy = y0 - 1 + r;
        int xx = x0 - 1 + q;
        xx = xx < 0 ? xx + W : (xx >= W ? xx - W : xx);
        bf16x8 v;
        if (yy >= 0 && yy < H) v = load8(in + (((size_t)b * H + yy) * W + xx) * CIN + c8);
        else
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = (bf16)0.0f;
        *reinterpret_cast<bf16x8*>(&sh.in[r][q][c8]) = v;
    }
    store_w(0);
    __syncthreads();
    floatx16 acc[2][2];
#pragma unroll
    for (int pb = 0; pb < 2; ++pb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) acc[pb][cb] = floatx16{0};
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        if (t + 1 < 9) fetch_w(t + 1);
        int dy = t / 3 - 1, dx = t % 3 - 1;
        if (MODE == MODE_BWD_DATA) { dy = -dy; dx = -dx; }
        const int lr = wave + 1 + dy;
        const int buf = t & 1;
#pragma unroll
        for (int kb = 0; kb < CIN / 16; ++kb) {
            const int co = kb * 16 + kh * 8;
            const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(&sh.in[lr][i + 1 + dx][co]);
            const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(&sh.in[lr][i + 33 + dx][co]);
            const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(&sh.w[buf][i][co]);
            const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(&sh.w[buf][i + 32][co]);
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc[1][1], 0, 0, 0);
        }
        if (t + 1 < 9) {
            store_w(buf ^ 1);  // the other buffer: last read in tap t-1, before the barrier below
            __syncthreads();
        }
    }
    const int y = y0 + wave;
    if (y >= H) return;
#pragma unroll
    for (int pb = 0; pb < 2; ++pb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int px = x0 + (pb << 5) + ((r >> 2) << 3) + (kh << 2) + (r & 3);
                const int oc = (cb << 5) + i;
                const size_t o = (((size_t)b * H + y) * W + px) * kC + oc;
                if (MODE == MODE_BWD_DATA) {
                    aout[o] = (bf16)(acc[pb][cb][r] * gelu_grad((float)zprev[o]));
                } else {
                    const float z = acc[pb][cb][r] + bias[oc];
                    const bf16 zb = (bf16)z;
                    if (zout) zout[o] = zb;
                    if (MODE == MODE_FWD_GELU) aout[o] = (bf16)gelu((float)zb);
                }
            }
}

// The persistent form: every weight of the layer (nine taps) is staged in LDS once per workgroup,
// which then loops over row tiles (tile = blockIdx.x, + gridDim.x, ...). Between tiles the next
// tile's inputs are fetched into registers while the current tile's outputs are written, so the
// loop has no per-tap barrier and no exposed global latency inside the MFMA sequence. The
// products and their order are those of conv_lds_kernel.
template <int CIN>
struct ConvLdsAll {
    static constexpr int CP = CIN + 8;
    // The tile's outputs (z and a) are staged in the input tile's space once the products are
    // done, [row][pixel][64 channels] with the 16-byte chunk c of pixel p at chunk position
    // c ^ (p & 7), and then stored with 16 contiguous bytes per lane (the accumulator layout gives
    // a lane four channels of a pixel, so direct stores wrote 8 bytes per lane, 128 bytes apart).
    // NOUT buffers: two at CIN = 64; one at CIN = 32, where two workgroups per CU leave room for
    // one, and z and a go out in two rounds.
    static constexpr int NOUT = CIN == kC ? 2 : 1;
    bf16 w[9][kC][CP];
    union {
        bf16 in[kTileRows + 2][66][CP];
        bf16 out[NOUT][kTileRows][64][kC];
    };
};

// 512 threads: wavefront w computes row y0 + (w >> 1), output channels 32 (w & 1) .. +31, as two
// 32x32 MFMA tiles (pixels 0..31 and 32..63 of the segment), so that two wavefronts share each
// SIMD and one's epilogue overlaps the other's MFMAs. The weights are the MFMA's first operand
// and the pixels its second, so a lane's accumulator holds four consecutive output channels of
// one pixel in each group of four registers: the epilogue reads zprev and writes zout / aout
// four channels (8 bytes) at a time.
constexpr int kAllThreads = 512;
template <int CIN, int MODE>
__global__ void __launch_bounds__(kAllThreads) conv_lds_all_kernel(const bf16* __restrict__ in,
                                                                   const bf16* __restrict__ w,
                                                                   const float* __restrict__ bias,
                                                                   bf16* __restrict__ zout, bf16* __restrict__ aout,
                                                                   const bf16* __restrict__ zprev, int H, int W,
                                                                   int ntiles) {
    using S = ConvLdsAll<CIN>;
    constexpr int C8 = CIN / 8;
    constexpr int NT = kAllThreads;
    constexpr int NIN = ((kTileRows + 2) * 66 * C8 + NT - 1) / NT;  // input chunks per thread
    __shared__ alignas(16) S sh;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // the two wavefronts of a row sit on different SIMDs, so that the last row tile of an odd
    // height (rows past the field skip their products) takes half the time of a full one
    const int wrow = wave >> 1, cb = wave & 1;
    const int nseg = W >> 6, nyt = (H + kTileRows - 1) / kTileRows;
    const int i = lane & 31, kh = lane >> 5;
    for (int e = tid; e < 9 * kC * C8; e += NT) {
        const int t = e / (kC * C8), rem = e % (kC * C8), n = rem / C8, c8 = (rem % C8) * 8;
        *reinterpret_cast<bf16x8*>(&sh.w[t][n][c8]) = load8(w + ((size_t)t * kC + n) * CIN + c8);
    }
    bf16x8 pre[NIN];
    auto fetch = [&](int tile) {
        const int seg = tile % nseg, yt = (tile / nseg) % nyt, b = tile / (nseg * nyt);
        const int x0 = seg << 6, y0 = yt * kTileRows;
#pragma unroll
        for (int q = 0; q < NIN; ++q) {
            const int e = tid + q * NT;
            if (e < (kTileRows + 2) * 66 * C8) {
                const int r = e / (66 * C8), rem = e % (66 * C8), px = rem / C8, c8 = (rem % C8) * 8;
                const int yy = y0 - 1 + r;
                int xx = x0 - 1 + px;
                xx = xx < 0 ? xx + W : (xx >= W ? xx - W : xx);
                if (yy >= 0 && yy < H) pre[q] = load8(in + (((size_t)b * H + yy) * W + xx) * CIN + c8);
                else
#pragma unroll
                    for (int k = 0; k < 8; ++k) pre[q][k] = (bf16)0.0f;
            }
        }
    };
    auto deposit = [&]() {
#pragma unroll
        for (int q = 0; q < NIN; ++q) {
            const int e = tid + q * NT;
            if (e < (kTileRows + 2) * 66 * C8) {
                const int r = e / (66 * C8), rem = e % (66 * C8), px = rem / C8, c8 = (rem % C8) * 8;
                *reinterpret_cast<bf16x8*>(&sh.in[r][px][c8]) = pre[q];
            }
        }
    };
    // the bias of the lane's output channels (forward modes), loaded once: inside the tile loop its
    // loads were issued after the next tile's prefetch and waited for it (loads complete in order)
    float bv[4][4];
    if (MODE != MODE_BWD_DATA) {
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[g][j] = bias[(cb << 5) + (g << 3) + (kh << 2) + j];
    }
    // DEFER (forward layers with 64 input channels): a tile's outputs are read back from the LDS
    // staging into registers and stored during the next tile's products, one 16-byte store every
    // four MFMA steps, instead of in a burst between tiles (every workgroup runs the same phases
    // at the same time). Measured at batch 32: 467 -> 452-456 us per layer, inference 0.1234 ->
    // 0.1211 ms; the input gradient (one output) measured the same with it (512 us), and at 32
    // input channels two workgroups per CU need the registers.
    constexpr bool DEFER = CIN == kC && MODE == MODE_FWD_GELU;
    const bool has_z = MODE != MODE_BWD_DATA && zout, has_a = MODE != MODE_FWD_LINEAR;
    bf16x8 pz[DEFER ? kTileRows : 1], pa[DEFER ? kTileRows : 1];
    size_t pbase = 0;
    int prow = 0;  // rows of the pending tile inside the field; 0: nothing pending
    // pending store k: row k / 2, z (even k) or a (odd k)
    auto put = [&](int k) {
        const int r = k >> 1;
        if (r < prow) {
            const size_t o = pbase + (size_t)r * W * kC;
            if (k & 1) {
                if (has_a) *reinterpret_cast<bf16x8*>(aout + o) = pa[DEFER ? r : 0];
            } else {
                if (has_z) *reinterpret_cast<bf16x8*>(zout + o) = pz[DEFER ? r : 0];
            }
        }
    };
    if (blockIdx.x < ntiles) fetch(blockIdx.x);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        deposit();
        __syncthreads();
        const int seg = tile % nseg, yt = (tile / nseg) % nyt, b = tile / (nseg * nyt);
        const int x0 = seg << 6, y = yt * kTileRows + wrow;
        // the input gradient's gelu'(z) operands, loaded before the products (issued after the
        // prefetch, their wait would include it)
        bf16x4 zp[2][4];
        if (MODE == MODE_BWD_DATA && y < H) {
#pragma unroll
            for (int pb = 0; pb < 2; ++pb) {
                const size_t pix = ((size_t)b * H + y) * W + x0 + (pb << 5) + i;
#pragma unroll
                for (int g = 0; g < 4; ++g)
                    zp[pb][g] = *reinterpret_cast<const bf16x4*>(zprev + pix * kC + (cb << 5) + (g << 3) + (kh << 2));
            }
        }
        floatx16 acc[2];
        acc[0] = floatx16{0};
        acc[1] = floatx16{0};
        const bool row_in = y < H;  // uniform over the wavefront
        if (row_in) {
            // products over (tap t, channel block kb), it = t KB + kb, with the three fragments of
            // step it + PF read from LDS before the MFMAs of step it are issued (a ring of PF + 1
            // fragment sets; the scheduling barriers keep that order): each step's reads then
            // have PF steps' MFMAs to arrive in, instead of one LDS latency between steps (PF = 1
            // at CIN = 32, where two workgroups per CU need at most 128 VGPRs)
            constexpr int KB = CIN / 16, NIT = 9 * KB, PF = CIN == 64 ? 2 : 1;
            auto ld = [&](int it, bf16x8& a0, bf16x8& a1, bf16x8& bw) {
                const int t = it / KB, kb = it % KB;
                int dy = t / 3 - 1, dx = t % 3 - 1;
                if (MODE == MODE_BWD_DATA) { dy = -dy; dx = -dx; }
                const int lr = wrow + 1 + dy, co = kb * 16 + kh * 8;
                a0 = *reinterpret_cast<const bf16x8*>(&sh.in[lr][i + 1 + dx][co]);
                a1 = *reinterpret_cast<const bf16x8*>(&sh.in[lr][i + 33 + dx][co]);
                bw = *reinterpret_cast<const bf16x8*>(&sh.w[t][(cb << 5) + i][co]);
            };
            bf16x8 fa0[PF + 1], fa1[PF + 1], fb[PF + 1];
#pragma unroll
            for (int it = 0; it < PF; ++it) ld(it, fa0[it], fa1[it], fb[it]);
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                if (it + PF < NIT) ld(it + PF, fa0[(it + PF) % (PF + 1)], fa1[(it + PF) % (PF + 1)], fb[(it + PF) % (PF + 1)]);
                __builtin_amdgcn_sched_barrier(0);
                const int s = it % (PF + 1);
                acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fb[s], fa0[s], acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fb[s], fa1[s], acc[1], 0, 0, 0);
                if (DEFER && (it & 3) == 3 && (it >> 2) < 2 * kTileRows) put(it >> 2);
                __builtin_amdgcn_sched_barrier(0);
            }
        } else if (DEFER) {
#pragma unroll
            for (int k = 0; k < 2 * kTileRows; ++k) put(k);
        }
        prow = 0;
        const int nxt = tile + gridDim.x;
        if (nxt < ntiles) fetch(nxt);  // in flight during the stores below
        {
            // acc[pb][4 g + j]: pixel x0 + 32 pb + i, channel 32 cb + 8 g + 4 kh + j
            constexpr int NOUT = S::NOUT;
            // part 0 is z, part 1 is a; NOUT = 2 stages both at once, NOUT = 1 one per round
            const int rounds = NOUT == 2 ? 1 : (has_z ? 1 : 0) + (has_a ? 1 : 0);
            const int y0 = yt * kTileRows;
            for (int rd = 0; rd < rounds; ++rd) {
                // rd = 0: every wavefront is done with sh.in, whose space the outputs take;
                // rd = 1: the previous round's stores have read the buffer
                __syncthreads();
                if (y < H) {
#pragma unroll
                    for (int pb = 0; pb < 2; ++pb)
#pragma unroll
                        for (int g = 0; g < 4; ++g) {
                            bf16x4 zv, ov;
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                if (MODE == MODE_BWD_DATA) {
                                    ov[j] = (bf16)(acc[pb][4 * g + j] * gelu_grad((float)zp[pb][g][j]));
                                } else {
                                    zv[j] = (bf16)(acc[pb][4 * g + j] + bv[g][j]);
                                    if (MODE == MODE_FWD_GELU) ov[j] = (bf16)gelu((float)zv[j]);
                                }
                            }
                            const int px = (pb << 5) + i, c = (cb << 2) + g;
                            const int l = (((c ^ (px & 7)) << 3) + (kh << 2));
                            if (NOUT == 2) {
                                if (has_z) *reinterpret_cast<bf16x4*>(&sh.out[0][wrow][px][l]) = zv;
                                if (has_a) *reinterpret_cast<bf16x4*>(&sh.out[NOUT - 1][wrow][px][l]) = ov;
                            } else {
                                const bool z_now = has_z && rd == 0;
                                *reinterpret_cast<bf16x4*>(&sh.out[0][wrow][px][l]) = z_now ? zv : ov;
                            }
                        }
                }
                __syncthreads();
                if constexpr (DEFER) {
                    // the thread's chunks (pixel px, channels 8c..8c+7 of every row), stored
                    // during the next tile's products (or after the loop)
                    const int px = (tid >> 3) & 63, c = tid & 7;
                    const int l = (c ^ (px & 7)) << 3;
#pragma unroll
                    for (int r = 0; r < kTileRows; ++r) {
                        if (has_z) pz[r] = *reinterpret_cast<const bf16x8*>(&sh.out[0][r][px][l]);
                        if (has_a) pa[r] = *reinterpret_cast<const bf16x8*>(&sh.out[NOUT - 1][r][px][l]);
                    }
                    pbase = (((size_t)b * H + y0) * W + x0 + px) * kC + (c << 3);
                    prow = min(kTileRows, H - y0);
                    continue;
                }
                const bool st_z = has_z && (NOUT == 2 || rd == 0);
                const bool st_a = has_a && (NOUT == 2 || rd == rounds - 1);
                bf16* zdst = st_z ? zout : nullptr;
                for (int q = tid; q < kTileRows * 64 * 8; q += NT) {
                    const int r = q >> 9, px = (q >> 3) & 63, c = q & 7;
                    if (y0 + r >= H) continue;
                    const size_t o = (((size_t)b * H + y0 + r) * W + x0 + px) * kC + (c << 3);
                    const int l = (c ^ (px & 7)) << 3;
                    if (zdst) *reinterpret_cast<bf16x8*>(zdst + o) = *reinterpret_cast<const bf16x8*>(&sh.out[0][r][px][l]);
                    if (st_a)
                        *reinterpret_cast<bf16x8*>(aout + o) =
                            *reinterpret_cast<const bf16x8*>(&sh.out[NOUT == 2 ? 1 : 0][r][px][l]);
                }
            }
        }
        __syncthreads();  // every wavefront is done with sh.in before the next deposit
    }
    if (DEFER) {
#pragma unroll
        for (int k = 0; k < 2 * kTileRows; ++k) put(k);
    }
}

// Last layer (64 -> 1): out(y, x) = b + sum over the 3x3 taps (dy, dx) and the channels c of
// a(y+dy, x+dx, c) w[3 (dy+1) + dx+1][c]; rows outside the field contribute nothing, x wraps.
// A workgroup takes a segment of 64 pixels of one row. For each of the three input rows and each
// of the 66 pixels the segment reaches, eight lanes (eight channels each, one 16-byte load) form
// the row's three partial sums s[dy][dx][px], reduced over the lanes and kept in LDS; the 64
// outputs then add nine of them. Each input pixel is read once per output row that needs it,
// instead of once per tap. closure != 0: the closure term of a coarse channel is written instead,
// (out) * scale + shift on rows 1..H-2 and 0 on the wall rows.
constexpr int kLastThreads = 256;
__global__ void __launch_bounds__(kLastThreads) last_fwd_rows_kernel(const bf16* __restrict__ a,
                                                                    const float* __restrict__ w,
                                                                    const float* __restrict__ bias,
                                                                    float* __restrict__ out, int H, int W,
                                                                    int closure, float scale, float shift) {
    __shared__ float sp[3][3][68];
    const int nseg = W >> 6;
    const int seg = blockIdx.x % nseg, y = (blockIdx.x / nseg) % H, b = blockIdx.x / (nseg * H);
    const int x0 = seg << 6;
    const int tid = threadIdx.x, sub = tid & 7, grp = tid >> 3;  // 32 octets
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int yy = y + r - 1;
        float wr[3][8];
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int q = 0; q < 8; ++q) wr[k][q] = w[(r * 3 + k) * kC + sub * 8 + q];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int px = grp + 32 * i;  // input pixel x0 - 1 + px
            float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
            if (px < 66 && yy >= 0 && yy < H) {
                int xx = x0 - 1 + px;
                xx = xx < 0 ? xx + W : (xx >= W ? xx - W : xx);
                const bf16x8 v = load8(a + (((size_t)b * H + yy) * W + xx) * kC + sub * 8);
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const float f = (float)v[q];
                    s0 += f * wr[0][q];
                    s1 += f * wr[1][q];
                    s2 += f * wr[2][q];
                }
            }
#pragma unroll
            for (int m = 1; m < 8; m <<= 1) {
                s0 += __shfl_xor(s0, m, 64);
                s1 += __shfl_xor(s1, m, 64);
                s2 += __shfl_xor(s2, m, 64);
            }
            if (sub == 0 && px < 66) {
                sp[r][0][px] = s0;
                sp[r][1][px] = s1;
                sp[r][2][px] = s2;
            }
        }
    }
    __syncthreads();
    if (tid < 64) {
        // output pixel x0 + tid takes tap (dy, dx) from input pixel x0 + tid + dx, i.e. px = tid + dx + 1
        float o = 0.0f;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int k = 0; k < 3; ++k) o += sp[r][k][tid + k];
        o += bias[0];
        if (closure) o = (y > 0 && y < H - 1) ? o * scale + shift : 0.0f;
        out[((size_t)b * H + y) * W + x0 + tid] = o;
    }
}

// Two bf16 values as one 32-bit word (lo in the low half).
__device__ inline unsigned pack_bf16x2(bf16 lo, bf16 hi) {
    return (unsigned)__builtin_bit_cast(unsigned short, lo) | ((unsigned)__builtin_bit_cast(unsigned short, hi) << 16);
}

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Elements s .. s+7 of the 16 bf16 values c0 (0..7), c1 (8..15), for s = 0, 1, 2.
template <int S>
__device__ inline bf16x8 window8(const bf16x8& c0, const bf16x8& c1) {
    if constexpr (S == 0) return c0;
    const u32x4 d = __builtin_bit_cast(u32x4, c0), e = __builtin_bit_cast(u32x4, c1);
    u32x4 r;
    if constexpr (S == 1) {
        r[0] = __builtin_amdgcn_alignbit(d[1], d[0], 16);
        r[1] = __builtin_amdgcn_alignbit(d[2], d[1], 16);
        r[2] = __builtin_amdgcn_alignbit(d[3], d[2], 16);
        r[3] = __builtin_amdgcn_alignbit(e[0], d[3], 16);
    } else {
        r[0] = d[1];
        r[1] = d[2];
        r[2] = d[3];
        r[3] = e[0];
    }
    return __builtin_bit_cast(bf16x8, r);
}

// Partial weight gradient of a 64-output layer over a strip of row segments:
//   dW[t][co][ci] += sum_pixels dZ[p][co] * X[p + off_t][ci]
// Workgroup: (CIN/32)*2 wavefronts, wavefront w owns co block w/(CIN/32), ci block w%(CIN/32), all 9
// taps (9 accumulators of 16). LDS holds dZ^T [64 ch][64 px] and the input rows y-1..y+1
// transposed, X^T[dy][ci][j] for the pixels x0 - 1 + j, j = 0..65. Both are staged a pixel pair at
// a time (two 16-byte loads, eight 4-byte LDS writes). The operand of tap dx starts at j = p + dx + 1:
// it is cut from two aligned 16-byte LDS reads by funnel shifts, one pair of reads for the three dx.
// The bias gradient (column sums of dZ) is summed from the staged dZ values on the way; the block's
// 64 sums go to bias_partial[blockIdx.x].
template <int CIN>
__global__ void __launch_bounds__(CIN * 4) __attribute__((amdgpu_waves_per_eu(CIN == 64 ? 2 : 1, 2))) conv_wgrad_kernel(const bf16* __restrict__ x, const bf16* __restrict__ dz,
                                                             float* __restrict__ partial,
                                                             float* __restrict__ bias_partial, int H, int W, int B,
                                                             int segs_per_block) {
    constexpr int NT = CIN * 4;  // threads
    constexpr int LP = 72;       // row length: 66 pixels used, rows 16-byte aligned
    __shared__ alignas(16) bf16 dzT[kC][LP];
    __shared__ alignas(16) bf16 xT[3][CIN][LP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int NCI = CIN / 32;
    const int cb = wave / NCI, ib = wave % NCI;
    const int i = lane & 31, kh = lane >> 5;
    floatx16 acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = floatx16{0};
    const int nseg = W >> 6, total = B * H * nseg;
    const int s0 = blockIdx.x * segs_per_block, s1 = min(total, s0 + segs_per_block);
    // Segment sg's global operands: a pixel pair x eight channels of dZ per element e < 32 (kC/8)
    // and of the three input rows per element e < 3 * 33 (CIN/8), the pixel pair fastest across
    // lanes: the transposed 4-byte LDS writes of a wavefront then fill consecutive words of one
    // channel row (with the channel octet fastest they fell 8 rows apart, on the same two banks).
    // They are fetched into registers one segment ahead: the next segment's loads are issued
    // before the current segment's products, so their latency runs under the MFMAs.
    constexpr int NDZ = (32 * (kC / 8) + NT - 1) / NT, NX = (3 * 33 * (CIN / 8) + NT - 1) / NT;
    static_assert((32 * (kC / 8)) % NT == 0, "every thread stages NDZ elements of dZ");
    bf16x8 pdz[NDZ][2], px[NX][2];
    float bsum[NDZ][8];  // element e = tid + n NT: channels (e / 32) * 8 + q
#pragma unroll
    for (int n = 0; n < NDZ; ++n)
#pragma unroll
        for (int q = 0; q < 8; ++q) bsum[n][q] = 0.0f;
    auto fetch = [&](int sg) {
        const int seg = sg % nseg, y = (sg / nseg) % H, b = sg / (nseg * H);
        const int x0 = seg << 6;
#pragma unroll
        for (int n = 0; n < NDZ; ++n) {
            const int e = tid + n * NT;
            if (e < 32 * (kC / 8)) {
                const int m = e % 32, c8 = (e / 32) * 8;
                const bf16* src = dz + (((size_t)b * H + y) * W + x0 + 2 * m) * kC + c8;
                pdz[n][0] = load8(src);
                pdz[n][1] = load8(src + kC);
            }
        }
        // unconditional loads (element index and row clamped into range; rows beyond the walls
        // are zeroed when staged): under a branch, the zero fill of the operand registers made
        // the compiler wait for every load issued before it, one memory latency per element
#pragma unroll
        for (int n = 0; n < NX; ++n) {
            const int e = min(tid + n * NT, 3 * 33 * (CIN / 8) - 1);
            const int dyi = e / (33 * (CIN / 8)), rem = e % (33 * (CIN / 8));
            const int m = rem % 33, c8 = (rem / 33) * 8;
            const int yy = min(max(y + dyi - 1, 0), H - 1);
            int xa = x0 - 1 + 2 * m, xb = x0 + 2 * m;
            xa = xa < 0 ? xa + W : xa;
            xb = xb >= W ? xb - W : xb;
            const bf16* row = x + ((size_t)b * H + yy) * W * CIN + c8;
            px[n][0] = load8(row + (size_t)xa * CIN);
            px[n][1] = load8(row + (size_t)xb * CIN);
        }
    };
    if (s0 < s1) fetch(s0);
    for (int sg = s0; sg < s1; ++sg) {
        // stage dZ^T (pixel pairs (2m, 2m+1) x 8 channels) and the input rows (pixel pairs
        // (j, j+1) = (2m, 2m+1), m = 0..32, of rows y-1..y+1), transposed
#pragma unroll
        for (int n = 0; n < NDZ; ++n) {
            const int e = tid + n * NT;
            if (e < 32 * (kC / 8)) {
                const int m = e % 32, c8 = (e / 32) * 8;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    *reinterpret_cast<unsigned*>(&dzT[c8 + q][2 * m]) = pack_bf16x2(pdz[n][0][q], pdz[n][1][q]);
                    bsum[n][q] += (float)pdz[n][0][q] + (float)pdz[n][1][q];
                }
            }
        }
        const int ycur = (sg / nseg) % H;
#pragma unroll
        for (int n = 0; n < NX; ++n) {
            const int e = tid + n * NT;
            if (e < 3 * 33 * (CIN / 8)) {
                const int dyi = e / (33 * (CIN / 8)), rem = e % (33 * (CIN / 8));
                const int m = rem % 33, c8 = (rem / 33) * 8;
                const int yy = ycur + dyi - 1;
                const bool in = yy >= 0 && yy < H;
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    *reinterpret_cast<unsigned*>(&xT[dyi][c8 + q][2 * m]) = in ? pack_bf16x2(px[n][0][q], px[n][1][q]) : 0u;
            }
        }
        __syncthreads();
        if (sg + 1 < s1) fetch(sg + 1);  // in flight during the products below
#pragma unroll
        for (int dyi = 0; dyi < 3; ++dyi)
#pragma unroll
            for (int pb = 0; pb < 4; ++pb) {
                const int p0 = pb * 16 + kh * 8;
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(&dzT[(cb << 5) + i][p0]);
                const bf16x8 c0 = *reinterpret_cast<const bf16x8*>(&xT[dyi][(ib << 5) + i][p0]);
                const bf16x8 c1 = *reinterpret_cast<const bf16x8*>(&xT[dyi][(ib << 5) + i][p0 + 8]);
                acc[dyi * 3 + 0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, window8<0>(c0, c1), acc[dyi * 3 + 0], 0, 0, 0);
                acc[dyi * 3 + 1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, window8<1>(c0, c1), acc[dyi * 3 + 1], 0, 0, 0);
                acc[dyi * 3 + 2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, window8<2>(c0, c1), acc[dyi * 3 + 2], 0, 0, 0);
            }
        __syncthreads();
    }
    {   // bias: the 32 pixel-pair lanes of each channel, added in a fixed order
        static_assert(sizeof(xT) >= sizeof(float) * 32 * kC, "bias sums fit in the input tile");
        float* bs = reinterpret_cast<float*>(&xT[0][0][0]);  // [32][64], free after the last barrier
#pragma unroll
        for (int n = 0; n < NDZ; ++n) {
            const int e = tid + n * NT;
#pragma unroll
            for (int q = 0; q < 8; ++q) bs[(e % 32) * kC + (e / 32) * 8 + q] = bsum[n][q];
        }
        __syncthreads();
        if (tid < kC) {
            float t = 0.0f;
            for (int r = 0; r < 32; ++r) t += bs[r * kC + tid];
            bias_partial[(size_t)blockIdx.x * kC + tid] = t;
        }
    }
    float* out = partial + (size_t)blockIdx.x * 9 * kC * CIN;
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int co = (cb << 5) + ((r >> 2) << 3) + (kh << 2) + (r & 3);
            const int ci = (ib << 5) + i;
            out[((size_t)t * kC + co) * CIN + ci] = acc[t][r];
        }
}

// Ordered sum of nb partial blocks of n floats: out[j] = sum_k partial[k][j]. A workgroup of 256
// threads takes cols consecutive columns (cols divides 256); its 256 / cols groups of threads add
// interleaved blocks k = g, g + groups, ... and the group sums are added in a fixed order, so the
// result does not depend on timing.
constexpr int kSumThreads = 256;
constexpr int kLastWgradBlocks = 1024;  // most workgroups of the last layer's weight gradient
__global__ void __launch_bounds__(kSumThreads) sum_partials_kernel(const float* __restrict__ partial, int nb, int n,
                                                                   int cols, float* __restrict__ out) {
    __shared__ float sm[kSumThreads];
    const int groups = kSumThreads / cols;
    const int c = threadIdx.x % cols, g = threadIdx.x / cols;
    const int j = blockIdx.x * cols + c;
    float s = 0.0f;
    if (j < n) {
        int k = g;
        float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
        for (; k + 3 * groups < nb; k += 4 * groups) {
            s += partial[(size_t)k * n + j];
            s1 += partial[(size_t)(k + groups) * n + j];
            s2 += partial[(size_t)(k + 2 * groups) * n + j];
            s3 += partial[(size_t)(k + 3 * groups) * n + j];
        }
        for (; k < nb; k += groups) s += partial[(size_t)k * n + j];
        s = (s + s1) + (s2 + s3);
    }
    sm[threadIdx.x] = s;
    __syncthreads();
    if (g == 0 && j < n) {
        float t = 0.0f;
        for (int q = 0; q < groups; ++q) t += sm[q * cols + c];
        out[j] = t;
    }
}

// Sum over the eight octets of a wavefront (lanes l, l^8, l^16, ... hold the same channels).
__device__ inline float octet_sum(float v) {
    v += __shfl_xor(v, 8, 64);
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// The output gradient g of the last layer around a row segment: gs[dyi][j] = g(y + 1 - dyi, x0 - 1 + j)
// for j = 0..65 (x wraps; rows outside the field give 0). Tap (dy, dx) = (dyi - 1, dxi - 1) of the
// pixel x0 + px takes g(y - dy, x0 + px - dx) = gs[dyi][px + 2 - dxi]. Threads 0..197 load one value.
__device__ inline float load_g_ring(const float* __restrict__ g, int b, int y, int x0, int H, int W, int tid) {
    if (tid >= 198) return 0.0f;
    const int dyi = tid / 66, j = tid % 66;
    const int yy = y + 1 - dyi;
    if (yy < 0 || yy >= H) return 0.0f;
    int xx = x0 - 1 + j;
    xx = xx < 0 ? xx + W : (xx >= W ? xx - W : xx);
    return g[((size_t)b * H + yy) * W + xx];
}

// Input gradient of the last layer, through the GELU of the layer before:
//   dz5[p][ci] = gelu'(z5[p][ci]) * sum_t g[p - off_t] w[t][ci]
// One workgroup per row segment of 64 pixels at a time (grid-stride over the segments): the
// segment's g ring goes to LDS, eight lanes take a pixel's 64 channels (16-byte loads and stores),
// each octet two pixels.
__global__ void __launch_bounds__(256) last_bwd_data_kernel(const float* __restrict__ g, const float* __restrict__ w,
                                                            const bf16* __restrict__ z5, bf16* __restrict__ dz5,
                                                            int H, int W, int B) {
    __shared__ float gs[198];
    const int tid = threadIdx.x, sub = tid & 7, grp = tid >> 3;
    float wr[9][8];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int q = 0; q < 8; ++q) wr[t][q] = w[t * kC + sub * 8 + q];
    const int nseg = W >> 6, total = B * H * nseg;
    // the loads of segment sg + gridDim.x are issued before segment sg is computed (clamped to the
    // last segment, so they need no branch)
    auto seg_row = [&](int sg, int& b, int& y, int& x0) {
        x0 = (sg % nseg) << 6;
        y = (sg / nseg) % H;
        b = sg / (nseg * H);
    };
    int b, y, x0;
    seg_row(blockIdx.x, b, y, x0);
    float gv = load_g_ring(g, b, y, x0, H, W, tid);
    bf16x8 z0 = load8(z5 + ((((size_t)b * H + y) * W + x0) + grp) * kC + sub * 8);
    bf16x8 z1 = load8(z5 + ((((size_t)b * H + y) * W + x0) + grp + 32) * kC + sub * 8);
    for (int sg = blockIdx.x; sg < total; sg += gridDim.x) {
        const size_t row = ((size_t)b * H + y) * W + x0;
        if (tid < 198) gs[tid] = gv;
        const bf16x8 zc0 = z0, zc1 = z1;
        seg_row(min(sg + (int)gridDim.x, total - 1), b, y, x0);
        gv = load_g_ring(g, b, y, x0, H, W, tid);
        z0 = load8(z5 + ((((size_t)b * H + y) * W + x0) + grp) * kC + sub * 8);
        z1 = load8(z5 + ((((size_t)b * H + y) * W + x0) + grp + 32) * kC + sub * 8);
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int px = grp + 32 * h;
            float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const float gt = gs[(t / 3) * 66 + px + 2 - (t % 3)];
#pragma unroll
                for (int q = 0; q < 8; ++q) s[q] += gt * wr[t][q];
            }
            const bf16x8 z = h ? zc1 : zc0;
            bf16x8 o;
#pragma unroll
            for (int q = 0; q < 8; ++q) o[q] = (bf16)(s[q] * gelu_grad((float)z[q]));
            *reinterpret_cast<bf16x8*>(dz5 + (row + px) * kC + sub * 8) = o;
        }
        __syncthreads();
    }
}

// Weight and bias gradient partials of the last layer over a run of row segments:
//   dW[t][ci] = sum_q a[q][ci] g[q - off_t],  db = sum_p g[p]
// taken over input pixels q, each read once (one 16-byte load per lane, eight lanes a pixel);
// the segment's g ring sits in LDS. Output per block: 9 * 64 weight sums, then the bias sum.
__global__ void __launch_bounds__(256) last_wgrad_kernel(const float* __restrict__ g, const bf16* __restrict__ a,
                                                         int H, int W, int B, int segs_per_block,
                                                         float* __restrict__ partial) {
    __shared__ float gs[198];
    __shared__ float sm[4][9 * kC + 1];
    const int tid = threadIdx.x, sub = tid & 7, grp = tid >> 3, wave = tid >> 6;
    float acc[9][8];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[t][q] = 0.0f;
    float gb = 0.0f;  // bias: threads 67..130 hold the segment's own row of g (dyi = 1, j = 1..64)
    const int nseg = W >> 6, total = B * H * nseg;
    const int s0 = blockIdx.x * segs_per_block, s1 = min(total, s0 + segs_per_block);
    // the loads of segment sg + 1 are issued before segment sg is computed (clamped to the block's
    // last segment, so they need no branch)
    auto seg_row = [&](int sg, int& b, int& y, int& x0) {
        x0 = (sg % nseg) << 6;
        y = (sg / nseg) % H;
        b = sg / (nseg * H);
    };
    int b, y, x0;
    seg_row(s0, b, y, x0);
    float gv = load_g_ring(g, b, y, x0, H, W, tid);
    bf16x8 an0 = load8(a + ((((size_t)b * H + y) * W + x0) + grp) * kC + sub * 8);
    bf16x8 an1 = load8(a + ((((size_t)b * H + y) * W + x0) + grp + 32) * kC + sub * 8);
    for (int sg = s0; sg < s1; ++sg) {
        if (tid < 198) gs[tid] = gv;
        if (tid >= 67 && tid <= 130) gb += gv;
        const bf16x8 a0 = an0, a1 = an1;
        seg_row(min(sg + 1, s1 - 1), b, y, x0);
        gv = load_g_ring(g, b, y, x0, H, W, tid);
        an0 = load8(a + ((((size_t)b * H + y) * W + x0) + grp) * kC + sub * 8);
        an1 = load8(a + ((((size_t)b * H + y) * W + x0) + grp + 32) * kC + sub * 8);
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int px = grp + 32 * h;
            const bf16x8 v = h ? a1 : a0;
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const float gt = gs[(t / 3) * 66 + px + 2 - (t % 3)];
#pragma unroll
                for (int q = 0; q < 8; ++q) acc[t][q] += (float)v[q] * gt;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[t][q] = octet_sum(acc[t][q]);
    if ((tid & 63) < 8)
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int q = 0; q < 8; ++q) sm[wave][t * kC + sub * 8 + q] = acc[t][q];
    // bias: wavefront sums, then the four wavefronts
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) gb += __shfl_xor(gb, m, 64);
    if ((tid & 63) == 0) sm[wave][9 * kC] = gb;
    __syncthreads();
    for (int j = tid; j < 9 * kC + 1; j += 256)
        partial[(size_t)blockIdx.x * (9 * kC + 1) + j] = (sm[0][j] + sm[1][j]) + (sm[2][j] + sm[3][j]);
}

// Inputs: fp32 [b][y][x][4] (already normalised) -> bf16 NHWC padded to 32 channels.
// One thread per pixel: its four inputs (one 16-byte load) go to channels 0..3 of its 32 bf16
// channels, zeros to the rest (four 16-byte stores).
__global__ void pack_input_kernel(const float* __restrict__ in, bf16* __restrict__ out, size_t npix) {
    static_assert(kInputs == 4 && kCin0 == 32, "one float4 in, four bf16x8 out");
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const float4 v = *reinterpret_cast<const float4*>(in + p * kInputs);
    bf16x8 first;
#pragma unroll
    for (int k = 0; k < 8; ++k) first[k] = (bf16)0.0f;
    first[0] = (bf16)v.x;
    first[1] = (bf16)v.y;
    first[2] = (bf16)v.z;
    first[3] = (bf16)v.w;
    bf16x8 zero;
#pragma unroll
    for (int k = 0; k < 8; ++k) zero[k] = (bf16)0.0f;
    bf16x8* o = reinterpret_cast<bf16x8*>(out + p * kCin0);
    o[0] = first;
    o[1] = zero;
    o[2] = zero;
    o[3] = zero;
}

// Mean squared error over the rows 1..H-2 (the wall rows carry no subfilter flux) and its gradient.
__global__ void __launch_bounds__(256) mse_kernel(const float* __restrict__ pred, const float* __restrict__ target,
                                                  float* __restrict__ grad, int H, int W, int B,
                                                  float* __restrict__ partial) {
    __shared__ float sm[256];
    const size_t np = (size_t)B * H * W;
    const double n_eff = (double)B * (H - 2) * W;
    float s = 0.0f;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < np; p += (size_t)gridDim.x * 256) {
        const int y = (int)((p / W) % H);
        float d = 0.0f;
        if (y > 0 && y < H - 1) d = pred[p] - target[p];
        grad[p] = (float)(2.0 * d / n_eff);
        s += d * d;
    }
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

// Adam (Kingma and Ba), fp32 master weights; also refreshes the bf16 copies of the conv layers.
__global__ void adam_kernel(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, size_t n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const float gj = g[j];
    const float mj = b1 * m[j] + (1.0f - b1) * gj;
    const float vj = b2 * v[j] + (1.0f - b2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    p[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
}

// bf16 weights of a 64-output conv layer from the fp32 master W[t][co][ci] (CIN channels), and the
// transposed copy WT[t][ci][co] (64-channel layers).
__global__ void weights_to_bf16_kernel(const float* __restrict__ w, int CIN, bf16* __restrict__ wb,
                                       bf16* __restrict__ wt) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 9 * kC * CIN) return;
    const int ci = e % CIN, co = (e / CIN) % kC, t = e / (CIN * kC);
    const float val = w[e];
    wb[e] = (bf16)val;
    if (wt) wt[((size_t)t * kC + ci) * kC + co] = (bf16)val;
}

}  // namespace cnn

// ------------------------------------------------------------------------------------------------
struct CnnImpl {
    int H = 0, W = 0, maxB = 0;
    hipStream_t s = nullptr;
    // parameters, fp32, one flat array: per layer W then b. Layer l in [0,5): W[9][64][cin_l], b[64];
    // layer 5: W[9][64], b[1].
    std::vector<size_t> off_w, off_b, n_w;
    size_t nparam = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    long long step = 0;
    cnn::bf16 *wb[5] = {nullptr}, *wt[5] = {nullptr};  // bf16 copies of the conv layers
    cnn::bf16* x0 = nullptr;                            // packed input [B][H][W][32]
    cnn::bf16 *z[5] = {nullptr}, *a[5] = {nullptr};     // pre-activations and activations
    cnn::bf16 *dz0 = nullptr, *dz1 = nullptr;           // gradient ping-pong [B][H][W][64]
    float *out = nullptr, *gout = nullptr, *partial = nullptr, *loss_part = nullptr;
    bool direct_conv = false;  // NSE2D_CNN_DIRECT=1: the unstaged convolution kernel
    bool persistent_conv = true;  // NSE2D_CNN_PERSISTENT=0: the per-tile staged kernel
    int cus = 256;                // compute units of the device (grid of the persistent kernel)
    size_t partial_floats = 0;
    int wgrad_blocks = 0;
};

namespace {

// input channels of conv layer l in the parameter layout (layer 0 keeps its 32 padded channels:
// the padded inputs are zero, so their weights get zero gradients and stay at their zero start)
int cinp_of(int l) { return l == 0 ? cnn::kCin0 : cnn::kC; }

void refresh_bf16(CnnImpl* c) {
    for (int l = 0; l < 5; ++l) {
        const int CIN = cinp_of(l), n = 9 * cnn::kC * CIN;
        cnn::weights_to_bf16_kernel<<<ceil_div(n, 256), 256, 0, c->s>>>(c->p + c->off_w[l], CIN, c->wb[l],
                                                                        l > 0 ? c->wt[l] : nullptr);
        NSE2D_LAUNCH_CHECK();
    }
}

template <int CIN, int MODE>
void launch_conv(CnnImpl* c, const cnn::bf16* in, const cnn::bf16* w, const float* bias, cnn::bf16* zo, cnn::bf16* ao,
                 const cnn::bf16* zp, int B) {
    const unsigned tiles = (unsigned)B * ceil_div(c->H, cnn::kTileRows) * (c->W / 64);
    if (c->direct_conv) {  // the unstaged kernel (kept as a cross-check of the staged ones)
        const unsigned blocks = (unsigned)B * c->H * (c->W / 64);
        cnn::conv_mfma_kernel<CIN, MODE><<<blocks, 256, 0, c->s>>>(in, w, bias, zo, ao, zp, c->H, c->W);
    } else if (c->persistent_conv) {
        // one workgroup per CU for the 64-channel layers (137 KB of LDS), two for the first layer
        const unsigned per_cu = CIN == cnn::kC ? 1 : 2;
        const unsigned blocks = std::min<unsigned>(tiles, per_cu * (unsigned)c->cus);
        cnn::conv_lds_all_kernel<CIN, MODE><<<blocks, cnn::kAllThreads, 0, c->s>>>(in, w, bias, zo, ao, zp, c->H,
                                                                                  c->W, (int)tiles);
    } else {
        cnn::conv_lds_kernel<CIN, MODE><<<tiles, 256, 0, c->s>>>(in, w, bias, zo, ao, zp, c->H, c->W);
    }
    NSE2D_LAUNCH_CHECK();
}

// keep = false (inference): the pre-activations z are not stored, so no backward may follow.
// The layers after the input packing. out: where the network's output goes (nullptr: the model's
// own buffer, which the loss reads); closure: the last layer writes the closure term instead
// (see last_fwd_rows_kernel).
void forward_packed(CnnImpl* c, int B, bool keep, float* out, int closure = 0, float scale = 1.0f,
                    float shift = 0.0f) {
    launch_conv<cnn::kCin0, cnn::MODE_FWD_GELU>(c, c->x0, c->wb[0], c->p + c->off_b[0], keep ? c->z[0] : nullptr,
                                                c->a[0], nullptr, B);
    for (int l = 1; l < 5; ++l)
        launch_conv<cnn::kC, cnn::MODE_FWD_GELU>(c, c->a[l - 1], c->wb[l], c->p + c->off_b[l],
                                                 keep ? c->z[l] : nullptr, c->a[l], nullptr, B);
    cnn::last_fwd_rows_kernel<<<B * c->H * (c->W / 64), cnn::kLastThreads, 0, c->s>>>(
        c->a[4], c->p + c->off_w[5], c->p + c->off_b[5], out ? out : c->out, c->H, c->W, closure, scale, shift);
    NSE2D_LAUNCH_CHECK();
}

void forward(CnnImpl* c, const float* d_in, int B, bool keep = true, float* out = nullptr) {
    const size_t npix = (size_t)B * c->H * c->W;
    cnn::pack_input_kernel<<<ceil_div(npix, 256), 256, 0, c->s>>>(d_in, c->x0, npix);
    NSE2D_LAUNCH_CHECK();
    forward_packed(c, B, keep, out);
}

// out[j] = sum over nb blocks of partial[k][j], j < n
void sum_partials(CnnImpl* c, int nb, int n, float* out) {
    const int cols = n >= 16384 ? 64 : 16;
    cnn::sum_partials_kernel<<<ceil_div(n, cols), cnn::kSumThreads, 0, c->s>>>(c->partial, nb, n, cols, out);
    NSE2D_LAUNCH_CHECK();
}

// Row segments of 64 pixels covering a batch
int row_segments(const CnnImpl* c, int B) { return B * c->H * (c->W / 64); }

template <int CIN>
void wgrad(CnnImpl* c, const cnn::bf16* x, const cnn::bf16* dz, float* gw, float* gb, int B) {
    const int total = B * c->H * (c->W / 64);
    // one round of resident workgroups, each over a long run of row segments (keeps the partial
    // sums small): as many workgroups per CU as the kernel's occupancy allows (two of four
    // wavefronts at CIN = 64; two of two at CIN = 32, whose accumulators and prefetched operands
    // hold one wavefront per SIMD)
    static const int per_cu = [] {
        int n = 0;
        NSE2D_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, reinterpret_cast<const void*>(cnn::conv_wgrad_kernel<CIN>), CIN * 4, 0));
        return std::max(1, n);
    }();
    const int nb = std::min(std::min(c->wgrad_blocks, c->cus * per_cu), total);
    const int per = (total + nb - 1) / nb;
    const int used = (total + per - 1) / per;
    float* bias_partial = c->partial + (size_t)c->wgrad_blocks * 9 * cnn::kC * cnn::kC;
    cnn::conv_wgrad_kernel<CIN><<<used, CIN * 4, 0, c->s>>>(x, dz, c->partial, bias_partial, c->H, c->W, B, per);
    NSE2D_LAUNCH_CHECK();
    sum_partials(c, used, 9 * cnn::kC * CIN, gw);
    cnn::sum_partials_kernel<<<ceil_div(cnn::kC, 16), cnn::kSumThreads, 0, c->s>>>(bias_partial, used, cnn::kC, 16, gb);
    NSE2D_LAUNCH_CHECK();
}

// Gradients of sum(gout * out) with respect to every parameter, after forward() on the same batch.
void backward(CnnImpl* c, const float* gout, int B) {
    {   // last layer
        const int total = row_segments(c, B);
        const int per = ceil_div(total, std::min(total, cnn::kLastWgradBlocks));
        const int used = ceil_div(total, per);
        cnn::last_wgrad_kernel<<<used, 256, 0, c->s>>>(gout, c->a[4], c->H, c->W, B, per, c->partial);
        NSE2D_LAUNCH_CHECK();
        sum_partials(c, used, 9 * cnn::kC + 1, c->g + c->off_w[5]);
        // four workgroups per CU are resident at the kernel's register count; each loops over segments
        cnn::last_bwd_data_kernel<<<std::min(total, 4 * c->cus), 256, 0, c->s>>>(gout, c->p + c->off_w[5], c->z[4],
                                                                                 c->dz0, c->H, c->W, B);
        NSE2D_LAUNCH_CHECK();
    }
    cnn::bf16* cur = c->dz0;
    cnn::bf16* nxt = c->dz1;
    for (int l = 4; l >= 1; --l) {
        wgrad<cnn::kC>(c, c->a[l - 1], cur, c->g + c->off_w[l], c->g + c->off_b[l], B);
        launch_conv<cnn::kC, cnn::MODE_BWD_DATA>(c, cur, c->wt[l], nullptr, nullptr, nxt, c->z[l - 1], B);
        std::swap(cur, nxt);
    }
    wgrad<cnn::kCin0>(c, c->x0, cur, c->g + c->off_w[0], c->g + c->off_b[0], B);
}

// Host normal deviates: splitmix64 -> uniform (0,1) -> Box-Muller.
struct HostNormal {
    unsigned long long s;
    bool have = false;
    double spare = 0;
    unsigned long long next() {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double operator()() {
        if (have) { have = false; return spare; }
        const double u1 = uniform(), u2 = uniform();
        const double r = std::sqrt(-2.0 * std::log(u1)), a = 2.0 * kPi * u2;
        spare = r * std::sin(a);
        have = true;
        return r * std::cos(a);
    }
};

void cnn_destroy(CnnImpl* c) {
    if (!c) return;
    for (int l = 0; l < 5; ++l) { (void)hipFree(c->wb[l]); (void)hipFree(c->wt[l]); (void)hipFree(c->z[l]); (void)hipFree(c->a[l]); }
    for (void* q : {(void*)c->p, (void*)c->g, (void*)c->m, (void*)c->v, (void*)c->x0, (void*)c->dz0, (void*)c->dz1,
                    (void*)c->out, (void*)c->gout, (void*)c->partial, (void*)c->loss_part})
        (void)hipFree(q);
    delete c;
}

CnnImpl* cnn_create(int H, int W, int maxB, unsigned long long seed, hipStream_t s) {
    if (H < 3 || W < 64 || W % 64 != 0 || maxB < 1) {
        set_last_error("nse2d_cnn_create: needs H >= 3, W a multiple of 64, max_batch >= 1");
        return nullptr;
    }
    auto* c = new CnnImpl;
    c->H = H; c->W = W; c->maxB = maxB; c->s = s;
    if (const char* e = std::getenv("NSE2D_CNN_DIRECT")) c->direct_conv = (e[0] == '1');
    if (const char* e = std::getenv("NSE2D_CNN_PERSISTENT")) c->persistent_conv = (e[0] != '0');
    {
        int dev = 0;
        NSE2D_HIP_CHECK(hipGetDevice(&dev));
        NSE2D_HIP_CHECK(hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, dev));
        if (c->cus < 1) c->cus = 1;
    }
    size_t o = 0;
    for (int l = 0; l < cnn::kLayers; ++l) {
        const size_t nw = l < 5 ? (size_t)9 * cnn::kC * cinp_of(l) : (size_t)9 * cnn::kC;
        c->off_w.push_back(o);
        c->n_w.push_back(nw);
        o += nw;
        c->off_b.push_back(o);
        o += l < 5 ? cnn::kC : 1;
    }
    c->nparam = o;
    // He initialisation over the real fan-in (padded input channels stay zero); zero biases
    std::vector<float> h(o, 0.0f);
    HostNormal rng{seed ? seed : 1ULL};
    for (int l = 0; l < cnn::kLayers; ++l) {
        const int cin = l == 0 ? cnn::kInputs : cnn::kC;
        const double sd = std::sqrt(2.0 / (9.0 * cin)) * (l == 5 ? 0.1 : 1.0);
        const int CIN = l < 5 ? cinp_of(l) : cnn::kC;
        const int nco = l < 5 ? cnn::kC : 1;
        for (int t = 0; t < 9; ++t)
            for (int co = 0; co < nco; ++co)
                for (int ci = 0; ci < CIN; ++ci) {
                    const double r = rng();
                    if (ci < cin) h[c->off_w[l] + ((size_t)t * nco + co) * CIN + ci] = (float)(sd * r);
                }
    }
    for (float** q : {&c->p, &c->g, &c->m, &c->v}) {
        NSE2D_HIP_CHECK(hipMalloc(q, o * sizeof(float)));
        NSE2D_HIP_CHECK(hipMemsetAsync(*q, 0, o * sizeof(float), s));
    }
    NSE2D_HIP_CHECK(hipMemcpyAsync(c->p, h.data(), o * sizeof(float), hipMemcpyHostToDevice, s));
    const size_t npix = (size_t)maxB * H * W;
    for (int l = 0; l < 5; ++l) {
        NSE2D_HIP_CHECK(hipMalloc(&c->wb[l], sizeof(cnn::bf16) * 9 * cnn::kC * cinp_of(l)));
        if (l > 0) NSE2D_HIP_CHECK(hipMalloc(&c->wt[l], sizeof(cnn::bf16) * 9 * cnn::kC * cnn::kC));
        NSE2D_HIP_CHECK(hipMalloc(&c->z[l], sizeof(cnn::bf16) * npix * cnn::kC));
        NSE2D_HIP_CHECK(hipMalloc(&c->a[l], sizeof(cnn::bf16) * npix * cnn::kC));
    }
    NSE2D_HIP_CHECK(hipMalloc(&c->x0, sizeof(cnn::bf16) * npix * cnn::kCin0));
    // channels kInputs..kCin0-1 stay zero (cnn_packed_input)
    NSE2D_HIP_CHECK(hipMemsetAsync(c->x0, 0, sizeof(cnn::bf16) * npix * cnn::kCin0, s));
    NSE2D_HIP_CHECK(hipMalloc(&c->dz0, sizeof(cnn::bf16) * npix * cnn::kC));
    NSE2D_HIP_CHECK(hipMalloc(&c->dz1, sizeof(cnn::bf16) * npix * cnn::kC));
    NSE2D_HIP_CHECK(hipMalloc(&c->out, sizeof(float) * npix));
    NSE2D_HIP_CHECK(hipMalloc(&c->gout, sizeof(float) * npix));
    // most weight-gradient workgroups of a launch (see wgrad)
    c->wgrad_blocks = 4 * c->cus;
    // weight-gradient partials of the hidden layers, then their bias partials
    const size_t pw = (size_t)c->wgrad_blocks * (9 * cnn::kC * cnn::kC + cnn::kC);
    const size_t pl = (size_t)std::min(row_segments(c, maxB), cnn::kLastWgradBlocks) * (9 * cnn::kC + 1);
    c->partial_floats = std::max(pw, pl);
    NSE2D_HIP_CHECK(hipMalloc(&c->partial, sizeof(float) * c->partial_floats));
    NSE2D_HIP_CHECK(hipMalloc(&c->loss_part, sizeof(float) * 1024));
    refresh_bf16(c);
    NSE2D_HIP_CHECK(hipStreamSynchronize(s));
    return c;
}

double cnn_loss_and_grad(CnnImpl* c, const float* target, int B) {
    cnn::mse_kernel<<<1024, 256, 0, c->s>>>(c->out, target, c->gout, c->H, c->W, B, c->loss_part);
    NSE2D_LAUNCH_CHECK();
    std::vector<float> lp(1024);
    NSE2D_HIP_CHECK(hipMemcpyAsync(lp.data(), c->loss_part, sizeof(float) * 1024, hipMemcpyDeviceToHost, c->s));
    NSE2D_HIP_CHECK(hipStreamSynchronize(c->s));
    double sum = 0.0;
    for (float v : lp) sum += v;
    return sum / ((double)B * (c->H - 2) * c->W);
}

void cnn_adam(CnnImpl* c, float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    c->step++;
    const float c1 = (float)(1.0 - std::pow((double)b1, (double)c->step));
    const float c2 = (float)(1.0 - std::pow((double)b2, (double)c->step));
    cnn::adam_kernel<<<ceil_div(c->nparam, 256), 256, 0, c->s>>>(c->p, c->m, c->v, c->g, c->nparam, lr, b1, b2, eps,
                                                                c1, c2);
    NSE2D_LAUNCH_CHECK();
    refresh_bf16(c);
}

}  // namespace
}  // namespace nse2d

struct nse2d_cnn {
    nse2d::CnnImpl* impl;
};

namespace nse2d {
__bf16* cnn_packed_input(nse2d_cnn_t c) { return c->impl->x0; }
void cnn_infer_packed_closure(nse2d_cnn_t c, float scale, float shift, float* dst) {
    forward_packed(c->impl, 1, false, dst, 1, scale, shift);
}
}  // namespace nse2d

using namespace nse2d;

extern "C" {

nse2d_cnn_t nse2d_cnn_create(int H, int W, int max_batch, unsigned long long seed, hipStream_t s) {
    CnnImpl* c = cnn_create(H, W, max_batch, seed, s);
    return c ? new nse2d_cnn{c} : nullptr;
}
void nse2d_cnn_destroy(nse2d_cnn_t c) {
    if (!c) return;
    cnn_destroy(c->impl);
    delete c;
}
size_t nse2d_cnn_num_params(nse2d_cnn_t c) { return c->impl->nparam; }
void nse2d_cnn_get_params(nse2d_cnn_t c, float* host) {
    NSE2D_HIP_CHECK(hipMemcpyAsync(host, c->impl->p, sizeof(float) * c->impl->nparam, hipMemcpyDeviceToHost, c->impl->s));
    NSE2D_HIP_CHECK(hipStreamSynchronize(c->impl->s));
}
void nse2d_cnn_set_params(nse2d_cnn_t c, const float* host) {
    CnnImpl* k = c->impl;
    NSE2D_HIP_CHECK(hipMemcpyAsync(k->p, host, sizeof(float) * k->nparam, hipMemcpyHostToDevice, k->s));
    NSE2D_HIP_CHECK(hipMemsetAsync(k->m, 0, sizeof(float) * k->nparam, k->s));
    NSE2D_HIP_CHECK(hipMemsetAsync(k->v, 0, sizeof(float) * k->nparam, k->s));
    k->step = 0;
    refresh_bf16(k);
    NSE2D_HIP_CHECK(hipStreamSynchronize(k->s));
}
void nse2d_cnn_get_grads(nse2d_cnn_t c, float* host) {
    NSE2D_HIP_CHECK(hipMemcpyAsync(host, c->impl->g, sizeof(float) * c->impl->nparam, hipMemcpyDeviceToHost, c->impl->s));
    NSE2D_HIP_CHECK(hipStreamSynchronize(c->impl->s));
}
// The input packing reads a pixel's four inputs as one 16-byte load.
static bool input_misaligned(const float* d_in, const char* who) {
    if (reinterpret_cast<uintptr_t>(d_in) % 16 == 0) return false;
    set_last_error(std::string(who) + ": d_in must be 16-byte aligned");
    return true;
}

int nse2d_cnn_forward(nse2d_cnn_t c, const float* d_in, float* d_out, int B) {
    CnnImpl* k = c->impl;
    if (B < 1 || B > k->maxB) { set_last_error("nse2d_cnn_forward: batch outside [1, max_batch]"); return -1; }
    if (input_misaligned(d_in, "nse2d_cnn_forward")) return -1;
    forward(k, d_in, B);
    if (d_out)
        NSE2D_HIP_CHECK(hipMemcpyAsync(d_out, k->out, sizeof(float) * (size_t)B * k->H * k->W, hipMemcpyDeviceToDevice,
                                       k->s));
    return 0;
}
int nse2d_cnn_infer(nse2d_cnn_t c, const float* d_in, float* d_out, int B) {
    CnnImpl* k = c->impl;
    if (B < 1 || B > k->maxB) { set_last_error("nse2d_cnn_infer: batch outside [1, max_batch]"); return -1; }
    if (input_misaligned(d_in, "nse2d_cnn_infer")) return -1;
    forward(k, d_in, B, false, d_out);  // the last layer writes d_out directly
    return 0;
}
int nse2d_cnn_set_direct(nse2d_cnn_t c, int on) {
    c->impl->direct_conv = on == 1;
    c->impl->persistent_conv = on != 2;
    return 0;
}
int nse2d_cnn_backward(nse2d_cnn_t c, const float* d_gout, int B) {
    CnnImpl* k = c->impl;
    if (B < 1 || B > k->maxB) { set_last_error("nse2d_cnn_backward: batch outside [1, max_batch]"); return -1; }
    backward(k, d_gout, B);
    return 0;
}
double nse2d_cnn_train_step(nse2d_cnn_t c, const float* d_in, const float* d_target, int B, float lr) {
    CnnImpl* k = c->impl;
    if (B < 1 || B > k->maxB) { set_last_error("nse2d_cnn_train_step: batch outside [1, max_batch]"); return -1.0; }
    if (input_misaligned(d_in, "nse2d_cnn_train_step")) return -1.0;
    forward(k, d_in, B);
    const double loss = cnn_loss_and_grad(k, d_target, B);
    backward(k, k->gout, B);
    cnn_adam(k, lr);
    return loss;
}
double nse2d_cnn_eval_loss(nse2d_cnn_t c, const float* d_in, const float* d_target, int B) {
    CnnImpl* k = c->impl;
    if (B < 1 || B > k->maxB) { set_last_error("nse2d_cnn_eval_loss: batch outside [1, max_batch]"); return -1.0; }
    if (input_misaligned(d_in, "nse2d_cnn_eval_loss")) return -1.0;
    forward(k, d_in, B);
    return cnn_loss_and_grad(k, d_target, B);
}
int nse2d_cnn_adam_step(nse2d_cnn_t c, float lr) {
    cnn_adam(c->impl, lr);
    return 0;
}

}  // extern "C"
