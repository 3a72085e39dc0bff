// Internal interface of the closure network (closure.hip) for the LES closure term (les.hip):
// the coarse state's inputs are formed straight into the network's packed input, and the
// inference writes the scaled closure term straight into the channel's buffer.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/nse2d_hip.h"

namespace nse2d {

// The network's packed bf16 input of batch item 0: [H][W][32], the four inputs in channels 0..3.
// Channels 4..31 are zero from creation on (every writer leaves them zero), so a writer that
// forms the inputs on the device stores channels 0..3 only.
__bf16* cnn_packed_input(nse2d_cnn_t c);

// Inference of batch item 0 from the packed input (no input packing), the last layer writing
//   dst[y][x] = 0 < y < H-1 ? (out(y, x)) * scale + shift : 0
// on the network's stream.
void cnn_infer_packed_closure(nse2d_cnn_t c, float scale, float shift, float* dst);

}  // namespace nse2d
