// 3D convolution on channels-last BF16 volumes with FP32 accumulation on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16).  Shared by the C entry points
// of conv3d.hip and the U-Net of unet.hip.
//
// Layout: activations x[n][z][y][x][c] (bf16, c fastest), weights
// w[co][tap][ci] (bf16), tap = (dz*K + dy)*K + dx for offsets -(K-1)/2..(K-1)/2,
// zero padding ("same" convolution, stride 1).  Channel counts are multiples
// of 16; depth, height and width are powers of two.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace nx::conv {

struct Shape {
    int n, d, h, w;      // batch, depth, height, width
    int cin, cout, ks;   // ks = 1 or 3
};

enum : int { kRelu = 1, kOutF32 = 2 };

// y = conv(x, w) + bias (bias may be null), ReLU if flags & kRelu; y is bf16
// or, with kOutF32, float.
int fwd(const Shape& s, const uint16_t* x, const uint16_t* w, const float* bias, void* y, int flags,
        hipStream_t st);
// w (bf16 [cout][T][cin]) -> wt (bf16 [cin][T][cout], taps reversed): the
// weights of the transposed convolution.
int flip(const Shape& s, const uint16_t* w, uint16_t* wt, hipStream_t st);
// dx = conv^T(dy, w) given the flipped weights wt; dx bf16 or float (kOutF32).
int bwd_data(const Shape& s, const uint16_t* dy, const uint16_t* wt, void* dx, int flags,
             hipStream_t st);
// dw[co][tap][ci] = sum_v dy[v][co] x[v + off(tap)][ci], db[co] = sum_v dy[v][co]
// (db may be null).  ws: workspace of bwd_weight_ws(s) bytes.
size_t bwd_weight_ws(const Shape& s);
int bwd_weight(const Shape& s, const uint16_t* x, const uint16_t* dy, float* dw, float* db, void* ws,
               hipStream_t st);
// argument checks shared by the entry points (returns NX_OK or an error code)
int check(const Shape& s);

}  // namespace nx::conv
