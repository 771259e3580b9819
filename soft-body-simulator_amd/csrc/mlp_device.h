// Device side of the warm-start network's matrix-core forward pass, shared by the inference kernels
// (mlp.hip) and the warm-start pass that runs the network inside its cluster kernel (kernels.hip).
//
// Matrix-core layout (v_mfma_f32_32x32x16_bf16, one wave): the A operand is a 32 x 16 tile,
// lane l holding row l%32 and the 8 k-values 8*(l/32) .. +7; the B operand is 16 x 32, lane l
// holding column l%32 and the same 8 k-values; the 32 x 32 result leaves lane l with column l%32
// and rows 8*(j/4) + 4*(l/32) + j%4 for its 16 accumulators j.
//
// Every layer is computed transposed, Z^T = W^T A^T: the output units are the rows, the data rows
// the columns. A lane then holds 16 units of one data row after each layer, and the 8 accumulators
// j = 8s .. 8s+7 of unit tile t are exactly the B-operand values of the next layer's k-step over
// units 32t + 16s .. +15, provided the weights of that k-step are ordered so that value i of lane
// half h is input unit 16q + (i < 4 ? 4h + i : 8 + 4h + i - 4). Activations never leave the
// registers. The first layer's B operand is the input row itself: value i of lane half h of k-step q
// is input unit 16q + 8h + i.
#pragma once
#include "mlp.h"

namespace sbs {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kT = kHidden / 32;  // unit tiles of a hidden layer
constexpr int kQ1 = kIn / 16;     // k-steps of the first layer
constexpr int kQH = kHidden / 16;  // k-steps of the hidden and output layers

// Offsets (in bf16x8 units) of the layers in the packed weights.
constexpr int kP1 = 0, kP2 = int(kPack1 / 8), kP3 = kP2 + int(kPack23 / 8), kP4 = kP3 + int(kPack23 / 8);
// Offsets of the biases in the contiguous bias array.
constexpr int kb1 = 0, kb2 = kHidden, kb3 = 2 * kHidden, kb4 = 3 * kHidden;

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Output unit of accumulator j of a lane in half h of a 32-unit tile.
__device__ __forceinline__ constexpr int acc_unit(int j, int h) { return 8 * (j / 4) + 4 * h + (j % 4); }

// bias + ReLU of one 32-unit tile of accumulators, rounded to BF16 (lane half h, bias of the
// layer starting at b).
__device__ __forceinline__ void relu_tile(f32x16 const& acc, const float* b, int T, int h, bf16x8& lo, bf16x8& hi)
{
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        float const z = fmaxf(acc[j] + b[32 * T + acc_unit(j, h)], 0.0f);
        if (j < 8)
            lo[j] = static_cast<bf16>(z);
        else
            hi[j - 8] = static_cast<bf16>(z);
    }
}

// Forward pass of R 32-row tiles that share every weight operand load: xin[r][q] is tile r's
// first-layer B operand of k-step q. W: packed weights (LDS or global), b: the biases. Returns the
// output accumulators y[r] (bias added): lane (c, h) holds output units acc_unit(j, h) of row c of
// tile r. A scheduling barrier after each layer keeps the next layer's weight loads from moving up:
// the scheduler would otherwise load all 26 weight operands (104 registers a lane) at the start.
template <int R, class W>
__device__ __forceinline__ void forward_tiles(W const* w, const float* b, bf16x8 const (&xin)[R][kQ1], int lane,
                                              f32x16 (&y)[R])
{
    int const h = lane >> 5;
    bf16x8 act[R][kT][2];
#pragma unroll
    for (int T = 0; T < kT; ++T) {
        f32x16 acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = f32x16{};
#pragma unroll
        for (int q = 0; q < kQ1; ++q) {
            bf16x8 const a = w[kP1 + (T * kQ1 + q) * 64 + lane];
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = mfma(a, xin[r][q], acc[r]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) relu_tile(acc[r], b + kb1, T, h, act[r][T][0], act[r][T][1]);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int L = 0; L < 2; ++L) {
        int const poff = L == 0 ? kP2 : kP3;
        const float* bl = b + (L == 0 ? kb2 : kb3);
        f32x16 acc[R][kT];
#pragma unroll
        for (int T = 0; T < kT; ++T) {
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r][T] = f32x16{};
#pragma unroll
            for (int q = 0; q < kQH; ++q) {
                bf16x8 const a = w[poff + (T * kQH + q) * 64 + lane];
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r][T] = mfma(a, act[r][q >> 1][q & 1], acc[r][T]);
            }
        }
#pragma unroll
        for (int T = 0; T < kT; ++T)
#pragma unroll
            for (int r = 0; r < R; ++r) relu_tile(acc[r][T], bl, T, h, act[r][T][0], act[r][T][1]);
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) y[r] = f32x16{};
#pragma unroll
    for (int q = 0; q < kQH; ++q) {
        bf16x8 const a = w[kP4 + q * 64 + lane];
#pragma unroll
        for (int r = 0; r < R; ++r) y[r] = mfma(a, act[r][q >> 1][q & 1], y[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int j = 0; j < 16; ++j) y[r][j] += b[kb4 + acc_unit(j, h)];
}

}  // namespace mlp
}  // namespace sbs
