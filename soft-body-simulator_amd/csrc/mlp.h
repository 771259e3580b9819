// Launch interface of the warm-start network kernels (implemented in mlp.hip).
//
// The network is a fixed-width MLP: 48 inputs (the first `num_inputs` used, the rest zero) ->
// 64 -> 64 -> 64 (ReLU) -> 32 outputs (the first `num_outputs` used). Parameters are kept in FP32
// (flat vector, layout below) and fed to the matrix cores as BF16 with FP32 accumulation.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace sbs {
namespace mlp {

constexpr int kIn = 48, kHidden = 64, kOut = 32;
// dataset row widths: features (raw FP32) and targets
constexpr int kMaxInputs = 40, kMaxOutputs = 12;

// Flat FP32 parameter vector: W1[48][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][32]
// b4[32]; W_l[i][o] multiplies input unit i into output unit o.
constexpr std::size_t kW1 = 0, kB1 = kW1 + kIn * kHidden, kW2 = kB1 + kHidden, kB2 = kW2 + kHidden * kHidden,
                      kW3 = kB2 + kHidden, kB3 = kW3 + kHidden * kHidden, kW4 = kB3 + kHidden,
                      kB4 = kW4 + kHidden * kOut, kParams = kB4 + kOut;
constexpr std::size_t kBiasTotal = 3 * kHidden + kOut;

// BF16 weights in the operand order of the matrix-core forward pass (mlp_device.h): per layer, per
// 32-unit output tile, per 16-unit k-step, 64 lanes x 8 values (16 bytes per lane).
constexpr std::size_t kPack1 = (kHidden / 32) * (kIn / 16) * 64 * 8, kPack23 = (kHidden / 32) * (kHidden / 16) * 64 * 8,
                      kPack4 = (kOut / 32) * (kHidden / 16) * 64 * 8;
constexpr std::size_t kPackTotal = kPack1 + 2 * kPack23 + kPack4;
// multiply-adds of one row (the padded layer widths)
constexpr std::size_t kMacsPerRow = std::size_t(kIn) * kHidden + 2 * std::size_t(kHidden) * kHidden +
                                    std::size_t(kHidden) * kOut;

using bf16 = __bf16;

// Input normalization x' = (x - mean) * inv_std and output scale y = scale * y'.
struct norm_dev {
    float mean[kIn];
    float inv_std[kIn];
    float out_scale[kMaxOutputs];
};

// params (FP32) -> pack (fused-kernel operand order), wrow (row-major BF16 W_l for the training
// GEMMs, same layout as the FP32 weights) and biases (b1 b2 b3 b4 contiguous, FP32).
void launch_pack(hipStream_t s, const float* params, bf16* pack, bf16* wrow, float* bias);

// Inference over n rows of BF16 inputs X[n][48]: Y[n][12] = out_scale * net(X) (the first
// num_outputs outputs; the rest of a row zero). grid: persistent, `blocks` workgroups of 256 threads.
void launch_infer(hipStream_t s, const bf16* pack, const float* bias, const bf16* X, uint32_t n, int num_outputs,
                  const float* out_scale, float* Y, int blocks);

// Raw FP32 rows (stride `ld`) -> normalized BF16 rows of 48 (unused inputs zero). rows = index
// list or nullptr for 0..n-1.
void launch_normalize(hipStream_t s, const float* raw, uint32_t ld, int num_inputs, const uint32_t* rows,
                      uint32_t n, uint32_t n_pad, const norm_dev* norm, bf16* X);
// Minibatch of n rows drawn with replacement from a dataset of `size` rows by a counter hash of
// (seed, step): X (normalized BF16, n_pad rows), T (targets divided by out_scale, FP32 [n][12]).
void launch_sample(hipStream_t s, const float* feats, const float* targets, uint64_t size, int num_inputs,
                   uint32_t n, uint32_t n_pad, uint64_t seed, uint64_t step, const norm_dev* norm, bf16* X,
                   float* T);

// C[M][N] (+)= A(M x K) B(K x N) on BF16 operands with FP32 accumulation; A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn]; K a multiple of 16 (rows and columns past M, N are neither read nor
// written, so M, N need not be multiples of 32). With splits > 1 the
// K range is split into `splits` chunks whose partial products go to C + z*M*N and are summed
// by launch_reduce_splits in a fixed order.
void launch_gemm(hipStream_t s, const bf16* A, long sam, long sak, const bf16* B, long sbk, long sbn, float* C,
                 int M, int N, int K, int splits);
void launch_reduce_splits(hipStream_t s, const float* parts, int splits, std::size_t count, float* out);

// A = bf16(relu(Z + b)) for Z[rows][cols] (FP32), b[cols].
void launch_bias_relu(hipStream_t s, const float* Z, const float* b, uint32_t rows, int cols, bf16* A);
// Y = Z + b (output layer), loss partials and dY = bf16(2 (Y - T) / (n * num_outputs)) on the
// used outputs of the n real rows (zero elsewhere). Z, dY: [n_pad][32]; T: [n][12].
void launch_output_grad(hipStream_t s, const float* Z, const float* b, const float* T, uint32_t n, uint32_t n_pad,
                        int num_outputs, bf16* dY, double* loss_parts, int loss_blocks);
// dZ = bf16(dA * (A > 0)) (ReLU backward through the stored BF16 activation).
void launch_relu_grad(hipStream_t s, const float* dA, const bf16* A, std::size_t count, bf16* dZ);
// db[c] = sum over rows of dZ[r][c] (rows summed in a fixed order).
void launch_colsum(hipStream_t s, const bf16* dZ, uint32_t rows, int cols, float* db);
// Adam on the flat parameter vector (bias-corrected with step t >= 1).
void launch_adam(hipStream_t s, float* params, const float* grad, float* m, float* v, std::size_t count, float lr,
                 float beta1, float beta2, float eps, int t);

}  // namespace mlp
}  // namespace sbs
