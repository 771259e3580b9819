// Launch interface of the warm-start network kernels (implemented in mlp.hip).
//
// The network is a fixed-width MLP: 16 inputs (the first `num_inputs` used, the rest zero) ->
// 128 -> 128 -> 128 (ReLU) -> 32 outputs (the first `num_outputs` used). Parameters are kept in
// FP32 (flat vector, layout below) and fed to the matrix cores as BF16 with FP32 accumulation.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace sbs {
namespace mlp {

constexpr int kIn = 16, kHidden = 128, kOut = 32;
constexpr int kMaxInputs = 13, kMaxOutputs = 4;

// Flat FP32 parameter vector: W1[16][128] b1[128] W2[128][128] b2[128] W3[128][128] b3[128]
// W4[128][32] b4[32]; W_l[i][o] multiplies input unit i into output unit o.
constexpr std::size_t kW1 = 0, kB1 = kW1 + kIn * kHidden, kW2 = kB1 + kHidden, kB2 = kW2 + kHidden * kHidden,
                      kW3 = kB2 + kHidden, kB3 = kW3 + kHidden * kHidden, kW4 = kB3 + kHidden,
                      kB4 = kW4 + kHidden * kOut, kParams = kB4 + kOut;
constexpr std::size_t kBiasTotal = 3 * kHidden + kOut;

// BF16 weights in the operand order of the fused inference kernel: per layer, per 32-unit output
// tile, per 16-unit k-step, 64 lanes x 8 values (16 bytes per lane).
constexpr std::size_t kPack1 = 4 * 1 * 64 * 8, kPack23 = 4 * 8 * 64 * 8, kPack4 = 1 * 8 * 64 * 8;
constexpr std::size_t kPackTotal = kPack1 + 2 * kPack23 + kPack4;

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

// Fused inference over n rows of BF16 inputs X[n][16]: Y[n][4] = out_scale * net(X) (first four
// outputs). grid: persistent, `blocks` workgroups of 256 threads.
void launch_infer(hipStream_t s, const bf16* pack, const float* bias, const bf16* X, uint32_t n, int num_outputs,
                  const float* out_scale, float* Y, int blocks);
// The same forward pass, but the output updates multipliers in place: lambda[r * num_outputs + o]
// += out_scale[o] * y[o] (the warm start: the network predicts the change of the multipliers).
void launch_infer_lambda(hipStream_t s, const bf16* pack, const float* bias, const bf16* X, uint32_t n,
                         int num_outputs, const float* out_scale, float* lambda, int blocks);

// Raw FP32 rows (stride `ld`) -> normalized BF16 rows of 16 (unused inputs zero). rows = index
// list or nullptr for 0..n-1.
void launch_normalize(hipStream_t s, const float* raw, uint32_t ld, int num_inputs, const uint32_t* rows,
                      uint32_t n, uint32_t n_pad, const norm_dev* norm, bf16* X);
// Minibatch of n rows drawn with replacement from a dataset of `size` rows by a counter hash of
// (seed, step): X (normalized BF16, n_pad rows), T (targets divided by out_scale, FP32 [n][4]).
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
// used outputs of the n real rows (zero elsewhere). Z, dY: [n_pad][32]; T: [n][4].
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
