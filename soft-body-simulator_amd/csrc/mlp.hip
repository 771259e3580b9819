// gfx950 kernels of the warm-start network: the BF16 matrix-core inference pass, the generic BF16
// GEMM of the training passes, and the elementwise training kernels (bias + ReLU, output loss
// gradient, ReLU backward, bias gradients, Adam). The forward pass and its operand layout are in
// mlp_device.h.
#include "mlp_device.h"

namespace sbs {
namespace mlp {

namespace {

// Row-major BF16 weight copies, same order as the FP32 parameter vector without the biases.
constexpr std::size_t kR1 = 0, kR2 = kR1 + kIn * kHidden, kR3 = kR2 + kHidden * kHidden,
                      kR4 = kR3 + kHidden * kHidden, kRTotal = kR4 + kHidden * kOut;

__global__ void k_pack(const float* __restrict__ P, bf16* __restrict__ pack, bf16* __restrict__ wrow,
                       float* __restrict__ bias)
{
    std::size_t const i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < kPackTotal) {
        std::size_t e = i;
        int L;
        if (e < kPack1) {
            L = 1;
        } else if (e < kPack1 + kPack23) {
            L = 2;
            e -= kPack1;
        } else if (e < kPack1 + 2 * kPack23) {
            L = 3;
            e -= kPack1 + kPack23;
        } else {
            L = 4;
            e -= kPack1 + 2 * kPack23;
        }
        int const Q = (L == 1) ? kQ1 : kQH;
        int const v = int(e % 8), lane = int((e / 8) % 64), q = int((e / 512) % Q), T = int(e / (512 * Q));
        int const h = lane / 32, c = lane % 32;
        int const in_unit = 16 * q + (L == 1 ? 8 * h + v : (v < 4 ? 4 * h + v : 8 + 4 * h + v - 4));
        int const out_unit = 32 * T + c;
        std::size_t const off = L == 1 ? kW1 : (L == 2 ? kW2 : (L == 3 ? kW3 : kW4));
        int const nout = (L == 4) ? kOut : kHidden;
        pack[i] = static_cast<bf16>(P[off + std::size_t(in_unit) * nout + out_unit]);
    }
    if (i < kRTotal) {
        std::size_t src;
        if (i < kR2)
            src = kW1 + i;
        else if (i < kR3)
            src = kW2 + (i - kR2);
        else if (i < kR4)
            src = kW3 + (i - kR3);
        else
            src = kW4 + (i - kR4);
        wrow[i] = static_cast<bf16>(P[src]);
    }
    if (i < kBiasTotal) {
        std::size_t const src = i < std::size_t(kb2) ? kB1 + i
                                : i < std::size_t(kb3) ? kB2 + (i - kb2)
                                : i < std::size_t(kb4) ? kB3 + (i - kb3)
                                                       : kB4 + (i - kb4);
        bias[i] = P[src];
    }
}

// Persistent inference over 32-row tiles of X[n][48] (BF16), weights staged in LDS; epi(row, y,
// h) receives the output accumulators of a real row for the lane of half h.
template <class Epilogue>
__device__ __forceinline__ void infer_rows(const bf16x8* __restrict__ pack, const float* __restrict__ bias,
                                           const bf16x8* __restrict__ X, uint32_t n, Epilogue&& epi)
{
    __shared__ bf16x8 sw[kPackTotal / 8];
    __shared__ float sb[kBiasTotal];
    for (uint32_t i = threadIdx.x; i < kPackTotal / 8; i += blockDim.x) sw[i] = pack[i];
    for (uint32_t i = threadIdx.x; i < kBiasTotal; i += blockDim.x) sb[i] = bias[i];
    __syncthreads();
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, c = lane & 31;
    uint32_t const waves = blockDim.x >> 6;
    uint32_t const tiles = (n + 31) / 32;
    for (uint32_t tile = blockIdx.x * waves + wave; tile < tiles; tile += gridDim.x * waves) {
        uint32_t const row = tile * 32 + uint32_t(c);
        bf16x8 xin[1][kQ1];
#pragma unroll
        for (int q = 0; q < kQ1; ++q) xin[0][q] = row < n ? X[std::size_t(row) * (kIn / 8) + 2 * q + h] : bf16x8{};
        f32x16 y[1];
        forward_tiles<1>(sw, sb, xin, lane, y);
        if (row < n) epi(row, y[0], h);
    }
}

__global__ void __launch_bounds__(256) k_infer(const bf16x8* __restrict__ pack, const float* __restrict__ bias,
                                               const bf16x8* __restrict__ X, uint32_t n, int nout,
                                               const float* __restrict__ scale, float* __restrict__ Y)
{
    infer_rows(pack, bias, X, n, [&](uint32_t row, f32x16 const& y, int h) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            int const o = acc_unit(j, h);
            if (o < kMaxOutputs) Y[std::size_t(row) * kMaxOutputs + o] = o < nout ? scale[o] * y[j] : 0.0f;
        }
    });
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ void normalize_row(const float* raw, int num_inputs, const norm_dev* norm, bf16* out)
{
#pragma unroll
    for (int k = 0; k < kIn; ++k)
        out[k] = static_cast<bf16>(k < num_inputs ? (raw[k] - norm->mean[k]) * norm->inv_std[k] : 0.0f);
}

__global__ void k_normalize(const float* __restrict__ raw, uint32_t ld, int num_inputs, const uint32_t* __restrict__ rows,
                            uint32_t n, uint32_t n_pad, const norm_dev* __restrict__ norm, bf16* __restrict__ X)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    bf16* out = X + std::size_t(i) * kIn;
    if (i >= n) {
        for (int k = 0; k < kIn; ++k) out[k] = static_cast<bf16>(0.0f);
        return;
    }
    uint32_t const r = rows ? rows[i] : i;
    normalize_row(raw + std::size_t(r) * ld, num_inputs, norm, out);
}

__global__ void k_sample(const float* __restrict__ feats, const float* __restrict__ targets, uint64_t size,
                         int num_inputs, uint32_t n, uint32_t n_pad, uint64_t seed, uint64_t step,
                         const norm_dev* __restrict__ norm, bf16* __restrict__ X, float* __restrict__ T)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    bf16* out = X + std::size_t(i) * kIn;
    if (i >= n) {
        for (int k = 0; k < kIn; ++k) out[k] = static_cast<bf16>(0.0f);
        return;
    }
    uint64_t const r = mix64(seed ^ mix64(step * 0x632BE59BD9B4E019ull + i)) % size;
    normalize_row(feats + r * kMaxInputs, num_inputs, norm, out);
    for (int o = 0; o < kMaxOutputs; ++o) T[std::size_t(i) * kMaxOutputs + o] = targets[r * kMaxOutputs + o] / norm->out_scale[o];
}

__global__ void __launch_bounds__(64) k_gemm(const bf16* __restrict__ A, long sam, long sak, const bf16* __restrict__ B,
                                             long sbk, long sbn, float* __restrict__ C, int M, int N, int K,
                                             int kchunk)
{
    int const lane = threadIdx.x, h = lane >> 5, c = lane & 31;
    int const n0 = blockIdx.x * 32, m0 = blockIdx.y * 32, z = blockIdx.z;
    int const k0 = z * kchunk, k1 = min(K, k0 + kchunk);
    f32x16 acc{};
    bool const mok = m0 + c < M, nok = n0 + c < N;
    for (int k = k0; k < k1; k += 16) {
        bf16x8 a{}, b{};
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            long const kk = k + 8 * h + v;
            if (mok) a[v] = A[long(m0 + c) * sam + kk * sak];
            if (nok) b[v] = B[kk * sbk + long(n0 + c) * sbn];
        }
        acc = mfma(a, b, acc);
    }
    float* out = C + std::size_t(z) * M * N;
    if (nok) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            int const m = m0 + 8 * (j / 4) + 4 * h + (j % 4);
            if (m < M) out[std::size_t(m) * N + n0 + c] = acc[j];
        }
    }
}

__global__ void k_reduce_splits(const float* __restrict__ parts, int splits, std::size_t count, float* __restrict__ out)
{
    std::size_t const i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    for (int z = 0; z < splits; ++z) s += parts[std::size_t(z) * count + i];
    out[i] = s;
}

__global__ void k_bias_relu(const float* __restrict__ Z, const float* __restrict__ b, uint32_t rows, int cols,
                            bf16* __restrict__ A)
{
    std::size_t const i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= std::size_t(rows) * cols) return;
    A[i] = static_cast<bf16>(fmaxf(Z[i] + b[i % cols], 0.0f));
}

__global__ void __launch_bounds__(256) k_output_grad(const float* __restrict__ Z, const float* __restrict__ b,
                                                     const float* __restrict__ T, uint32_t n, uint32_t n_pad,
                                                     int nout, bf16* __restrict__ dY, double* __restrict__ loss_parts)
{
    __shared__ double red[256];
    std::size_t const total = std::size_t(n_pad) * kOut;
    double acc = 0.0;
    float const g = 2.0f / (float(n) * float(nout));
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += std::size_t(gridDim.x) * blockDim.x) {
        uint32_t const r = uint32_t(i / kOut);
        int const o = int(i % kOut);
        float d = 0.0f;
        if (r < n && o < nout) {
            d = Z[i] + b[o] - T[std::size_t(r) * kMaxOutputs + o];
            acc += double(d) * double(d);
        }
        dY[i] = static_cast<bf16>(g * d);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_parts[blockIdx.x] = red[0];
}

__global__ void k_relu_grad(const float* __restrict__ dA, const bf16* __restrict__ A, std::size_t count,
                            bf16* __restrict__ dZ)
{
    std::size_t const i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    dZ[i] = static_cast<bf16>(static_cast<float>(A[i]) > 0.0f ? dA[i] : 0.0f);
}

__global__ void __launch_bounds__(256) k_colsum(const bf16* __restrict__ dZ, uint32_t rows, int cols,
                                                float* __restrict__ db)
{
    __shared__ float red[256];
    int const c = blockIdx.x;
    float s = 0.0f;
    for (uint32_t r = threadIdx.x; r < rows; r += blockDim.x) s += static_cast<float>(dZ[std::size_t(r) * cols + c]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       std::size_t count, float lr, float b1, float b2, float eps, float c1, float c2)
{
    std::size_t const i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float const gi = g[i];
    float const mi = b1 * m[i] + (1.0f - b1) * gi;
    float const vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

unsigned blocks_for(std::size_t n, int block) { return unsigned((n + std::size_t(block) - 1) / std::size_t(block)); }

}  // namespace

void launch_pack(hipStream_t s, const float* params, bf16* pack, bf16* wrow, float* bias)
{
    std::size_t const n = kPackTotal > kRTotal ? kPackTotal : kRTotal;
    k_pack<<<blocks_for(n, 256), 256, 0, s>>>(params, pack, wrow, bias);
}

void launch_infer(hipStream_t s, const bf16* pack, const float* bias, const bf16* X, uint32_t n, int num_outputs,
                  const float* out_scale, float* Y, int blocks)
{
    if (n == 0) return;
    k_infer<<<blocks, 256, 0, s>>>(reinterpret_cast<const bf16x8*>(pack), bias, reinterpret_cast<const bf16x8*>(X), n,
                                   num_outputs, out_scale, Y);
}

void launch_normalize(hipStream_t s, const float* raw, uint32_t ld, int num_inputs, const uint32_t* rows, uint32_t n,
                      uint32_t n_pad, const norm_dev* norm, bf16* X)
{
    if (n_pad == 0) return;
    k_normalize<<<blocks_for(n_pad, 256), 256, 0, s>>>(raw, ld, num_inputs, rows, n, n_pad, norm, X);
}

void launch_sample(hipStream_t s, const float* feats, const float* targets, uint64_t size, int num_inputs, uint32_t n,
                   uint32_t n_pad, uint64_t seed, uint64_t step, const norm_dev* norm, bf16* X, float* T)
{
    k_sample<<<blocks_for(n_pad, 256), 256, 0, s>>>(feats, targets, size, num_inputs, n, n_pad, seed, step, norm, X, T);
}

void launch_gemm(hipStream_t s, const bf16* A, long sam, long sak, const bf16* B, long sbk, long sbn, float* C, int M,
                 int N, int K, int splits)
{
    int const chunk = ((K + splits - 1) / splits + 15) / 16 * 16;
    int const z = (K + chunk - 1) / chunk;
    k_gemm<<<dim3((N + 31) / 32, (M + 31) / 32, z), 64, 0, s>>>(A, sam, sak, B, sbk, sbn, C, M, N, K, chunk);
}

void launch_reduce_splits(hipStream_t s, const float* parts, int splits, std::size_t count, float* out)
{
    k_reduce_splits<<<blocks_for(count, 256), 256, 0, s>>>(parts, splits, count, out);
}

void launch_bias_relu(hipStream_t s, const float* Z, const float* b, uint32_t rows, int cols, bf16* A)
{
    k_bias_relu<<<blocks_for(std::size_t(rows) * cols, 256), 256, 0, s>>>(Z, b, rows, cols, A);
}

void launch_output_grad(hipStream_t s, const float* Z, const float* b, const float* T, uint32_t n, uint32_t n_pad,
                        int num_outputs, bf16* dY, double* loss_parts, int loss_blocks)
{
    k_output_grad<<<loss_blocks, 256, 0, s>>>(Z, b, T, n, n_pad, num_outputs, dY, loss_parts);
}

void launch_relu_grad(hipStream_t s, const float* dA, const bf16* A, std::size_t count, bf16* dZ)
{
    k_relu_grad<<<blocks_for(count, 256), 256, 0, s>>>(dA, A, count, dZ);
}

void launch_colsum(hipStream_t s, const bf16* dZ, uint32_t rows, int cols, float* db)
{
    k_colsum<<<cols, 256, 0, s>>>(dZ, rows, cols, db);
}

void launch_adam(hipStream_t s, float* params, const float* grad, float* m, float* v, std::size_t count, float lr,
                 float beta1, float beta2, float eps, int t)
{
    float const c1 = 1.0f - powf(beta1, float(t)), c2 = 1.0f - powf(beta2, float(t));
    k_adam<<<blocks_for(count, 256), 256, 0, s>>>(params, grad, m, v, count, lr, beta1, beta2, eps, c1, c2);
}

}  // namespace mlp
}  // namespace sbs
