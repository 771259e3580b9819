// C ABI of the warm-start network (declared in include/sbs_gpu.h): parameters, normalization,
// the fused inference pass, one-batch loss gradients, the device-resident training set and Adam
// training on the GPU. All device work goes to the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <memory>
#include <random>
#include <vector>

#include "abi_util.h"
#include "mlp.h"
#include "mlp_handle.h"
#include "sbs_gpu.h"

using namespace sbs;
using namespace sbs::mlp;

namespace {

constexpr int kLossBlocks = 64;
constexpr int kInferBlocks = 512;  // 2 workgroups (of 4 waves) per CU, persistent
constexpr uint32_t kSplitRows = 512;  // rows per K-chunk of the weight-gradient products

uint32_t pad32(uint32_t n) { return (n + 31u) / 32u * 32u; }

}  // namespace

void sbs_mlp::reserve_work(uint32_t rows)
{
    rows = pad32(std::max<uint32_t>(rows, 32));
    if (rows <= work_rows) return;
    std::size_t const r = rows;
    X.alloc(r * kIn);
    A1.alloc(r * kHidden);
    A2.alloc(r * kHidden);
    A3.alloc(r * kHidden);
    dY.alloc(r * kOut);
    dZa.alloc(r * kHidden);
    dZb.alloc(r * kHidden);
    Z.alloc(r * kHidden);
    dA.alloc(r * kHidden);
    T.alloc(r * kMaxOutputs);
    Y.alloc(r * kMaxOutputs);
    raw.alloc(r * kMaxInputs);
    std::size_t const splits = (r + kSplitRows - 1) / kSplitRows;
    parts.alloc(splits * kHidden * kHidden);
    loss_parts.alloc(kLossBlocks);
    work_rows = rows;
}

double sbs_mlp::loss_grad(uint32_t n, uint32_t n_pad, bool want_loss)
{
    hipStream_t const s = stream;
    bf16 const* W1 = wrow.p;
    bf16 const* W2 = W1 + kIn * kHidden;
    bf16 const* W3 = W2 + kHidden * kHidden;
    bf16 const* W4 = W3 + kHidden * kHidden;
    float const* b = bias.p;
    int const M = int(n_pad);
    int const splits = int((n_pad + kSplitRows - 1) / kSplitRows);
    // forward: Z_l = A_{l-1} W_l, A_l = bf16(relu(Z_l + b_l))
    launch_gemm(s, X.p, kIn, 1, W1, kHidden, 1, Z.p, M, kHidden, kIn, 1);
    launch_bias_relu(s, Z.p, b, n_pad, kHidden, A1.p);
    launch_gemm(s, A1.p, kHidden, 1, W2, kHidden, 1, Z.p, M, kHidden, kHidden, 1);
    launch_bias_relu(s, Z.p, b + kHidden, n_pad, kHidden, A2.p);
    launch_gemm(s, A2.p, kHidden, 1, W3, kHidden, 1, Z.p, M, kHidden, kHidden, 1);
    launch_bias_relu(s, Z.p, b + 2 * kHidden, n_pad, kHidden, A3.p);
    launch_gemm(s, A3.p, kHidden, 1, W4, kOut, 1, Z.p, M, kOut, kHidden, 1);
    launch_output_grad(s, Z.p, b + 3 * kHidden, T.p, n, n_pad, int(nout), dY.p, loss_parts.p, kLossBlocks);
    // backward: dW_l = A_{l-1}^T dZ_l, db_l = column sums of dZ_l, dZ_{l-1} = (dZ_l W_l^T) * relu'
    float* g = grad.p;
    auto weight_grad = [&](bf16 const* Aprev, int in, bf16 const* dZ, int out, float* dst) {
        launch_gemm(s, Aprev, 1, in, dZ, out, 1, parts.p, in, out, M, splits);
        launch_reduce_splits(s, parts.p, splits, std::size_t(in) * out, dst);
    };
    weight_grad(A3.p, kHidden, dY.p, kOut, g + kW4);
    launch_colsum(s, dY.p, n_pad, kOut, g + kB4);
    launch_gemm(s, dY.p, kOut, 1, W4, 1, kOut, dA.p, M, kHidden, kOut, 1);
    launch_relu_grad(s, dA.p, A3.p, std::size_t(n_pad) * kHidden, dZa.p);
    weight_grad(A2.p, kHidden, dZa.p, kHidden, g + kW3);
    launch_colsum(s, dZa.p, n_pad, kHidden, g + kB3);
    launch_gemm(s, dZa.p, kHidden, 1, W3, 1, kHidden, dA.p, M, kHidden, kHidden, 1);
    launch_relu_grad(s, dA.p, A2.p, std::size_t(n_pad) * kHidden, dZb.p);
    weight_grad(A1.p, kHidden, dZb.p, kHidden, g + kW2);
    launch_colsum(s, dZb.p, n_pad, kHidden, g + kB2);
    launch_gemm(s, dZb.p, kHidden, 1, W2, 1, kHidden, dA.p, M, kHidden, kHidden, 1);
    launch_relu_grad(s, dA.p, A1.p, std::size_t(n_pad) * kHidden, dZa.p);
    weight_grad(X.p, kIn, dZa.p, kHidden, g + kW1);
    launch_colsum(s, dZa.p, n_pad, kHidden, g + kB1);
    hip_check(hipGetLastError(), "training kernels");
    if (!want_loss) return 0.0;
    std::vector<double> h(kLossBlocks);
    hip_check(hipMemcpyAsync(h.data(), loss_parts.p, kLossBlocks * sizeof(double), hipMemcpyDeviceToHost, s),
              "read loss");
    hip_check(hipStreamSynchronize(s), "loss");
    double acc = 0.0;
    for (double x : h) acc += x;
    return acc / (double(n) * double(nout));
}

extern "C" {

int sbs_mlp_num_params(uint32_t* count)
{
    return guarded([&] {
        require(count != nullptr, "count is null");
        *count = uint32_t(kParams);
    });
}

int sbs_mlp_create(int device, uint32_t num_inputs, uint32_t num_outputs, uint64_t seed, sbs_mlp** out)
{
    return guarded([&] {
        require(out != nullptr, "out is null");
        require(num_inputs >= 1 && num_inputs <= uint32_t(kMaxInputs), "num_inputs must be 1..40");
        require(num_outputs >= 1 && num_outputs <= uint32_t(kMaxOutputs), "num_outputs must be 1..12");
        *out = nullptr;
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        require(device >= 0 && device < count, "no such device", SBS_ERR_NO_DEVICE);
        auto net = std::make_unique<sbs_mlp>();
        net->device = device;
        net->nin = num_inputs;
        net->nout = num_outputs;
        net->set_current();
        hip_check(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking), "hipStreamCreate");
        // He initialization of the weights feeding ReLUs; the output layer starts small so that
        // an untrained network predicts (almost) no change.
        std::vector<float> p(kParams, 0.0f);
        std::mt19937_64 rng(seed);
        // rows of unused inputs and columns of unused outputs stay zero
        auto fill = [&](std::size_t off, int in, int outw, int used_out, double gain) {
            std::normal_distribution<double> nd(0.0, gain * std::sqrt(2.0 / double(in)));
            for (int i = 0; i < in; ++i)
                for (int o = 0; o < used_out; ++o) p[off + std::size_t(i) * outw + o] = float(nd(rng));
        };
        fill(kW1, int(num_inputs), kHidden, kHidden, 1.0);
        fill(kW2, kHidden, kHidden, kHidden, 1.0);
        fill(kW3, kHidden, kHidden, kHidden, 1.0);
        fill(kW4, kHidden, kOut, int(num_outputs), 0.1);
        net->params.upload(p.data(), p.size(), net->stream);
        net->grad.alloc(kParams);
        net->adam_m.alloc(kParams);
        net->adam_v.alloc(kParams);
        hip_check(hipMemsetAsync(net->grad.p, 0, kParams * sizeof(float), net->stream), "memset");
        hip_check(hipMemsetAsync(net->adam_m.p, 0, kParams * sizeof(float), net->stream), "memset");
        hip_check(hipMemsetAsync(net->adam_v.p, 0, kParams * sizeof(float), net->stream), "memset");
        net->pack.alloc(kPackTotal);
        net->wrow.alloc(kPackTotal);
        net->bias.alloc(kBiasTotal);
        net->norm.alloc(1);
        for (int k = 0; k < kIn; ++k) {
            net->hnorm.mean[k] = 0.0f;
            net->hnorm.inv_std[k] = 1.0f;
        }
        for (int o = 0; o < kMaxOutputs; ++o) net->hnorm.out_scale[o] = 1.0f;
        net->upload_norm();
        net->repack();
        hip_check(hipStreamSynchronize(net->stream), "create");
        *out = net.release();
    });
}

int sbs_mlp_destroy(sbs_mlp* net)
{
    return guarded([&] { delete net; });
}

int sbs_mlp_get_params(sbs_mlp* net, float* params)
{
    return guarded([&] {
        require(net != nullptr && params != nullptr, "null argument");
        net->set_current();
        hip_check(hipMemcpyAsync(params, net->params.p, kParams * sizeof(float), hipMemcpyDeviceToHost, net->stream),
                  "read params");
        hip_check(hipStreamSynchronize(net->stream), "read params");
    });
}

int sbs_mlp_set_params(sbs_mlp* net, const float* params)
{
    return guarded([&] {
        require(net != nullptr && params != nullptr, "null argument");
        net->set_current();
        hip_check(hipMemcpyAsync(net->params.p, params, kParams * sizeof(float), hipMemcpyHostToDevice, net->stream),
                  "write params");
        net->repack();
        hip_check(hipStreamSynchronize(net->stream), "write params");
    });
}

int sbs_mlp_set_normalization(sbs_mlp* net, const float* mean, const float* inv_std, const float* out_scale)
{
    return guarded([&] {
        require(net != nullptr && mean != nullptr && inv_std != nullptr && out_scale != nullptr, "null argument");
        for (uint32_t o = 0; o < net->nout; ++o) require(out_scale[o] != 0.0f, "out_scale must be nonzero");
        for (uint32_t k = 0; k < net->nin; ++k) {
            net->hnorm.mean[k] = mean[k];
            net->hnorm.inv_std[k] = inv_std[k];
        }
        for (uint32_t o = 0; o < net->nout; ++o) net->hnorm.out_scale[o] = out_scale[o];
        net->set_current();
        net->upload_norm();
        hip_check(hipStreamSynchronize(net->stream), "normalization");
    });
}

int sbs_mlp_get_normalization(sbs_mlp* net, float* mean, float* inv_std, float* out_scale)
{
    return guarded([&] {
        require(net != nullptr, "net is null");
        for (uint32_t k = 0; k < net->nin; ++k) {
            if (mean) mean[k] = net->hnorm.mean[k];
            if (inv_std) inv_std[k] = net->hnorm.inv_std[k];
        }
        for (uint32_t o = 0; o < net->nout; ++o)
            if (out_scale) out_scale[o] = net->hnorm.out_scale[o];
    });
}

int sbs_mlp_forward(sbs_mlp* net, const float* x, uint32_t n, float* y)
{
    return guarded([&] {
        require(net != nullptr && (n == 0 || (x != nullptr && y != nullptr)), "null argument");
        if (n == 0) return;
        net->set_current();
        net->reserve_work(n);
        uint32_t const n_pad = pad32(n);
        dev_buf<float> xin;
        xin.upload(x, std::size_t(n) * net->nin, net->stream);
        launch_normalize(net->stream, xin.p, net->nin, int(net->nin), nullptr, n, n_pad, net->norm.p, net->X.p);
        launch_infer(net->stream, net->pack.p, net->bias.p, net->X.p, n, int(net->nout), net->out_scale_dev(), net->Y.p,
                     kInferBlocks);
        hip_check(hipGetLastError(), "inference kernel");
        std::vector<float> h(std::size_t(n) * kMaxOutputs);
        hip_check(hipMemcpyAsync(h.data(), net->Y.p, h.size() * sizeof(float), hipMemcpyDeviceToHost, net->stream),
                  "read outputs");
        hip_check(hipStreamSynchronize(net->stream), "forward");
        for (uint32_t r = 0; r < n; ++r)
            for (uint32_t o = 0; o < net->nout; ++o) y[std::size_t(r) * net->nout + o] = h[std::size_t(r) * kMaxOutputs + o];
    });
}

int sbs_mlp_loss_grad(sbs_mlp* net, const float* x, const float* t, uint32_t n, double* loss, float* grad)
{
    return guarded([&] {
        require(net != nullptr && x != nullptr && t != nullptr && n > 0, "null or empty argument");
        net->set_current();
        net->reserve_work(n);
        uint32_t const n_pad = pad32(n);
        dev_buf<float> xin;
        xin.upload(x, std::size_t(n) * net->nin, net->stream);
        launch_normalize(net->stream, xin.p, net->nin, int(net->nin), nullptr, n, n_pad, net->norm.p, net->X.p);
        std::vector<float> th(std::size_t(n) * kMaxOutputs, 0.0f);
        for (uint32_t r = 0; r < n; ++r)
            for (uint32_t o = 0; o < net->nout; ++o)
                th[std::size_t(r) * kMaxOutputs + o] = t[std::size_t(r) * net->nout + o] / net->hnorm.out_scale[o];
        hip_check(hipMemcpyAsync(net->T.p, th.data(), th.size() * sizeof(float), hipMemcpyHostToDevice, net->stream),
                  "upload targets");
        double const l = net->loss_grad(n, n_pad, true);
        if (loss) *loss = l;
        if (grad) {
            hip_check(hipMemcpyAsync(grad, net->grad.p, kParams * sizeof(float), hipMemcpyDeviceToHost, net->stream),
                      "read gradient");
            hip_check(hipStreamSynchronize(net->stream), "gradient");
        }
    });
}

int sbs_mlp_dataset_reserve(sbs_mlp* net, uint64_t rows)
{
    return guarded([&] {
        require(net != nullptr, "net is null");
        if (rows <= net->ds_cap) return;
        net->set_current();
        dev_buf<float> nx, nt;
        nx.alloc(rows * kMaxInputs);
        nt.alloc(rows * kMaxOutputs);
        if (net->ds_size) {
            hip_check(hipMemcpyAsync(nx.p, net->ds_x.p, net->ds_size * kMaxInputs * sizeof(float),
                                     hipMemcpyDeviceToDevice, net->stream), "grow dataset");
            hip_check(hipMemcpyAsync(nt.p, net->ds_t.p, net->ds_size * kMaxOutputs * sizeof(float),
                                     hipMemcpyDeviceToDevice, net->stream), "grow dataset");
            hip_check(hipStreamSynchronize(net->stream), "grow dataset");
        }
        net->ds_x = std::move(nx);
        net->ds_t = std::move(nt);
        net->ds_cap = rows;
    });
}

int sbs_mlp_dataset_append(sbs_mlp* net, const float* x, const float* t, uint64_t rows)
{
    return guarded([&] {
        require(net != nullptr && (rows == 0 || (x != nullptr && t != nullptr)), "null argument");
        require(net->ds_size + rows <= net->ds_cap, "dataset capacity exceeded (sbs_mlp_dataset_reserve)");
        net->set_current();
        std::vector<float> hx(rows * kMaxInputs, 0.0f), ht(rows * kMaxOutputs, 0.0f);
        for (uint64_t r = 0; r < rows; ++r) {
            for (uint32_t k = 0; k < net->nin; ++k) hx[r * kMaxInputs + k] = x[r * net->nin + k];
            for (uint32_t o = 0; o < net->nout; ++o) ht[r * kMaxOutputs + o] = t[r * net->nout + o];
        }
        hip_check(hipMemcpyAsync(net->ds_x.p + net->ds_size * kMaxInputs, hx.data(), hx.size() * sizeof(float),
                                 hipMemcpyHostToDevice, net->stream), "append dataset");
        hip_check(hipMemcpyAsync(net->ds_t.p + net->ds_size * kMaxOutputs, ht.data(), ht.size() * sizeof(float),
                                 hipMemcpyHostToDevice, net->stream), "append dataset");
        hip_check(hipStreamSynchronize(net->stream), "append dataset");
        net->ds_size += rows;
    });
}

int sbs_mlp_dataset_size(sbs_mlp* net, uint64_t* rows)
{
    return guarded([&] {
        require(net != nullptr && rows != nullptr, "null argument");
        *rows = net->ds_size;
    });
}

int sbs_mlp_dataset_read(sbs_mlp* net, uint64_t first, uint64_t rows, float* x, float* t)
{
    return guarded([&] {
        require(net != nullptr, "net is null");
        require(first + rows <= net->ds_size, "rows past the end of the dataset");
        if (rows == 0) return;
        net->set_current();
        std::vector<float> hx(rows * kMaxInputs), ht(rows * kMaxOutputs);
        hip_check(hipMemcpyAsync(hx.data(), net->ds_x.p + first * kMaxInputs, hx.size() * sizeof(float),
                                 hipMemcpyDeviceToHost, net->stream), "read dataset");
        hip_check(hipMemcpyAsync(ht.data(), net->ds_t.p + first * kMaxOutputs, ht.size() * sizeof(float),
                                 hipMemcpyDeviceToHost, net->stream), "read dataset");
        hip_check(hipStreamSynchronize(net->stream), "read dataset");
        for (uint64_t r = 0; r < rows; ++r) {
            if (x)
                for (uint32_t k = 0; k < net->nin; ++k) x[r * net->nin + k] = hx[r * kMaxInputs + k];
            if (t)
                for (uint32_t o = 0; o < net->nout; ++o) t[r * net->nout + o] = ht[r * kMaxOutputs + o];
        }
    });
}

int sbs_mlp_dataset_clear(sbs_mlp* net)
{
    return guarded([&] {
        require(net != nullptr, "net is null");
        net->ds_size = 0;
    });
}

int sbs_mlp_fit_normalization(sbs_mlp* net)
{
    return guarded([&] {
        require(net != nullptr, "net is null");
        require(net->ds_size > 0, "dataset is empty", SBS_ERR_STATE);
        net->set_current();
        // statistics over an evenly strided sample of at most 2^20 rows
        uint64_t const m = std::min<uint64_t>(net->ds_size, 1u << 20);
        uint64_t const stride = net->ds_size / m;
        std::vector<float> sx(m * kMaxInputs), st(m * kMaxOutputs);
        hip_check(hipMemcpy2DAsync(sx.data(), kMaxInputs * sizeof(float), net->ds_x.p, stride * kMaxInputs * sizeof(float),
                                   kMaxInputs * sizeof(float), m, hipMemcpyDeviceToHost, net->stream), "read dataset");
        hip_check(hipMemcpy2DAsync(st.data(), kMaxOutputs * sizeof(float), net->ds_t.p,
                                   stride * kMaxOutputs * sizeof(float), kMaxOutputs * sizeof(float), m,
                                   hipMemcpyDeviceToHost, net->stream), "read dataset");
        hip_check(hipStreamSynchronize(net->stream), "read dataset");
        for (uint32_t k = 0; k < net->nin; ++k) {
            double s = 0, s2 = 0;
            for (uint64_t i = 0; i < m; ++i) {
                double const v = sx[i * kMaxInputs + k];
                s += v;
                s2 += v * v;
            }
            double const mean = s / double(m);
            double const var = std::max(0.0, s2 / double(m) - mean * mean);
            double const sd = std::sqrt(var);
            net->hnorm.mean[k] = float(mean);
            net->hnorm.inv_std[k] = float(sd > 1e-12 * std::max(1.0, std::fabs(mean)) && sd > 0 ? 1.0 / sd : 1.0);
        }
        for (uint32_t o = 0; o < net->nout; ++o) {
            double s2 = 0;
            for (uint64_t i = 0; i < m; ++i) s2 += double(st[i * kMaxOutputs + o]) * st[i * kMaxOutputs + o];
            double const rms = std::sqrt(s2 / double(m));
            net->hnorm.out_scale[o] = float(rms > 0 ? rms : 1.0);
        }
        net->upload_norm();
        hip_check(hipStreamSynchronize(net->stream), "normalization");
    });
}

int sbs_mlp_train(sbs_mlp* net, uint32_t steps, uint32_t batch, float lr, uint64_t seed, double* loss)
{
    return guarded([&] {
        require(net != nullptr, "net is null");
        require(net->ds_size > 0, "dataset is empty", SBS_ERR_STATE);
        require(batch >= 1 && batch <= (1u << 20), "batch must be 1..2^20");
        require(lr > 0.0f, "lr must be positive");
        net->set_current();
        net->reserve_work(batch);
        uint32_t const n_pad = pad32(batch);
        double last = 0.0;
        for (uint32_t it = 0; it < steps; ++it) {
            launch_sample(net->stream, net->ds_x.p, net->ds_t.p, net->ds_size, int(net->nin), batch, n_pad, seed,
                          uint64_t(net->adam_t), net->norm.p, net->X.p, net->T.p);
            bool const want = loss != nullptr && it + 1 == steps;
            last = net->loss_grad(batch, n_pad, want);
            ++net->adam_t;
            launch_adam(net->stream, net->params.p, net->grad.p, net->adam_m.p, net->adam_v.p, kParams, lr, 0.9f, 0.999f,
                        1e-8f, net->adam_t);
            net->repack();
        }
        hip_check(hipGetLastError(), "training kernels");
        hip_check(hipStreamSynchronize(net->stream), "train");
        if (loss) *loss = last;
    });
}

}  // extern "C"
