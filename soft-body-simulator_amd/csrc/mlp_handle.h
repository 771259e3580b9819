// The sbs_mlp handle (shared by the network's C ABI in mlp.cpp and the warm start of the time
// step in sbs_gpu.cpp).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "abi_util.h"
#include "mlp.h"

struct sbs_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t nin = 0, nout = 0;

    sbs::dev_buf<float> params, grad, adam_m, adam_v;
    sbs::dev_buf<sbs::mlp::bf16> pack, wrow;
    sbs::dev_buf<float> bias;
    sbs::dev_buf<sbs::mlp::norm_dev> norm;
    sbs::mlp::norm_dev hnorm{};
    int adam_t = 0;

    // training rows: features [cap][13], targets [cap][4] (FP32, raw units)
    sbs::dev_buf<float> ds_x, ds_t;
    uint64_t ds_size = 0, ds_cap = 0;

    // work buffers of the training passes, sized for `work_rows` (a multiple of 32) rows
    uint32_t work_rows = 0;
    sbs::dev_buf<sbs::mlp::bf16> X, A1, A2, A3, dY, dZa, dZb;
    sbs::dev_buf<float> Z, dA, T, parts, Y, raw;
    sbs::dev_buf<double> loss_parts;

    ~sbs_mlp()
    {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }

    const float* out_scale_dev() const
    {
        return reinterpret_cast<const float*>(reinterpret_cast<const char*>(norm.p) +
                                              offsetof(sbs::mlp::norm_dev, out_scale));
    }
    void set_current() const { sbs::hip_check(hipSetDevice(device), "hipSetDevice"); }
    // refreshes the BF16 copies after the FP32 parameters changed
    void repack() { sbs::mlp::launch_pack(stream, params.p, pack.p, wrow.p, bias.p); }
    void upload_norm()
    {
        sbs::hip_check(hipMemcpyAsync(norm.p, &hnorm, sizeof(hnorm), hipMemcpyHostToDevice, stream), "upload norm");
    }
    void reserve_work(uint32_t rows);
    // One forward + backward pass over X/T rows [0, n) (n_pad = work rows in use): fills grad and
    // returns the mean squared error (on the host, after a stream sync when want_loss).
    double loss_grad(uint32_t n, uint32_t n_pad, bool want_loss);
};
