// Host side of the warm-start MLP: parameters, activations, full-batch Adam
// training and inference on the FP32 kernels of rie_mlp.hip.
//
// Layout of the packed parameters (row-major, one sample per row):
//   W1 [in x hidden], b1 [hidden], W2 [hidden x hidden], b2 [hidden],
//   W3 [hidden x out], b3 [out].
// Initialisation follows torch.nn.Linear's default, U(-1/sqrt(fan_in),
// 1/sqrt(fan_in)) for weights and biases, from a seeded host generator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <random>
#include <vector>

#include "rie.h"
#include "rie_engine_internal.hpp"
#include "rie_kernels.hpp"

namespace rie {

struct Mlp {
    Mlp(int in, int hidden, int out, uint64_t seed);
    ~Mlp();
    Mlp(const Mlp &) = delete;
    Mlp &operator=(const Mlp &) = delete;

    int in;
    int hidden;
    int out;
    size_t count = 0;
    size_t o_w1 = 0, o_b1 = 0, o_w2 = 0, o_b2 = 0, o_w3 = 0, o_b3 = 0;
    float *d_p = nullptr;
    float *d_g = nullptr;
    float *d_m = nullptr;
    float *d_v = nullptr;
    hipStream_t s = nullptr;

    int cap = 0;  // rows the activation buffers hold
    float *d_x = nullptr, *d_t = nullptr;
    float *d_z1 = nullptr, *d_a1 = nullptr, *d_z2 = nullptr, *d_a2 = nullptr, *d_y = nullptr;
    float *d_dy = nullptr, *d_d2 = nullptr, *d_d1 = nullptr;
    float *d_partials = nullptr;
    float *d_loss = nullptr;
    int mse_blocks = 0;

    void ensure_rows(int rows);
    void upload_rows(const float *x, const float *t, int rows);
    void forward(int rows);
    void backward(int rows);  // after forward; loss into d_loss, gradient into d_g
    void free_rows();
};

Mlp::Mlp(int in_, int hidden_, int out_, uint64_t seed) : in(in_), hidden(hidden_), out(out_) {
    if (in < 1 || hidden < 1 || out < 1 || in > 4096 || hidden > 4096 || out > 4096) {
        fail(RIE_INVALID_ARGUMENT, "MLP sizes must be in [1, 4096].");
    }
    o_w1 = 0;
    o_b1 = o_w1 + (size_t)in * hidden;
    o_w2 = o_b1 + hidden;
    o_b2 = o_w2 + (size_t)hidden * hidden;
    o_w3 = o_b2 + hidden;
    o_b3 = o_w3 + (size_t)hidden * out;
    count = o_b3 + out;
    std::vector<float> h(count);
    std::mt19937_64 gen(seed);
    auto fill = [&](size_t off, size_t len, int fan_in) {
        const float bound = 1.0f / std::sqrt((float)fan_in);
        std::uniform_real_distribution<float> u(-bound, bound);
        for (size_t i = 0; i < len; ++i) {
            h[off + i] = u(gen);
        }
    };
    fill(o_w1, (size_t)in * hidden, in);
    fill(o_b1, hidden, in);
    fill(o_w2, (size_t)hidden * hidden, hidden);
    fill(o_b2, hidden, hidden);
    fill(o_w3, (size_t)hidden * out, hidden);
    fill(o_b3, out, hidden);
    check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    d_p = (float *)device_alloc(sizeof(float) * count, "alloc MLP params");
    d_g = (float *)device_alloc(sizeof(float) * count, "alloc MLP gradient");
    d_m = (float *)device_alloc(sizeof(float) * count, "alloc MLP adam m");
    d_v = (float *)device_alloc(sizeof(float) * count, "alloc MLP adam v");
    d_loss = (float *)device_alloc(sizeof(float), "alloc MLP loss");
    check(hipMemcpy(d_p, h.data(), sizeof(float) * count, hipMemcpyHostToDevice), "upload MLP");
}

void Mlp::free_rows() {
    for (float *p : {d_x, d_t, d_z1, d_a1, d_z2, d_a2, d_y, d_dy, d_d2, d_d1, d_partials}) {
        device_free(p);
    }
    d_x = d_t = d_z1 = d_a1 = d_z2 = d_a2 = d_y = d_dy = d_d2 = d_d1 = d_partials = nullptr;
    cap = 0;
}

Mlp::~Mlp() {
    if (s != nullptr) {
        (void)hipStreamSynchronize(s);
    }
    free_rows();
    for (float *p : {d_p, d_g, d_m, d_v, d_loss}) {
        device_free(p);
    }
    if (s != nullptr) {
        (void)hipStreamDestroy(s);
    }
}

void Mlp::ensure_rows(int rows) {
    if (rows < 1 || rows > (1 << 22)) {
        fail(RIE_INVALID_ARGUMENT, "MLP batches need 1 <= rows <= 2^22.");
    }
    if (rows <= cap) {
        return;
    }
    check(hipStreamSynchronize(s), "sync");
    free_rows();
    const size_t r = (size_t)rows;
    d_x = (float *)device_alloc(sizeof(float) * r * in, "alloc MLP rows");
    d_t = (float *)device_alloc(sizeof(float) * r * out, "alloc MLP rows");
    d_z1 = (float *)device_alloc(sizeof(float) * r * hidden, "alloc MLP rows");
    d_a1 = (float *)device_alloc(sizeof(float) * r * hidden, "alloc MLP rows");
    d_z2 = (float *)device_alloc(sizeof(float) * r * hidden, "alloc MLP rows");
    d_a2 = (float *)device_alloc(sizeof(float) * r * hidden, "alloc MLP rows");
    d_y = (float *)device_alloc(sizeof(float) * r * out, "alloc MLP rows");
    d_dy = (float *)device_alloc(sizeof(float) * r * out, "alloc MLP rows");
    d_d2 = (float *)device_alloc(sizeof(float) * r * hidden, "alloc MLP rows");
    d_d1 = (float *)device_alloc(sizeof(float) * r * hidden, "alloc MLP rows");
    mse_blocks = (int)std::min<size_t>(1024, (r * out + 255) / 256);
    d_partials = (float *)device_alloc(sizeof(float) * mse_blocks, "alloc MLP partials");
    cap = rows;
}

void Mlp::upload_rows(const float *x, const float *t, int rows) {
    ensure_rows(rows);
    check(hipMemcpyAsync(d_x, x, sizeof(float) * rows * in, hipMemcpyHostToDevice, s),
          "upload MLP inputs");
    if (t != nullptr) {
        check(hipMemcpyAsync(d_t, t, sizeof(float) * rows * out, hipMemcpyHostToDevice, s),
              "upload MLP targets");
    }
}

void Mlp::forward(int rows) {
    launch_mlp_gemm(false, false, kMlpEpiBiasGelu, d_x, d_p + o_w1, d_z1, rows, hidden, in,
                    d_p + o_b1, d_a1, nullptr, s);
    launch_mlp_gemm(false, false, kMlpEpiBiasGelu, d_a1, d_p + o_w2, d_z2, rows, hidden, hidden,
                    d_p + o_b2, d_a2, nullptr, s);
    launch_mlp_gemm(false, false, kMlpEpiBias, d_a2, d_p + o_w3, d_y, rows, out, hidden,
                    d_p + o_b3, nullptr, nullptr, s);
    check_launch("MLP forward");
}

void Mlp::backward(int rows) {
    // loss = mean over all rows * out elements of (y - t)^2
    launch_mlp_mse(d_y, d_t, rows * out, d_dy, d_partials, mse_blocks, d_loss, s);
    // layer 3: dW3 = A2^T dY, db3 = colsum dY, dZ2 = (dY W3^T) * GELU'(Z2)
    launch_mlp_gemm(true, false, kMlpEpiNone, d_a2, d_dy, d_g + o_w3, hidden, out, rows, nullptr,
                    nullptr, nullptr, s);
    launch_mlp_colsum(d_dy, rows, out, d_g + o_b3, s);
    launch_mlp_gemm(false, true, kMlpEpiDGelu, d_dy, d_p + o_w3, d_d2, rows, hidden, out, nullptr,
                    nullptr, d_z2, s);
    // layer 2
    launch_mlp_gemm(true, false, kMlpEpiNone, d_a1, d_d2, d_g + o_w2, hidden, hidden, rows,
                    nullptr, nullptr, nullptr, s);
    launch_mlp_colsum(d_d2, rows, hidden, d_g + o_b2, s);
    launch_mlp_gemm(false, true, kMlpEpiDGelu, d_d2, d_p + o_w2, d_d1, rows, hidden, hidden,
                    nullptr, nullptr, d_z1, s);
    // layer 1
    launch_mlp_gemm(true, false, kMlpEpiNone, d_x, d_d1, d_g + o_w1, in, hidden, rows, nullptr,
                    nullptr, nullptr, s);
    launch_mlp_colsum(d_d1, rows, hidden, d_g + o_b1, s);
    check_launch("MLP backward");
}

}  // namespace rie

// ===========================================================================
// C ABI
// ===========================================================================

using rie::guard;
using rie::Mlp;

namespace {

Mlp *as_mlp(rie_mlp *m) {
    if (m == nullptr) {
        rie::fail(RIE_INVALID_ARGUMENT, "null MLP");
    }
    return reinterpret_cast<Mlp *>(m);
}

}  // namespace

extern "C" {

rie_status rie_mlp_create(int inputs, int hidden, int outputs, uint64_t seed, rie_mlp **mlp_out) {
    return guard([&]() {
        if (mlp_out == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        *mlp_out = nullptr;
        auto m = std::make_unique<Mlp>(inputs, hidden, outputs, seed);
        *mlp_out = reinterpret_cast<rie_mlp *>(m.release());
    });
}

void rie_mlp_destroy(rie_mlp *mlp) {
    if (mlp != nullptr) {
        delete reinterpret_cast<Mlp *>(mlp);
    }
}

size_t rie_mlp_num_params(const rie_mlp *mlp) {
    return mlp == nullptr ? 0 : reinterpret_cast<const Mlp *>(mlp)->count;
}

rie_status rie_mlp_get_params(rie_mlp *mlp, float *params_out, size_t count) {
    return guard([&]() {
        Mlp *m = as_mlp(mlp);
        if (params_out == nullptr || count != m->count) {
            rie::fail(RIE_INVALID_ARGUMENT, "count does not match the MLP.");
        }
        rie::check(hipStreamSynchronize(m->s), "sync");
        rie::check(hipMemcpy(params_out, m->d_p, sizeof(float) * count, hipMemcpyDeviceToHost),
                   "download MLP params");
    });
}

rie_status rie_mlp_set_params(rie_mlp *mlp, const float *params, size_t count) {
    return guard([&]() {
        Mlp *m = as_mlp(mlp);
        if (params == nullptr || count != m->count) {
            rie::fail(RIE_INVALID_ARGUMENT, "count does not match the MLP.");
        }
        rie::check(hipStreamSynchronize(m->s), "sync");
        rie::check(hipMemcpy(m->d_p, params, sizeof(float) * count, hipMemcpyHostToDevice),
                   "upload MLP params");
    });
}

rie_status rie_mlp_forward(rie_mlp *mlp, const float *inputs, int rows, float *outputs) {
    return guard([&]() {
        Mlp *m = as_mlp(mlp);
        if (inputs == nullptr || outputs == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        m->upload_rows(inputs, nullptr, rows);
        m->forward(rows);
        rie::check(hipMemcpyAsync(outputs, m->d_y, sizeof(float) * rows * m->out,
                                  hipMemcpyDeviceToHost, m->s),
                   "download MLP outputs");
        rie::check(hipStreamSynchronize(m->s), "sync");
    });
}

rie_status rie_mlp_loss_and_grad(rie_mlp *mlp, const float *inputs, const float *targets,
                                 int rows, float *loss_out, float *grad_out) {
    return guard([&]() {
        Mlp *m = as_mlp(mlp);
        if (inputs == nullptr || targets == nullptr || loss_out == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        m->upload_rows(inputs, targets, rows);
        m->forward(rows);
        m->backward(rows);
        rie::check(hipMemcpyAsync(loss_out, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, m->s),
                   "download loss");
        if (grad_out != nullptr) {
            rie::check(hipMemcpyAsync(grad_out, m->d_g, sizeof(float) * m->count,
                                      hipMemcpyDeviceToHost, m->s),
                       "download gradient");
        }
        rie::check(hipStreamSynchronize(m->s), "sync");
    });
}

rie_status rie_mlp_train(rie_mlp *mlp, const float *inputs, const float *targets, int rows,
                         int steps, float learning_rate, float *loss_history_out) {
    return guard([&]() {
        Mlp *m = as_mlp(mlp);
        if (inputs == nullptr || targets == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        if (steps < 0 || !std::isfinite(learning_rate)) {
            rie::fail(RIE_INVALID_ARGUMENT, "steps must be >= 0 and the learning rate finite.");
        }
        m->upload_rows(inputs, targets, rows);
        rie::check(hipMemsetAsync(m->d_m, 0, sizeof(float) * m->count, m->s), "clear adam m");
        rie::check(hipMemsetAsync(m->d_v, 0, sizeof(float) * m->count, m->s), "clear adam v");
        std::vector<float> hist;
        float *d_hist = nullptr;
        if (loss_history_out != nullptr && steps > 0) {
            d_hist = (float *)rie::device_alloc(sizeof(float) * steps, "alloc loss history");
        }
        try {
            for (int t = 1; t <= steps; ++t) {
                m->forward(rows);
                m->backward(rows);
                if (d_hist != nullptr) {
                    rie::check(hipMemcpyAsync(d_hist + (t - 1), m->d_loss, sizeof(float),
                                              hipMemcpyDeviceToDevice, m->s),
                               "record loss");
                }
                rie::launch_mlp_adam(m->d_p, m->d_g, m->d_m, m->d_v, (int)m->count,
                                     learning_rate, 0.9f, 0.999f, 1e-8f, t, m->s);
                rie::check_launch("MLP adam");
            }
            if (d_hist != nullptr) {
                rie::check(hipMemcpyAsync(loss_history_out, d_hist, sizeof(float) * steps,
                                          hipMemcpyDeviceToHost, m->s),
                           "download loss history");
            }
            rie::check(hipStreamSynchronize(m->s), "sync");
        } catch (...) {
            (void)hipStreamSynchronize(m->s);
            rie::device_free(d_hist);
            throw;
        }
        rie::device_free(d_hist);
    });
}

}  // extern "C"
