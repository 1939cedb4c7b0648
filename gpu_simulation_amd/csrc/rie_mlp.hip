// FP32 multilayer perceptron kernels for the learned warm start: training
// (forward, backward, Adam) and inference of x -> GELU(x W1 + b1) ->
// GELU(. W2 + b2) -> . W3 + b3, all row-major with one sample per row.
//
// No counterpart in the reference.  The GEMMs are small (a few thousand rows
// by 256), so one LDS-tiled FP32 kernel with fused epilogues serves every
// product of the forward and backward pass:
//   forward  Z = X W + b, A = GELU(Z)           (epilogue kBiasGelu / kBias)
//   backward dA = dZn W^T, dZ = dA * GELU'(Z)     (transposed B, epilogue kDGelu)
//            dW = A^T dZ                          (transposed A)
// GELU is the exact form 0.5 x (1 + erf(x / sqrt 2)) (torch.nn.GELU's
// default), evaluated with the device library's erff / expf.
#include <hip/hip_runtime.h>

#include "rie_kernels.hpp"

namespace rie {

namespace {

constexpr int kTile = 64;
constexpr int kTK = 16;

__device__ __forceinline__ float gelu(float x) {
    return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = expf(-0.5f * x * x) * 0.39894228040143268f;
    return cdf + x * pdf;
}

}  // namespace

// C[M x N] = op(A)[M x K] * op(B)[K x N]; op(A) = A (M x K, lda) or A^T (A is
// K x M); op(B) = B (K x N) or B^T (B is N x K).  256 threads, 4 x 4 outputs
// each over a 64 x 64 tile.
template <bool kTA, bool kTB, int kEpi>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(const float *__restrict__ A,
                                                       const float *__restrict__ B,
                                                       float *__restrict__ C, int M, int N, int K,
                                                       const float *__restrict__ bias,
                                                       float *__restrict__ act,
                                                       const float *__restrict__ zin) {
    __shared__ float as[kTK][kTile + 1];  // as[k][m]
    __shared__ float bs[kTK][kTile + 1];  // bs[k][n]
    const int t = threadIdx.x;
    const int tm = (t / 16) * 4;
    const int tn = (t % 16) * 4;
    const int m0 = blockIdx.y * kTile;
    const int n0 = blockIdx.x * kTile;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            acc[i][j] = 0.0f;
        }
    }
    for (int k0 = 0; k0 < K; k0 += kTK) {
        for (int e = t; e < kTK * kTile; e += 256) {
            const int kk = e / kTile;
            const int mm = e % kTile;
            const int gm = m0 + mm;
            const int gk = k0 + kk;
            float va = 0.0f;
            if (gm < M && gk < K) {
                va = kTA ? A[(size_t)gk * M + gm] : A[(size_t)gm * K + gk];
            }
            as[kk][mm] = va;
            const int gn = n0 + mm;
            float vb = 0.0f;
            if (gn < N && gk < K) {
                vb = kTB ? B[(size_t)gn * K + gk] : B[(size_t)gk * N + gn];
            }
            bs[kk][mm] = vb;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTK; ++kk) {
            float a[4];
            float b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = as[kk][tm + i];
                b[i] = bs[kk][tn + i];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gm = m0 + tm + i;
        if (gm >= M) {
            continue;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gn = n0 + tn + j;
            if (gn >= N) {
                continue;
            }
            const size_t o = (size_t)gm * N + gn;
            float v = acc[i][j];
            if (kEpi == kMlpEpiBias || kEpi == kMlpEpiBiasGelu) {
                v += bias[gn];
            }
            if (kEpi == kMlpEpiDGelu) {
                v *= gelu_grad(zin[o]);
            }
            C[o] = v;
            if (kEpi == kMlpEpiBiasGelu) {
                act[o] = gelu(v);
            }
        }
    }
}

void launch_mlp_gemm(bool ta, bool tb, int epi, const float *A, const float *B, float *C, int M,
                     int N, int K, const float *bias, float *act, const float *zin,
                     hipStream_t s) {
    dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
#define RIE_MLP_GEMM(TA, TB, EPI)                                                           \
    mlp_gemm_kernel<TA, TB, EPI><<<grid, 256, 0, s>>>(A, B, C, M, N, K, bias, act, zin)
    if (!ta && !tb) {
        if (epi == kMlpEpiBias) {
            RIE_MLP_GEMM(false, false, kMlpEpiBias);
        } else if (epi == kMlpEpiBiasGelu) {
            RIE_MLP_GEMM(false, false, kMlpEpiBiasGelu);
        } else {
            RIE_MLP_GEMM(false, false, kMlpEpiNone);
        }
    } else if (!ta && tb) {
        if (epi == kMlpEpiDGelu) {
            RIE_MLP_GEMM(false, true, kMlpEpiDGelu);
        } else {
            RIE_MLP_GEMM(false, true, kMlpEpiNone);
        }
    } else {
        RIE_MLP_GEMM(true, false, kMlpEpiNone);
    }
#undef RIE_MLP_GEMM
}

// out[j] = sum_i X[i][j] over rows (fixed order: one thread per column,
// rows in order), for bias gradients.
__global__ void mlp_colsum_kernel(const float *__restrict__ X, int rows, int cols,
                                  float *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) {
        return;
    }
    float acc = 0.0f;
    for (int i = 0; i < rows; ++i) {
        acc += X[(size_t)i * cols + j];
    }
    out[j] = acc;
}

void launch_mlp_colsum(const float *X, int rows, int cols, float *out, hipStream_t s) {
    mlp_colsum_kernel<<<(cols + 127) / 128, 128, 0, s>>>(X, rows, cols, out);
}

// dY = (2 / count) (Y - T) and per-block partial sums of (Y - T)^2.
__global__ __launch_bounds__(256) void mlp_mse_kernel(const float *__restrict__ Y,
                                                      const float *__restrict__ T, int count,
                                                      float *__restrict__ dY,
                                                      float *__restrict__ partials) {
    __shared__ float red[4];
    float acc = 0.0f;
    const float scale = 2.0f / (float)count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        dY[i] = scale * d;
        acc += d * d;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        acc += __shfl_xor(acc, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    }
}

__global__ void mlp_loss_kernel(const float *__restrict__ partials, int n, int count,
                                float *__restrict__ loss_out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float acc = 0.0f;
        for (int i = 0; i < n; ++i) {
            acc += partials[i];
        }
        loss_out[0] = acc / (float)count;
    }
}

void launch_mlp_mse(const float *Y, const float *T, int count, float *dY, float *partials,
                    int blocks, float *loss_out, hipStream_t s) {
    mlp_mse_kernel<<<blocks, 256, 0, s>>>(Y, T, count, dY, partials);
    mlp_loss_kernel<<<1, 64, 0, s>>>(partials, blocks, count, loss_out);
}

// Adam on the packed FP32 parameters (torch.optim.Adam arithmetic), step t.
__global__ void mlp_adam_kernel(float *__restrict__ p, const float *__restrict__ g,
                                float *__restrict__ m, float *__restrict__ v, int count, float lr,
                                float b1, float b2, float eps, int t) {
    const float bc1 = 1.0f - powf(b1, (float)t);
    const float bc2 = 1.0f - powf(b2, (float)t);
    const float step_size = lr / bc1;
    const float bc2_sqrt = sqrtf(bc2);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] = p[i] - step_size * (mi / (sqrtf(vi) / bc2_sqrt + eps));
    }
}

void launch_mlp_adam(float *p, const float *g, float *m, float *v, int count, float lr, float b1,
                     float b2, float eps, int t, hipStream_t s) {
    const int blocks = (count + 255) / 256;
    mlp_adam_kernel<<<blocks < 1024 ? blocks : 1024, 256, 0, s>>>(p, g, m, v, count, lr, b1, b2,
                                                                  eps, t);
}

}  // namespace rie
