// dense_scan strategy (num_qubits <= 8): layer matrices and a parallel
// product tree of complex128 GEMMs on the FP64 matrix cores.
//
// Reference behaviour: cpp/ising_cuda_dense_modes.inc:96-369 and
// cpp/ising_cuda_dense_kernels.cu:12-305 build [R_0, C_0, R_1, C_1, ...],
// multiply them pairwise with cuBLAS ZGEMM into a product tree, derive the
// prefix states with a down-sweep, and form the gradients layer by layer.
//
// This build:
//  * merges each rotation layer with the following ring-CNOT permutation,
//    M_l = C R_l (a row permutation of R_l), so the tree has L leaves;
//  * pads every matrix to D = max(2^n, 16) with an identity block so all
//    products run on v_mfma_f64_16x16x4f64 tiles;
//  * multiplies split-complex (re / im planes) matrices with four real MFMAs
//    per k-step (Cr += Ar Br - Ai Bi, Ci += Ar Bi + Ai Br), 64x64 output
//    tiles staged through LDS;
//  * obtains every gradient from the prefix vectors x_l = M_{l-1}..M_0 e0 and
//    y_l = M_{l-1}..M_0 T^dagger H psi:
//      dE/dtheta_{l,w} = Re<y_l | -i Y_w | x_l>
//      dE/dphi_{l,w}   = Re<C^-1 y_{l+1} | -i Z_w | C^-1 x_{l+1}>
//    with x_L = psi and y_L = H psi.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "rie_engine_internal.hpp"

namespace rie {

typedef double double4_t __attribute__((ext_vector_type(4)));

struct DenseState {
    int D = 0;         // padded dimension
    int P = 0;         // leaves (power of two >= L)
    int levels = 0;    // log2(P)
    double *tree = nullptr;   // (2P - 1) nodes, each 2 * D * D doubles (re plane, im plane)
    double2 *vec_a = nullptr; // P * 2 * D incoming vectors (ping)
    double2 *vec_b = nullptr; // (pong)
    double2 *head = nullptr;  // psi, lambda, eta: 3 * D
    double *epart = nullptr;  // energy scratch
    size_t node_doubles = 0;
};

// ---------------------------------------------------------------------------
// Layer matrices: M_l[r][c] = R_l[P^-1(r)][c], R_l = (x)_w U_w with
// U = RZ(phi) RY(theta) = [[c e-, -s e-], [s e+, c e+]], e-+ = e^{-+i phi/2}
// (reference: fill_ryrz_local_u_entries_device, cpp/ising_cuda_dense_kernels.cu:43-55).
// ---------------------------------------------------------------------------

__device__ __forceinline__ double2 u_entry(const WireCoef &k, int rb, int cb) {
    // row bit rb (output), column bit cb (input)
    const double mag = (rb == cb) ? k.c : (rb == 0 ? -k.s : k.s);
    const double ph = rb == 0 ? -k.sh : k.sh;
    return make_double2(mag * k.ch, mag * ph);
}

__global__ void dense_layers_kernel(const WireCoef *__restrict__ coef, int n, int L, int D,
                                    double *__restrict__ tree, size_t node_doubles) {
    const int leaf = blockIdx.x / D;
    const int r = blockIdx.x % D;
    const int c = threadIdx.x;
    const int d = 1 << n;
    double *re = tree + (size_t)leaf * node_doubles;
    double *im = re + (size_t)D * D;
    double2 v;
    if (leaf >= L || r >= d || c >= d) {
        v = make_double2((r == c) ? 1.0 : 0.0, 0.0);
    } else {
        const uint32_t src = ring_inv32((uint32_t)r, n);  // (C R)[r] = R[P^-1 r]
        v = make_double2(1.0, 0.0);
        const WireCoef *lc = coef + (size_t)leaf * n;
        for (int w = 0; w < n; ++w) {
            v = cmul(v, u_entry(lc[w], (src >> w) & 1, (c >> w) & 1));
        }
    }
    re[(size_t)r * D + c] = v.x;
    im[(size_t)r * D + c] = v.y;
}

// ---------------------------------------------------------------------------
// Batched split-complex GEMM on FP64 MFMA: C_b = A_b * B_b, D x D row-major.
// v_mfma_f64_16x16x4f64 lane layout: A[i][k] in lane i + 16k, B[k][j] in
// lane j + 16k, D[i][j] in lane j + 16(i%4), element i/4 (one output row per
// 16-lane group and register, unlike the FP32 16x16 forms).
// ---------------------------------------------------------------------------

template <int TILE>
__global__ __launch_bounds__(256) void zgemm_mfma_kernel(const double *__restrict__ A,
                                                         const double *__restrict__ B,
                                                         double *__restrict__ C, int D,
                                                         size_t stride_a, size_t stride_b,
                                                         size_t stride_c) {
    constexpr int KB = 16;
    constexpr int SUB = TILE / 16;                    // 16x16 sub-tiles per side
    constexpr int PER_WAVE = (SUB * SUB + 3) / 4;     // sub-tiles per wave
    __shared__ double as_r[TILE][KB + 1];
    __shared__ double as_i[TILE][KB + 1];
    __shared__ double bs_r[KB][TILE];
    __shared__ double bs_i[KB][TILE];
    const int tiles_per_side = D / TILE;
    const int bt = blockIdx.x;
    const int batch = blockIdx.y;
    const int tr = (bt / tiles_per_side) * TILE;
    const int tc = (bt % tiles_per_side) * TILE;
    const size_t DD = (size_t)D * D;
    const double *ar = A + batch * stride_a;
    const double *ai = ar + DD;
    const double *br = B + batch * stride_b;
    const double *bi = br + DD;
    double *cr = C + batch * stride_c;
    double *ci = cr + DD;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;

    double4_t acc_r[PER_WAVE];
    double4_t acc_i[PER_WAVE];
#pragma unroll
    for (int s = 0; s < PER_WAVE; ++s) {
        acc_r[s] = (double4_t){0.0, 0.0, 0.0, 0.0};
        acc_i[s] = (double4_t){0.0, 0.0, 0.0, 0.0};
    }

    for (int k0 = 0; k0 < D; k0 += KB) {
        __syncthreads();
        for (int e = t; e < TILE * KB; e += 256) {
            const int r = e / KB;
            const int kk = e % KB;
            as_r[r][kk] = ar[(size_t)(tr + r) * D + k0 + kk];
            as_i[r][kk] = ai[(size_t)(tr + r) * D + k0 + kk];
            const int kb = e / TILE;
            const int cc = e % TILE;
            bs_r[kb][cc] = br[(size_t)(k0 + kb) * D + tc + cc];
            bs_i[kb][cc] = bi[(size_t)(k0 + kb) * D + tc + cc];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KB; kk += 4) {
#pragma unroll
            for (int s = 0; s < PER_WAVE; ++s) {
                const int sub = wave * PER_WAVE + s;
                if (sub < SUB * SUB) {
                    const int sr = (sub / SUB) * 16;
                    const int sc = (sub % SUB) * 16;
                    const double a_r = as_r[sr + (lane & 15)][kk + (lane >> 4)];
                    const double a_i = as_i[sr + (lane & 15)][kk + (lane >> 4)];
                    const double b_r = bs_r[kk + (lane >> 4)][sc + (lane & 15)];
                    const double b_i = bs_i[kk + (lane >> 4)][sc + (lane & 15)];
                    acc_r[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r, b_r, acc_r[s], 0, 0, 0);
                    acc_r[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a_i, b_i, acc_r[s], 0, 0, 0);
                    acc_i[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r, b_i, acc_i[s], 0, 0, 0);
                    acc_i[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i, b_r, acc_i[s], 0, 0, 0);
                }
            }
        }
    }
#pragma unroll
    for (int s = 0; s < PER_WAVE; ++s) {
        const int sub = wave * PER_WAVE + s;
        if (sub < SUB * SUB) {
            const int sr = (sub / SUB) * 16;
            const int sc = (sub % SUB) * 16;
            const int col = tc + sc + (lane & 15);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = tr + sr + 4 * i + (lane >> 4);
                cr[(size_t)row * D + col] = acc_r[s][i];
                ci[(size_t)row * D + col] = acc_i[s][i];
            }
        }
    }
}

static void launch_zgemm(const double *A, const double *B, double *C, int D, int batch,
                         size_t sa, size_t sb, size_t sc, hipStream_t s) {
    if (D >= 64) {
        dim3 grid((D / 64) * (D / 64), batch);
        zgemm_mfma_kernel<64><<<grid, 256, 0, s>>>(A, B, C, D, sa, sb, sc);
    } else if (D == 32) {
        dim3 grid(1, batch);
        zgemm_mfma_kernel<32><<<grid, 256, 0, s>>>(A, B, C, D, sa, sb, sc);
    } else {
        dim3 grid(1, batch);
        zgemm_mfma_kernel<16><<<grid, 256, 0, s>>>(A, B, C, D, sa, sb, sc);
    }
}

// ---------------------------------------------------------------------------
// Root: psi = T e0, lambda = H psi, E = Re<psi|lambda>, eta = T^dagger lambda.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void dense_head_kernel(const double *__restrict__ root, int n,
                                                         int D, double field,
                                                         double2 *__restrict__ head,
                                                         double *__restrict__ energy) {
    __shared__ double2 psi[256];
    __shared__ double2 lam[256];
    __shared__ double red[4];
    const int t = threadIdx.x;
    const int d = 1 << n;
    const size_t DD = (size_t)D * D;
    const double *re = root;
    const double *im = root + DD;
    if (t < D) {
        psi[t] = (t < d) ? make_double2(re[(size_t)t * D], im[(size_t)t * D]) : make_double2(0, 0);
    }
    __syncthreads();
    double e = 0.0;
    if (t < D) {
        double2 hv = make_double2(0.0, 0.0);
        if (t < d) {
            const double dg = ham_diag((uint64_t)t, n);
            const double2 v = psi[t];
            double fx = 0.0;
            double fy = 0.0;
            for (int w = 0; w < n; ++w) {
                const double2 p = psi[t ^ (1 << w)];
                fx += p.x;
                fy += p.y;
            }
            hv = make_double2(dg * v.x - field * fx, dg * v.y - field * fy);
            e = v.x * hv.x + v.y * hv.y;
        }
        lam[t] = hv;
    }
    e = wave_sum(e);
    if ((t & 63) == 0) {
        red[t >> 6] = e;
    }
    __syncthreads();
    if (t == 0) {
        *energy = red[0] + red[1] + red[2] + red[3];
    }
    if (t < D) {
        // eta_c = sum_r conj(T[r][c]) lambda_r
        double ex = 0.0;
        double ey = 0.0;
        for (int r = 0; r < D; ++r) {
            const double tr = re[(size_t)r * D + t];
            const double ti = im[(size_t)r * D + t];
            const double2 l = lam[r];
            ex += tr * l.x + ti * l.y;
            ey += tr * l.y - ti * l.x;
        }
        head[t] = psi[t];
        head[D + t] = lam[t];
        head[2 * D + t] = make_double2(ex, ey);
    }
}

// One down-sweep level: parent node i (level j) hands (x, y) to its left
// child and (N x, N y) to its right child, N = left child's product.
__global__ __launch_bounds__(256) void dense_down_kernel(const double *__restrict__ child_level,
                                                         size_t node_doubles, int D,
                                                         const double2 *__restrict__ vin,
                                                         double2 *__restrict__ vout,
                                                         int parents, const double2 *__restrict__ seed) {
    const int p = blockIdx.x;
    const int r = threadIdx.x;
    if (p >= parents || r >= D) {
        return;
    }
    const double2 *x = seed != nullptr ? seed : vin + (size_t)p * 2 * D;
    const double2 *y = seed != nullptr ? seed + 2 * D : x + D;  // seed: [e0 | ... | eta]
    const double *re = child_level + (size_t)(2 * p) * node_doubles;
    const double *im = re + (size_t)D * D;
    double2 nx = make_double2(0.0, 0.0);
    double2 ny = make_double2(0.0, 0.0);
    for (int c = 0; c < D; ++c) {
        const double mr = re[(size_t)r * D + c];
        const double mi = im[(size_t)r * D + c];
        const double2 xv = x[c];
        const double2 yv = y[c];
        nx.x += mr * xv.x - mi * xv.y;
        nx.y += mr * xv.y + mi * xv.x;
        ny.x += mr * yv.x - mi * yv.y;
        ny.y += mr * yv.y + mi * yv.x;
    }
    double2 *left = vout + (size_t)(2 * p) * 2 * D;
    double2 *right = vout + (size_t)(2 * p + 1) * 2 * D;
    left[r] = x[r];
    left[D + r] = y[r];
    right[r] = nx;
    right[D + r] = ny;
}

// e0 and eta as a seed pair for the root: [e0 (D) | psi (D, unused) | eta (D)] is
// laid out so that seed[0..D) = e0 and seed[2D..3D) = eta.
__global__ void dense_seed_kernel(double2 *__restrict__ seed, const double2 *__restrict__ head,
                                  int D) {
    const int t = threadIdx.x;
    if (t < D) {
        seed[t] = make_double2(t == 0 ? 1.0 : 0.0, 0.0);
        seed[D + t] = make_double2(0.0, 0.0);
        seed[2 * D + t] = head[2 * D + t];
    }
}

// Gradients of layer l from (x_l, y_l) and (x_{l+1}, y_{l+1}).
__global__ __launch_bounds__(256) void dense_tail_kernel(const double2 *__restrict__ leaves,
                                                         const double2 *__restrict__ head, int n,
                                                         int L, int D, double *__restrict__ grad,
                                                         const double *__restrict__ scale) {
    __shared__ double red[4][2 * 8];
    const int l = blockIdx.x;
    const int t = threadIdx.x;
    const int d = 1 << n;
    const double2 *x = leaves + (size_t)l * 2 * D;
    const double2 *y = x + D;
    const double2 *xn = (l + 1 < L) ? leaves + (size_t)(l + 1) * 2 * D : head;      // psi
    const double2 *yn = (l + 1 < L) ? xn + D : head + D;                              // lambda
    double dt[8];
    double dp[8];
    for (int w = 0; w < 8; ++w) {
        dt[w] = 0.0;
        dp[w] = 0.0;
    }
    if (t < d) {
        const double2 xv = x[t];
        const double2 yv = y[t];
        const uint32_t src = ring_fwd32((uint32_t)t, n);  // (C^-1 v)[t] = v[P(t)]
        const double2 a = xn[src];
        const double2 b = yn[src];
        const double imba = b.x * a.y - b.y * a.x;
        for (int w = 0; w < n; ++w) {
            const int bit = (t >> w) & 1;
            const double2 xp = x[t ^ (1 << w)];
            // (-iY x)[t] = -x[t^w] for bit 0, +x[t^w] for bit 1
            const double re = yv.x * xp.x + yv.y * xp.y;
            dt[w] = bit ? re : -re;
            dp[w] = bit ? -imba : imba;
        }
    }
    const int lane = t & 63;
    const int wave = t >> 6;
    for (int w = 0; w < n; ++w) {
        const double a = wave_sum(dt[w]);
        const double b = wave_sum(dp[w]);
        if (lane == 0) {
            red[wave][2 * w] = a;
            red[wave][2 * w + 1] = b;
        }
    }
    __syncthreads();
    if (t < 2 * n) {
        const double f = scale != nullptr ? *scale : 1.0;
        grad[(size_t)l * n * 2 + t] = (red[0][t] + red[1][t] + red[2][t] + red[3][t]) * f;
    }
}

// ---------------------------------------------------------------------------
// Engine glue
// ---------------------------------------------------------------------------

void Engine::dense_init() {
    dense = new DenseState();
    DenseState &ds = *dense;
    ds.D = std::max(1 << n, 16);
    ds.P = 1;
    while (ds.P < L) {
        ds.P <<= 1;
    }
    ds.levels = 0;
    while ((1 << ds.levels) < ds.P) {
        ds.levels++;
    }
    ds.node_doubles = (size_t)2 * ds.D * ds.D;
    ds.tree = (double *)device_alloc(sizeof(double) * ds.node_doubles * (2 * (size_t)ds.P - 1),
                                     "alloc dense tree");
    ds.vec_a = (double2 *)device_alloc(sizeof(double2) * (size_t)ds.P * 2 * ds.D, "alloc dense vec");
    ds.vec_b = (double2 *)device_alloc(sizeof(double2) * (size_t)ds.P * 2 * ds.D, "alloc dense vec");
    ds.head = (double2 *)device_alloc(sizeof(double2) * 3 * ds.D, "alloc dense head");
}

void Engine::dense_free() {
    if (dense == nullptr) {
        return;
    }
    device_free(dense->tree);
    device_free(dense->vec_a);
    device_free(dense->vec_b);
    device_free(dense->head);
    delete dense;
    dense = nullptr;
}

void Engine::dense_run(hipStream_t s, bool want_grad, const double *scale, double *energy_out,
                       double *grad_out) {
    DenseState &ds = *dense;
    const int D = ds.D;
    // leaves: level 0 nodes [0, P)
    dense_layers_kernel<<<ds.P * D, D, 0, s>>>(d_coef, n, L, D, ds.tree, ds.node_doubles);
    check_launch("dense_layers");
    // up-sweep: level j nodes follow level j-1; parent = right * left
    size_t level_off = 0;
    int count = ds.P;
    std::vector<size_t> offsets;
    offsets.push_back(0);
    while (count > 1) {
        const double *level = ds.tree + level_off * ds.node_doubles;
        double *parent = ds.tree + (level_off + count) * ds.node_doubles;
        launch_zgemm(level + ds.node_doubles, level, parent, D, count / 2, 2 * ds.node_doubles,
                     2 * ds.node_doubles, ds.node_doubles, s);
        check_launch("zgemm_mfma");
        level_off += count;
        count /= 2;
        offsets.push_back(level_off);
    }
    const double *root = ds.tree + level_off * ds.node_doubles;
    dense_head_kernel<<<1, 256, 0, s>>>(root, n, D, h, ds.head, energy_out);
    check_launch("dense_head");
    if (!want_grad) {
        return;
    }
    // down-sweep from the root (level `levels`) to the leaves (level 0)
    double2 *vin = ds.vec_a;
    double2 *vout = ds.vec_b;
    if (ds.levels == 0) {
        dense_seed_kernel<<<1, 256, 0, s>>>(vout, ds.head, D);
        // single leaf: x_0 = e0, y_0 = eta with layout [x | y]
        hipMemcpyAsync(vout + D, ds.head + 2 * D, sizeof(double2) * D, hipMemcpyDeviceToDevice, s);
        vin = vout;
    } else {
        dense_seed_kernel<<<1, 256, 0, s>>>(vin, ds.head, D);
        check_launch("dense_seed");
        for (int j = ds.levels; j >= 1; --j) {
            const int parents = ds.P >> j;
            const double *child_level = ds.tree + offsets[(size_t)j - 1] * ds.node_doubles;
            dense_down_kernel<<<parents, 256, 0, s>>>(child_level, ds.node_doubles, D, vin, vout,
                                                      parents, j == ds.levels ? vin : nullptr);
            check_launch("dense_down");
            std::swap(vin, vout);
        }
    }
    dense_tail_kernel<<<L, 256, 0, s>>>(vin, ds.head, n, L, D, grad_out, scale);
    check_launch("dense_tail");
}

}  // namespace rie
