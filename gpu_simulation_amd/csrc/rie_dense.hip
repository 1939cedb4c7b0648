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
//  * multiplies split-complex (re / im planes) matrices on FP64 MFMA: 64x64
//    output tiles (D >= 64) with three real products per complex one and the
//    next k-block prefetched into registers; 16x16 / 32x32 tiles (D <= 32)
//    with four real products per k-step;
//  * obtains every gradient from the prefix vectors x_l = M_{l-1}..M_0 e0 and
//    y_l = M_{l-1}..M_0 T^dagger H psi:
//      dE/dtheta_{l,w} = Re<y_l | -i Y_w | x_l>
//      dE/dphi_{l,w}   = Re<C^-1 y_{l+1} | -i Z_w | C^-1 x_{l+1}>
//    with x_L = psi and y_L = H psi.
//
// Every kernel is driven by a job table (DenseJob) and a work table of
// (job, first block) pairs, so one launch serves one circuit or a whole
// batch of circuits of different sizes: the product-tree level j of every
// circuit deep enough to have one is one GEMM launch per tile class.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "rie_dense.hpp"
#include "rie_engine_internal.hpp"

namespace rie {

typedef double double4_t __attribute__((ext_vector_type(4)));

// Entry of a launch's work table whose block range holds block b:
// tab[k].x = job index, tab[k].y = first block of that job in the launch.
__device__ __forceinline__ int find_item(const int2 *__restrict__ tab, int count, int b) {
    int lo = 0;
    int hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].y <= b) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    return lo;
}

__device__ __forceinline__ size_t level_node(int P, int level) {
    return (size_t)(2 * P - 2 * (P >> level));
}

// ---------------------------------------------------------------------------
// Layer matrices: M_l[r][c] = R_l[P^-1(r)][c], R_l = (x)_w U_w with
// U = RZ(phi) RY(theta) = [[c e-, -s e-], [s e+, c e+]], e-+ = e^{-+i phi/2}
// (reference: fill_ryrz_local_u_entries_device, cpp/ising_cuda_dense_kernels.cu:43-55).
// Leaves l >= L and the padding rows / columns are identity.
// ---------------------------------------------------------------------------

__device__ __forceinline__ double2 u_entry(const WireCoef &k, int rb, int cb) {
    // row bit rb (output), column bit cb (input)
    const double mag = (rb == cb) ? k.c : (rb == 0 ? -k.s : k.s);
    const double ph = rb == 0 ? -k.sh : k.sh;
    return make_double2(mag * k.ch, mag * ph);
}

// Each block fills kLayerChunk consecutive entries of one circuit's P * D * D
// leaf matrices (D and D * D are powers of two: indices by shifts and masks).
// A row of R_l is the tensor product of the rows U_w[src_w][.] of the wire
// gates, so for n >= 6 every row's entries are one product of two 16-entry
// tables (wires 0..3 and 4..n-1) built in LDS; smaller circuits multiply the
// n factors directly.
constexpr int kLayerChunk = 4096;

__global__ __launch_bounds__(256) void dense_layers_kernel(const DenseJob *__restrict__ jobs,
                                                           const int2 *__restrict__ tab,
                                                           int count) {
    __shared__ double2 tables[64][2][16];  // rows of the chunk (D >= 64: at most 64)
    const int item = find_item(tab, count, blockIdx.x);
    const DenseJob &J = jobs[tab[item].x];
    const int D = J.D;
    const int logD = J.log_d;
    const uint32_t maskD = (uint32_t)D - 1u;
    const size_t DD = (size_t)D * D;
    const size_t total = (size_t)J.P * DD;
    const int n = J.n;
    const int d = 1 << n;
    const int t = threadIdx.x;
    const size_t begin = (size_t)(blockIdx.x - tab[item].y) * kLayerChunk;
    const bool factored = D >= 64;
    const size_t row0 = begin >> logD;  // first global row (leaf * D + r) of the chunk
    if (factored) {
        const int rows = kLayerChunk >> logD;
        for (int q = t; q < rows * 32; q += 256) {
            const int lr = q >> 5;
            const int half = (q >> 4) & 1;
            const int idx = q & 15;
            const size_t grow = row0 + lr;
            const int leaf = (int)(grow >> logD);
            const int r = (int)(grow & maskD);
            double2 v = make_double2(1.0, 0.0);
            if (leaf < J.L && r < d) {
                const uint32_t src = ring_inv32((uint32_t)r, n);
                const WireCoef *lc = J.coef + (size_t)leaf * n;
                const int w0 = half ? 4 : 0;
                const int w1 = half ? n : 4;
                for (int w = w0; w < w1; ++w) {
                    v = cmul(v, u_entry(lc[w], (src >> w) & 1, (idx >> (w - w0)) & 1));
                }
            }
            tables[lr][half][idx] = v;
        }
        __syncthreads();
    }
    for (int q = t; q < kLayerChunk; q += 256) {
        const size_t e = begin + q;
        if (e >= total) {
            break;
        }
        const int leaf = (int)(e >> (2 * logD));
        const uint32_t rc = (uint32_t)(e & (DD - 1));
        const int r = (int)(rc >> logD);
        const int c = (int)(rc & maskD);
        double2 v;
        if (leaf >= J.L || r >= d || c >= d) {
            v = make_double2((r == c) ? 1.0 : 0.0, 0.0);
        } else if (factored) {
            const int lr = q >> logD;
            v = cmul(tables[lr][0][c & 15], tables[lr][1][c >> 4]);
        } else {
            const uint32_t src = ring_inv32((uint32_t)r, n);  // (C R)[r] = R[P^-1 r]
            v = make_double2(1.0, 0.0);
            const WireCoef *lc = J.coef + (size_t)leaf * n;
            for (int w = 0; w < n; ++w) {
                v = cmul(v, u_entry(lc[w], (src >> w) & 1, (c >> w) & 1));
            }
        }
        double *re = J.tree + (size_t)leaf * J.node_doubles;
        re[rc] = v.x;
        re[DD + rc] = v.y;
    }
}

// ---------------------------------------------------------------------------
// Product-tree level on FP64 MFMA: parent q = right child (2q+1) * left child
// (2q), D x D row-major split-complex, one TILE x TILE output tile per block.
// v_mfma_f64_16x16x4f64 lane layout: A[i][k] in lane i + 16k, B[k][j] in
// lane j + 16k, D[i][j] in lane j + 16(i%4), element i/4 (one output row per
// 16-lane group and register, unlike the FP32 16x16 forms).
// ---------------------------------------------------------------------------

template <int TILE>
__global__ __launch_bounds__(256) void zgemm_level_kernel(const DenseJob *__restrict__ jobs,
                                                          const int2 *__restrict__ tab, int count,
                                                          int level) {
    constexpr int KB = 16;
    constexpr int SUB = TILE / 16;                 // 16x16 sub-tiles per side
    constexpr int PER_WAVE = (SUB * SUB + 3) / 4;  // sub-tiles per wave
    __shared__ double as_r[TILE][KB + 1];
    __shared__ double as_i[TILE][KB + 1];
    __shared__ double bs_r[KB][TILE];
    __shared__ double bs_i[KB][TILE];
    const int item = find_item(tab, count, blockIdx.x);
    const DenseJob &J = jobs[tab[item].x];
    const int D = J.D;
    const int local = blockIdx.x - tab[item].y;
    const int tiles_per_side = D / TILE;
    const int q = local / (tiles_per_side * tiles_per_side);
    const int bt = local % (tiles_per_side * tiles_per_side);
    const int tr = (bt / tiles_per_side) * TILE;
    const int tc = (bt % tiles_per_side) * TILE;
    const size_t DD = (size_t)D * D;
    const size_t nd = J.node_doubles;
    const double *left = J.tree + (level_node(J.P, level - 1) + 2 * (size_t)q) * nd;
    const double *ar = left + nd;  // right child
    const double *ai = ar + DD;
    const double *br = left;
    const double *bi = br + DD;
    double *cr = J.tree + (level_node(J.P, level) + (size_t)q) * nd;
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

// 64 x 64 output tiles for D >= 64: each of the four waves computes a 32 x 32
// quadrant (2 x 2 MFMA sub-tiles, A and B fragments shared across them) with
// three real products per complex one (P1 = Ar Br, P2 = Ai Bi,
// P3 = (Ar + Ai)(Br + Bi); Cr = P1 - P2, Ci = P3 - P1 - P2), and the next
// 16-wide k-block is loaded into registers while the current one is
// multiplied from LDS.
__global__ __launch_bounds__(256) void zgemm64_level_kernel(const DenseJob *__restrict__ jobs,
                                                            const int2 *__restrict__ tab,
                                                            int count, int level) {
    constexpr int TILE = 64;
    constexpr int KB = 16;
    __shared__ double as[3][TILE][KB + 1];  // re, im, re + im
    __shared__ double bs[3][KB][TILE + 2];
    const int item = find_item(tab, count, blockIdx.x);
    const DenseJob &J = jobs[tab[item].x];
    const int D = J.D;
    const int local = blockIdx.x - tab[item].y;
    const int tiles_per_side = D / TILE;
    const int q = local / (tiles_per_side * tiles_per_side);
    const int bt = local % (tiles_per_side * tiles_per_side);
    const int tr = (bt / tiles_per_side) * TILE;
    const int tc = (bt % tiles_per_side) * TILE;
    const size_t DD = (size_t)D * D;
    const size_t nd = J.node_doubles;
    const double *left = J.tree + (level_node(J.P, level - 1) + 2 * (size_t)q) * nd;
    const double *ar = left + nd;  // right child
    const double *ai = ar + DD;
    const double *br = left;
    const double *bi = br + DD;
    double *cr = J.tree + (level_node(J.P, level) + (size_t)q) * nd;
    double *ci = cr + DD;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int wr = (wave >> 1) * 32;  // quadrant of this wave
    const int wc = (wave & 1) * 32;
    // global -> register staging: A rows t/4, cols 4(t%4)..+3; B rows t/16, cols 4(t%16)..+3
    const int a_row = t >> 2;
    const int a_col = (t & 3) * 4;
    const int b_row = t >> 4;
    const int b_col = (t & 15) * 4;
    double2 ra[4];
    double2 rb[4];
    auto load_block = [&](int k0) {
        const size_t ao = (size_t)(tr + a_row) * D + k0 + a_col;
        const size_t bo = (size_t)(k0 + b_row) * D + tc + b_col;
        ra[0] = *reinterpret_cast<const double2 *>(ar + ao);
        ra[1] = *reinterpret_cast<const double2 *>(ar + ao + 2);
        ra[2] = *reinterpret_cast<const double2 *>(ai + ao);
        ra[3] = *reinterpret_cast<const double2 *>(ai + ao + 2);
        rb[0] = *reinterpret_cast<const double2 *>(br + bo);
        rb[1] = *reinterpret_cast<const double2 *>(br + bo + 2);
        rb[2] = *reinterpret_cast<const double2 *>(bi + bo);
        rb[3] = *reinterpret_cast<const double2 *>(bi + bo + 2);
    };
    double4_t p1[2][2];
    double4_t p2[2][2];
    double4_t p3[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            p1[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
            p2[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
            p3[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
        }
    }
    load_block(0);
    for (int k0 = 0; k0 < D; k0 += KB) {
        __syncthreads();
        {
            const double a_re[4] = {ra[0].x, ra[0].y, ra[1].x, ra[1].y};
            const double a_im[4] = {ra[2].x, ra[2].y, ra[3].x, ra[3].y};
            const double b_re[4] = {rb[0].x, rb[0].y, rb[1].x, rb[1].y};
            const double b_im[4] = {rb[2].x, rb[2].y, rb[3].x, rb[3].y};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                as[0][a_row][a_col + e] = a_re[e];
                as[1][a_row][a_col + e] = a_im[e];
                as[2][a_row][a_col + e] = a_re[e] + a_im[e];
                bs[0][b_row][b_col + e] = b_re[e];
                bs[1][b_row][b_col + e] = b_im[e];
                bs[2][b_row][b_col + e] = b_re[e] + b_im[e];
            }
        }
        __syncthreads();
        if (k0 + KB < D) {
            load_block(k0 + KB);  // in flight during the products below
        }
#pragma unroll 2
        for (int kk = 0; kk < KB; kk += 4) {
            double fa[3][2];
            double fb[3][2];
#pragma unroll
            for (int p = 0; p < 3; ++p) {
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    fa[p][i] = as[p][wr + 16 * i + (lane & 15)][kk + (lane >> 4)];
                    fb[p][i] = bs[p][kk + (lane >> 4)][wc + 16 * i + (lane & 15)];
                }
            }
#pragma unroll
            for (int i = 0; i < 2; ++i) {
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    p1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[0][i], fb[0][j], p1[i][j],
                                                                     0, 0, 0);
                    p2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[1][i], fb[1][j], p2[i][j],
                                                                     0, 0, 0);
                    p3[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[2][i], fb[2][j], p3[i][j],
                                                                     0, 0, 0);
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = tc + wc + 16 * j + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = tr + wr + 16 * i + 4 * r + (lane >> 4);
                cr[(size_t)row * D + col] = p1[i][j][r] - p2[i][j][r];
                ci[(size_t)row * D + col] = p3[i][j][r] - p1[i][j][r] - p2[i][j][r];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Root: psi = T e0, lambda = H psi, E = Re<psi|lambda>, eta = T^dagger lambda;
// also seeds the down-sweep with the root pair [x = e0 | y = eta] in vec_a.
// One block per circuit.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void dense_head_kernel(const DenseJob *__restrict__ jobs,
                                                         const int2 *__restrict__ tab,
                                                         double *energy_override) {
    __shared__ double2 psi[256];
    __shared__ double2 lam[256];
    __shared__ double red[4];
    const DenseJob &J = jobs[tab[blockIdx.x].x];
    const int t = threadIdx.x;
    const int n = J.n;
    const int D = J.D;
    const int d = 1 << n;
    const size_t DD = (size_t)D * D;
    const double *re = J.tree + (size_t)(2 * J.P - 2) * J.node_doubles;
    const double *im = re + DD;
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
            hv = make_double2(dg * v.x - J.field * fx, dg * v.y - J.field * fy);
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
        double *out = energy_override != nullptr ? energy_override : J.energy;
        *out = red[0] + red[1] + red[2] + red[3];
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
        const double2 eta = make_double2(ex, ey);
        J.head[t] = psi[t];
        J.head[D + t] = lam[t];
        J.head[2 * D + t] = eta;
        J.vec_a[t] = make_double2(t == 0 ? 1.0 : 0.0, 0.0);
        J.vec_a[D + t] = eta;
    }
}

// One down-sweep level: parent p hands (x, y) to its left child and
// (N x, N y) to its right child, N = the left child's product.  A circuit's
// step at level j reads vec_a when (levels - j) is even, else vec_b.  One
// block per parent.  N is row-major; 16-column strips of it are staged
// through LDS with coalesced loads, and thread r accumulates row r.
__global__ __launch_bounds__(256) void dense_down_kernel(const DenseJob *__restrict__ jobs,
                                                         const int2 *__restrict__ tab, int count,
                                                         int level) {
    constexpr int CB = 16;
    __shared__ double mre[256][CB + 1];
    __shared__ double mim[256][CB + 1];
    __shared__ double2 xs[256];
    __shared__ double2 ys[256];
    const int item = find_item(tab, count, blockIdx.x);
    const DenseJob &J = jobs[tab[item].x];
    const int D = J.D;
    const int t = threadIdx.x;
    const bool even = ((J.levels - level) & 1) == 0;
    const double2 *vin = even ? J.vec_a : J.vec_b;
    double2 *vout = even ? J.vec_b : J.vec_a;
    const size_t DD = (size_t)D * D;
    {
        const int p = blockIdx.x - tab[item].y;
        const double2 *x = vin + (size_t)p * 2 * D;
        const double2 *y = x + D;
        const double *re = J.tree + (level_node(J.P, level - 1) + 2 * (size_t)p) * J.node_doubles;
        const double *im = re + DD;
        if (t < D) {
            xs[t] = x[t];
            ys[t] = y[t];
        }
        double2 nx = make_double2(0.0, 0.0);
        double2 ny = make_double2(0.0, 0.0);
        // strip c0 of N: D rows x CB columns, at most 16 entries per thread
        // and plane, loaded into registers (all loads in flight at once; the
        // next strip's while this one is multiplied)
        constexpr int PER = 256 * CB / 256;
        const int strip = D * CB;
        double vr[PER];
        double vi[PER];
        auto load_strip = [&](int c0) {
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int e = t + 256 * i;
                if (e < strip) {
                    const size_t g = (size_t)(e / CB) * D + c0 + (e % CB);
                    vr[i] = re[g];
                    vi[i] = im[g];
                }
            }
        };
        load_strip(0);
        for (int c0 = 0; c0 < D; c0 += CB) {
            __syncthreads();
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int e = t + 256 * i;
                if (e < strip) {
                    mre[e / CB][e % CB] = vr[i];
                    mim[e / CB][e % CB] = vi[i];
                }
            }
            __syncthreads();
            if (c0 + CB < D) {
                load_strip(c0 + CB);
            }
            if (t < D) {
#pragma unroll
                for (int c = 0; c < CB; ++c) {
                    const double mr = mre[t][c];
                    const double mi = mim[t][c];
                    const double2 xv = xs[c0 + c];
                    const double2 yv = ys[c0 + c];
                    nx.x += mr * xv.x - mi * xv.y;
                    nx.y += mr * xv.y + mi * xv.x;
                    ny.x += mr * yv.x - mi * yv.y;
                    ny.y += mr * yv.y + mi * yv.x;
                }
            }
        }
        if (t < D) {
            double2 *left = vout + (size_t)(2 * p) * 2 * D;
            double2 *right = vout + (size_t)(2 * p + 1) * 2 * D;
            left[t] = xs[t];
            left[D + t] = ys[t];
            right[t] = nx;
            right[D + t] = ny;
        }
    }
}

// Gradients of layer l from (x_l, y_l) and (x_{l+1}, y_{l+1}).  One block per
// (circuit, layer).
__global__ __launch_bounds__(256) void dense_tail_kernel(const DenseJob *__restrict__ jobs,
                                                         const int2 *__restrict__ tab, int count,
                                                         double *grad_override,
                                                         const double *scale_override) {
    __shared__ double red[4][2 * 8];
    const int item = find_item(tab, count, blockIdx.x);
    const DenseJob &J = jobs[tab[item].x];
    const int l = blockIdx.x - tab[item].y;
    const int t = threadIdx.x;
    const int n = J.n;
    const int L = J.L;
    const int D = J.D;
    const int d = 1 << n;
    // the last down-sweep step (levels - 1) wrote vec_b when levels is odd
    const double2 *leaves =
        (J.levels == 0) ? J.vec_a : ((((J.levels - 1) & 1) == 0) ? J.vec_b : J.vec_a);
    const double2 *x = leaves + (size_t)l * 2 * D;
    const double2 *y = x + D;
    const double2 *xn = (l + 1 < L) ? leaves + (size_t)(l + 1) * 2 * D : J.head;  // psi
    const double2 *yn = (l + 1 < L) ? xn + D : J.head + D;                          // lambda
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
            const double rr = yv.x * xp.x + yv.y * xp.y;
            dt[w] = bit ? rr : -rr;
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
        const double *sp = scale_override != nullptr ? scale_override : J.scale;
        const double f = sp != nullptr ? *sp : 1.0;
        double *g = grad_override != nullptr ? grad_override : J.grad;
        g[(size_t)l * n * 2 + t] = (red[0][t] + red[1][t] + red[2][t] + red[3][t]) * f;
    }
}

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------

namespace {

struct Geometry {
    int D;
    int P;
    int levels;
    size_t nd;
    size_t tree_bytes;
    size_t vec_bytes;
    size_t head_bytes;
};

Geometry geometry(int n, int L) {
    Geometry g{};
    g.D = std::max(1 << n, 16);
    g.P = 1;
    g.levels = 0;
    while (g.P < L) {
        g.P <<= 1;
        g.levels++;
    }
    g.nd = (size_t)2 * g.D * g.D;
    g.tree_bytes = sizeof(double) * g.nd * (2 * (size_t)g.P - 1);
    g.vec_bytes = sizeof(double2) * (size_t)g.P * 2 * g.D;
    g.head_bytes = sizeof(double2) * 3 * (size_t)g.D;
    return g;
}

size_t align256(size_t b) { return (b + 255) & ~size_t{255}; }

int tile_for(int D, int max_tile) {
    if (D >= 64) {
        return max_tile;
    }
    return D;  // 16 or 32
}

}  // namespace

size_t DensePlan::bytes_for(int n, int L) {
    const Geometry g = geometry(n, L);
    return align256(g.tree_bytes) + 2 * align256(g.vec_bytes) + align256(g.head_bytes);
}

void DensePlan::release() {
    device_free(d_jobs_);
    device_free(d_tab_);
    device_free(arena_);
    d_jobs_ = nullptr;
    d_tab_ = nullptr;
    arena_ = nullptr;
    arena_bytes_ = 0;
    jobs_.clear();
    gemm_.clear();
    down_.clear();
}

DensePlan::~DensePlan() { release(); }

void DensePlan::build(const std::vector<DenseSpec> &specs, int max_tile) {
    release();
    if (max_tile != 32 && max_tile != 64) {
        fail(RIE_INVALID_ARGUMENT, "dense tile must be 32 or 64");
    }
    if (specs.empty()) {
        return;
    }
    if (specs.size() > 65535) {
        fail(RIE_INVALID_ARGUMENT, "at most 65535 dense_scan circuits per batch");
    }
    size_t total = 0;
    for (const auto &sp : specs) {
        if (sp.n < 1 || sp.n > 8 || sp.L < 1) {
            fail(RIE_INVALID_ARGUMENT, "dense_scan requires 1 <= num_qubits <= 8 and layers >= 1.");
        }
        total += bytes_for(sp.n, sp.L);
    }
    arena_ = device_alloc(total, "alloc dense_scan buffers");
    arena_bytes_ = total;
    char *cursor = (char *)arena_;
    int max_levels = 0;
    for (const auto &sp : specs) {
        const Geometry g = geometry(sp.n, sp.L);
        DenseJob J{};
        J.coef = sp.coef;
        J.tree = (double *)cursor;
        cursor += align256(g.tree_bytes);
        J.vec_a = (double2 *)cursor;
        cursor += align256(g.vec_bytes);
        J.vec_b = (double2 *)cursor;
        cursor += align256(g.vec_bytes);
        J.head = (double2 *)cursor;
        cursor += align256(g.head_bytes);
        J.energy = sp.energy;
        J.grad = sp.grad;
        J.scale = sp.scale;
        J.node_doubles = g.nd;
        J.field = sp.field;
        J.n = sp.n;
        J.L = sp.L;
        J.D = g.D;
        J.log_d = 0;
        while ((1 << J.log_d) < g.D) {
            J.log_d++;
        }
        J.P = g.P;
        J.levels = g.levels;
        jobs_.push_back(J);
        max_levels = std::max(max_levels, g.levels);
    }

    std::vector<int2> tab;
    auto add_launch = [&](Launch &la, auto &&blocks_of) {
        la.tab_off = tab.size();
        la.blocks = 0;
        la.count = 0;
        for (int j = 0; j < (int)jobs_.size(); ++j) {
            const long b = blocks_of(jobs_[(size_t)j]);
            if (b <= 0) {
                continue;
            }
            if ((long)la.blocks + b > 0x7fffffffL) {
                fail(RIE_INVALID_ARGUMENT, "dense_scan batch too large for one launch");
            }
            tab.push_back(make_int2(j, la.blocks));
            la.blocks += (int)b;
            la.count += 1;
        }
    };
    add_launch(layers_, [](const DenseJob &J) {
        return (long)(((size_t)J.P * J.D * J.D + kLayerChunk - 1) / kLayerChunk);
    });
    for (int level = 1; level <= max_levels; ++level) {
        for (int tile : {16, 32, 64}) {
            Launch la;
            la.level = level;
            la.tile = tile;
            add_launch(la, [&](const DenseJob &J) {
                if (J.levels < level || tile_for(J.D, max_tile) != tile) {
                    return 0L;
                }
                const long side = J.D / tile;
                return (long)(J.P >> level) * side * side;
            });
            if (la.count > 0) {
                gemm_.push_back(la);
            }
        }
    }
    add_launch(head_, [](const DenseJob &) { return 1L; });
    for (int level = max_levels; level >= 1; --level) {
        Launch la;
        la.level = level;
        add_launch(la, [&](const DenseJob &J) {
            return J.levels >= level ? (long)(J.P >> level) : 0L;
        });
        if (la.count > 0) {
            down_.push_back(la);
        }
    }
    add_launch(tail_, [](const DenseJob &J) { return (long)J.L; });

    d_jobs_ = (DenseJob *)device_alloc(sizeof(DenseJob) * jobs_.size(), "alloc dense jobs");
    d_tab_ = (int2 *)device_alloc(sizeof(int2) * tab.size(), "alloc dense work tables");
    check(hipMemcpy(d_jobs_, jobs_.data(), sizeof(DenseJob) * jobs_.size(),
                    hipMemcpyHostToDevice),
          "upload dense jobs");
    check(hipMemcpy(d_tab_, tab.data(), sizeof(int2) * tab.size(), hipMemcpyHostToDevice),
          "upload dense work tables");
}

void DensePlan::run(hipStream_t s, bool want_grad, double *energy_override, double *grad_override,
                    const double *scale_override) const {
    if (jobs_.empty()) {
        return;
    }
    if ((energy_override != nullptr || grad_override != nullptr || scale_override != nullptr) &&
        jobs_.size() != 1) {
        fail(RIE_DEVICE_ERROR, "internal: dense overrides need a single-circuit plan");
    }
    dense_layers_kernel<<<layers_.blocks, 256, 0, s>>>(d_jobs_, d_tab_ + layers_.tab_off,
                                                       layers_.count);
    check_launch("dense_layers");
    for (const Launch &la : gemm_) {
        const int2 *tab = d_tab_ + la.tab_off;
        if (la.tile == 64) {
            zgemm64_level_kernel<<<la.blocks, 256, 0, s>>>(d_jobs_, tab, la.count, la.level);
        } else if (la.tile == 32) {
            zgemm_level_kernel<32><<<la.blocks, 256, 0, s>>>(d_jobs_, tab, la.count, la.level);
        } else {
            zgemm_level_kernel<16><<<la.blocks, 256, 0, s>>>(d_jobs_, tab, la.count, la.level);
        }
        check_launch("zgemm_level");
    }
    dense_head_kernel<<<head_.blocks, 256, 0, s>>>(d_jobs_, d_tab_ + head_.tab_off, energy_override);
    check_launch("dense_head");
    if (!want_grad) {
        return;
    }
    for (const Launch &la : down_) {
        dense_down_kernel<<<la.blocks, 256, 0, s>>>(d_jobs_, d_tab_ + la.tab_off, la.count,
                                                    la.level);
        check_launch("dense_down");
    }
    dense_tail_kernel<<<tail_.blocks, 256, 0, s>>>(d_jobs_, d_tab_ + tail_.tab_off, tail_.count,
                                                   grad_override, scale_override);
    check_launch("dense_tail");
}

// ---------------------------------------------------------------------------
// Engine glue (one circuit)
// ---------------------------------------------------------------------------

void Engine::dense_init() {
    dense = new DensePlan();
    DenseSpec sp{n, L, h, d_coef, d_energy, d_grad, nullptr};
    dense->build({sp}, 64);
}

void Engine::dense_free() {
    delete dense;
    dense = nullptr;
}

void Engine::dense_run(hipStream_t s, bool want_grad, const double *scale, double *energy_out,
                       double *grad_out) {
    dense->run(s, want_grad, energy_out, want_grad ? grad_out : nullptr, scale);
}

}  // namespace rie
