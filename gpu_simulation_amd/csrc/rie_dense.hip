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

__global__ __launch_bounds__(256) void dense_layers_kernel(const DenseJob *__restrict__ jobs,
                                                           const int2 *__restrict__ tab,
                                                           int count) {
    const int k = find_item(tab, count, blockIdx.x);
    const DenseJob &J = jobs[tab[k].x];
    const int D = J.D;
    const size_t DD = (size_t)D * D;
    const size_t e = (size_t)(blockIdx.x - tab[k].y) * 256 + threadIdx.x;
    if (e >= (size_t)J.P * DD) {
        return;
    }
    const int leaf = (int)(e / DD);
    const int r = (int)((e % DD) / D);
    const int c = (int)(e % D);
    const int n = J.n;
    const int d = 1 << n;
    double2 v;
    if (leaf >= J.L || r >= d || c >= d) {
        v = make_double2((r == c) ? 1.0 : 0.0, 0.0);
    } else {
        const uint32_t src = ring_inv32((uint32_t)r, n);  // (C R)[r] = R[P^-1 r]
        v = make_double2(1.0, 0.0);
        const WireCoef *lc = J.coef + (size_t)leaf * n;
        for (int w = 0; w < n; ++w) {
            v = cmul(v, u_entry(lc[w], (src >> w) & 1, (c >> w) & 1));
        }
    }
    double *re = J.tree + (size_t)leaf * J.node_doubles;
    re[(size_t)r * D + c] = v.x;
    re[DD + (size_t)r * D + c] = v.y;
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
// step at level j reads vec_a when (levels - j) is even, else vec_b.
__global__ __launch_bounds__(256) void dense_down_kernel(const DenseJob *__restrict__ jobs,
                                                         const int2 *__restrict__ tab, int count,
                                                         int level) {
    const int item = find_item(tab, count, blockIdx.x);
    const DenseJob &J = jobs[tab[item].x];
    const int p = blockIdx.x - tab[item].y;
    const int r = threadIdx.x;
    const int D = J.D;
    if (r >= D) {
        return;
    }
    const bool even = ((J.levels - level) & 1) == 0;
    const double2 *vin = even ? J.vec_a : J.vec_b;
    double2 *vout = even ? J.vec_b : J.vec_a;
    const double2 *x = vin + (size_t)p * 2 * D;
    const double2 *y = x + D;
    const double *re = J.tree + (level_node(J.P, level - 1) + 2 * (size_t)p) * J.node_doubles;
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
        return (long)(((size_t)J.P * J.D * J.D + 255) / 256);
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
            zgemm_level_kernel<64><<<la.blocks, 256, 0, s>>>(d_jobs_, tab, la.count, la.level);
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
