// Whole-circuit structured adjoint kernels for small circuits (n <= 14), one
// workgroup per circuit at a time, and the optimizer-step kernels of a batch.
//
// One launch runs a list of circuits of different qubit and layer counts:
// workgroup b evaluates circuits order[wg_begin[b] .. wg_begin[b+1]) one after
// the other (a persistent grid with a host-side longest-first schedule), each
// circuit entirely on chip: psi and lambda in LDS for n <= 12, in a per-
// workgroup global scratch of 2 * 2^n amplitudes (L2 / MALL resident) for
// n = 13, 14.  The same kernel with one circuit serves an engine's n <= 12
// path.  The ring-CNOT layers cost no data movement: they only change the
// linear frame in which the amplitudes are stored (FrameTables below).
//
// Per layer, the rotations are applied in chunks of K consecutive wires:
// each thread holds the 2^K amplitudes of one group in registers, applies the
// K fused RZ*RY gates there, and writes back, so a layer costs ceil(n / K)
// sweeps of the state instead of n (the rotation-chunk width of the
// reference's structured kernels, cpp/ising_cuda_rotation_kernels.cu:95-308,
// and its backward counterpart, cpp/ising_cuda_statevector_grad_kernels.cu:
// 440-587).  LDS slots are padded by one amplitude per 16 to spread the
// strided accesses of a group over the LDS banks.
#include <hip/hip_runtime.h>

#include "rie_kernels.hpp"

namespace rie {

constexpr int kRedCols = 2 * kBatchMaxQubits + 2;
constexpr int kFrameCols = 16;

template <int kThreads>
constexpr int max_chunk() {
    return batch_max_chunk(kThreads);
}

template <int kThreads>
constexpr int clamp_chunk(int k) {
    return k < max_chunk<kThreads>() ? k : max_chunk<kThreads>();
}

template <bool kLds>
__device__ __forceinline__ uint32_t slot(uint32_t i) {
    return kLds ? i + (i >> 4) : i;
}

__device__ __forceinline__ uint32_t parity(uint32_t x) { return (uint32_t)__popc(x) & 1u; }

// The state is kept in a linear frame F over GF(2)^n: storage index j holds
// the logical amplitude F(j).  A ring-CNOT layer only changes the frame
// (P is GF(2)-linear, cpp/ising_cuda_kernels.cu:178-205), so no amplitude
// moves.  F is held as its rows r_w (bit w of F(j) = parity(j & r_w)) and
// the columns m_w = F^-1(e_w) of its inverse: logical wire w pairs storage
// indices j and j ^ m_w.
struct FrameTables {
    uint32_t r[2][kFrameCols];  // double-buffered rows of F
    uint32_t m[2][kFrameCols];  // double-buffered columns of F^-1
    uint32_t row_p[kFrameCols];  // rows of P
    uint32_t row_q[kFrameCols];  // rows of P^-1
    uint32_t col_p[kFrameCols];  // P(e_w)
    uint32_t col_q[kFrameCols];  // P^-1(e_w)
};

// F <- P F (forward ring CNOT) or P^-1 F (inverse); fb selects the current buffer.
__device__ __forceinline__ void frame_step(FrameTables &ft, int n, int &fb, bool forward) {
    const int t = threadIdx.x;
    if (t < n) {
        const uint32_t sel_r = forward ? ft.row_p[t] : ft.row_q[t];
        const uint32_t sel_m = forward ? ft.col_q[t] : ft.col_p[t];
        uint32_t nr = 0;
        uint32_t nm = 0;
        for (int b = 0; b < n; ++b) {
            nr ^= ((sel_r >> b) & 1u) ? ft.r[fb][b] : 0u;
            nm ^= ((sel_m >> b) & 1u) ? ft.m[fb][b] : 0u;
        }
        ft.r[fb ^ 1][t] = nr;
        ft.m[fb ^ 1][t] = nm;
    }
    __syncthreads();
    fb ^= 1;
}

// Group geometry of one K-wire chunk in the current frame: the pair masks
// m_j of wires w0 + j, the rows r_j that give their logical bits, and K pivot
// bit positions (ascending) such that indices with zeros there represent
// every coset of span{m_j} exactly once.
template <int K>
struct ChunkGeom {
    uint32_t m[K];
    uint32_t r[K];
    uint32_t piv[K];
};

template <int K>
__device__ __forceinline__ ChunkGeom<K> chunk_geom(const FrameTables &ft, int fb, int w0) {
    ChunkGeom<K> cg;
    uint32_t red[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        cg.m[j] = ft.m[fb][w0 + j];
        cg.r[j] = ft.r[fb][w0 + j];
        uint32_t x = cg.m[j];
#pragma unroll
        for (int i = 0; i < j; ++i) {
            x ^= ((x >> cg.piv[i]) & 1u) ? red[i] : 0u;
        }
        red[j] = x;
        cg.piv[j] = (uint32_t)__builtin_ctz(x);  // x != 0: the m_j are independent
    }
    // ascending order for zero insertion
#pragma unroll
    for (int a = 0; a < K; ++a) {
#pragma unroll
        for (int b = 0; b + 1 < K - a; ++b) {
            const uint32_t lo = cg.piv[b] < cg.piv[b + 1] ? cg.piv[b] : cg.piv[b + 1];
            const uint32_t hi = cg.piv[b] < cg.piv[b + 1] ? cg.piv[b + 1] : cg.piv[b];
            cg.piv[b] = lo;
            cg.piv[b + 1] = hi;
        }
    }
    return cg;
}

// Storage index of the member of group g whose chunk wires are all logical 0.
template <int K>
__device__ __forceinline__ uint32_t group_base(const ChunkGeom<K> &cg, uint32_t g) {
    uint32_t x = g;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t p = cg.piv[j];
        x = ((x >> p) << (p + 1)) | (x & ((1u << p) - 1u));
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
        x ^= parity(x & cg.r[j]) ? cg.m[j] : 0u;
    }
    return x;
}

template <int K>
__device__ __forceinline__ uint32_t member_offset(const ChunkGeom<K> &cg, int o) {
    uint32_t off = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        off ^= ((o >> j) & 1) ? cg.m[j] : 0u;
    }
    return off;
}

// RZ factors of one layer's chunks (RZ on one wire commutes with RY on the
// others, so a chunk's rotations are its RY followed by one diagonal phase):
// rzt[c * 2^K + o] = prod over the chunk's wires j of e^{-i phi/2} (bit j of
// o clear) or e^{+i phi/2} (set), chunk c = wires [c K, c K + K).  Callers
// synchronize before the first chunk reads it.
__device__ __forceinline__ void build_chunk_rz(double2 *rzt, int n, int K, const WireCoef *lc) {
    const int A = 1 << K;
    const int chunks = (n + K - 1) / K;
    for (int i = threadIdx.x; i < chunks * A; i += blockDim.x) {
        const int w0 = (i / A) * K;
        const int o = i % A;
        const int kw = K < n - w0 ? K : n - w0;
        double2 q = make_double2(1.0, 0.0);
        for (int j = 0; j < kw; ++j) {
            const WireCoef w = lc[w0 + j];
            q = cmul(q, make_double2(w.ch, ((o >> j) & 1) ? w.sh : -w.sh));
        }
        rzt[i] = q;
    }
}

// ceil(n / K) * 2^K <= 64 for n <= 14, K <= 4
constexpr int kBatchRzEntries = 64;

template <int K, int kThreads, bool kLds>
__device__ __forceinline__ void chunk_fwd(double2 *st, int n, int w0, const WireCoef *lc,
                                          const FrameTables &ft, int fb, const double2 *rz) {
    constexpr int A = 1 << K;
    const ChunkGeom<K> cg = chunk_geom<K>(ft, fb, w0);
    WireCoef k[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        k[j] = lc[w0 + j];
    }
    const uint32_t groups = 1u << (n - K);
    for (uint32_t g = threadIdx.x; g < groups; g += kThreads) {
        const uint32_t base = group_base<K>(cg, g);
        double2 a[A];
#pragma unroll
        for (int o = 0; o < A; ++o) {
            a[o] = st[slot<kLds>(base ^ member_offset<K>(cg, o))];
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
#pragma unroll
            for (int o = 0; o < A; ++o) {
                if (((o >> j) & 1) == 0) {
                    ry_pair(a[o], a[o | (1 << j)], k[j]);
                }
            }
        }
#pragma unroll
        for (int o = 0; o < A; ++o) {
            st[slot<kLds>(base ^ member_offset<K>(cg, o))] = cmul(a[o], rz[o]);
        }
    }
}

// Un-rotates wires [w0, w0 + K) on (cur, lam) and leaves each wire's
// per-wave gradient sums in red[wave][2w], red[wave][2w + 1].
template <int K, int kThreads, bool kLds>
__device__ __forceinline__ void chunk_bwd(double2 *cur, double2 *lam, int n, int w0,
                                          const WireCoef *lc, const FrameTables &ft, int fb,
                                          double (*red)[kRedCols], const double2 *rz) {
    constexpr int A = 1 << K;
    const ChunkGeom<K> cg = chunk_geom<K>(ft, fb, w0);
    WireCoef k[K];
    double dt[K];
    double dp[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        k[j] = lc[w0 + j];
        dt[j] = 0.0;
        dp[j] = 0.0;
    }
    const uint32_t groups = 1u << (n - K);
    for (uint32_t g = threadIdx.x; g < groups; g += kThreads) {
        const uint32_t base = group_base<K>(cg, g);
        double2 c[A];
        double2 l[A];
#pragma unroll
        for (int o = 0; o < A; ++o) {
            const uint32_t i = slot<kLds>(base ^ member_offset<K>(cg, o));
            const double2 z = rz[o];
            const double2 zc = make_double2(z.x, -z.y);
            c[o] = cmul(cur[i], zc);
            l[o] = cmul(lam[i], zc);
        }
        // the chunk's RZ^dagger is applied; per wire dE/dphi (phase-invariant),
        // RY^dagger on both states, then dE/dtheta (the quantities of bwd_pair)
#pragma unroll
        for (int j = K - 1; j >= 0; --j) {
#pragma unroll
            for (int o = 0; o < A; ++o) {
                if (((o >> j) & 1) == 0) {
                    double2 &c0 = c[o];
                    double2 &c1 = c[o | (1 << j)];
                    double2 &l0 = l[o];
                    double2 &l1 = l[o | (1 << j)];
                    dp[j] += (l0.x * c0.y - l0.y * c0.x) - (l1.x * c1.y - l1.y * c1.x);
                    ry_inv_pair(c0, c1, k[j]);
                    ry_inv_pair(l0, l1, k[j]);
                    dt[j] += (l1.x * c0.x + l1.y * c0.y) - (l0.x * c1.x + l0.y * c1.y);
                }
            }
        }
#pragma unroll
        for (int o = 0; o < A; ++o) {
            const uint32_t i = slot<kLds>(base ^ member_offset<K>(cg, o));
            cur[i] = c[o];
            lam[i] = l[o];
        }
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const double a = wave_sum(dt[j]);
        const double b = wave_sum(dp[j]);
        if (lane == 0) {
            red[wave][2 * (w0 + j)] = a;
            red[wave][2 * (w0 + j) + 1] = b;
        }
    }
}

template <int kThreads, bool kLds>
__device__ __forceinline__ void chunk_fwd_any(int K, double2 *st, int n, int w0,
                                              const WireCoef *lc, const FrameTables &ft,
                                              int fb, const double2 *rz) {
    switch (K) {
        case 1:
            chunk_fwd<1, kThreads, kLds>(st, n, w0, lc, ft, fb, rz);
            break;
        case 2:
            chunk_fwd<clamp_chunk<kThreads>(2), kThreads, kLds>(st, n, w0, lc, ft, fb, rz);
            break;
        case 3:
            chunk_fwd<clamp_chunk<kThreads>(3), kThreads, kLds>(st, n, w0, lc, ft, fb, rz);
            break;
        default:
            chunk_fwd<max_chunk<kThreads>(), kThreads, kLds>(st, n, w0, lc, ft, fb, rz);
            break;
    }
}

template <int kThreads, bool kLds>
__device__ __forceinline__ void chunk_bwd_any(int K, double2 *cur, double2 *lam, int n, int w0,
                                              const WireCoef *lc, const FrameTables &ft, int fb,
                                              double (*red)[kRedCols], const double2 *rz) {
    switch (K) {
        case 1:
            chunk_bwd<1, kThreads, kLds>(cur, lam, n, w0, lc, ft, fb, red, rz);
            break;
        case 2:
            chunk_bwd<clamp_chunk<kThreads>(2), kThreads, kLds>(cur, lam, n, w0, lc, ft, fb, red, rz);
            break;
        case 3:
            chunk_bwd<clamp_chunk<kThreads>(3), kThreads, kLds>(cur, lam, n, w0, lc, ft, fb, red, rz);
            break;
        default:
            chunk_bwd<max_chunk<kThreads>(), kThreads, kLds>(cur, lam, n, w0, lc, ft, fb, red, rz);
            break;
    }
}

template <int kThreads, bool kLds>
__device__ __forceinline__ void run_job(const BatchJob &job, double2 *cur, double2 *lam,
                                        FrameTables &ft, double (*red)[kRedCols], double2 *rzt) {
    const int n = job.n;
    const int L = job.layers;
    const uint32_t N = 1u << n;
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wave = t / kWave;
    constexpr int kWaves = kThreads / kWave;
    const WireCoef *__restrict__ coef = job.coef;
    int K = job.chunk < 1 ? 1 : job.chunk;
    K = K > max_chunk<kThreads>() ? max_chunk<kThreads>() : K;

    // Ring-CNOT tables, and the frame F = P of layer 0's ring CNOT.
    if (t < n) {
        uint32_t rp = 0;
        uint32_t rq = 0;
        for (int b = 0; b < n; ++b) {
            rp |= ((ring_fwd32(1u << b, n) >> t) & 1u) << b;
            rq |= ((ring_inv32(1u << b, n) >> t) & 1u) << b;
        }
        ft.row_p[t] = rp;
        ft.row_q[t] = rq;
        ft.col_p[t] = ring_fwd32(1u << t, n);
        ft.col_q[t] = ring_inv32(1u << t, n);
        ft.r[0][t] = rp;
        ft.m[0][t] = ring_inv32(1u << t, n);
    }
    int fb = 0;
    // Layer 0 on |0...0> is the product state chi[j] = prod_w u_w[bit_w(j)]; in
    // frame P it is stored as is.
    for (uint32_t j = t; j < N; j += kThreads) {
        double2 v = make_double2(1.0, 0.0);
        for (int w = 0; w < n; ++w) {
            const WireCoef k = coef[w];
            v = cmul(v, ((j >> w) & 1) ? u1_of(k) : u0_of(k));
        }
        cur[slot<kLds>(j)] = v;
    }
    __syncthreads();

    for (int l = 1; l < L; ++l) {
        const WireCoef *lc = coef + (size_t)l * n;
        build_chunk_rz(rzt, n, K, lc);
        __syncthreads();
        for (int w0 = 0; w0 < n; w0 += K) {
            chunk_fwd_any<kThreads, kLds>(K < n - w0 ? K : n - w0, cur, n, w0, lc, ft, fb,
                                          rzt + (w0 / K) * (1 << K));
            __syncthreads();
        }
        frame_step(ft, n, fb, true);
    }

    // lam = H psi and the energy, in the final frame.
    double e = 0.0;
    for (uint32_t j = t; j < N; j += kThreads) {
        uint32_t i = 0;
        for (int w = 0; w < n; ++w) {
            i |= parity(j & ft.r[fb][w]) << w;
        }
        const double2 v = cur[slot<kLds>(j)];
        const double d = ham_diag(i, n);
        double fx = 0.0;
        double fy = 0.0;
        for (int w = 0; w < n; ++w) {
            const double2 p = cur[slot<kLds>(j ^ ft.m[fb][w])];
            fx += p.x;
            fy += p.y;
        }
        const double2 hv = make_double2(d * v.x - job.field * fx, d * v.y - job.field * fy);
        lam[slot<kLds>(j)] = hv;
        e += v.x * hv.x + v.y * hv.y;
    }
    e = wave_sum(e);
    if (lane == 0) {
        red[wave][0] = e;
    }
    __syncthreads();
    if (t == 0) {
        double acc = 0.0;
        for (int w = 0; w < kWaves; ++w) {
            acc += red[w][0];
        }
        *job.energy = acc;
    }
    __syncthreads();
    if (job.grad == nullptr) {
        return;
    }
    const double gscale = job.scale != nullptr ? *job.scale : 1.0;

    for (int l = L - 1; l >= 0; --l) {
        frame_step(ft, n, fb, false);  // inverse ring CNOT on both states
        const WireCoef *lc = coef + (size_t)l * n;
        build_chunk_rz(rzt, n, K, lc);
        __syncthreads();
        for (int w0 = 0; w0 < n; w0 += K) {
            chunk_bwd_any<kThreads, kLds>(K < n - w0 ? K : n - w0, cur, lam, n, w0, lc, ft, fb,
                                          red, rzt + (w0 / K) * (1 << K));
            __syncthreads();
        }
        if (t < 2 * n) {
            double acc = 0.0;
            for (int w = 0; w < kWaves; ++w) {
                acc += red[w][t];
            }
            job.grad[(size_t)l * n * 2 + t] = acc * gscale;
        }
        __syncthreads();
    }

    if (job.drift != nullptr) {
        // the frame is the identity again
        double acc = 0.0;
        for (uint32_t i = t; i < N; i += kThreads) {
            const double2 v = cur[slot<kLds>(i)];
            const double re = v.x - (i == 0 ? 1.0 : 0.0);
            acc += re * re + v.y * v.y;
        }
        acc = wave_sum(acc);
        if (lane == 0) {
            red[wave][0] = acc;
        }
        __syncthreads();
        if (t == 0) {
            double s = 0.0;
            for (int w = 0; w < kWaves; ++w) {
                s += red[w][0];
            }
            *job.drift = s;
        }
        __syncthreads();
    }
}

// Workgroup b runs circuits order[wg_begin[b] .. wg_begin[b+1]); without
// wg_begin it runs circuit b (jobs[b], or `single` when jobs is null).
template <int kThreads>
__global__ __launch_bounds__(kThreads) void batch_structured_kernel(
    const BatchJob *__restrict__ jobs, BatchJob single, const int *__restrict__ order,
    const int *__restrict__ wg_begin, double2 *scratch, uint64_t scratch_stride,
    uint32_t lds_state) {
    extern __shared__ double2 smem[];
    __shared__ double red[kThreads / kWave][kRedCols];
    __shared__ FrameTables ft;
    __shared__ double2 rzt[kBatchRzEntries];
    const int b = blockIdx.x;
    const int q0 = wg_begin != nullptr ? wg_begin[b] : b;
    const int q1 = wg_begin != nullptr ? wg_begin[b + 1] : b + 1;
    for (int q = q0; q < q1; ++q) {
        const int idx = order != nullptr ? order[q] : q;
        const BatchJob job = jobs != nullptr ? jobs[idx] : single;
        if (job.n <= kBatchMaxLdsQubits) {
            run_job<kThreads, true>(job, smem, smem + lds_state, ft, red, rzt);
        } else {
            const uint64_t N = uint64_t{1} << job.n;
            double2 *gs = scratch + (uint64_t)b * scratch_stride;
            run_job<kThreads, false>(job, gs, gs + N, ft, red, rzt);
        }
        __syncthreads();
    }
}

uint32_t batch_lds_state_amps(int max_lds_n) {
    if (max_lds_n <= 0) {
        return 0;
    }
    const uint32_t N = 1u << max_lds_n;
    return N + (N >> 4);
}

void launch_batch_structured(const BatchJob *d_jobs, const BatchJob &single, const int *d_order,
                             const int *d_wg_begin, int grid, int threads, int max_lds_n,
                             double2 *scratch, uint64_t scratch_stride, hipStream_t s) {
    const uint32_t lds_state = batch_lds_state_amps(max_lds_n);
    const size_t bytes = sizeof(double2) * 2 * (size_t)lds_state;
    switch (threads) {
        case 256:
            batch_structured_kernel<256><<<grid, 256, bytes, s>>>(d_jobs, single, d_order,
                                                                  d_wg_begin, scratch,
                                                                  scratch_stride, lds_state);
            break;
        case 512:
            batch_structured_kernel<512><<<grid, 512, bytes, s>>>(d_jobs, single, d_order,
                                                                  d_wg_begin, scratch,
                                                                  scratch_stride, lds_state);
            break;
        default:
            batch_structured_kernel<1024><<<grid, 1024, bytes, s>>>(d_jobs, single, d_order,
                                                                    d_wg_begin, scratch,
                                                                    scratch_stride, lds_state);
            break;
    }
}

// ---------------------------------------------------------------------------
// Optimizer steps over a packed parameter array.
// ---------------------------------------------------------------------------

__global__ void gd_update_kernel(double *__restrict__ p, const double *__restrict__ g,
                                 uint64_t count, double lr) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        p[i] = p[i] - lr * g[i];
    }
}

// Adam as torch.optim.Adam computes it (no weight decay, no amsgrad):
// m = b1 m + (1 - b1) g; v = b2 v + (1 - b2) g^2;
// p -= (lr / (1 - b1^t)) * m / (sqrt(v) / sqrt(1 - b2^t) + eps), t = step + 1.
__global__ void adam_update_kernel(double *__restrict__ p, const double *__restrict__ g,
                                   double *__restrict__ m, double *__restrict__ v, uint64_t count,
                                   double lr, double b1, double b2, double eps,
                                   const int *__restrict__ step) {
    const double tt = (double)(*step + 1);
    const double bc1 = 1.0 - pow(b1, tt);
    const double bc2 = 1.0 - pow(b2, tt);
    const double step_size = lr / bc1;
    const double bc2_sqrt = sqrt(bc2);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const double gi = g[i];
        const double mi = b1 * m[i] + (1.0 - b1) * gi;
        const double vi = b2 * v[i] + (1.0 - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const double denom = sqrt(vi) / bc2_sqrt + eps;
        p[i] = p[i] - step_size * (mi / denom);
    }
}

// history[step][j] = energies[j] while step < capacity.
__global__ void record_energies_kernel(const double *__restrict__ energies, int count,
                                       double *__restrict__ history, const int *__restrict__ step,
                                       int capacity) {
    const int s = *step;
    if (s >= capacity) {
        return;
    }
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += gridDim.x * blockDim.x) {
        history[(size_t)s * count + j] = energies[j];
    }
}

__global__ void step_advance_kernel(int *__restrict__ step) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        step[0] = step[0] + 1;
    }
}

__global__ void fill_kernel(double *__restrict__ p, uint64_t count, double value) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        p[i] = value;
    }
}

__global__ void set_int_kernel(int *__restrict__ p, int value) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        p[0] = value;
    }
}

void launch_gd_update(double *p, const double *g, uint64_t count, double lr, hipStream_t s) {
    gd_update_kernel<<<elementwise_grid(count), 256, 0, s>>>(p, g, count, lr);
}

void launch_adam_update(double *p, const double *g, double *m, double *v, uint64_t count,
                        double lr, double b1, double b2, double eps, const int *step,
                        hipStream_t s) {
    adam_update_kernel<<<elementwise_grid(count), 256, 0, s>>>(p, g, m, v, count, lr, b1, b2, eps,
                                                               step);
}

void launch_record_energies(const double *energies, int count, double *history, const int *step,
                            int capacity, hipStream_t s) {
    const int grid = (count + 255) / 256;
    record_energies_kernel<<<grid > 0 ? grid : 1, 256, 0, s>>>(energies, count, history, step,
                                                               capacity);
}

void launch_step_advance(int *step, hipStream_t s) { step_advance_kernel<<<1, 64, 0, s>>>(step); }

void launch_fill(double *p, uint64_t count, double value, hipStream_t s) {
    fill_kernel<<<elementwise_grid(count), 256, 0, s>>>(p, count, value);
}

void launch_set_int(int *p, int value, hipStream_t s) { set_int_kernel<<<1, 64, 0, s>>>(p, value); }

}  // namespace rie
