// HIP kernels of the ring-Ising statevector engine for gfx950 (MI355X).
//
// Design (see the header of each kernel):
//  * Large states stream through HBM in tile passes.  A workgroup loads a
//    tile of 2^T amplitudes (T = 12) into registers, 2^R per thread, applies
//    the rotations of the R tile bits held in registers, and exchanges the
//    tile through LDS between its T / R register mappings so every tile bit
//    is register-resident once.
//    Global loads/stores are 16-byte-per-lane and contiguous across lanes.
//  * The ring-CNOT layer is a GF(2)-linear index permutation.  The image of a
//    low tile (index bits [0, T)) is two contiguous runs, so the permutation
//    rides on the last LDS exchange of the forward low pass (scatter) and on
//    the first exchange of the backward low pass (gather).
//  * Small circuits (n <= 12) run entirely in one workgroup's LDS.
//  * Reductions (energy, gradients) are per-workgroup partials followed by a
//    fixed-order reduction kernel: results are bitwise reproducible.
#include "rie_kernels.hpp"

#include <algorithm>
#include <stdexcept>

namespace rie {

// ---------------------------------------------------------------------------
// Tile helpers
// ---------------------------------------------------------------------------

// A tile is 2^12 amplitudes (T = 12) held as 2^R registers by 2^(T - R)
// threads, visited in P = T / R register mappings:
//   R = 4: 256 threads x 16 registers, P = 3;
//   R = 3: 512 threads x 8 registers, P = 4 (half the registers per thread,
//          so twice the waves per SIMD, at one more LDS exchange per pass).
template <int R>
struct Tile {
    static constexpr int T = 12;
    static constexpr int P = T / R;
    static constexpr uint32_t SIZE = 1u << T;
    static constexpr int THREADS = 1 << (T - R);
    static constexpr int REGS = 1 << R;
    static constexpr int WAVES = THREADS / kWave;
    static constexpr uint32_t SLOTS = SIZE + (SIZE >> R);
    static_assert(P * R == T, "tile bits must split into whole mappings");
};

// Register/thread mappings of a tile offset o in [0, 2^T): mapping M holds
// offset bits [lo, lo + R) in registers, lo = T - R for M = 0 (threads hold
// the contiguous low bits: loads), lo = (M - 1) R for M >= 1; the threads hold
// the other bits in order.  Mapping P - 1 stores.
template <int R, int M>
constexpr int phase_first_bit() {
    return M == 0 ? Tile<R>::T - R : (M - 1) * R;
}

template <int R, int M>
__device__ __forceinline__ uint32_t omap(uint32_t t, int k) {
    constexpr int lo = phase_first_bit<R, M>();
    return ((t >> lo) << (lo + R)) | ((uint32_t)k << lo) | (t & ((1u << lo) - 1u));
}

// LDS slot of offset o: one pad amplitude every 2^R keeps the strided
// mapping B conflict-free.
template <int R>
__device__ __forceinline__ uint32_t lds_slot(uint32_t o) {
    return o + (o >> R);
}

__device__ __forceinline__ uint64_t tile_base(uint64_t g, const PassGeom &G, int T) {
    const int gap = G.h0 - G.a;
    const uint64_t lo = g & ((uint64_t{1} << gap) - 1);
    const uint64_t hi = g >> gap;
    return (lo << G.a) | (hi << (G.h0 + T - G.a));
}

__device__ __forceinline__ uint64_t tile_index(uint64_t base, uint32_t o, const PassGeom &G) {
    const uint64_t lo = o & ((1u << G.a) - 1u);
    const uint64_t hi = (uint64_t)(o >> G.a) << G.h0;
    return base | lo | hi;
}

__device__ __forceinline__ int tile_wire(int j, const PassGeom &G) {
    return j < G.a ? j : G.h0 + (j - G.a);
}

// Scattered tiles (sc_n > 0): offset bits [0, a) are index bits [0, a), offset
// bit a + i is index bit sc[i]; the tile number fills the other index bits in
// order.
__device__ __forceinline__ uint64_t scat_index(uint32_t o, const PassGeom &G) {
    uint64_t r = o & ((1u << G.a) - 1u);
    for (int i = 0; i < G.sc_n; ++i) {
        r |= (uint64_t)((o >> (G.a + i)) & 1u) << G.sc[i];
    }
    return r;
}

__device__ __forceinline__ uint64_t scat_base(uint64_t g, const PassGeom &G) {
    uint64_t b = g << G.a;
    for (int i = 0; i < G.sc_n; ++i) {  // open a zero bit at each sc[i], ascending
        const int p = G.sc[i];
        b = (b & ((uint64_t{1} << p) - 1)) | ((b >> p) << (p + 1));
    }
    return b;
}

// Per-thread global index offsets of the load mapping (0, "A") and the store
// mapping (P - 1, "C") and the uniform index step of their register index k
// (k covers offset bits [T - R, T) in A and [(P - 2) R, (P - 1) R) in C, which
// lie entirely in the low or the high part of the tile).
struct TileAddr {
    uint64_t offA;
    uint64_t offC;
    uint64_t strideA;
    uint64_t strideC;
};

template <int R>
__device__ __forceinline__ TileAddr tile_addr(uint32_t t, const PassGeom &G) {
    TileAddr ta;
    constexpr int T = Tile<R>::T;
    constexpr int lo_c = phase_first_bit<R, Tile<R>::P - 1>();
    ta.offA = tile_index(0, omap<R, 0>(t, 0), G);
    ta.offC = tile_index(0, omap<R, Tile<R>::P - 1>(t, 0), G);
    if (G.a >= T) {
        ta.strideA = uint64_t{1} << (T - R);
        ta.strideC = uint64_t{1} << lo_c;
    } else {
        ta.strideA = uint64_t{1} << (G.h0 + T - R - G.a);
        ta.strideC = uint64_t{1} << (G.h0 + lo_c - G.a);
    }
    return ta;
}

template <int R, int FROM, int TO>
__device__ __forceinline__ void exchange(double2 (&v)[Tile<R>::REGS], double2 *lds, uint32_t t) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < Tile<R>::REGS; ++k) {
        lds[lds_slot<R>(omap<R, FROM>(t, k))] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < Tile<R>::REGS; ++k) {
        v[k] = lds[lds_slot<R>(omap<R, TO>(t, k))];
    }
}

// RZ factors of the register windows.  RZ on one wire commutes with RY on
// the others, so a window's rotations RZ RY (per wire) are all its RY
// followed by one diagonal phase per register: tab[M * 2^R + k] is the
// product over the window's rotated bits of e^{-i phi/2} (bit of k clear) or
// e^{+i phi/2} (set).  The adjoint applies the conjugate first.  Built once
// per workgroup (the table does not depend on the tile).
template <int R>
__device__ __forceinline__ void build_rz_tables(double2 *tab, const WireCoef *__restrict__ lc,
                                                const PassGeom &G, uint32_t t) {
    using TL = Tile<R>;
    for (uint32_t i = t; i < (uint32_t)(TL::P * TL::REGS); i += TL::THREADS) {
        const int M = (int)(i / TL::REGS);
        const int k = (int)(i % TL::REGS);
        const int lo = M == 0 ? TL::T - R : (M - 1) * R;
        double2 q = make_double2(1.0, 0.0);
        for (int rb = 0; rb < R; ++rb) {
            const int j = lo + rb;
            if (j >= G.rot_lo && j < G.rot_hi) {
                const WireCoef w = lc[tile_wire(j, G)];
                q = cmul(q, make_double2(w.ch, ((k >> rb) & 1) ? w.sh : -w.sh));
            }
        }
        tab[i] = q;
    }
}

template <int R, int M>
__device__ __forceinline__ bool window_rotates(const PassGeom &G) {
    constexpr int P0 = phase_first_bit<R, M>();
    return G.rot_lo < P0 + R && G.rot_hi > P0;
}

// Branch-free: tile bits outside [rot_lo, rot_hi) carry the identity in
// tile_coef and phase 1, so the register array needs no copies at control-flow
// joins.  kHigh: tile bits 0..2 of a high pass never rotate (the contiguous low
// index bits that ride along) and are skipped at compile time.
template <int R, int M, bool kHigh>
__device__ __forceinline__ void phase_fwd(double2 (&v)[Tile<R>::REGS], const WireCoef *tile_coef,
                                          const double2 *rz_tab) {
    constexpr int P0 = phase_first_bit<R, M>();
    constexpr int kSkip = kHigh ? (3 - P0 > 0 ? (3 - P0 < R ? 3 - P0 : R) : 0) : 0;
    if constexpr (kSkip == R) {
        return;
    }
#pragma unroll
    for (int rb = kSkip; rb < R; ++rb) {
        const WireCoef k = tile_coef[P0 + rb];
#pragma unroll
        for (int q = 0; q < Tile<R>::REGS; ++q) {
            if (!(q & (1 << rb))) {
                ry_pair(v[q], v[q | (1 << rb)], k);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < Tile<R>::REGS; ++q) {
        v[q] = cmul(v[q], rz_tab[M * Tile<R>::REGS + q]);
    }
}

// Per-wire gradient sums of one tile are reduced across each row of 16 lanes
// right away (DPP) and added by the row's first lane to that row's LDS slot
// (fixed order, no atomics), so no per-thread accumulators stay live across
// phases and tiles.
// The window's RZ^dagger as one phase per register comes first.  dE/dphi of
// a wire, sum of Im(l* c) signed by the wire's bit (Im(l0* c0) - Im(l1* c1)
// per pair), does not change under the phases or under RY^dagger of the
// window's other wires, so one sweep over the registers forms it for every
// wire of the window.  Then per wire RY^dagger on both states and
// dE/dtheta = Re(l1* c0) - Re(l0* c1) just before the wire's rotations (the
// quantities of bwd_pair).
// kSweepDp: form dE/dphi of all the window's wires in the phase sweep (fewer
// FP64 operations for windows of several rotated wires, four more live
// registers); otherwise per pair next to RY^dagger.
template <int R, int M, bool kSweepDp, bool kHigh = false>
__device__ __forceinline__ void phase_bwd(double2 (&c)[Tile<R>::REGS], double2 (&l)[Tile<R>::REGS],
                                          const WireCoef *tile_coef, double *__restrict__ row_acc,
                                          int lane, const double2 *rz_tab) {
    constexpr int P0 = phase_first_bit<R, M>();
    double dps[kSweepDp ? R : 1];
#pragma unroll
    for (int rb = 0; rb < (kSweepDp ? R : 1); ++rb) {
        dps[rb] = 0.0;
    }
    // branch-free: unrotated tile bits carry phase 1 and the identity in
    // tile_coef, and their gradient slots are not read.  kHigh: tile bits 0..2
    // of a high pass never rotate (the contiguous low index bits that ride
    // along) and are skipped at compile time.
    constexpr int kSkip = kHigh ? (3 - P0 > 0 ? (3 - P0 < R ? 3 - P0 : R) : 0) : 0;
    if constexpr (kSkip == R) {
        return;
    }
#pragma unroll
    for (int q = 0; q < Tile<R>::REGS; ++q) {
        const double2 z = rz_tab[M * Tile<R>::REGS + q];
        const double2 zc = make_double2(z.x, -z.y);
        c[q] = cmul(c[q], zc);
        l[q] = cmul(l[q], zc);
        if constexpr (kSweepDp) {
            const double im = l[q].x * c[q].y - l[q].y * c[q].x;
#pragma unroll
            for (int rb = 0; rb < R; ++rb) {
                dps[rb] += ((q >> rb) & 1) ? -im : im;
            }
        }
    }
#pragma unroll
    for (int rb = R - 1; rb >= kSkip; --rb) {
        const int j = P0 + rb;
        const WireCoef k = tile_coef[j];
        double dt = 0.0;
        double dp = 0.0;
#pragma unroll
        for (int q = 0; q < Tile<R>::REGS; ++q) {
            if (!(q & (1 << rb))) {
                double2 &c0 = c[q];
                double2 &c1 = c[q | (1 << rb)];
                double2 &l0 = l[q];
                double2 &l1 = l[q | (1 << rb)];
                if constexpr (!kSweepDp) {
                    dp += (l0.x * c0.y - l0.y * c0.x) - (l1.x * c1.y - l1.y * c1.x);
                }
                ry_inv_pair(c0, c1, k);
                ry_inv_pair(l0, l1, k);
                dt += (l1.x * c0.x + l1.y * c0.y) - (l0.x * c1.x + l0.y * c1.y);
            }
        }
        if constexpr (kSweepDp) {
            dp = dps[rb];
        }
        dt = row_sum(dt);
        dp = row_sum(dp);
        if ((lane & 15) == 0) {
            row_acc[2 * j] += dt;
            row_acc[2 * j + 1] += dp;
        }
    }
}

// ---------------------------------------------------------------------------
// Parameters -> coefficient table
// ---------------------------------------------------------------------------

__global__ void coef_kernel(const double *__restrict__ params, WireCoef *__restrict__ coef,
                            int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) {
        return;
    }
    WireCoef k;
    sincos(0.5 * params[2 * i], &k.s, &k.c);
    sincos(0.5 * params[2 * i + 1], &k.sh, &k.ch);
    coef[i] = k;
}

void launch_coef(const double *d_params, WireCoef *d_coef, int count, hipStream_t s) {
    const int threads = 128;
    coef_kernel<<<(count + threads - 1) / threads, threads, 0, s>>>(d_params, d_coef, count);
}

// ---------------------------------------------------------------------------
// Layer 0 + ring CNOT as one write pass: out[j] = prod_w u_w[bit_w(P^-1 j)].
// The product over up to 40 wires is split into 5 byte-wide LDS tables.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void init_product_cnot_kernel(double2 *__restrict__ out,
                                                                const WireCoef *__restrict__ lc0,
                                                                int n, uint64_t size) {
    __shared__ double2 table[5][256];
    const int t = threadIdx.x;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
        double2 v = make_double2(1.0, 0.0);
        for (int bit = 0; bit < 8; ++bit) {
            const int w = 8 * b + bit;
            if (w < n) {
                const WireCoef k = lc0[w];
                v = cmul(v, ((t >> bit) & 1) ? u1_of(k) : u0_of(k));
            }
        }
        table[b][t] = v;
    }
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + t; j < size; j += stride) {
        const uint64_t x = ring_inv(j, n);
        double2 v = cmul(table[0][x & 255], table[1][(x >> 8) & 255]);
        v = cmul(v, table[2][(x >> 16) & 255]);
        v = cmul(v, table[3][(x >> 24) & 255]);
        v = cmul(v, table[4][(x >> 32) & 255]);
        out[j] = v;
    }
}

int elementwise_grid(uint64_t work_items) {
    const uint64_t blocks = (work_items + 255) / 256;
    return (int)std::min<uint64_t>(std::max<uint64_t>(blocks, 1), 8192);
}

void launch_init_product_cnot(double2 *out, const WireCoef *coef_layer0, int n, hipStream_t s) {
    const uint64_t size = uint64_t{1} << n;
    init_product_cnot_kernel<<<elementwise_grid(size), 256, 0, s>>>(out, coef_layer0, n, size);
}

// ---------------------------------------------------------------------------
// Forward tile pass.
// ---------------------------------------------------------------------------

// Waves per SIMD to budget registers for: two 512-thread workgroups per CU
// (four waves per SIMD) for the R = 3 shape, whose LDS allows two.
template <int R>
constexpr int tile_min_blocks() {
    return R == 3 ? 4 : 2;
}

// Stores a low tile (a == T) held in mapping M through the ring-CNOT
// permutation: P(base ^ o) = P(base) ^ P_T(o) ^ (parity(o) ? hi : 0), and bit
// T-1 of P_T(o) is parity(o), so the image is two contiguous runs.  The tile
// is placed at P_T(o) in LDS and written out run by run.
template <int R, int M>
__device__ __forceinline__ void store_ring_scatter(const double2 (&v)[Tile<R>::REGS],
                                                   double2 *lds, uint32_t t, uint64_t base,
                                                   const PassGeom &G, double2 *out) {
    using TL = Tile<R>;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TL::REGS; ++k) {
        lds[lds_slot<R>(ring_fwd32(omap<R, M>(t, k), TL::T))] = v[k];
    }
    __syncthreads();
    // q = t | k << (T - R) is contiguous in the image run; P(base) has zero
    // bits [1, T) and its bit 0 only swaps neighbours inside the run.
    const uint64_t pbase = ring_fwd(base, G.n);
    const uint64_t lo_flip = pbase & 1u;
    const uint64_t run = (pbase & ~uint64_t{TL::SIZE - 1}) + (t ^ lo_flip);
    double2 *dst0 = out + run;
    double2 *dst1 = out + (run ^ G.hi_mask);
#pragma unroll
    for (int k = 0; k < TL::REGS; ++k) {
        const double2 val = lds[lds_slot<R>(omap<R, 0>(t, k))];
        if (k < TL::REGS / 2) {
            dst0[(uint64_t)k << (TL::T - R)] = val;
        } else {
            dst1[(uint64_t)k << (TL::T - R)] = val;
        }
    }
}

// With kGenInit the pass does not read its input: the tile of the layer-0
// state after its ring CNOT, chi(i) = prod_w u_w(bit w of P^-1(i)), is
// generated from nibble tables of the layer-0 factors (lc0).
constexpr int kInitNibbles = 9;  // covers n <= 36

template <int R, bool kScatter, bool kGenInit>
__global__ __launch_bounds__(Tile<R>::THREADS, tile_min_blocks<R>()) void fwd_pass_kernel(
    const double2 *in, double2 *out, const WireCoef *__restrict__ lc,
    const WireCoef *__restrict__ lc0, PassGeom G) {
    using TL = Tile<R>;
    __shared__ double2 lds[TL::SLOTS];
    __shared__ double2 nib[kGenInit ? kInitNibbles : 1][16];
    __shared__ double2 rz_tab[TL::P * TL::REGS];
    __shared__ WireCoef tile_coef[TL::T];
    const uint32_t t = threadIdx.x;
    const TileAddr ta = tile_addr<R>(t, G);
    build_rz_tables<R>(rz_tab, lc, G, t);
    if (t < (uint32_t)TL::T) {
        const bool rot = (int)t >= G.rot_lo && (int)t < G.rot_hi;
        tile_coef[t] = rot ? lc[tile_wire((int)t, G)] : WireCoef{1.0, 0.0, 1.0, 0.0};
    }
    if constexpr (!kGenInit) {
        __syncthreads();  // gen_init syncs after its nibble tables
    }
    if constexpr (kGenInit) {
        for (uint32_t e = t; e < kInitNibbles * 16; e += TL::THREADS) {
            const int q = (int)(e >> 4);
            const int bits = (int)(e & 15);
            double2 v = make_double2(1.0, 0.0);
            for (int j = 0; j < 4; ++j) {
                const int w = 4 * q + j;
                if (w < G.n) {
                    const WireCoef k = lc0[w];
                    v = cmul(v, ((bits >> j) & 1) ? u1_of(k) : u0_of(k));
                }
            }
            nib[q][bits] = v;
        }
        __syncthreads();
    }
    // The next tile's loads are issued before this tile's phases, so they
    // stay in flight across the LDS exchanges (whose barriers wait on LDS
    // traffic only).
    double2 v[TL::REGS];
    double2 nv[TL::REGS];
    if constexpr (!kGenInit) {
        if ((uint64_t)blockIdx.x < G.num_tiles) {
            const double2 *src = in + tile_base(blockIdx.x, G, TL::T) + ta.offA;
#pragma unroll
            for (int k = 0; k < TL::REGS; ++k) {
                nv[k] = src[k * ta.strideA];
            }
        }
    }
    // Generation: x = P^-1(index) is GF(2)-linear and the base, thread and
    // register parts of an index are disjoint, so x = P^-1(base) ^ xt ^ xk[k].
    // Nibbles of x that no tile bit reaches are the same for the whole tile:
    // their product is formed once per tile.
    uint32_t vary = 0;  // bit q: nibble q of x depends on the tile offset
    uint64_t xt = 0;
    uint64_t xk[kGenInit ? TL::REGS : 1];
    if constexpr (kGenInit) {
        for (int b = 0; b < TL::T; ++b) {
            const uint64_t vb = ring_inv(tile_index(0, 1u << b, G), G.n);
            for (int q = 0; q < kInitNibbles; ++q) {
                vary |= ((vb >> (4 * q)) & 15u) != 0 ? (1u << q) : 0u;
            }
        }
        xt = ring_inv(ta.offA, G.n);
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            xk[k] = ring_inv((uint64_t)k * ta.strideA, G.n);
        }
    }
    for (uint64_t g = blockIdx.x; g < G.num_tiles; g += gridDim.x) {
        const uint64_t base = tile_base(g, G, TL::T);
        if constexpr (kGenInit) {
            const uint64_t xb = ring_inv(base, G.n) ^ xt;
            double2 cb = make_double2(1.0, 0.0);
#pragma unroll
            for (int q = 0; q < kInitNibbles; ++q) {
                if (((vary >> q) & 1u) == 0) {
                    cb = cmul(cb, nib[q][(xb >> (4 * q)) & 15]);
                }
            }
#pragma unroll
            for (int k = 0; k < TL::REGS; ++k) {
                const uint64_t x = xb ^ xk[k];
                double2 p = cb;
#pragma unroll
                for (int q = 0; q < kInitNibbles; ++q) {
                    if ((vary >> q) & 1u) {
                        p = cmul(p, nib[q][(x >> (4 * q)) & 15]);
                    }
                }
                v[k] = p;
            }
        } else {
#pragma unroll
            for (int k = 0; k < TL::REGS; ++k) {
                v[k] = nv[k];
            }
            const uint64_t gn = g + gridDim.x;
            if (gn < G.num_tiles) {
                const double2 *src = in + tile_base(gn, G, TL::T) + ta.offA;
#pragma unroll
                for (int k = 0; k < TL::REGS; ++k) {
                    nv[k] = src[k * ta.strideA];
                }
            }
        }
        constexpr bool kHigh = !kScatter;  // high passes: tile bits 0..2 ride along
        phase_fwd<R, 0, kHigh>(v, tile_coef, rz_tab);
        if constexpr (TL::P == 3) {
            exchange<R, 0, 1>(v, lds, t);
            phase_fwd<R, 1, kHigh>(v, tile_coef, rz_tab);
            exchange<R, 1, 2>(v, lds, t);
            phase_fwd<R, 2, kHigh>(v, tile_coef, rz_tab);
        } else {
            if constexpr (!kHigh) {  // the low tile: offset bits [0, R) rotate too
                exchange<R, 0, 1>(v, lds, t);
                phase_fwd<R, 1, kHigh>(v, tile_coef, rz_tab);
                exchange<R, 1, 2>(v, lds, t);
            } else {
                exchange<R, 0, 2>(v, lds, t);
            }
            phase_fwd<R, 2, kHigh>(v, tile_coef, rz_tab);
            exchange<R, 2, 3>(v, lds, t);
            phase_fwd<R, 3, kHigh>(v, tile_coef, rz_tab);
        }
        if constexpr (!kScatter) {
            double2 *dst = out + base + ta.offC;
#pragma unroll
            for (int k = 0; k < TL::REGS; ++k) {
                dst[k * ta.strideC] = v[k];
            }
        } else {
            store_ring_scatter<R, TL::P - 1>(v, lds, t, base, G, out);
        }
    }
}

template <int R>
static void launch_fwd_pass_r(bool scatter, bool gen_init, const double2 *in, double2 *out,
                              const WireCoef *lc, const WireCoef *lc0, const PassGeom &g,
                              int grid, hipStream_t s) {
    constexpr int TH = Tile<R>::THREADS;
    if (scatter ? g.a != Tile<R>::T || g.rot_lo != 0 : g.a != 3 || g.rot_lo < 3) {
        // the kernel fixes which tile bits rotate at compile time
        throw std::invalid_argument("forward pass geometry: low tile or 3 ride-along bits");
    }
    if (gen_init) {
        if (scatter) {
            fwd_pass_kernel<R, true, true><<<grid, TH, 0, s>>>(in, out, lc, lc0, g);
        } else {
            fwd_pass_kernel<R, false, true><<<grid, TH, 0, s>>>(in, out, lc, lc0, g);
        }
    } else if (scatter) {
        fwd_pass_kernel<R, true, false><<<grid, TH, 0, s>>>(in, out, lc, lc0, g);
    } else {
        fwd_pass_kernel<R, false, false><<<grid, TH, 0, s>>>(in, out, lc, lc0, g);
    }
}

void launch_fwd_pass(int R, bool scatter, bool gen_init, const double2 *in, double2 *out,
                     const WireCoef *layer_coef, const WireCoef *layer0_coef, const PassGeom &g,
                     int grid, hipStream_t s) {
    if (R == 4) {
        launch_fwd_pass_r<4>(scatter, gen_init, in, out, layer_coef, layer0_coef, g, grid, s);
    } else {
        launch_fwd_pass_r<3>(scatter, gen_init, in, out, layer_coef, layer0_coef, g, grid, s);
    }
}

// ---------------------------------------------------------------------------
// Layer-1 states generated from the parameters (chain_site): tables.
// ---------------------------------------------------------------------------

// One workgroup per table set: the 2n site matrices in LDS, then the
// register-run products tab[0, 2^R) and the chunk products C0 at tab[16 + v]
// and C1 at tab[32 + v].
__global__ __launch_bounds__(256) void chain_tables_kernel(const WireCoef *__restrict__ lc0,
                                                           const WireCoef *__restrict__ lc1,
                                                           ChainSetup cs, CMat2 *__restrict__ tabs) {
    __shared__ CMat2 site[2 * 64];
    const ChainSet st = cs.set[blockIdx.x];
    const int n = cs.n;
    for (int i = (int)threadIdx.x; i < 2 * n; i += (int)blockDim.x) {
        const int w = i >> 1;
        site[i] = chain_site(w, i & 1, lc0[w], lc1[w], ((st.rot >> w) & 1u) != 0,
                             ((st.ry_only >> w) & 1u) == 0);
    }
    __syncthreads();
    CMat2 *tab = tabs + (size_t)blockIdx.x * kChainTabEntries;
    for (int e = (int)threadIdx.x; e < kChainTabEntries; e += (int)blockDim.x) {
        int first;
        int len;
        int v;
        if (e < kChainC0) {
            first = st.r0;
            len = st.R;
            v = e;
        } else if (e < kChainC1) {
            first = st.c0_first;
            len = st.c0_len;
            v = e - kChainC0;
        } else {
            first = st.c1_first;
            len = st.c1_len;
            v = e - kChainC1;
        }
        if (len == 0 || v >= (1 << len)) {
            continue;
        }
        CMat2 p = site[2 * first + (v & 1)];
        for (int b = 1; b < len; ++b) {
            const int w = first + b;
            p = cmat_mul(p, site[2 * w + ((v >> b) & 1)]);
        }
        tab[e] = p;
    }
}

// Product of the site matrices of wires [first, first + len) at the bits of
// index i (the identity when len is 0).
__device__ __forceinline__ CMat2 chain_segment(const CMat2 *site, int first, int len,
                                               uint64_t i) {
    CMat2 p = {{1.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {1.0, 0.0}};
    for (int b = 0; b < len; ++b) {
        const int w = first + b;
        const CMat2 &m = site[2 * w + (int)((i >> w) & 1u)];
        p = (b == 0) ? m : cmat_mul(p, m);
    }
    return p;
}

// The tile-uniform segments of every tile: utab[2g] = U_A (wires [r0 + R, n)),
// utab[2g + 1] = U_B, at the bits of tile g's base index.  grid.y: table set.
__global__ __launch_bounds__(256) void chain_uniform_kernel(const WireCoef *__restrict__ lc0,
                                                            const WireCoef *__restrict__ lc1,
                                                            ChainSetup cs,
                                                            CMat2 *__restrict__ utab) {
    __shared__ CMat2 site[2 * 64];
    const ChainSet st = cs.set[blockIdx.y];
    const int n = cs.n;
    for (int i = (int)threadIdx.x; i < 2 * n; i += (int)blockDim.x) {
        const int w = i >> 1;
        site[i] = chain_site(w, i & 1, lc0[w], lc1[w], ((st.rot >> w) & 1u) != 0,
                             ((st.ry_only >> w) & 1u) == 0);
    }
    __syncthreads();
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= cs.tiles) {
        return;
    }
    PassGeom G{};
    G.a = st.ta;
    G.h0 = st.th0;
    const uint64_t base = tile_base(g, G, Tile<4>::T);
    CMat2 *u = utab + ((size_t)blockIdx.y * cs.tiles + g) * 2;
    u[0] = chain_segment(site, st.r0 + st.R, n - st.r0 - st.R, base);
    u[1] = chain_segment(site, st.ub_first, st.ub_len, base);
}

void launch_chain_tables(const WireCoef *lc0, const WireCoef *lc1, const ChainSetup &cs,
                         CMat2 *tabs, CMat2 *utab, hipStream_t s) {
    chain_tables_kernel<<<cs.count, 256, 0, s>>>(lc0, lc1, cs, tabs);
    const dim3 grid((unsigned)((cs.tiles + 255) / 256), (unsigned)cs.count);
    chain_uniform_kernel<<<grid, 256, 0, s>>>(lc0, lc1, cs, utab);
}

// The flipped segments of table set 0 per tile (see launch_chain_flips in the
// header).  The sum over the pairs above the tile is one left-to-right walk
// over the sites T .. n-1 with three running products: A the plain prefix,
// C the prefix whose last site is flipped (a pair waiting for its partner),
// B the sum of prefixes with one completed flipped pair.
__global__ __launch_bounds__(256) void chain_flip_kernel(const WireCoef *__restrict__ lc0,
                                                         const WireCoef *__restrict__ lc1,
                                                         ChainSetup cs, CMat2 *__restrict__ xtab) {
    __shared__ CMat2 site[2 * 64];
    const ChainSet st = cs.set[0];
    const int n = cs.n;
    for (int i = (int)threadIdx.x; i < 2 * n; i += (int)blockDim.x) {
        const int w = i >> 1;
        site[i] = chain_site(w, i & 1, lc0[w], lc1[w], ((st.rot >> w) & 1u) != 0,
                             ((st.ry_only >> w) & 1u) == 0);
    }
    __syncthreads();
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= cs.tiles) {
        return;
    }
    PassGeom G{};
    G.a = st.ta;
    G.h0 = st.th0;
    const int lo = st.r0 + st.R;  // the low tile's top: U_A covers wires [lo, n)
    const uint64_t base = tile_base(g, G, Tile<4>::T);
    CMat2 *x = xtab + g * kFlipSegments;
    x[0] = chain_segment(site, lo, n - lo, base ^ (uint64_t{1} << lo));
    x[1] = chain_segment(site, lo, n - lo, base ^ (uint64_t{1} << (n - 1)));
    const CMat2 zero = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    CMat2 A = site[2 * lo + (int)((base >> lo) & 1u)];
    CMat2 C = site[2 * lo + (int)(((base >> lo) & 1u) ^ 1u)];
    CMat2 B = zero;
    for (int v = lo + 1; v < n; ++v) {
        const int b = (int)((base >> v) & 1u);
        const CMat2 &sv = site[2 * v + b];
        const CMat2 &sf = site[2 * v + (b ^ 1)];
        B = cmat_add(cmat_mul(B, sv), cmat_mul(C, sf));
        C = cmat_mul(A, sf);
        A = cmat_mul(A, sv);
    }
    x[2] = B;
}

void launch_chain_flips(const WireCoef *lc0, const WireCoef *lc1, const ChainSetup &cs,
                        CMat2 *xtab, hipStream_t s) {
    chain_flip_kernel<<<(unsigned)((cs.tiles + 255) / 256), 256, 0, s>>>(lc0, lc1, cs, xtab);
}

// Y of the forward's low tile at row i (run bits zero): U_A . C0 . C1 with C0
// the index bits 0..3 and C1 the bits 4 .. T - R - 1.
template <int R>
__device__ __forceinline__ CMat2 chain_y_low(const CMat2 *__restrict__ tab,
                                             const CMat2 *__restrict__ u, uint64_t i) {
    constexpr int kC1Bits = Tile<R>::T - R - 4;
    const CMat2 c0 = tab[kChainC0 + (int)(i & 15u)];
    const CMat2 c1 = tab[kChainC1 + (int)((i >> 4) & ((1u << kC1Bits) - 1u))];
    return cmat_mul(cmat_mul(u[0], c0), c1);
}

// Y of an adjoint high pass's tile at row i (run bits zero):
// U_A . C0 . U_B . C1 with C0 the index bits [0, a) and C1 the bits
// [h0, h0 + T - R - a); chunks: C0 at chunks[0, 16), C1 at chunks[16, 80).
template <int R>
__device__ __forceinline__ CMat2 chain_y_high(const CMat2 *chunks, const CMat2 *__restrict__ u,
                                              uint64_t i, const PassGeom &G) {
    const uint32_t v0 = (uint32_t)i & ((1u << G.a) - 1u);
    const uint32_t v1 = (uint32_t)(i >> G.h0) & ((1u << (Tile<R>::T - R - G.a)) - 1u);
    CMat2 y = cmat_mul(u[0], chunks[v0]);
    y = cmat_mul(y, u[1]);
    return cmat_mul(y, chunks[kChainC1 - kChainC0 + v1]);
}

// v[k] = Tr(run[k] Y): register bit b is the run's wire r0 + b.
template <int R>
__device__ __forceinline__ void chain_fill(double2 (&v)[Tile<R>::REGS],
                                           const CMat2 *__restrict__ tab, const CMat2 &y) {
#pragma unroll
    for (int k = 0; k < Tile<R>::REGS; ++k) {
        v[k] = cmat_trace_mul(tab[k], y);
    }
}

// Layer 1 in one write pass: the low tile of V_1 P chi (every wire of layer 1
// rotated; tables with r0 = T - R) generated in mapping 0, then the layer's
// ring CNOT through LDS.
template <int R>
__global__ __launch_bounds__(Tile<R>::THREADS, tile_min_blocks<R>()) void gen_layer1_kernel(
    double2 *out, const CMat2 *__restrict__ tab, const CMat2 *__restrict__ utab, PassGeom G) {
    using TL = Tile<R>;
    __shared__ double2 lds[TL::SLOTS];
    const uint32_t t = threadIdx.x;
    for (uint64_t g = blockIdx.x; g < G.num_tiles; g += gridDim.x) {
        uint32_t tl = t;
        asm volatile("" : "+v"(tl));
        const uint64_t base = tile_base(g, G, TL::T);
        const CMat2 y = chain_y_low<R>(tab, utab + 2 * g, base + omap<R, 0>(tl, 0));
        double2 v[TL::REGS];
        chain_fill<R>(v, tab, y);
        store_ring_scatter<R, 0>(v, lds, tl, base, G, out);
    }
}

void launch_gen_layer1(int R, double2 *out, const CMat2 *tab, const CMat2 *utab,
                       const PassGeom &g, int grid, hipStream_t s) {
    if (R == 4) {
        gen_layer1_kernel<4><<<grid, Tile<4>::THREADS, 0, s>>>(out, tab, utab, g);
    } else {
        gen_layer1_kernel<3><<<grid, Tile<3>::THREADS, 0, s>>>(out, tab, utab, g);
    }
}

// ---------------------------------------------------------------------------
// Adjoint tile pass.
// ---------------------------------------------------------------------------


// Register window of mapping M as a mask of tile offset bits.
template <int R, int M>
constexpr uint32_t window_mask() {
    return ((1u << R) - 1u) << phase_first_bit<R, M>();
}

// Transverse-field terms of the adjacent tile-bit pairs (j, j + 1),
// pair_lo <= j < pair_hi - 1: lam += -field * X_j X_{j+1} cur (their energy is
// part of the layer-0 contraction's).  Called before un-rotating the window of
// mapping M, it adds the pairs that touch the window and no bit of `done`
// (the windows un-rotated earlier in the pass, whose pairs were added then):
// each pair is added once, while both of its bits are still rotated, so it
// commutes with everything un-rotated before.  Pairs inside the window read
// registers; pairs with one bit outside read the LDS copy of the tile, which
// must hold cur at this point in offset order.
// The low (gather) pass's pairs start at tile bit T - 4 or above (the
// scattered Hamiltonian tile takes at most four crossing wires), so windows
// wholly below bit T - 5 have none.
template <int R, int M>
constexpr bool gather_pairs_reach() {
    return phase_first_bit<R, M>() + R - 1 >= Tile<R>::T - 4;
}

// kFullTile: pair_lo = 0 and pair_hi = T (the generating low pass), so the
// range tests are resolved at compile time instead of held as uniform masks.
template <int R, int M, bool kFullTile = false>
__device__ __forceinline__ void add_pair_terms(const double2 (&c)[Tile<R>::REGS],
                                               double2 (&l)[Tile<R>::REGS], const double2 *lds,
                                               uint32_t t, const PassGeom &G, uint32_t done,
                                               double field) {
    constexpr int T = Tile<R>::T;
    constexpr int lo = phase_first_bit<R, M>();
    // the two pairs across the window's edges: one bit in registers, the
    // partner from LDS
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int j = side == 0 ? lo - 1 : lo + R - 1;
        if (j < 0 || j + 1 >= T) {
            continue;
        }
        const uint32_t m = 3u << j;
        const bool in_range = kFullTile || (j >= G.pair_lo && j + 1 < G.pair_hi);
        if (in_range && (m & done) == 0) {
#pragma unroll
            for (int k = 0; k < Tile<R>::REGS; ++k) {
                const double2 p = lds[lds_slot<R>(omap<R, M>(t, k) ^ m)];
                l[k].x -= field * p.x;
                l[k].y -= field * p.y;
            }
        }
    }
    // pairs inside the window
#pragma unroll
    for (int rb = 0; rb + 1 < R; ++rb) {
        const int j = lo + rb;
        const bool in_range = kFullTile || (j >= G.pair_lo && j + 1 < G.pair_hi);
        if (in_range && ((3u << j) & done) == 0) {
#pragma unroll
            for (int k = 0; k < Tile<R>::REGS; ++k) {
                const double2 p = c[k ^ (3 << rb)];
                l[k].x -= field * p.x;
                l[k].y -= field * p.y;
            }
        }
    }
}

// Loads one tile of (cur, lam) into the first mapping of the adjoint pass:
// the low pass gathers through the ring CNOT (pre-CNOT frame, mapping 1),
// a high pass reads mapping 0.  With kHam the pair terms of that first
// window are added.
// kGen (high passes): the state is generated in mapping 0 from the chain table
// set gen_tab (its chunk products copied to gen_chunks in LDS) and the tile's
// uniform segments gen_u instead of read.
// kReady, high passes: lam of the tile is already in l (prefetched); low
// (gather) pass: c already holds the tile's state as loaded from the two
// image runs (load_cur_gather), before its placement through LDS.
template <int R, bool kGather, bool kHam, bool kGen = false, bool kReady = false>
__device__ __forceinline__ void bwd_load_tile(double2 (&c)[Tile<R>::REGS],
                                              double2 (&l)[Tile<R>::REGS],
                                              const double2 *cur_in, const double2 *lam_in,
                                              uint64_t base, const TileAddr &ta,
                                              const PassGeom &G, double2 *lds, uint32_t t,
                                              double field,
                                              const CMat2 *__restrict__ gen_tab = nullptr,
                                              const CMat2 *gen_chunks = nullptr,
                                              const CMat2 *__restrict__ gen_u = nullptr,
                                              uint64_t lam_base = 0,
                                              const TileAddr *lam_ta = nullptr) {
    using TL = Tile<R>;
    constexpr int T = TL::T;
    if constexpr (kGather) {
        // Low tile in the pre-CNOT frame: phi[base ^ o] = psi[P(base) ^ P_T(o) ^ ...].
        // Load contiguous q = P_T(o) runs, place at o = P_T^-1(q) in LDS.
        const uint64_t pbase = ring_fwd(base, G.n);
        const uint64_t lo_flip = pbase & 1u;
        const uint64_t run = (pbase & ~uint64_t{TL::SIZE - 1}) + (t ^ lo_flip);
        const uint64_t run1 = run ^ G.hi_mask;
        static_assert(!kGen, "the gather pass reads the state");
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            const uint64_t src = (k < TL::REGS / 2 ? run : run1) + ((uint64_t)k << (T - R));
            l[k] = lam_in[src];
            if constexpr (!kReady) {
                c[k] = cur_in[src];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            lds[lds_slot<R>(ring_inv32(omap<R, 0>(t, k), T))] = l[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            l[k] = lds[lds_slot<R>(omap<R, 1>(t, k))];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            lds[lds_slot<R>(ring_inv32(omap<R, 0>(t, k), T))] = c[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            c[k] = lds[lds_slot<R>(omap<R, 1>(t, k))];
        }
        if constexpr (kHam && gather_pairs_reach<R, 1>()) {
            // the LDS still holds the gathered tile of cur in offset order
            add_pair_terms<R, 1>(c, l, lds, t, G, 0u, field);
        }
    } else {
        // row: the index of register 0 (the state's address, the chain's
        // argument); lambda may sit in another layout (lam_ta, lam_base)
        const uint64_t row = base + ta.offA;
        const TileAddr &la = lam_ta != nullptr ? *lam_ta : ta;
        const uint64_t lrow = (lam_ta != nullptr ? lam_base : base) + la.offA;
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            if constexpr (!kReady) {
                l[k] = load_stream(lam_in + lrow + k * la.strideA);
            }
            if constexpr (!kGen) {
                c[k] = cur_in[row + k * ta.strideA];
            }
        }
        if constexpr (kGen) {
            const CMat2 y = chain_y_high<R>(gen_chunks, gen_u, row, G);
            chain_fill<R>(c, gen_tab, y);
        }
        if constexpr (kHam) {
            // only the pair straddling the load window's lower edge needs
            // a copy of the tile in LDS
            if (G.pair_lo <= T - R - 1 && G.pair_hi >= T - R + 1) {
                __syncthreads();
#pragma unroll
                for (int k = 0; k < TL::REGS; ++k) {
                    lds[lds_slot<R>(omap<R, 0>(t, k))] = c[k];
                }
                __syncthreads();
            }
            add_pair_terms<R, 0>(c, l, lds, t, G, 0u, field);
        }
    }
}

template <int R, bool kGather, bool kWriteCur, bool kWriteLam, bool kHam>
__global__ __launch_bounds__(Tile<R>::THREADS, tile_min_blocks<R>()) void bwd_pass_kernel(
    const double2 *cur_in, const double2 *lam_in, double2 *cur_out, double2 *lam_out,
    const WireCoef *__restrict__ lc, PassGeom G, double *__restrict__ partials, double field) {
    using TL = Tile<R>;
    constexpr int T = TL::T;
    __shared__ double2 lds[TL::SLOTS];
    // gradient sums per row of 16 lanes (DPP row reductions, one lane of each
    // row adds into its own slots)
    constexpr int ROWS = TL::THREADS / 16;
    __shared__ double acc[ROWS][2 * T];
    __shared__ double2 rz_tab[TL::P * TL::REGS];
    __shared__ WireCoef tile_coef[T];
    const uint32_t t = threadIdx.x;
    const int lane = t & (kWave - 1);
    double *row_acc = acc[t / 16];
    for (uint32_t i = t; i < (uint32_t)(ROWS * 2 * T); i += TL::THREADS) {
        acc[i / (2 * T)][i % (2 * T)] = 0.0;
    }
    build_rz_tables<R>(rz_tab, lc, G, t);
    if (t < (uint32_t)T) {
        const bool rot = (int)t >= G.rot_lo && (int)t < G.rot_hi;
        tile_coef[t] = rot ? lc[tile_wire((int)t, G)] : WireCoef{1.0, 0.0, 1.0, 0.0};
    }
    __syncthreads();

    const TileAddr ta = tile_addr<R>(t, G);
    for (uint64_t g = blockIdx.x; g < G.num_tiles; g += gridDim.x) {
        // per-tile opaque thread index: LDS slots are formed where used
        uint32_t tl = t;
        asm volatile("" : "+v"(tl));
        const uint64_t base = tile_base(g, G, T);
        double2 c[TL::REGS];
        double2 l[TL::REGS];
        bwd_load_tile<R, kGather, kHam>(c, l, cur_in, lam_in, base, ta, G, lds, tl, field);

        // lam is exchanged before cur, so the LDS holds cur for the pair terms
        if constexpr (kGather) {
            constexpr uint32_t w1 = window_mask<R, 1>();
            constexpr uint32_t w2 = window_mask<R, 2>();
            phase_bwd<R, 1, kGather, !kGather>(c, l, tile_coef, row_acc, lane, rz_tab);
            exchange<R, 1, 2>(l, lds, tl);
            exchange<R, 1, 2>(c, lds, tl);
            if constexpr (kHam && gather_pairs_reach<R, 2>()) {
                add_pair_terms<R, 2>(c, l, lds, tl, G, w1, field);
            }
            phase_bwd<R, 2, kGather, !kGather>(c, l, tile_coef, row_acc, lane, rz_tab);
            if constexpr (TL::P == 4) {
                exchange<R, 2, 3>(l, lds, tl);
                exchange<R, 2, 3>(c, lds, tl);
                if constexpr (kHam && gather_pairs_reach<R, 3>()) {
                    add_pair_terms<R, 3>(c, l, lds, tl, G, w1 | w2, field);
                }
                phase_bwd<R, 3, kGather, !kGather>(c, l, tile_coef, row_acc, lane, rz_tab);
                exchange<R, 3, 0>(l, lds, tl);
                exchange<R, 3, 0>(c, lds, tl);
                if constexpr (kHam) {
                    add_pair_terms<R, 0>(c, l, lds, tl, G, w1 | w2 | window_mask<R, 3>(), field);
                }
            } else {
                exchange<R, 2, 0>(l, lds, tl);
                exchange<R, 2, 0>(c, lds, tl);
                if constexpr (kHam) {
                    add_pair_terms<R, 0>(c, l, lds, tl, G, w1 | w2, field);
                }
            }
            phase_bwd<R, 0, kGather, !kGather>(c, l, tile_coef, row_acc, lane, rz_tab);
            const uint64_t row = base + ta.offA;
#pragma unroll
            for (int k = 0; k < TL::REGS; ++k) {
                const uint64_t idx = row + k * ta.strideA;
                if constexpr (kWriteCur) {
                    cur_out[idx] = c[k];
                }
                if constexpr (kWriteLam) {
                    lam_out[idx] = l[k];
                }
            }
        } else {
            constexpr uint32_t w0 = window_mask<R, 0>();
            phase_bwd<R, 0, kGather, !kGather>(c, l, tile_coef, row_acc, lane, rz_tab);
            if constexpr (TL::P == 3) {
                exchange<R, 0, 1>(l, lds, tl);
                exchange<R, 0, 1>(c, lds, tl);
                if constexpr (kHam) {
                    add_pair_terms<R, 1>(c, l, lds, tl, G, w0, field);
                }
                phase_bwd<R, 1, kGather, !kGather>(c, l, tile_coef, row_acc, lane, rz_tab);
                exchange<R, 1, 2>(l, lds, tl);
                exchange<R, 1, 2>(c, lds, tl);
                if constexpr (kHam) {
                    add_pair_terms<R, 2>(c, l, lds, tl, G, w0 | window_mask<R, 1>(), field);
                }
                phase_bwd<R, 2, kGather, !kGather>(c, l, tile_coef, row_acc, lane, rz_tab);
            } else {
                uint32_t done = w0;
                if (G.rot_lo < R) {
                    exchange<R, 0, 1>(l, lds, tl);
                    exchange<R, 0, 1>(c, lds, tl);
                    if constexpr (kHam) {
                        add_pair_terms<R, 1>(c, l, lds, tl, G, done, field);
                    }
                    phase_bwd<R, 1, kGather, !kGather>(c, l, tile_coef, row_acc, lane, rz_tab);
                    done |= window_mask<R, 1>();
                    exchange<R, 1, 2>(l, lds, tl);
                    exchange<R, 1, 2>(c, lds, tl);
                } else {
                    exchange<R, 0, 2>(l, lds, tl);
                    exchange<R, 0, 2>(c, lds, tl);
                }
                if constexpr (kHam) {
                    add_pair_terms<R, 2>(c, l, lds, tl, G, done, field);
                }
                phase_bwd<R, 2, kGather, !kGather>(c, l, tile_coef, row_acc, lane, rz_tab);
                done |= window_mask<R, 2>();
                exchange<R, 2, 3>(l, lds, tl);
                exchange<R, 2, 3>(c, lds, tl);
                if constexpr (kHam) {
                    add_pair_terms<R, 3>(c, l, lds, tl, G, done, field);
                }
                phase_bwd<R, 3, kGather, !kGather>(c, l, tile_coef, row_acc, lane, rz_tab);
            }
            const uint64_t row = base + ta.offC;
#pragma unroll
            for (int k = 0; k < TL::REGS; ++k) {
                const uint64_t idx = row + k * ta.strideC;
                if constexpr (kWriteCur) {
                    cur_out[idx] = c[k];
                }
                if constexpr (kWriteLam) {
                    lam_out[idx] = l[k];
                }
            }
        }
    }
    __syncthreads();
    if (t < 2 * T) {
        double total = 0.0;
#pragma unroll
        for (int w = 0; w < ROWS; ++w) {
            total += acc[w][t];
        }
        partials[(uint64_t)blockIdx.x * (2 * T) + t] = total;
    }
}

// Adjoint pass that does not rewrite the state (stored schedule).  After a
// window's RZ^dagger phase both gradients of its wires are fixed:
// dE/dphi = Im<l|Z|c> and dE/dtheta = Re<l|-iY|c> do not change under RY of
// the same wire (RY commutes with Y) or under rotations of other wires.  So a
// first sweep over the windows applies the phases to both states and takes
// the gradients, and cur never gets RY^dagger; a second sweep applies
// RY^dagger to lam alone.  Against the one-sweep pass: no RY on cur (8 of
// 28 FP64 operations per amplitude pair and wire), two more exchanges of lam.
// kPhased = false: the window's RZ phases are already divided out of both
// arrays (a state generated without them and lambda multiplied by the tile's
// combined phase, apply_tile_phase), so the window takes its gradients only.
// kAllBits: the window's gradients of every bit, without the uniform
// rot_lo / rot_hi tests: the generating low pass rotates every tile bit, and
// the generated start's high passes leave only gradient slots that
// reduce_grad never reads (it takes the columns of [rot_lo, rot_hi)).
template <int R, int M, bool kPhased = true, bool kAllBits = false>
__device__ __forceinline__ void grad_window(double2 (&c)[Tile<R>::REGS],
                                            double2 (&l)[Tile<R>::REGS],
                                            const WireCoef *__restrict__ lc, const PassGeom &G,
                                            double *__restrict__ row_acc, int lane,
                                            const double2 *rz_tab) {
    constexpr int P0 = phase_first_bit<R, M>();
    // no early exit for a window without rotated bits: its phases are 1 and
    // its gradients are not read, and branch-free code keeps the register
    // arrays free of copies at control-flow joins
    if constexpr (R == 4) {
        // dps[b] = sum_q (-1)^{q_b} im_q from the partial sums of each quad
        // of registers (q_3 q_2 fixed): 40 adds against 64
        double dps[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            double im[4];
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                const int q = 4 * x + y;
                if constexpr (kPhased) {
                    const double2 z = rz_tab[M * Tile<R>::REGS + q];
                    const double2 zc = make_double2(z.x, -z.y);
                    c[q] = cmul(c[q], zc);
                    l[q] = cmul(l[q], zc);
                }
                im[y] = l[q].x * c[q].y - l[q].y * c[q].x;
            }
            const double a0 = im[0] + im[1];
            const double a1 = im[2] + im[3];
            const double d = (im[0] - im[1]) + (im[2] - im[3]);
            const double s = a0 + a1;
            const double e = a0 - a1;
            if (x == 0) {
                dps[0] = d;
                dps[1] = e;
                dps[2] = s;
                dps[3] = s;
            } else {
                dps[0] += d;
                dps[1] += e;
                dps[2] += (x & 1) ? -s : s;
                dps[3] += (x & 2) ? -s : s;
            }
        }
        // dt per bit, each (dt, dp) pair folded into the first step of the
        // transposed row sums at once
        const bool b0 = (lane & 1) != 0;
        double u[4];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            const int j = P0 + rb;
            double dt = 0.0;
            if (kAllBits || (j >= G.rot_lo && j < G.rot_hi)) {
#pragma unroll
                for (int q = 0; q < Tile<R>::REGS; ++q) {
                    if (!(q & (1 << rb))) {
                        const double2 c0 = c[q];
                        const double2 c1 = c[q | (1 << rb)];
                        const double2 l0 = l[q];
                        const double2 l1 = l[q | (1 << rb)];
                        dt += (l1.x * c0.x + l1.y * c0.y) - (l0.x * c1.x + l0.y * c1.y);
                    }
                }
            }
            u[rb] = row_sums8_first(dt, dps[rb], b0);
        }
        // lane s (s < 8) of each row holds slot 2 * P0 + s: bit P0 + s / 2,
        // dt for s even, dp for s odd.  The slots of unrotated bits are
        // accumulated too and never read (reduce_grad takes the columns of
        // [rot_lo, rot_hi) only).
        const double r = row_sums8_rest(u, lane);
        if ((lane & 15) < 8) {
            row_acc[2 * P0 + (lane & 7)] += r;
        }
        return;
    }
    double dps[R];
#pragma unroll
    for (int rb = 0; rb < R; ++rb) {
        dps[rb] = 0.0;
    }
#pragma unroll
    for (int q = 0; q < Tile<R>::REGS; ++q) {
        if constexpr (kPhased) {
            const double2 z = rz_tab[M * Tile<R>::REGS + q];
            const double2 zc = make_double2(z.x, -z.y);
            c[q] = cmul(c[q], zc);
            l[q] = cmul(l[q], zc);
        }
        const double im = l[q].x * c[q].y - l[q].y * c[q].x;
#pragma unroll
        for (int rb = 0; rb < R; ++rb) {
            dps[rb] += ((q >> rb) & 1) ? -im : im;
        }
    }
#pragma unroll
    for (int rb = R - 1; rb >= 0; --rb) {
        const int j = P0 + rb;
        if (kAllBits || (j >= G.rot_lo && j < G.rot_hi)) {
            double dt = 0.0;
#pragma unroll
            for (int q = 0; q < Tile<R>::REGS; ++q) {
                if (!(q & (1 << rb))) {
                    const double2 c0 = c[q];
                    const double2 c1 = c[q | (1 << rb)];
                    const double2 l0 = l[q];
                    const double2 l1 = l[q | (1 << rb)];
                    dt += (l1.x * c0.x + l1.y * c0.y) - (l0.x * c1.x + l0.y * c1.y);
                }
            }
            dt = row_sum(dt);
            const double dp = row_sum(dps[rb]);
            if ((lane & 15) == 0) {
                row_acc[2 * j] += dt;
                row_acc[2 * j + 1] += dp;
            }
        }
    }
}

// tile_coef: the coefficients of the tile's wires by tile bit, in LDS (tile
// bits outside [rot_lo, rot_hi) hold the identity: no branch, so the register
// array needs no copies at control-flow joins)
template <int R, int M>
__device__ __forceinline__ void ry_inv_window(double2 (&l)[Tile<R>::REGS],
                                              const WireCoef *tile_coef) {
    constexpr int P0 = phase_first_bit<R, M>();
#pragma unroll
    for (int rb = 0; rb < R; ++rb) {
        const WireCoef k = tile_coef[P0 + rb];
#pragma unroll
        for (int q = 0; q < Tile<R>::REGS; ++q) {
            if (!(q & (1 << rb))) {
                ry_inv_pair(l[q], l[q | (1 << rb)], k);
            }
        }
    }
}

// RY^dagger of a window in tangent form (ry_inv_pair_tan_a / _b): tan_t[j]
// the tangent of tile bit j (0 for an unrotated bit: the identity, factor 1),
// bit j of form_b set where that wire takes form B.  form_b is uniform, so
// each wire's choice is a scalar branch.
template <int R, int M>
__device__ __forceinline__ void ry_inv_window_tan(double2 (&l)[Tile<R>::REGS],
                                                  const double *tan_t, uint32_t form_b) {
    constexpr int P0 = phase_first_bit<R, M>();
#pragma unroll
    for (int rb = 0; rb < R; ++rb) {
        const double tt = tan_t[P0 + rb];
        if ((form_b >> (P0 + rb)) & 1u) {
#pragma unroll
            for (int q = 0; q < Tile<R>::REGS; ++q) {
                if (!(q & (1 << rb))) {
                    ry_inv_pair_tan_b(l[q], l[q | (1 << rb)], tt);
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < Tile<R>::REGS; ++q) {
                if (!(q & (1 << rb))) {
                    ry_inv_pair_tan_a(l[q], l[q | (1 << rb)], tt);
                }
            }
        }
    }
}

// The tangent form's tables for the tile bits [lo, T) of a pass (tile_coef
// by tile bit, the identity outside the rotated bits): thread j < T writes
// tan_t[j] for bits in [lo, T) (bits below lo keep the full form); every
// thread returns the uniform form mask, thread 0 writes the scale sigma (the
// product of the factors taken out, |sigma| >= 2^(-(T - lo) / 2)) to *sigma.
// tile_coef is read once the caller's barrier has published it; the caller
// places a barrier after this call as well.
template <int T>
__device__ __forceinline__ uint32_t ry_tan_tables(const WireCoef *tile_coef, int lo,
                                                  double *tan_t, double *sigma, uint32_t t) {
    if (t < (uint32_t)T && (int)t >= lo) {
        const WireCoef k = tile_coef[t];
        tan_t[t] = ry_tan_form_b(k) ? k.c / k.s : k.s / k.c;
    }
    uint32_t form_b = 0;
    double sg = 1.0;
    for (int j = lo; j < T; ++j) {
        const WireCoef k = tile_coef[j];
        const bool fb = ry_tan_form_b(k);
        form_b |= (fb ? 1u : 0u) << j;
        sg *= fb ? k.s : k.c;
    }
    if (t == 0) {
        *sigma = sg;
    }
    return __builtin_amdgcn_readfirstlane(form_b);
}

// lambda times the tangent form's scale
template <int R>
__device__ __forceinline__ void scale_regs(double2 (&l)[Tile<R>::REGS], double sg) {
#pragma unroll
    for (int k = 0; k < Tile<R>::REGS; ++k) {
        l[k].x *= sg;
        l[k].y *= sg;
    }
}

// High passes: window 1 of the 256-thread shape holds tile bits 0..3, of which
// only bit 3 can be a rotated wire (bits 0..2 are the contiguous low index
// bits that ride along).  In mappings 0 and 2 tile bit 3 is lane bit 3, so
// its pair partner is the lane 8 apart in the same row of 16: DPP row_ror:8
// brings it over without an LDS exchange.
__device__ __forceinline__ double2 lane_xor8(double2 v) {
    return make_double2(dpp_move<0x128>(v.x), dpp_move<0x128>(v.y));
}

// Tile bit 3: RZ^dagger on both states (window 1's phase table at register
// index 8 or 0; none with kPhased = false), its gradients as in grad_window,
// then RY^dagger of bit 3 on lam (before window 2's: they commute).  One lane
// move of lam serves both: a pair's lanes hold (c0, l0) and (c1, l1), and
// Re(l1* c0) - Re(l0* c1) is summed as Re(p* c) on the lane with bit 3 clear
// and -Re(p* c) on the other, p the partner's lam, which is also the partner
// value RY^dagger takes.
template <int R, bool kPhased = true>
__device__ __forceinline__ void grad_ry_lane_bit3(double2 (&c)[Tile<R>::REGS],
                                                  double2 (&l)[Tile<R>::REGS],
                                                  double *__restrict__ row_acc, int lane,
                                                  const double2 *rz_tab, const WireCoef &k) {
    const bool hi = ((lane >> 3) & 1) != 0;
    const double2 z = rz_tab[Tile<R>::REGS + (hi ? 8 : 0)];
    const double2 zc = make_double2(z.x, -z.y);
    const double s = hi ? -k.s : k.s;
    double dt = 0.0;
    double dp = 0.0;
#pragma unroll
    for (int q = 0; q < Tile<R>::REGS; ++q) {
        if constexpr (kPhased) {
            c[q] = cmul(c[q], zc);
            l[q] = cmul(l[q], zc);
        }
        const double im = l[q].x * c[q].y - l[q].y * c[q].x;
        dp += hi ? -im : im;
        const double2 p = lane_xor8(l[q]);
        const double re = p.x * c[q].x + p.y * c[q].y;
        dt += hi ? -re : re;
        l[q].x = k.c * l[q].x + s * p.x;
        l[q].y = k.c * l[q].y + s * p.y;
    }
    const double r = row_sums2(dt, dp, lane);
    if ((lane & 15) < 2) {
        row_acc[2 * 3 + (lane & 1)] += r;
    }
}

// The load mapping's addresses in lambda layout L (a bit permutation: the
// offset and the register step map bit by bit; the host checks that the
// register bits stay contiguous).
__device__ __forceinline__ TileAddr lay_tile_addr(const TileAddr &ta, const LamLayout &L) {
    TileAddr r = ta;
    r.offA = lay_addr(ta.offA, L);
    r.strideA = lay_addr(ta.strideA, L);
    return r;
}

// Stores lam from mapping M: the load mapping of a high pass, any mapping of
// the low tile (contiguous: offset = index - base).
template <int R, int M, bool kGather>
__device__ __forceinline__ void store_lam(const double2 (&l)[Tile<R>::REGS], double2 *lam_out,
                                          uint64_t base, const TileAddr &ta, uint32_t t) {
    if constexpr (kGather) {
#pragma unroll
        for (int k = 0; k < Tile<R>::REGS; ++k) {
            lam_out[base + omap<R, M>(t, k)] = l[k];
        }
    } else {
        static_assert(M == 0, "high passes store from the load mapping");
        // 64-bit offsets: at 33 and 34 qubits the register step of the top
        // high group is 2^29 or more, and 15 steps leave 32 bits
        double2 *dst = lam_out + base + ta.offA;
#pragma unroll
        for (int k = 0; k < Tile<R>::REGS; ++k) {
            store_stream(dst + (uint64_t)k * ta.strideA, l[k]);
        }
    }
}

// Loads the state of a low (gather) pass's tile from its two image runs under
// the ring CNOT, in the order bwd_load_tile places them through LDS.
template <int R>
__device__ __forceinline__ void load_cur_gather(double2 (&c)[Tile<R>::REGS],
                                                const double2 *cur_in, uint64_t base,
                                                const PassGeom &G, uint32_t t) {
    using TL = Tile<R>;
    const uint64_t pbase = ring_fwd(base, G.n);
    const uint64_t run = (pbase & ~uint64_t{TL::SIZE - 1}) + (t ^ (pbase & 1u));
    const uint64_t run1 = run ^ G.hi_mask;
#pragma unroll
    for (int k = 0; k < TL::REGS; ++k) {
        c[k] = cur_in[(k < TL::REGS / 2 ? run : run1) + ((uint64_t)k << (TL::T - R))];
    }
}

// Loads lam of a high pass's tile in the load mapping (mapping 0).
template <int R>
__device__ __forceinline__ void load_lam_high(double2 (&l)[Tile<R>::REGS],
                                              const double2 *lam_in, uint64_t base,
                                              const TileAddr &ta) {
    const uint64_t row = base + ta.offA;
#pragma unroll
    for (int k = 0; k < Tile<R>::REGS; ++k) {
        l[k] = load_stream(lam_in + row + k * ta.strideA);
    }
}

// lambda of a high pass's tile (mapping 0, R = 4) times the conjugate of the
// combined RZ phase of every rotated tile bit: window 0's (register bits
// 8..11), window 1's (thread bits 0..3) and window 2's (thread bits 4..7)
// factors from rz_tab.  With a state generated without those phases, the
// windows then take their gradients with no phase work (kPhased = false).
template <int R>
__device__ __forceinline__ void apply_tile_phase(double2 (&l)[Tile<R>::REGS], const double2 *rz_tab,
                                                 uint32_t t) {
    static_assert(R == 4, "three register windows: bits 8..11, 0..3, 4..7");
    const double2 pt = cmul(rz_tab[Tile<R>::REGS + (t & 15u)],
                            rz_tab[2 * Tile<R>::REGS + ((t >> 4) & 15u)]);
#pragma unroll
    for (int k = 0; k < Tile<R>::REGS; ++k) {
        const double2 z = cmul(rz_tab[k], pt);
        l[k] = cmul(l[k], make_double2(z.x, -z.y));
    }
}

// kPrefetch: once the tile's state is no longer needed (after its
// gradients), the registers it held receive the workgroup's next tile --
// lam for a high pass, the state for the low (gather) pass -- so that load is
// in flight during the lam-only sweep, the store and the next tile's start
// instead of stalling it (a register double buffer: the state and the
// prefetched array share registers).
// kPhaseFree (generated high passes, R = 4): the table set generates the
// state without the RZ of the pass's own wires, so lambda takes the tile's
// combined phase once (apply_tile_phase) and the windows none.
template <int R, bool kGather, bool kHam, bool kGen, bool kPrefetch = false, bool kPhaseFree = false>
__global__ __launch_bounds__(Tile<R>::THREADS, tile_min_blocks<R>()) void bwd_lam_kernel(
    const double2 *cur_in, const double2 *lam_in, double2 *lam_out,
    const WireCoef *__restrict__ lc, PassGeom G, double *__restrict__ partials, double field,
    const CMat2 *__restrict__ gen_tab, const CMat2 *__restrict__ gen_utab) {
    using TL = Tile<R>;
    constexpr int T = TL::T;
    __shared__ double2 lds[TL::SLOTS];
    // gradient sums per row of 16 lanes: the per-tile reductions stop at the
    // row (DPP only) and one lane of each row adds into its own slots
    constexpr int ROWS = TL::THREADS / 16;
    __shared__ double acc[ROWS][2 * T];
    __shared__ double2 rz_tab[TL::P * TL::REGS];
    __shared__ WireCoef tile_coef[T];
    // phase-free sweep: RY^dagger of lambda's window bits (4..11) in tangent
    // form, its scale applied once before the store
    __shared__ double tan_t[kPhaseFree ? T : 1];
    __shared__ double ry_sigma;
    // chunk products of the generating table set: read at every tile start,
    // where LDS latency is short next to L2 behind the streaming traffic
    __shared__ CMat2 gen_chunks[kGen ? kChainTabEntries - kChainC0 : 1];
    const uint32_t t = threadIdx.x;
    const int lane = t & (kWave - 1);
    double *row_acc = acc[t / 16];
    for (uint32_t i = t; i < (uint32_t)(ROWS * 2 * T); i += TL::THREADS) {
        acc[i / (2 * T)][i % (2 * T)] = 0.0;
    }
    build_rz_tables<R>(rz_tab, lc, G, t);
    if (t < (uint32_t)T) {
        const bool rot = (int)t >= G.rot_lo && (int)t < G.rot_hi;
        tile_coef[t] = rot ? lc[tile_wire((int)t, G)] : WireCoef{1.0, 0.0, 1.0, 0.0};
    }
    if constexpr (kGen) {
        for (int i = (int)t; i < kChainTabEntries - kChainC0; i += TL::THREADS) {
            gen_chunks[i] = gen_tab[kChainC0 + i];
        }
    }
    __syncthreads();
    uint32_t form_b = 0;
    if constexpr (kPhaseFree) {
        form_b = ry_tan_tables<T>(tile_coef, 4, tan_t, &ry_sigma, t);
        __syncthreads();
    }

    const TileAddr ta = tile_addr<R>(t, G);
    // lambda's layouts (high passes; the gather pass reads and writes the
    // natural order)
    const TileAddr ta_in = lay_tile_addr(ta, kGather ? LamLayout{} : G.lay_in);
    const TileAddr ta_out = lay_tile_addr(ta, kGather ? LamLayout{} : G.lay_out);
    static_assert(!(kPrefetch && kGather && kGen), "the gather pass reads the state");
    static_assert(!kPhaseFree || (kGen && !kGather && !kHam && TL::P == 3),
                  "phase-free sweep: generated high passes of register shape 4 without pairs");
    constexpr bool kPhased = !kPhaseFree;

    double2 x_next[kPrefetch ? TL::REGS : 1];
    auto prefetch_tile = [&](uint64_t gn, const TileAddr &a, uint32_t tt) {
        if constexpr (kPrefetch) {
            if constexpr (kGather) {
                load_cur_gather<R>(x_next, cur_in, tile_base(gn, G, T), G, tt);
            } else {
                load_lam_high<R>(x_next, lam_in,
                                 kPhaseFree ? G.tile_in[gn] : lay_addr(tile_base(gn, G, T), G.lay_in),
                                 a);
            }
        }
    };
    if constexpr (kPrefetch) {
        if ((uint64_t)blockIdx.x < G.num_tiles) {
            prefetch_tile(blockIdx.x, ta_in, t);
        }
    }
    for (uint64_t g = blockIdx.x; g < G.num_tiles; g += gridDim.x) {
        // the thread index is made opaque per tile so the LDS addresses of the
        // exchanges are formed where they are used, not held in
        // registers across the whole loop
        uint32_t tl = t;
        asm volatile("" : "+v"(tl));
        // with the prefetch the per-register addresses of the tile's loads and
        // stores are formed per tile as well (hoisted out of the loop they
        // would be spilled)
        TileAddr tla = ta;
        TileAddr tla_in = ta_in;
        if constexpr (kPrefetch) {
            asm volatile("" : "+v"(tla.offA));
            asm volatile("" : "+v"(tla_in.offA));
        }
        const uint64_t base = tile_base(g, G, T);
        double2 c[TL::REGS];
        double2 l[TL::REGS];
        if constexpr (kPrefetch) {
#pragma unroll
            for (int k = 0; k < TL::REGS; ++k) {
                if constexpr (kGather) {
                    c[k] = x_next[k];
                } else {
                    l[k] = x_next[k];
                }
            }
        }
        bwd_load_tile<R, kGather, kHam, kGen, kPrefetch>(c, l, cur_in, lam_in, base, tla, G, lds,
                                                         tl, field, gen_tab, gen_chunks,
                                                         gen_utab + 2 * g,
                                                         kPhaseFree ? G.tile_in[g]
                                                                    : lay_addr(base, G.lay_in),
                                                         &tla_in);
        // the next tile, into the registers the state frees (called right
        // after the state's last use; scheduling barriers keep the loads from
        // being hoisted above it, where both would need registers)
        auto prefetch_next = [&]() {
            if constexpr (kPrefetch) {
                __builtin_amdgcn_sched_barrier(0);
                const uint64_t gn = g + gridDim.x;
                if (gn < G.num_tiles) {
                    prefetch_tile(gn, tla_in, tl);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        };
        // sweep 1: phases and gradients (lam exchanged before cur, so the LDS
        // holds cur for the pair terms); sweep 2: RY^dagger on lam
        if constexpr (kGather) {
            constexpr uint32_t w1 = window_mask<R, 1>();
            constexpr uint32_t w2 = window_mask<R, 2>();
            grad_window<R, 1>(c, l, lc, G, row_acc, lane, rz_tab);
            if constexpr (TL::P == 4) {
                exchange<R, 1, 2>(l, lds, tl);
                exchange<R, 1, 2>(c, lds, tl);
                if constexpr (kHam && gather_pairs_reach<R, 2>()) {
                    add_pair_terms<R, 2>(c, l, lds, tl, G, w1, field);
                }
                grad_window<R, 2>(c, l, lc, G, row_acc, lane, rz_tab);
                exchange<R, 2, 3>(l, lds, tl);
                exchange<R, 2, 3>(c, lds, tl);
                if constexpr (kHam && gather_pairs_reach<R, 3>()) {
                    add_pair_terms<R, 3>(c, l, lds, tl, G, w1 | w2, field);
                }
                grad_window<R, 3>(c, l, lc, G, row_acc, lane, rz_tab);
                exchange<R, 3, 0>(l, lds, tl);
                exchange<R, 3, 0>(c, lds, tl);
                if constexpr (kHam) {
                    add_pair_terms<R, 0>(c, l, lds, tl, G, w1 | w2 | window_mask<R, 3>(), field);
                }
                grad_window<R, 0>(c, l, lc, G, row_acc, lane, rz_tab);
                prefetch_next();
                ry_inv_window<R, 0>(l, tile_coef);
                exchange<R, 0, 1>(l, lds, tl);
                ry_inv_window<R, 1>(l, tile_coef);
                exchange<R, 1, 2>(l, lds, tl);
                ry_inv_window<R, 2>(l, tile_coef);
                exchange<R, 2, 3>(l, lds, tl);
                ry_inv_window<R, 3>(l, tile_coef);
                store_lam<R, 3, true>(l, lam_out, base, tla, tl);
            } else {
                // windows 1, 0, 2, then lam back through 1 to the contiguous
                // mapping 0 for the store
                exchange<R, 1, 0>(l, lds, tl);
                exchange<R, 1, 0>(c, lds, tl);
                if constexpr (kHam) {
                    add_pair_terms<R, 0>(c, l, lds, tl, G, w1, field);
                }
                grad_window<R, 0>(c, l, lc, G, row_acc, lane, rz_tab);
                exchange<R, 0, 2>(l, lds, tl);
                exchange<R, 0, 2>(c, lds, tl);
                if constexpr (kHam && gather_pairs_reach<R, 2>()) {
                    add_pair_terms<R, 2>(c, l, lds, tl, G, w1 | window_mask<R, 0>(), field);
                }
                grad_window<R, 2>(c, l, lc, G, row_acc, lane, rz_tab);
                prefetch_next();
                ry_inv_window<R, 2>(l, tile_coef);
                exchange<R, 2, 1>(l, lds, tl);
                ry_inv_window<R, 1>(l, tile_coef);
                exchange<R, 1, 0>(l, lds, tl);
                ry_inv_window<R, 0>(l, tile_coef);
                store_lam<R, 0, true>(l, lam_out, base, tla, tl);
            }
        } else {
            constexpr uint32_t w0 = window_mask<R, 0>();
            if constexpr (kPhaseFree) {
                apply_tile_phase<R>(l, rz_tab, tl);
            }
            grad_window<R, 0, kPhased, kPhaseFree>(c, l, lc, G, row_acc, lane, rz_tab);
            if constexpr (TL::P == 3) {
                // windows 0 and 2, tile bit 3 across lanes (window 1 has no
                // other rotated bit: its pairs with bit 4 were added at window
                // 2), then lam back to mapping 0 for the store: three LDS
                // exchanges
                exchange<R, 0, 2>(l, lds, tl);
                exchange<R, 0, 2>(c, lds, tl);
                if constexpr (kHam) {
                    add_pair_terms<R, 2>(c, l, lds, tl, G, w0, field);
                }
                grad_window<R, 2, kPhased, kPhaseFree>(c, l, lc, G, row_acc, lane, rz_tab);
                // unconditional: an unrotated bit 3 has phase 1 and the
                // identity in tile_coef, and its gradient slot is not read
                grad_ry_lane_bit3<R, kPhased>(c, l, row_acc, lane, rz_tab, tile_coef[3]);
                prefetch_next();
                if constexpr (kPhaseFree) {
                    ry_inv_window_tan<R, 2>(l, tan_t, form_b);
                } else {
                    ry_inv_window<R, 2>(l, tile_coef);
                }
                exchange<R, 2, 0>(l, lds, tl);
                if constexpr (kPhaseFree) {
                    ry_inv_window_tan<R, 0>(l, tan_t, form_b);
                    scale_regs<R>(l, ry_sigma);
                } else {
                    ry_inv_window<R, 0>(l, tile_coef);
                }
                store_lam<R, 0, false>(l, lam_out,
                                      kPhaseFree ? G.tile_out[g] : lay_addr(base, G.lay_out),
                                      ta_out, tl);
            } else {
                uint32_t done = w0;
                const bool w1_rot = G.rot_lo < R;
                if (w1_rot) {
                    exchange<R, 0, 1>(l, lds, tl);
                    exchange<R, 0, 1>(c, lds, tl);
                    if constexpr (kHam) {
                        add_pair_terms<R, 1>(c, l, lds, tl, G, done, field);
                    }
                    grad_window<R, 1>(c, l, lc, G, row_acc, lane, rz_tab);
                    done |= window_mask<R, 1>();
                    exchange<R, 1, 2>(l, lds, tl);
                    exchange<R, 1, 2>(c, lds, tl);
                } else {
                    exchange<R, 0, 2>(l, lds, tl);
                    exchange<R, 0, 2>(c, lds, tl);
                }
                if constexpr (kHam) {
                    add_pair_terms<R, 2>(c, l, lds, tl, G, done, field);
                }
                grad_window<R, 2>(c, l, lc, G, row_acc, lane, rz_tab);
                done |= window_mask<R, 2>();
                exchange<R, 2, 3>(l, lds, tl);
                exchange<R, 2, 3>(c, lds, tl);
                if constexpr (kHam) {
                    add_pair_terms<R, 3>(c, l, lds, tl, G, done, field);
                }
                grad_window<R, 3>(c, l, lc, G, row_acc, lane, rz_tab);
                prefetch_next();
                ry_inv_window<R, 3>(l, tile_coef);
                exchange<R, 3, 2>(l, lds, tl);
                ry_inv_window<R, 2>(l, tile_coef);
                if (w1_rot) {
                    exchange<R, 2, 1>(l, lds, tl);
                    ry_inv_window<R, 1>(l, tile_coef);
                    exchange<R, 1, 0>(l, lds, tl);
                } else {
                    exchange<R, 2, 0>(l, lds, tl);
                }
                ry_inv_window<R, 0>(l, tile_coef);
                store_lam<R, 0, false>(l, lam_out,
                                      kPhaseFree ? G.tile_out[g] : lay_addr(base, G.lay_out),
                                      ta_out, tl);
            }
        }
    }
    __syncthreads();
    if (t < 2 * T) {
        double total = 0.0;
#pragma unroll
        for (int w = 0; w < ROWS; ++w) {
            total += acc[w][t];
        }
        partials[(uint64_t)blockIdx.x * (2 * T) + t] = total;
    }
}

// Generating low pass of a two-layer circuit (see launch_bwd_low_gen).  The
// tile is generated in mapping 0 and swept 0, 1, .., P - 1 (phases, pair
// terms and gradients), then lambda alone goes back to mapping 0 under RY^dagger
// and is stored there (contiguous rows).
// Waves per SIMD to budget registers for (the second launch-bounds value):
// R = 3 aims at two 512-thread workgroups per CU, four waves per SIMD with
// 128 VGPRs; the chunk tables stay in global memory so that two workgroups'
// LDS fits.
template <int R>
constexpr int low_gen_min_waves() {
    return R == 3 ? 4 : 2;
}

template <int R>
__global__ __launch_bounds__(Tile<R>::THREADS, low_gen_min_waves<R>()) void bwd_low_gen_kernel(
    double2 *lam_out, const WireCoef *__restrict__ lc, PassGeom G, double *__restrict__ partials,
    double field, const CMat2 *__restrict__ tab, const CMat2 *__restrict__ utab,
    const CMat2 *__restrict__ xtab) {
    using TL = Tile<R>;
    constexpr int T = TL::T;
    constexpr int kC1Bits = T - R - 4;
    __shared__ double2 lds[TL::SLOTS];
    constexpr int ROWS = TL::THREADS / 16;
    __shared__ double acc[ROWS][2 * T];
    __shared__ double2 rz_tab[TL::P * TL::REGS];
    __shared__ WireCoef tile_coef[T];
    // RY^dagger of lambda in tangent form for the 4-register shape
    constexpr bool kTan = R == 4;
    __shared__ double tan_t[kTan ? T : 1];
    __shared__ double ry_sigma;
    // lambda's store addresses in layout G.lay_out for the 4-register shape:
    // each thread's row offset and the register step (tile independent: read
    // per tile instead of keeping the layout live across the loop; the
    // 512-thread shape keeps its LDS within two workgroups per CU instead)
    constexpr bool kLayRows = R == 4;
    __shared__ uint32_t lay_row[kLayRows ? TL::THREADS : 1];
    __shared__ uint64_t lay_step;
    // chunk products C0 (index bits 0..3) and C1 (bits 4 .. T - R - 1): in LDS
    // for R = 4, read from the table in global memory (L2) for R = 3
    constexpr bool kChunksLds = R == 4;
    __shared__ CMat2 chunk_lds[kChunksLds ? kChainTabEntries - kChainC0 : 1];
    const CMat2 *chunks = kChunksLds ? chunk_lds : tab + kChainC0;
    const uint32_t t = threadIdx.x;
    const int lane = t & (kWave - 1);
    double *row_acc = acc[t / 16];
    for (uint32_t i = t; i < (uint32_t)(ROWS * 2 * T); i += TL::THREADS) {
        acc[i / (2 * T)][i % (2 * T)] = 0.0;
    }
    build_rz_tables<R>(rz_tab, lc, G, t);
    if (t < (uint32_t)T) {
        tile_coef[t] = lc[t];  // every bit of the low tile is a rotated wire
    }
    if constexpr (kLayRows) {
        // (offsets within the low tile map to address bits below n <= 32)
        lay_row[t] = (uint32_t)lay_addr(omap<R, 0>(t, 0), G.lay_out);
        if (t == 0) {
            lay_step = lay_addr(uint64_t{1} << (T - R), G.lay_out);
        }
    }
    if constexpr (kChunksLds) {
        for (int i = (int)t; i < kChainTabEntries - kChainC0; i += TL::THREADS) {
            chunk_lds[i] = tab[kChainC0 + i];
        }
    }
    __syncthreads();
    uint32_t form_b = 0;
    if constexpr (kTan) {
        form_b = ry_tan_tables<T>(tile_coef, 0, tan_t, &ry_sigma, t);
        __syncthreads();
    }
    for (uint64_t g = blockIdx.x; g < G.num_tiles; g += gridDim.x) {
        uint32_t tl = t;
        asm volatile("" : "+v"(tl));
        const uint64_t base = tile_base(g, G, T);
        double2 c[TL::REGS];
        double2 l[TL::REGS];
        // phi of the tile in mapping 0, c[k] = Tr(run[k] U_A C0 C1), and
        // lambda's diagonal and out-of-tile X terms.  The thread's row holds
        // tile bits 0 .. T - R - 1 (register bits zero).  Wire T - 1's flip
        // moves the register index's top bit: its amplitude at register
        // k ^ 2^(R-1) is Tr(run[k] Y_0).  The other flips leaving the tile keep
        // the register index and share one set of traces, Tr(run[k] Y_sum):
        // wire n - 1's (thread index ^ 3) and the pairs above the tile, whose
        // segments are summed per tile.  Each run matrix serves all three traces.
        const uint32_t row = omap<R, 0>(tl, 0);
        {
            auto y_of = [&](const CMat2 &u, uint32_t r2) {
                return cmat_mul(cmat_mul(u, chunks[r2 & 15u]),
                                chunks[kChainC1 - kChainC0 + ((r2 >> 4) & ((1u << kC1Bits) - 1u))]);
            };
            const CMat2 *xg = xtab + g * kFlipSegments;
            const CMat2 y = y_of(utab[2 * g], row);
            const CMat2 y0 = y_of(xg[0], row);
            const CMat2 ysum = cmat_add(y_of(xg[1], row ^ 3u), y_of(xg[2], row));
#pragma unroll
            for (int k = 0; k < TL::REGS; ++k) {
                l[k] = make_double2(0.0, 0.0);
            }
            // H's diagonal -sum_w Z_w Z_{w+1} at P(i) from the pre-CNOT index
            // i: bits k, k + 1 of P(i) differ where i_{k+1} = 1
            // (1 <= k <= n - 2), bits n - 1 and 0 where i_0 = 1, bits 0 and 1
            // where i_0 xor parity(i >> 2) = 1, so d = 2 walls - n with
            // walls = popcount(i >> 2) + (i_0 ^ parity(i >> 2)) + i_0.  The
            // register bits T - R .. T - 1 (all >= 2) of register k add
            // pk = popcount(k) to popcount(i >> 2) and flip that parity pk
            // times: d = dpar[pk & 1] + 2 pk with two values per thread
            const uint64_t i0 = base | row;
            const int hi0 = (int)__popcll(i0 >> 2);
            const int b0 = (int)(i0 & 1u);
            const int e0 = b0 ^ (hi0 & 1);
            const double dpar0 = (double)(2 * (hi0 + b0 + e0) - G.n);
            const double dpar1 = (double)(2 * (hi0 + b0 + (e0 ^ 1)) - G.n);
#pragma unroll
            for (int k = 0; k < TL::REGS; ++k) {
                const CMat2 rk = tab[k];
                c[k] = cmat_trace_mul(rk, y);
                const int pk = __builtin_popcount((unsigned)k);
                const double d = ((pk & 1) ? dpar1 : dpar0) + (double)(2 * pk);
                const double2 v1 = cmat_trace_mul(rk, ysum);
                const double2 v0 = cmat_trace_mul(rk, y0);
                l[k].x += d * c[k].x - field * v1.x;
                l[k].y += d * c[k].y - field * v1.y;
                l[k ^ (1 << (R - 1))].x -= field * v0.x;
                l[k ^ (1 << (R - 1))].y -= field * v0.y;
            }
        }
        // pair terms of window 0 (its lower edge from an LDS copy), sweep 1
        if constexpr (T - R - 1 >= 0) {  // every pair in the tile (pair_lo = 0, pair_hi = T)
            __syncthreads();
#pragma unroll
            for (int k = 0; k < TL::REGS; ++k) {
                lds[lds_slot<R>(omap<R, 0>(tl, k))] = c[k];
            }
            __syncthreads();
        }
        add_pair_terms<R, 0, true>(c, l, lds, tl, G, 0u, field);
        grad_window<R, 0, true, true>(c, l, lc, G, row_acc, lane, rz_tab);
        constexpr uint32_t w0 = window_mask<R, 0>();
        constexpr uint32_t w1 = window_mask<R, 1>();
        exchange<R, 0, 1>(l, lds, tl);
        exchange<R, 0, 1>(c, lds, tl);
        add_pair_terms<R, 1, true>(c, l, lds, tl, G, w0, field);
        grad_window<R, 1, true, true>(c, l, lc, G, row_acc, lane, rz_tab);
        exchange<R, 1, 2>(l, lds, tl);
        exchange<R, 1, 2>(c, lds, tl);
        add_pair_terms<R, 2, true>(c, l, lds, tl, G, w0 | w1, field);
        grad_window<R, 2, true, true>(c, l, lc, G, row_acc, lane, rz_tab);
        if constexpr (TL::P == 3) {
            // sweep 2: lambda back to mapping 0 under RY^dagger (tangent
            // form, its scale applied once before the store)
            ry_inv_window_tan<R, 2>(l, tan_t, form_b);
            exchange<R, 2, 1>(l, lds, tl);
            ry_inv_window_tan<R, 1>(l, tan_t, form_b);
            exchange<R, 1, 0>(l, lds, tl);
            ry_inv_window_tan<R, 0>(l, tan_t, form_b);
            scale_regs<R>(l, ry_sigma);
        } else {
            constexpr uint32_t w2 = window_mask<R, 2>();
            exchange<R, 2, 3>(l, lds, tl);
            exchange<R, 2, 3>(c, lds, tl);
            add_pair_terms<R, 3, true>(c, l, lds, tl, G, w0 | w1 | w2, field);
            grad_window<R, 3, true, true>(c, l, lc, G, row_acc, lane, rz_tab);
            ry_inv_window<R, 3>(l, tile_coef);
            exchange<R, 3, 2>(l, lds, tl);
            ry_inv_window<R, 2>(l, tile_coef);
            exchange<R, 2, 1>(l, lds, tl);
            ry_inv_window<R, 1>(l, tile_coef);
            exchange<R, 1, 0>(l, lds, tl);
            ry_inv_window<R, 0>(l, tile_coef);
        }
        // lambda in layout G.lay_out (a bit permutation: mapping 0's offset and
        // its register step, tile bit T - R, map bit by bit)
        TileAddr ta{};
        if constexpr (kLayRows) {
            ta.offA = lay_row[tl];
            ta.strideA = lay_step;
        } else {
            ta.offA = lay_addr(omap<R, 0>(tl, 0), G.lay_out);
            ta.strideA = lay_addr(uint64_t{1} << (T - R), G.lay_out);
        }
        store_lam<R, 0, false>(l, lam_out, G.tile_out[g], ta, tl);
    }
    __syncthreads();
    if (t < 2 * T) {
        double total = 0.0;
#pragma unroll
        for (int w = 0; w < ROWS; ++w) {
            total += acc[w][t];
        }
        partials[(uint64_t)blockIdx.x * (2 * T) + t] = total;
    }
}

void launch_bwd_low_gen(int R, double2 *lam_out, const WireCoef *layer_coef, const PassGeom &g,
                        double *partials, int grid, hipStream_t s, double field,
                        const CMat2 *tab, const CMat2 *utab, const CMat2 *xtab) {
    if (g.a != Tile<4>::T || g.h0 != Tile<4>::T || g.rot_lo != 0 || g.rot_hi != Tile<4>::T ||
        g.n <= Tile<4>::T || g.lay_in.count != 0 || !lay_fits(g.lay_out, g.n, Tile<4>::T - R, R)) {
        throw std::invalid_argument("generating low pass: geometry outside the low tile");
    }
    if (g.tile_out == nullptr) {
        throw std::invalid_argument("generating low pass: no per-tile lambda address table");
    }
    if (g.pair_lo != 0 || g.pair_hi != Tile<4>::T) {
        // the kernel adds every in-tile pair (resolved at compile time)
        throw std::invalid_argument("generating low pass: the pair range must be the whole tile");
    }
    if (g.n > 32) {
        // the threads' store offsets in the layout are held in 32 bits
        throw std::invalid_argument("generating low pass: at most 32 qubits");
    }
    if (R == 4) {
        bwd_low_gen_kernel<4><<<grid, Tile<4>::THREADS, 0, s>>>(lam_out, layer_coef, g, partials,
                                                              field, tab, utab, xtab);
    } else {
        bwd_low_gen_kernel<3><<<grid, Tile<3>::THREADS, 0, s>>>(lam_out, layer_coef, g, partials,
                                                              field, tab, utab, xtab);
    }
}

template <int R>
static void launch_bwd_pass_r(bool gather, bool write_cur, bool write_lam, const double2 *cur_in,
                              const double2 *lam_in, double2 *cur_out, double2 *lam_out,
                              const WireCoef *lc, const PassGeom &g, double *partials, int grid,
                              hipStream_t s, bool ham, double field, const CMat2 *gen_tab,
                              const CMat2 *gen_utab, bool prefetch, bool phase_free) {
    constexpr int TH = Tile<R>::THREADS;
    const bool gen = gen_tab != nullptr;
    if (gen && (gen_utab == nullptr || gather || g.a > 4 || Tile<R>::T - R - g.a > 6 ||
                Tile<R>::T - R - g.a < 0)) {
        // chunk C0 holds up to 4 tile bits, C1 up to 6
        throw std::invalid_argument("generated state: high pass geometry outside the chain tables");
    }
    if (g.lay_in.count != 0 || g.lay_out.count != 0) {
        // lambda layouts: lambda-only high passes whose register bits the
        // layouts keep contiguous
        const int first = g.h0 + (Tile<R>::T - R - g.a);
        if (gather || write_cur || !write_lam || g.a != 3 || !lay_fits(g.lay_in, g.n, first, R) ||
            !lay_fits(g.lay_out, g.n, first, R)) {
            throw std::invalid_argument("lambda layouts: pass or geometry outside their scope");
        }
    }
    if (phase_free && (R != 4 || gather || !gen || ham || write_cur || !write_lam)) {
        // the tables leave the pass's own phases out only for this sweep
        throw std::invalid_argument(
            "phase-free sweep: generated lambda-only high pass of shape 4 without pair terms");
    }
    if (phase_free && (g.tile_in == nullptr || g.tile_out == nullptr)) {
        throw std::invalid_argument("phase-free sweep: no per-tile lambda address tables");
    }
    if (!write_cur && write_lam) {  // the state is not rewritten: lambda-only sweep
        if (!gather && R == 4 && (g.a != 3 || g.rot_lo < 3)) {
            // the sweep takes tile bit 3 across lanes and no other bit of window 1
            throw std::invalid_argument("high pass geometry: tile bits 0..2 must ride along");
        }
        if constexpr (R == 4) {
            if (phase_free && prefetch) {
                bwd_lam_kernel<R, false, false, true, true, true><<<grid, TH, 0, s>>>(
                    cur_in, lam_in, lam_out, lc, g, partials, field, gen_tab, gen_utab);
                return;
            }
            if (phase_free) {
                bwd_lam_kernel<R, false, false, true, false, true><<<grid, TH, 0, s>>>(
                    cur_in, lam_in, lam_out, lc, g, partials, field, gen_tab, gen_utab);
                return;
            }
        }

#define RIE_BWD_LAM(GA, HM, GN, PF)                                                          \
    if (gather == GA && ham == HM && gen == GN && prefetch == PF) {                          \
        bwd_lam_kernel<R, GA, HM, GN, PF><<<grid, TH, 0, s>>>(                               \
            cur_in, lam_in, lam_out, lc, g, partials, field, gen_tab, gen_utab);             \
        return;                                                                              \
    }
        RIE_BWD_LAM(true, false, false, false)
        RIE_BWD_LAM(false, false, false, false)
        RIE_BWD_LAM(true, true, false, false)
        RIE_BWD_LAM(false, true, false, false)
        RIE_BWD_LAM(false, false, true, false)
        RIE_BWD_LAM(false, true, true, false)
        RIE_BWD_LAM(true, false, false, true)
        RIE_BWD_LAM(true, true, false, true)
        RIE_BWD_LAM(false, false, false, true)
        RIE_BWD_LAM(false, true, false, true)
        RIE_BWD_LAM(false, false, true, true)
        RIE_BWD_LAM(false, true, true, true)
#undef RIE_BWD_LAM
    }
    if (gen) {
        throw std::invalid_argument("a generated state needs a high pass on the lambda-only sweep");
    }
    if (!gather && (g.a != 3 || g.rot_lo < 3)) {
        // the one-sweep high pass skips tile bits 0..2 at compile time
        throw std::invalid_argument("high pass geometry: tile bits 0..2 must ride along");
    }
#define RIE_BWD(GA, WC, WL, HM)                                                              \
    if (gather == GA && write_cur == WC && write_lam == WL && ham == HM) {                   \
        bwd_pass_kernel<R, GA, WC, WL, HM><<<grid, TH, 0, s>>>(                              \
            cur_in, lam_in, cur_out, lam_out, lc, g, partials, field);                       \
        return;                                                                              \
    }
    RIE_BWD(true, true, true, false)
    RIE_BWD(false, true, true, false)
    RIE_BWD(true, false, true, false)
    RIE_BWD(false, false, true, false)
    RIE_BWD(true, true, true, true)
    RIE_BWD(false, true, true, true)
    RIE_BWD(true, false, true, true)
    RIE_BWD(false, false, true, true)
#undef RIE_BWD
}

void launch_bwd_pass(int R, bool gather, bool write_cur, bool write_lam, const double2 *cur_in,
                     const double2 *lam_in, double2 *cur_out, double2 *lam_out,
                     const WireCoef *layer_coef, const PassGeom &g, double *partials, int grid,
                     hipStream_t s, bool ham_pairs, double field, const CMat2 *gen_tab,
                     const CMat2 *gen_utab, bool prefetch, bool phase_free) {
    if (R == 4) {
        launch_bwd_pass_r<4>(gather, write_cur, write_lam, cur_in, lam_in, cur_out, lam_out,
                             layer_coef, g, partials, grid, s, ham_pairs, field, gen_tab,
                             gen_utab, prefetch, phase_free);
    } else {
        launch_bwd_pass_r<3>(gather, write_cur, write_lam, cur_in, lam_in, cur_out, lam_out,
                             layer_coef, g, partials, grid, s, ham_pairs, field, gen_tab,
                             gen_utab, prefetch, phase_free);
    }
}

// ---------------------------------------------------------------------------
// Hamiltonian tile pass.
// ---------------------------------------------------------------------------

template <int R, bool kDiag, bool kReadLam, bool kWriteLam, bool kScat>
__global__ __launch_bounds__(Tile<R>::THREADS, tile_min_blocks<R>()) void ham_pass_kernel(
    const double2 *__restrict__ psi, const double2 *lam_in, double2 *lam_out, PassGeom G,
    double field, double *__restrict__ e_partials) {
    using TL = Tile<R>;
    constexpr int T = TL::T;
    __shared__ double2 lds[TL::SLOTS];
    // the scattered pass serves gradient evaluations only, which take the
    // energy from the layer-0 contraction: no energy partials there
    constexpr bool kEnergy = !kScat;
    __shared__ double red[kEnergy ? TL::WAVES : 1];
    const uint32_t t = threadIdx.x;
    // Index of register k of the load mapping relative to the thread's row:
    // a uniform stride for contiguous high bits, a deposit of k's bits into
    // the scattered positions otherwise (the offset bits are disjoint).
    TileAddr ta{};
    uint64_t koff[kScat ? TL::REGS : 1];
    if constexpr (kScat) {
        ta.offA = scat_index(omap<R, 0>(t, 0), G);
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            koff[k] = scat_index((uint32_t)k << (T - R), G);
        }
    } else {
        ta = tile_addr<R>(t, G);
    }
    auto eidx = [&](uint64_t row, int k) -> uint64_t {
        if constexpr (kScat) {
            return row + koff[k];
        } else {
            return row + k * ta.strideA;
        }
    };
    auto tbase = [&](uint64_t g) -> uint64_t {
        if constexpr (kScat) {
            return scat_base(g, G);
        } else {
            return tile_base(g, G, T);
        }
    };
    double e = 0.0;
    // psi of the next tile is loaded while this tile is processed; flips of
    // the register-resident bits [T - R, T) come from registers, the others
    // from the LDS copy of the tile.
    double2 nv[TL::REGS];
    if ((uint64_t)blockIdx.x < G.num_tiles) {
        const uint64_t row0 = tbase(blockIdx.x) + ta.offA;
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            nv[k] = psi[eidx(row0, k)];
        }
    }
    const int lds_hi = G.rot_hi < T - R ? G.rot_hi : T - R;
    for (uint64_t g = blockIdx.x; g < G.num_tiles; g += gridDim.x) {
        // per-tile opaque thread index: LDS slots are formed where used
        uint32_t tl = t;
        asm volatile("" : "+v"(tl));
        const uint64_t base = tbase(g);
        const uint64_t row = base + ta.offA;
        double2 v[TL::REGS];
        double2 prev[kReadLam ? TL::REGS : 1];
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            v[k] = nv[k];
            if constexpr (kReadLam) {
                prev[k] = lam_in[eidx(row, k)];
            }
        }
        const uint64_t gn = g + gridDim.x;
        if (gn < G.num_tiles) {
            const uint64_t rown = tbase(gn) + ta.offA;
#pragma unroll
            for (int k = 0; k < TL::REGS; ++k) {
                nv[k] = psi[eidx(rown, k)];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            lds[lds_slot<R>(omap<R, 0>(tl, k))] = v[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            const uint32_t o = omap<R, 0>(tl, k);
            const uint64_t idx = eidx(row, k);
            double2 flips = make_double2(0.0, 0.0);
            for (int j = G.rot_lo; j < lds_hi; ++j) {
                const double2 p = lds[lds_slot<R>(o ^ (1u << j))];
                flips.x += p.x;
                flips.y += p.y;
            }
#pragma unroll
            for (int rb = 0; rb < R; ++rb) {
                const int j = T - R + rb;
                if (j >= G.rot_lo && j < G.rot_hi) {
                    const double2 p = v[k ^ (1 << rb)];
                    flips.x += p.x;
                    flips.y += p.y;
                }
            }
            double2 contrib = make_double2(-field * flips.x, -field * flips.y);
            if constexpr (kDiag) {
                const double d = ham_diag(idx, G.n);
                contrib.x += d * v[k].x;
                contrib.y += d * v[k].y;
            }
            if constexpr (kEnergy) {
                e += v[k].x * contrib.x + v[k].y * contrib.y;
            }
            if constexpr (kWriteLam) {
                double2 out = contrib;
                if constexpr (kReadLam) {
                    out.x += prev[k].x;
                    out.y += prev[k].y;
                }
                lam_out[idx] = out;
            }
        }
    }
    if constexpr (kEnergy) {
        const double ws = wave_sum(e);
        if ((t & (kWave - 1)) == 0) {
            red[t / kWave] = ws;
        }
        __syncthreads();
        if (t == 0) {
            double acc = 0.0;
            for (int w = 0; w < TL::WAVES; ++w) {
                acc += red[w];
            }
            e_partials[blockIdx.x] = acc;
        }
    }
}

template <int R>
static void launch_ham_pass_r(bool diag, bool read_lam, bool write_lam, const double2 *psi,
                              const double2 *lam_in, double2 *lam_out, const PassGeom &g,
                              double field, double *e_partials, int grid, hipStream_t s) {
    constexpr int TH = Tile<R>::THREADS;
#define RIE_HAM(DG, RL, WL)                                                                  \
    if (diag == DG && read_lam == RL && write_lam == WL) {                                   \
        ham_pass_kernel<R, DG, RL, WL, false><<<grid, TH, 0, s>>>(psi, lam_in, lam_out, g,   \
                                                                  field, e_partials);        \
        return;                                                                              \
    }
    RIE_HAM(true, false, true)
    RIE_HAM(false, true, true)
    RIE_HAM(true, false, false)
    RIE_HAM(false, false, false)
#undef RIE_HAM
}

void launch_ham_pass_scattered(int R, const double2 *psi, double2 *lam_out, const PassGeom &g,
                               double field, int grid, hipStream_t s) {
    if (R == 4) {
        ham_pass_kernel<4, true, false, true, true>
            <<<grid, Tile<4>::THREADS, 0, s>>>(psi, nullptr, lam_out, g, field, nullptr);
    } else {
        ham_pass_kernel<3, true, false, true, true>
            <<<grid, Tile<3>::THREADS, 0, s>>>(psi, nullptr, lam_out, g, field, nullptr);
    }
}

void launch_ham_pass(int R, bool diag, bool read_lam, bool write_lam, const double2 *psi,
                     const double2 *lam_in, double2 *lam_out, const PassGeom &g, double field,
                     double *e_partials, int grid, hipStream_t s) {
    if (R == 4) {
        launch_ham_pass_r<4>(diag, read_lam, write_lam, psi, lam_in, lam_out, g, field,
                             e_partials, grid, s);
    } else {
        launch_ham_pass_r<3>(diag, read_lam, write_lam, psi, lam_in, lam_out, g, field,
                             e_partials, grid, s);
    }
}

// ---------------------------------------------------------------------------
// Fixed-order reductions.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(64) void reduce_grad_kernel(const double *__restrict__ partials,
                                                         int num_blocks, int T, PassGeom G,
                                                         int layer, double *__restrict__ grad,
                                                         const double *__restrict__ scale) {
    const int col = 2 * G.rot_lo + blockIdx.x;  // 2 * tile bit + component
    const int j = col >> 1;
    const int comp = col & 1;
    double acc = 0.0;
    for (int b = threadIdx.x; b < num_blocks; b += 64) {
        acc += partials[(uint64_t)b * (2 * T) + col];
    }
    acc = wave_sum(acc);
    if (threadIdx.x == 0) {
        const int wire = tile_wire(j, G);
        const double f = scale != nullptr ? *scale : 1.0;
        grad[((uint64_t)layer * G.n + wire) * 2 + comp] = acc * f;
    }
}

void launch_reduce_grad(const double *partials, int num_blocks, int T, const PassGeom &g,
                        int layer, double *grad, const double *scale, hipStream_t s) {
    const int cols = 2 * (g.rot_hi - g.rot_lo);
    reduce_grad_kernel<<<cols, 64, 0, s>>>(partials, num_blocks, T, g, layer, grad, scale);
}

__global__ __launch_bounds__(256) void reduce_sum_kernel(const double *__restrict__ values,
                                                         int count, double *__restrict__ out,
                                                         const double *__restrict__ scale) {
    __shared__ double red[4];
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int i = t; i < count; i += 256) {
        acc += values[i];
    }
    acc = wave_sum(acc);
    if ((t & 63) == 0) {
        red[t >> 6] = acc;
    }
    __syncthreads();
    if (t == 0) {
        const double f = scale != nullptr ? *scale : 1.0;
        *out = (red[0] + red[1] + red[2] + red[3]) * f;
    }
}

void launch_reduce_sum(const double *values, int count, double *out, hipStream_t s) {
    reduce_sum_kernel<<<1, 256, 0, s>>>(values, count, out, nullptr);
}

void launch_reduce_to(const double *partials, int count, double *out, const double *scale,
                      hipStream_t s) {
    reduce_sum_kernel<<<1, 256, 0, s>>>(partials, count, out, scale);
}

__global__ __launch_bounds__(256) void drift_kernel(const double2 *__restrict__ state,
                                                    uint64_t size, double *__restrict__ partials) {
    __shared__ double red[4];
    const int t = threadIdx.x;
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + t; i < size; i += stride) {
        const double2 v = state[i];
        const double re = v.x - (i == 0 ? 1.0 : 0.0);
        acc += re * re + v.y * v.y;
    }
    acc = wave_sum(acc);
    if ((t & 63) == 0) {
        red[t >> 6] = acc;
    }
    __syncthreads();
    if (t == 0) {
        partials[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    }
}

void launch_drift(const double2 *state, uint64_t size, double *partials, int grid, hipStream_t s) {
    drift_kernel<<<grid, 256, 0, s>>>(state, size, partials);
}

// ---------------------------------------------------------------------------
// Layer-0 gradients by contraction.
//
// Layer 0 acts on |0...0>, so the state right after its rotations is the
// product state (x)_w u_w with u_w = U_w|0>, and every layer-0 derivative is
// that product state with one factor replaced by d u_w.  With mu the adjoint
// state at the same point (lambda after the inverse ring CNOT of layer 0),
//   dE/dp_w = 2 Re sum_b E_w[b] (d u_w / dp)[b],
//   E_w[b]  = sum_{i : i_w = b} conj(mu_i) prod_{v != w} u_v(i_v).
// One read pass over lambda (gathered through P^-1 per low tile) produces
// the environments E_w of the tile wires (accumulated with the tile's high
// product B(g)) and the per-tile contraction S(g) of all tile wires; a small
// second kernel contracts S with the high wires.  This replaces the layer-0
// inverse walk (which would read and rewrite lambda once per tile pass).
// ---------------------------------------------------------------------------

__device__ __forceinline__ double2 cconj_mul(double2 a, double2 b) {
    // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ void cmac(double2 &acc, double2 a, double2 b) {
    acc.x += a.x * b.x - a.y * b.y;
    acc.y += a.x * b.y + a.y * b.x;
}

// conj(a) * b and acc += conj(a) * b
__device__ __forceinline__ double2 cmul_ca(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ void cmac_ca(double2 &acc, double2 a, double2 b) {
    acc.x += a.x * b.x + a.y * b.y;
    acc.y += a.x * b.y - a.y * b.x;
}

__device__ __forceinline__ double2 u_of(const WireCoef &k, int bit) {
    return bit ? u1_of(k) : u0_of(k);
}

// Contraction of one thread's 16 registers m[k] (k = k0 + 2 k1 + 4 k2 + 8 k3)
// against conj(m) and the product state uf[j][k_j] of the four register wires:
// pk = sum_k conj(m_k) prod_j uf[j][k_j], and env[e][b] the same sum over
// k_e = b without wire e's factor.  Hierarchical: A contracts k0, AB k0..k1,
// ABC k0..k2; E, F, G keep one bit uncontracted for the environments.
__device__ __forceinline__ void contract_reg4(const double2 (&m)[16], const double2 (&uf)[4][2],
                                              double2 &pk, double2 (&env)[4][2]) {
    double2 A[8], F[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {  // r = k1 + 2 k2 + 4 k3  /  k0 + 2 k2 + 4 k3
        A[r] = cmul_ca(m[2 * r], uf[0][0]);
        cmac_ca(A[r], m[2 * r + 1], uf[0][1]);
        const int k0 = r & 1;
        const int hi = r >> 1;  // k2 + 2 k3
        F[r] = cmul_ca(m[k0 + 4 * hi], uf[1][0]);
        cmac_ca(F[r], m[k0 + 2 + 4 * hi], uf[1][1]);
    }
    double2 AB[4], E[4], G[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        // AB[k2 + 2 k3]
        AB[r] = cmul(A[2 * r], uf[1][0]);
        cmac(AB[r], A[2 * r + 1], uf[1][1]);
        // E[k1 + 2 k3] = contract k2 of A[k1, k2, k3]
        const int k1 = r & 1;
        const int k3 = r >> 1;
        E[r] = cmul(A[k1 + 4 * k3], uf[2][0]);
        cmac(E[r], A[k1 + 2 + 4 * k3], uf[2][1]);
        // G[k0 + 2 k3] = contract k2 of F[k0, k2, k3]
        G[r] = cmul(F[k1 + 4 * k3], uf[2][0]);
        cmac(G[r], F[k1 + 2 + 4 * k3], uf[2][1]);
    }
    double2 ABC[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        ABC[b] = cmul(AB[2 * b], uf[2][0]);          // contract k2 of AB[k2, k3=b]
        cmac(ABC[b], AB[2 * b + 1], uf[2][1]);
        env[3][b] = ABC[b];
        env[2][b] = cmul(AB[b], uf[3][0]);           // contract k3 of AB[k2=b, k3]
        cmac(env[2][b], AB[b + 2], uf[3][1]);
        env[1][b] = cmul(E[b], uf[3][0]);            // contract k3 of E[k1=b, k3]
        cmac(env[1][b], E[b + 2], uf[3][1]);
        env[0][b] = cmul(G[b], uf[3][0]);            // contract k3 of G[k0=b, k3]
        cmac(env[0][b], G[b + 2], uf[3][1]);
    }
    pk = cmul(ABC[0], uf[3][0]);
    cmac(pk, ABC[1], uf[3][1]);
}


// R = 4 (T = 12) low tiles; 256 threads, thread t holds offsets t | k << 8.
__global__ __launch_bounds__(256) void layer0_contract_kernel(
    const double2 *__restrict__ lam, const WireCoef *__restrict__ lc0, PassGeom G,
    double2 *__restrict__ tile_sums, double *__restrict__ partials) {
    constexpr int R = 4;
    using TL = Tile<R>;
    constexpr int T = TL::T;
    __shared__ double2 lds[TL::SLOTS];
    __shared__ double2 red2[TL::WAVES];
    __shared__ double red[TL::WAVES][48];
    const uint32_t t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    double2 uf[4][2];  // u_{8+j}(b) for the register wires
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const WireCoef kj = lc0[8 + j];
        uf[j][0] = u0_of(kj);
        uf[j][1] = u1_of(kj);
    }
    // thread constant over offset bits 0..7 (wires 0..7)
    double2 a_lo = make_double2(1.0, 0.0);
    for (int v = 0; v < 8; ++v) {
        a_lo = cmul(a_lo, u_of(lc0[v], (t >> v) & 1));
    }
    // Accumulated over this workgroup's tiles; the thread constants over
    // wires 0..7 factor out and are applied once at the end.
    double2 bpk_sum = make_double2(0.0, 0.0);   // sum_g B(g) pk(g)
    double2 benv[4][2];                          // sum_g B(g) env_e[b](g)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        benv[e][0] = make_double2(0.0, 0.0);
        benv[e][1] = make_double2(0.0, 0.0);
    }
    __syncthreads();

    for (uint64_t g = blockIdx.x; g < G.num_tiles; g += gridDim.x) {
        // per-tile opaque thread index: the LDS slots are formed where used,
        // not held (or spilled) across the loop
        uint32_t tl = t;
        asm volatile("" : "+v"(tl));
        const uint64_t base = g << T;
        const uint64_t pbase = ring_fwd(base, G.n);
        const uint64_t lo_flip = pbase & 1u;
        const uint64_t run = (pbase & ~uint64_t{TL::SIZE - 1}) + (tl ^ lo_flip);
        const uint64_t run1 = run ^ G.hi_mask;
        double2 m[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            m[k] = load_stream(lam + (k < 8 ? run : run1) + ((uint64_t)k << 8));
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            lds[lds_slot<R>(ring_inv32(omap<R, 0>(tl, k), T))] = m[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            m[k] = lds[lds_slot<R>(omap<R, 0>(tl, k))];
        }
        // high-wire product of this tile
        double2 bg = make_double2(1.0, 0.0);
        for (int v = T; v < G.n; ++v) {
            bg = cmul(bg, u_of(lc0[v], (int)((base >> v) & 1U)));
        }
        // the register bits k0..k3 (wires 8..11) against conj(mu)
        double2 pk, env[4][2];
        contract_reg4(m, uf, pk, env);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            cmac(benv[e][0], bg, env[e][0]);
            cmac(benv[e][1], bg, env[e][1]);
        }
        cmac(bpk_sum, bg, pk);
        // per-tile contraction of all tile wires: S(g) = sum_t pk * a_lo
        const double2 st = cmul(pk, a_lo);
        const double sx = wave_sum(st.x);
        const double sy = wave_sum(st.y);
        if (lane == 0) {
            red2[wave] = make_double2(sx, sy);
        }
        __syncthreads();
        if (t == 0) {
            double2 tot = red2[0];
            for (int w = 1; w < TL::WAVES; ++w) {
                tot.x += red2[w].x;
                tot.y += red2[w].y;
            }
            tile_sums[g] = tot;
        }
    }
    // reduce the environments: 8 low wires x 2 entries + 4 register wires x 2
    // entries, complex -> 48 doubles.  The thread index is made opaque here so
    // the per-thread factors below are not formed before the tile loop and
    // held in registers across it.
    uint32_t te = t;
    asm volatile("" : "+v"(te));
#pragma unroll
    for (int w = 0; w < 8; ++w) {
        double2 skip = make_double2(1.0, 0.0);  // prod_{v < 8, v != w} u_v(t_v)
        for (int v = 0; v < 8; ++v) {
            if (v != w) {
                skip = cmul(skip, u_of(lc0[v], (te >> v) & 1));
            }
        }
        const double2 val = cmul(bpk_sum, skip);
        const int b = (te >> w) & 1;
        const double vals[4] = {b ? 0.0 : val.x, b ? 0.0 : val.y, b ? val.x : 0.0,
                                b ? val.y : 0.0};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double r = wave_sum(vals[c]);
            if (lane == 0) {
                red[wave][4 * w + c] = r;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const double2 v0 = cmul(a_lo, benv[e][0]);
        const double2 v1 = cmul(a_lo, benv[e][1]);
        const double vals[4] = {v0.x, v0.y, v1.x, v1.y};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double r = wave_sum(vals[c]);
            if (lane == 0) {
                red[wave][32 + 4 * e + c] = r;
            }
        }
    }
    __syncthreads();
    if (t < 48) {
        double acc = 0.0;
        for (int w = 0; w < TL::WAVES; ++w) {
            acc += red[w][t];
        }
        partials[(uint64_t)blockIdx.x * 48 + t] = acc;
    }
}

// The same contraction with 512 threads x 8 registers: threads hold tile
// offset bits 0..8 (wires 0..8), registers bits 9..11 (wires 9..11); twice
// the waves of the 256-thread kernel keep more of the read stream in flight.
// Partials use the same layout (4 doubles per wire w at column 4 w).
__global__ __launch_bounds__(512, 4) void layer0_contract_r3_kernel(
    const double2 *__restrict__ lam, const WireCoef *__restrict__ lc0, PassGeom G,
    double2 *__restrict__ tile_sums, double *__restrict__ partials) {
    constexpr int R = 3;
    using TL = Tile<R>;
    constexpr int T = TL::T;
    constexpr int TB = T - R;  // thread bits: wires 0..8
    __shared__ double2 lds[TL::SLOTS];
    __shared__ double2 red2[TL::WAVES];
    __shared__ double red[TL::WAVES][48];
    const uint32_t t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    double2 uf[3][2];  // u_{9+j}(b) for the register wires
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const WireCoef kj = lc0[TB + j];
        uf[j][0] = u0_of(kj);
        uf[j][1] = u1_of(kj);
    }
    double2 a_lo = make_double2(1.0, 0.0);  // thread constant over wires 0..8
    for (int v = 0; v < TB; ++v) {
        a_lo = cmul(a_lo, u_of(lc0[v], (t >> v) & 1));
    }
    double2 bpk_sum = make_double2(0.0, 0.0);
    double2 benv[3][2];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        benv[e][0] = make_double2(0.0, 0.0);
        benv[e][1] = make_double2(0.0, 0.0);
    }
    __syncthreads();

    for (uint64_t g = blockIdx.x; g < G.num_tiles; g += gridDim.x) {
        // per-tile opaque thread index: the LDS slots are formed where used,
        // not held (or spilled) across the loop
        uint32_t tl = t;
        asm volatile("" : "+v"(tl));
        const uint64_t base = g << T;
        const uint64_t pbase = ring_fwd(base, G.n);
        const uint64_t lo_flip = pbase & 1u;
        const uint64_t run = (pbase & ~uint64_t{TL::SIZE - 1}) + (tl ^ lo_flip);
        const uint64_t run1 = run ^ G.hi_mask;
        double2 m[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            m[k] = load_stream(lam + (k < 4 ? run : run1) + ((uint64_t)k << TB));
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            lds[lds_slot<R>(ring_inv32(omap<R, 0>(tl, k), T))] = m[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            m[k] = lds[lds_slot<R>(omap<R, 0>(tl, k))];
        }
        double2 bg = make_double2(1.0, 0.0);  // high-wire product of this tile
        for (int v = T; v < G.n; ++v) {
            bg = cmul(bg, u_of(lc0[v], (int)((base >> v) & 1U)));
        }
        // against conj(mu), k = k0 + 2 k1 + 4 k2 (wires 9, 10, 11)
        double2 A[4];  // k0 contracted: A[k1 + 2 k2]
        double2 F[4];  // k1 contracted: F[k0 + 2 k2]
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            A[r] = cmul_ca(m[2 * r], uf[0][0]);
            cmac_ca(A[r], m[2 * r + 1], uf[0][1]);
            const int k0 = r & 1;
            const int k2 = r >> 1;
            F[r] = cmul_ca(m[k0 + 4 * k2], uf[1][0]);
            cmac_ca(F[r], m[k0 + 2 + 4 * k2], uf[1][1]);
        }
        double2 env[3][2];
        double2 AB[2];  // k0, k1 contracted: AB[k2]
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            AB[b] = cmul(A[2 * b], uf[1][0]);
            cmac(AB[b], A[2 * b + 1], uf[1][1]);
            env[2][b] = AB[b];
            env[1][b] = cmul(A[b], uf[2][0]);      // contract k2 of A[k1 = b, k2]
            cmac(env[1][b], A[b + 2], uf[2][1]);
            env[0][b] = cmul(F[b], uf[2][0]);      // contract k2 of F[k0 = b, k2]
            cmac(env[0][b], F[b + 2], uf[2][1]);
        }
        double2 pk = cmul(AB[0], uf[2][0]);
        cmac(pk, AB[1], uf[2][1]);
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            cmac(benv[e][0], bg, env[e][0]);
            cmac(benv[e][1], bg, env[e][1]);
        }
        cmac(bpk_sum, bg, pk);
        const double2 st = cmul(pk, a_lo);
        const double sx = wave_sum(st.x);
        const double sy = wave_sum(st.y);
        if (lane == 0) {
            red2[wave] = make_double2(sx, sy);
        }
        __syncthreads();
        if (t == 0) {
            double2 tot = red2[0];
            for (int w = 1; w < TL::WAVES; ++w) {
                tot.x += red2[w].x;
                tot.y += red2[w].y;
            }
            tile_sums[g] = tot;
        }
    }
    // environments: 9 thread wires and 3 register wires, 4 doubles each (the
    // thread index is opaque so the per-thread factors are formed here)
    uint32_t te = t;
    asm volatile("" : "+v"(te));
    for (int w = 0; w < TB; ++w) {
        double2 skip = make_double2(1.0, 0.0);  // prod_{v < 9, v != w} u_v(t_v)
        for (int v = 0; v < TB; ++v) {
            if (v != w) {
                skip = cmul(skip, u_of(lc0[v], (te >> v) & 1));
            }
        }
        const double2 val = cmul(bpk_sum, skip);
        const int b = (te >> w) & 1;
        const double vals[4] = {b ? 0.0 : val.x, b ? 0.0 : val.y, b ? val.x : 0.0,
                                b ? val.y : 0.0};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double r = wave_sum(vals[c]);
            if (lane == 0) {
                red[wave][4 * w + c] = r;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const double2 v0 = cmul(a_lo, benv[e][0]);
        const double2 v1 = cmul(a_lo, benv[e][1]);
        const double vals[4] = {v0.x, v0.y, v1.x, v1.y};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double r = wave_sum(vals[c]);
            if (lane == 0) {
                red[wave][4 * (TB + e) + c] = r;
            }
        }
    }
    __syncthreads();
    if (t < 48) {
        double acc = 0.0;
        for (int w = 0; w < TL::WAVES; ++w) {
            acc += red[w][t];
        }
        partials[(uint64_t)blockIdx.x * 48 + t] = acc;
    }
}

// Environments of the high wires v in [12, n) from the per-tile sums S(g):
// E_v[b] = sum_{g : bit v of (g << 12) = b} S(g) prod_{u >= 12, u != v} u_u(bit).
// Block (v - 12, chunk) writes 4 doubles (E_v[0], E_v[1]) into partials.
__global__ __launch_bounds__(256) void layer0_high_env_kernel(
    const double2 *__restrict__ tile_sums, uint64_t num_tiles, const WireCoef *__restrict__ lc0,
    int n, double *__restrict__ partials) {
    constexpr int T = 12;
    __shared__ double red[4][4];
    const int v = T + blockIdx.y;
    const int t = threadIdx.x;
    double2 e0 = make_double2(0.0, 0.0);
    double2 e1 = make_double2(0.0, 0.0);
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t g = (uint64_t)blockIdx.x * 256 + t; g < num_tiles; g += stride) {
        double2 f = tile_sums[g];
        for (int u = T; u < n; ++u) {
            if (u != v) {
                f = cmul(f, u_of(lc0[u], (int)((g >> (u - T)) & 1U)));
            }
        }
        if ((g >> (v - T)) & 1U) {
            e1.x += f.x;
            e1.y += f.y;
        } else {
            e0.x += f.x;
            e0.y += f.y;
        }
    }
    const double vals[4] = {e0.x, e0.y, e1.x, e1.y};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double r = wave_sum(vals[c]);
        if ((t & 63) == 0) {
            red[t >> 6][c] = r;
        }
    }
    __syncthreads();
    if (t < 4) {
        partials[((uint64_t)blockIdx.y * gridDim.x + blockIdx.x) * 4 + t] =
            red[0][t] + red[1][t] + red[2][t] + red[3][t];
    }
}

// Gradients of layer-0 wire w from its environments e = (E_w[0], E_w[1]) as
// {re, im, re, im}: d u / d theta = (-s/2 e-, c/2 e+),
// d u / d phi = (-i/2 c e-, +i/2 s e+), e-+ = e^{-+i phi/2}.  Wire 0 also
// writes the energy: with lambda = H psi carried back to layer 0,
// <chi|mu> = <psi|H|psi>, so E = Re sum_b E_w[b] u_w(b) for any wire w.
__device__ __forceinline__ void layer0_wire_result(const WireCoef &k, const double (&e)[4], int w,
                                                   double *__restrict__ grad,
                                                   const double *__restrict__ scale,
                                                   double *__restrict__ energy_out) {
    const double2 em = make_double2(k.ch, -k.sh);
    const double2 ep = make_double2(k.ch, k.sh);
    const double2 dth0 = make_double2(-0.5 * k.s * em.x, -0.5 * k.s * em.y);
    const double2 dth1 = make_double2(0.5 * k.c * ep.x, 0.5 * k.c * ep.y);
    // -i/2 c e- and +i/2 s e+
    const double2 dph0 = make_double2(0.5 * k.c * em.y, -0.5 * k.c * em.x);
    const double2 dph1 = make_double2(-0.5 * k.s * ep.y, 0.5 * k.s * ep.x);
    const double2 E0 = make_double2(e[0], e[1]);
    const double2 E1 = make_double2(e[2], e[3]);
    const double dth = 2.0 * (cmul(E0, dth0).x + cmul(E1, dth1).x);
    const double dph = 2.0 * (cmul(E0, dph0).x + cmul(E1, dph1).x);
    const double f = scale != nullptr ? *scale : 1.0;
    grad[2 * w] = dth * f;
    grad[2 * w + 1] = dph * f;
    if (w == 0 && energy_out != nullptr) {
        *energy_out = cmul(E0, u0_of(k)).x + cmul(E1, u1_of(k)).x;
    }
}

// Layer-0 gradients from the environments of the low-tile contraction
// (wires 0..11, per workgroup) and of layer0_high_env (wires 12..n-1).
__global__ __launch_bounds__(64) void layer0_finish_kernel(
    const double *__restrict__ low_partials, int low_blocks, const double *__restrict__ high_partials,
    int high_chunks, const WireCoef *__restrict__ lc0, int n, double *__restrict__ grad,
    const double *__restrict__ scale, double *__restrict__ energy_out) {
    const int w = blockIdx.x;  // wire
    const int t = threadIdx.x;
    double e[4] = {0.0, 0.0, 0.0, 0.0};
    if (w < 12) {
        const int col = (w < 8) ? 4 * w : 32 + 4 * (w - 8);
        for (int b = t; b < low_blocks; b += 64) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                e[c] += low_partials[(uint64_t)b * 48 + col + c];
            }
        }
    } else {
        for (int b = t; b < high_chunks; b += 64) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                e[c] += high_partials[((uint64_t)(w - 12) * high_chunks + b) * 4 + c];
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        e[c] = wave_sum(e[c]);
    }
    if (t == 0) {
        layer0_wire_result(lc0[w], e, w, grad, scale, energy_out);
    }
}

void launch_layer0_contract(int R, const double2 *lam, const WireCoef *lc0, const PassGeom &g,
                            double2 *tile_sums, double *partials, int grid, hipStream_t s) {
    if (R == 3) {
        layer0_contract_r3_kernel<<<grid, 512, 0, s>>>(lam, lc0, g, tile_sums, partials);
    } else {
        layer0_contract_kernel<<<grid, 256, 0, s>>>(lam, lc0, g, tile_sums, partials);
    }
}

void launch_layer0_high_env(const double2 *tile_sums, uint64_t num_tiles, const WireCoef *lc0,
                            int n, double *partials, int chunks, hipStream_t s) {
    if (n <= 12) {
        return;
    }
    dim3 grid(chunks, n - 12);
    layer0_high_env_kernel<<<grid, 256, 0, s>>>(tile_sums, num_tiles, lc0, n, partials);
}

void launch_layer0_finish(const double *low_partials, int low_blocks, const double *high_partials,
                          int high_chunks, const WireCoef *lc0, int n, double *grad,
                          const double *scale, hipStream_t s, double *energy_out) {
    layer0_finish_kernel<<<n, 64, 0, s>>>(low_partials, low_blocks, high_partials, high_chunks,
                                          lc0, n, grad, scale, energy_out);
}

// ---------------------------------------------------------------------------
// The last high pass of a two-layer evaluation, with the layer-0 contraction.
//
// When the generated start runs its high passes from the lowest group up, the
// last one holds the top group's tiles: index bits {0, 1, 2} and [h, n)
// (h = n - 9) of the frame after layer 0's ring CNOT, the bits [3, h) fixed by
// the tile.  Once the group is un-rotated, lambda of the tile is mu = P^-1 lambda
// at the layer-0 coordinates i = P^-1 j:
//   i_0 = j_0 ^ j_{n-1},  i_1 = j_0 ^ j_1 ^ j_{n-1},  i_2 = j_1 ^ j_2,
//   i_w = j_w ^ j_{w-1} (w >= 3),
// so q = (i_0, i_1, i_2, i_h, .., i_{n-1}) runs over all 2^12 values in the tile,
// i_4 .. i_{h-1} are fixed by it, and i_3 = j_3 ^ i_0 ^ i_1 ^ i_2 (j_2 =
// i_0 ^ i_1 ^ i_2).  Instead of storing lambda the pass permutes it through LDS
// to q (threads: q bits 0..7, registers: q bits 8..11) and contracts it with
// the layer-0 product state as layer0_contract_kernel does with a low tile:
// the register wires n-4 .. n-1 per thread (contract_reg4), the thread wires
// from per-thread sums split by x = j_3 ^ parity(q_0, q_1, q_2) (wire 3's
// factor is u_3(x)), the fixed wires 4 .. h-1 from per-tile sums S(g)
// (layer0_fixed_env_kernel) and the per-tile product phi_c(g) of their factors.
// ---------------------------------------------------------------------------

// q of top-tile offset o (bits 0..2: j_0..j_2, bits 3..11: j_h..j_{n-1})
// without the tile's j_{h-1}, which flips q bit 3.
__device__ __forceinline__ uint32_t top_q_of_offset(uint32_t o) {
    const uint32_t d = (o ^ (o << 1)) & 0xFFFu;
    const uint32_t q = (d & ~8u) | (o & 8u);
    return q ^ (((o >> 11) & 1u) * 3u);
}

// Wire of q bit b in the top tile.
__device__ __forceinline__ int top_wire(int b, int h) {
    return b < 3 ? b : h + (b - 3);
}

// Per top tile g: phi_c[g] = prod_{v = 4}^{h-1} u_v(j_v ^ j_{v-1}) (j_{3+m} =
// bit m of g) and lam_tile[g], the address of the tile's base index in the
// layout the pass reads.
__global__ __launch_bounds__(256) void lam_tile_table_kernel(PassGeom G, LamLayout L,
                                                             uint64_t *__restrict__ out) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < G.num_tiles) {
        out[g] = lay_addr(tile_base(g, G, Tile<4>::T), L);
    }
}

void launch_lam_tile_table(const PassGeom &g, const LamLayout &L, uint64_t *out, hipStream_t s) {
    lam_tile_table_kernel<<<(unsigned)((g.num_tiles + 255) / 256), 256, 0, s>>>(g, L, out);
}

__global__ __launch_bounds__(256) void layer0_top_factor_kernel(const WireCoef *__restrict__ lc0,
                                                                PassGeom G,
                                                                double2 *__restrict__ phi_c,
                                                                uint64_t *__restrict__ lam_tile) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G.num_tiles) {
        return;
    }
    double2 f = make_double2(1.0, 0.0);
    for (int v = 4; v < G.h0; ++v) {
        f = cmul(f, u_of(lc0[v], (int)(((g >> (v - 3)) ^ (g >> (v - 4))) & 1u)));
    }
    phi_c[g] = f;
    lam_tile[g] = lay_addr(tile_base(g, G, Tile<4>::T), G.lay_in);
}

// Per-block partials of the fused pass: layer-1 gradients as bwd_lam_kernel
// (2T per block), then kTopL0Cols layer-0 values per block: 4 doubles
// (E_w[0], E_w[1]) per thread wire at column 4b (q bit b = 0..7), per register
// wire at 32 + 4e (wire n - 4 + e), wire 3 at 48.
static_assert(kTopL0Cols == 52 && Tile4Waves == Tile<4>::WAVES, "fused top pass layout");

__global__ __launch_bounds__(256, 2) void bwd_top_l0_kernel(
    const double2 *lam_in, const WireCoef *__restrict__ lc, PassGeom G,
    double *__restrict__ partials, const CMat2 *__restrict__ gen_tab,
    const CMat2 *__restrict__ gen_utab, const WireCoef *__restrict__ lc0,
    const double2 *__restrict__ phi_c, const uint64_t *__restrict__ lam_tile,
    double2 *__restrict__ tile_sums, double *__restrict__ l0_partials) {
    constexpr int R = 4;
    using TL = Tile<R>;
    constexpr int T = TL::T;
    __shared__ double2 lds[TL::SLOTS];
    constexpr int ROWS = TL::THREADS / 16;
    __shared__ double acc[ROWS][2 * T];
    __shared__ double2 rz_tab[TL::P * TL::REGS];
    __shared__ WireCoef tile_coef[T];
    // RY^dagger of lambda's window bits (4..11) in tangent form, its scale
    // applied once before the contraction
    __shared__ double tan_t[T];
    __shared__ double ry_sigma;
    __shared__ CMat2 gen_chunks[kChainTabEntries - kChainC0];
    // u_{n-4+j}(b) at 2j + b (the register wires), u_3(b) at 8 + b
    __shared__ double2 l0_u[10];
    const uint32_t t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wave = (int)(t / kWave);
    const int n = G.n;
    const int h = G.h0;
    double *row_acc = acc[t / 16];
    for (uint32_t i = t; i < (uint32_t)(ROWS * 2 * T); i += TL::THREADS) {
        acc[i / (2 * T)][i % (2 * T)] = 0.0;
    }
    build_rz_tables<R>(rz_tab, lc, G, t);
    if (t < (uint32_t)T) {
        const bool rot = (int)t >= G.rot_lo && (int)t < G.rot_hi;
        tile_coef[t] = rot ? lc[tile_wire((int)t, G)] : WireCoef{1.0, 0.0, 1.0, 0.0};
    }
    for (int i = (int)t; i < kChainTabEntries - kChainC0; i += TL::THREADS) {
        gen_chunks[i] = gen_tab[kChainC0 + i];
    }
    if (t < 10) {
        const int w = t < 8 ? n - 4 + (int)(t >> 1) : 3;
        l0_u[t] = u_of(lc0[w], (int)(t & 1u));
    }
    __syncthreads();
    const uint32_t form_b = ry_tan_tables<T>(tile_coef, 4, tan_t, &ry_sigma, t);
    __syncthreads();

    const TileAddr ta = tile_addr<R>(t, G);
    const TileAddr ta_in = lay_tile_addr(ta, G.lay_in);
    // the thread's q bits 0..7 are its index: parity of q_0..q_2 and the
    // product a_t over its wires 0, 1, 2, h .. h + 4 (formed where used, so it
    // holds no registers across the sweep)
    auto thread_factor = [&](uint32_t tt) {
        double2 a = make_double2(1.0, 0.0);
        for (int b = 0; b < 8; ++b) {
            a = cmul(a, u_of(lc0[top_wire(b, h)], (int)((tt >> b) & 1u)));
        }
        return a;
    };
    // per-thread sums over this workgroup's tiles: zc[x] of phi_c pk with
    // x = j_3 ^ par, benv[e][b] of u_3(x) phi_c env[e][b]
    double2 zc0 = make_double2(0.0, 0.0);
    double2 zc1 = make_double2(0.0, 0.0);
    double2 benv[4][2];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        benv[e][0] = make_double2(0.0, 0.0);
        benv[e][1] = make_double2(0.0, 0.0);
    }
    for (uint64_t g = blockIdx.x; g < G.num_tiles; g += gridDim.x) {
        uint32_t tl = t;
        asm volatile("" : "+v"(tl));
        const uint64_t base = tile_base(g, G, T);
        double2 c[TL::REGS];
        double2 l[TL::REGS];
        // lambda's address of the tile in its layout from the per-tile table
        // (layer0_top_factor_kernel), so the layout is not evaluated here
        bwd_load_tile<R, false, false, true>(c, l, nullptr, lam_in, base, ta, G, lds, tl, 0.0,
                                             gen_tab, gen_chunks, gen_utab + 2 * g, lam_tile[g],
                                             &ta_in);
        // the sweep of the lambda-only high pass (bwd_lam_kernel, R = 4), the
        // state generated without the group's phases
        apply_tile_phase<R>(l, rz_tab, tl);
        grad_window<R, 0, false, true>(c, l, lc, G, row_acc, lane, rz_tab);
        exchange<R, 0, 2>(l, lds, tl);
        exchange<R, 0, 2>(c, lds, tl);
        grad_window<R, 2, false, true>(c, l, lc, G, row_acc, lane, rz_tab);
        grad_ry_lane_bit3<R, false>(c, l, row_acc, lane, rz_tab, tile_coef[3]);
        ry_inv_window_tan<R, 2>(l, tan_t, form_b);
        exchange<R, 2, 0>(l, lds, tl);
        ry_inv_window_tan<R, 0>(l, tan_t, form_b);
        scale_regs<R>(l, ry_sigma);
        // lambda to the layer-0 coordinates q (mapping 0 of q: thread bits 0..7)
        const uint32_t cflip = (uint32_t)((base >> (h - 1)) & 1u) << 3;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            lds[lds_slot<R>(top_q_of_offset(omap<R, 0>(tl, k)) ^ cflip)] = l[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TL::REGS; ++k) {
            l[k] = lds[lds_slot<R>(omap<R, 0>(tl, k))];
        }
        double2 uf[4][2];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uf[j][0] = l0_u[2 * j];
            uf[j][1] = l0_u[2 * j + 1];
        }
        double2 pk, env[4][2];
        contract_reg4(l, uf, pk, env);
        const uint32_t x = (uint32_t)(((base >> 3) ^ tl ^ (tl >> 1) ^ (tl >> 2)) & 1u);
        const double2 phi = phi_c[g];
        const double2 u3x = l0_u[8 + x];
        const double2 pp = cmul(phi, pk);
        zc0.x += x ? 0.0 : pp.x;
        zc0.y += x ? 0.0 : pp.y;
        zc1.x += x ? pp.x : 0.0;
        zc1.y += x ? pp.y : 0.0;
        const double2 bgw = cmul(phi, u3x);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            cmac(benv[e][0], bgw, env[e][0]);
            cmac(benv[e][1], bgw, env[e][1]);
        }
        // the tile's sum without phi_c, per wave
        const double2 st = cmul(cmul(thread_factor(tl), u3x), pk);
        const double sx = wave_sum(st.x);
        const double sy = wave_sum(st.y);
        if (lane == 0) {
            tile_sums[g * TL::WAVES + (uint64_t)wave] = make_double2(sx, sy);
        }
    }
    __syncthreads();
    if (t < 2 * T) {
        double total = 0.0;
#pragma unroll
        for (int w = 0; w < ROWS; ++w) {
            total += acc[w][t];
        }
        partials[(uint64_t)blockIdx.x * (2 * T) + t] = total;
    }
    // layer-0 environments of this workgroup's tiles; the exchange buffer
    // holds the per-wave sums (the thread index is opaque here so the factors
    // below are formed after the loop)
    double *red = reinterpret_cast<double *>(lds);
    uint32_t te = t;
    asm volatile("" : "+v"(te));
    const double2 a_t = thread_factor(te);
    const double2 u30 = l0_u[8];
    const double2 u31 = l0_u[9];
    double2 zsum = cmul(u30, zc0);
    cmac(zsum, u31, zc1);
#pragma unroll
    for (int bw = 0; bw < 8; ++bw) {
        double2 skip = make_double2(1.0, 0.0);
        for (int b = 0; b < 8; ++b) {
            if (b != bw) {
                skip = cmul(skip, u_of(lc0[top_wire(b, h)], (int)((te >> b) & 1u)));
            }
        }
        const double2 val = cmul(zsum, skip);
        const int bit = (int)((te >> bw) & 1u);
        const double vals[4] = {bit ? 0.0 : val.x, bit ? 0.0 : val.y, bit ? val.x : 0.0,
                                bit ? val.y : 0.0};
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
            const double r = wave_sum(vals[cc]);
            if (lane == 0) {
                red[wave * kTopL0Cols + 4 * bw + cc] = r;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < 5; ++e) {
        // e < 4: register wire n - 4 + e; e == 4: wire 3
        const double2 v0 = cmul(a_t, e < 4 ? benv[e & 3][0] : zc0);
        const double2 v1 = cmul(a_t, e < 4 ? benv[e & 3][1] : zc1);
        const double vals[4] = {v0.x, v0.y, v1.x, v1.y};
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
            const double r = wave_sum(vals[cc]);
            if (lane == 0) {
                red[wave * kTopL0Cols + 32 + 4 * e + cc] = r;
            }
        }
    }
    __syncthreads();
    if (t < (uint32_t)kTopL0Cols) {
        double sum = 0.0;
        for (int w = 0; w < TL::WAVES; ++w) {
            sum += red[w * kTopL0Cols + t];
        }
        l0_partials[(uint64_t)blockIdx.x * kTopL0Cols + t] = sum;
    }
}

// Environments of the fixed wires v in [4, h) of the top-group tiles from the
// per-tile sums (waves partials per tile):
// E_v[b] = sum_{g : i_v(g) = b} S(g) prod_{u in [4, h), u != v} u_u(i_u(g)),
// i_u(g) = j_u ^ j_{u-1}, j_{3+m} = bit m of g.  One thread per tile forms all
// of them from prefix and suffix products of the tile's factors (one read of
// the tile sums, O(h) products); block `chunk` writes 4 doubles per wire at
// ((v - 4) * chunks + chunk) * 4.
constexpr int kMaxFixedWires = 18;  // h - 4 = n - 13 for n <= 31

__global__ __launch_bounds__(256) void layer0_fixed_env_kernel(
    const double2 *__restrict__ tile_sums, int waves, uint64_t num_tiles,
    const WireCoef *__restrict__ lc0, int h, double *__restrict__ partials) {
    constexpr int M = kMaxFixedWires;
    __shared__ double red[4][4 * M];
    const int m = h - 4;
    const int t = threadIdx.x;
    double2 e0[M];
    double2 e1[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        e0[i] = make_double2(0.0, 0.0);
        e1[i] = make_double2(0.0, 0.0);
    }
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t g = (uint64_t)blockIdx.x * 256 + t; g < num_tiles; g += stride) {
        double2 f = tile_sums[g * waves];
        for (int w = 1; w < waves; ++w) {
            const double2 p = tile_sums[g * waves + w];
            f.x += p.x;
            f.y += p.y;
        }
        // prefix[i] = f * prod_{k < i} u_{4+k}, then the suffix products
        // from the top: E_{4+i} takes prefix[i] * prod_{k > i} u_{4+k}
        double2 uu[M];
        double2 pre[M];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            if (i < m) {
                uu[i] = u_of(lc0[4 + i], (int)(((g >> (i + 1)) ^ (g >> i)) & 1u));
                pre[i] = f;
                f = cmul(f, uu[i]);
            }
        }
        double2 suf = make_double2(1.0, 0.0);
#pragma unroll
        for (int i = M - 1; i >= 0; --i) {
            if (i < m) {
                const double2 val = cmul(pre[i], suf);
                const bool bit = (((g >> (i + 1)) ^ (g >> i)) & 1u) != 0;
                e0[i].x += bit ? 0.0 : val.x;
                e0[i].y += bit ? 0.0 : val.y;
                e1[i].x += bit ? val.x : 0.0;
                e1[i].y += bit ? val.y : 0.0;
                suf = cmul(suf, uu[i]);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < M; ++i) {
        if (i < m) {
            const double vals[4] = {e0[i].x, e0[i].y, e1[i].x, e1[i].y};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double r = wave_sum(vals[c]);
                if ((t & 63) == 0) {
                    red[t >> 6][4 * i + c] = r;
                }
            }
        }
    }
    __syncthreads();
    for (int k = t; k < 4 * m; k += 256) {
        partials[((uint64_t)(k >> 2) * gridDim.x + blockIdx.x) * 4 + (k & 3)] =
            red[0][k] + red[1][k] + red[2][k] + red[3][k];
    }
}

// Layer-0 gradients (and the energy) from the fused pass's per-block partials
// and the fixed wires' chunks.
__global__ __launch_bounds__(64) void layer0_finish_top_kernel(
    const double *__restrict__ blk_partials, int blocks, const double *__restrict__ fixed_partials,
    int chunks, const WireCoef *__restrict__ lc0, int n, double *__restrict__ grad,
    const double *__restrict__ scale, double *__restrict__ energy_out) {
    const int w = blockIdx.x;
    const int t = threadIdx.x;
    const int h = n - 9;
    double e[4] = {0.0, 0.0, 0.0, 0.0};
    if (w >= 4 && w < h) {
        for (int b = t; b < chunks; b += 64) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                e[c] += fixed_partials[((uint64_t)(w - 4) * chunks + b) * 4 + c];
            }
        }
    } else {
        const int col = w < 3        ? 4 * w
                        : w == 3     ? 48
                        : w < h + 5  ? 4 * (3 + w - h)
                                     : 32 + 4 * (w - h - 5);
        for (int b = t; b < blocks; b += 64) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                e[c] += blk_partials[(uint64_t)b * kTopL0Cols + col + c];
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        e[c] = wave_sum(e[c]);
    }
    if (t == 0) {
        layer0_wire_result(lc0[w], e, w, grad, scale, energy_out);
    }
}

void launch_bwd_top_l0(const double2 *lam_in, const WireCoef *layer_coef, const PassGeom &g,
                       double *partials, int grid, hipStream_t s, const CMat2 *gen_tab,
                       const CMat2 *gen_utab, const WireCoef *lc0, double2 *phi_c,
                       uint64_t *lam_tile, double2 *tile_sums, double *l0_partials) {
    if (g.a != 3 || g.h0 + 9 != g.n || g.h0 < 4 || g.rot_lo < 3 || g.lay_out.count != 0 ||
        !lay_fits(g.lay_in, g.n, g.h0 + 5, 4)) {
        throw std::invalid_argument("layer-0 contraction in a high pass: top group tiles only");
    }
    layer0_top_factor_kernel<<<(unsigned)((g.num_tiles + 255) / 256), 256, 0, s>>>(lc0, g, phi_c,
                                                                                  lam_tile);
    bwd_top_l0_kernel<<<grid, Tile<4>::THREADS, 0, s>>>(lam_in, layer_coef, g, partials, gen_tab,
                                                        gen_utab, lc0, phi_c, lam_tile, tile_sums,
                                                        l0_partials);
}

void launch_layer0_top_finish(const double2 *tile_sums, uint64_t num_tiles, const double *blk_partials,
                              int blocks, const WireCoef *lc0, int n, double *fixed_partials,
                              int chunks, double *grad, const double *scale, hipStream_t s,
                              double *energy_out) {
    const int h = n - 9;
    if (h - 4 > kMaxFixedWires) {
        throw std::invalid_argument("layer-0 contraction in the top pass: too many fixed wires");
    }
    if (h > 4) {
        layer0_fixed_env_kernel<<<chunks, 256, 0, s>>>(tile_sums, Tile<4>::WAVES, num_tiles, lc0, h,
                                                       fixed_partials);
    }
    layer0_finish_top_kernel<<<n, 64, 0, s>>>(blk_partials, blocks, fixed_partials, chunks, lc0, n,
                                              grad, scale, energy_out);
}

// ---------------------------------------------------------------------------
// Gate-level kernels (PennyLane's gate order: RY, RZ per wire, then n CNOTs).
// ---------------------------------------------------------------------------

__global__ void zero_state_kernel(double2 *__restrict__ s, uint64_t size) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) {
        s[i] = make_double2(i == 0 ? 1.0 : 0.0, 0.0);
    }
}

void launch_zero_state(double2 *state, uint64_t size, hipStream_t s) {
    zero_state_kernel<<<elementwise_grid(size), 256, 0, s>>>(state, size);
}

__device__ __forceinline__ uint64_t pair_low(uint64_t p, int w) {
    return ((p >> w) << (w + 1)) | (p & ((uint64_t{1} << w) - 1));
}

__global__ void gate_ry_kernel(double2 *__restrict__ st, uint64_t pairs, int w,
                               const WireCoef *__restrict__ kp, int adjoint) {
    const double c = kp->c;
    const double sn = adjoint ? -kp->s : kp->s;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < pairs; p += stride) {
        const uint64_t i0 = pair_low(p, w);
        const uint64_t i1 = i0 | (uint64_t{1} << w);
        const double2 a0 = st[i0];
        const double2 a1 = st[i1];
        st[i0] = make_double2(c * a0.x - sn * a1.x, c * a0.y - sn * a1.y);
        st[i1] = make_double2(sn * a0.x + c * a1.x, sn * a0.y + c * a1.y);
    }
}

__global__ void gate_rz_kernel(double2 *__restrict__ st, uint64_t pairs, int w,
                               const WireCoef *__restrict__ kp, int adjoint) {
    const double ch = kp->ch;
    const double sh = adjoint ? -kp->sh : kp->sh;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < pairs; p += stride) {
        const uint64_t i0 = pair_low(p, w);
        const uint64_t i1 = i0 | (uint64_t{1} << w);
        const double2 a0 = st[i0];
        const double2 a1 = st[i1];
        // bit 0: e^{-i phi/2}; bit 1: e^{+i phi/2}
        st[i0] = make_double2(ch * a0.x + sh * a0.y, ch * a0.y - sh * a0.x);
        st[i1] = make_double2(ch * a1.x - sh * a1.y, ch * a1.y + sh * a1.x);
    }
}

__global__ void gate_cnot_kernel(double2 *__restrict__ st, uint64_t pairs, int control,
                                 int target) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < pairs; p += stride) {
        const uint64_t i0 = pair_low(p, target);
        if ((i0 >> control) & 1U) {
            const uint64_t i1 = i0 | (uint64_t{1} << target);
            const double2 a0 = st[i0];
            st[i0] = st[i1];
            st[i1] = a0;
        }
    }
}

void launch_gate_ry(double2 *state, int n, int wire, const WireCoef *k, bool adjoint,
                    hipStream_t s) {
    const uint64_t pairs = uint64_t{1} << (n - 1);
    gate_ry_kernel<<<elementwise_grid(pairs), 256, 0, s>>>(state, pairs, wire, k, adjoint ? 1 : 0);
}

void launch_gate_rz(double2 *state, int n, int wire, const WireCoef *k, bool adjoint,
                    hipStream_t s) {
    const uint64_t pairs = uint64_t{1} << (n - 1);
    gate_rz_kernel<<<elementwise_grid(pairs), 256, 0, s>>>(state, pairs, wire, k, adjoint ? 1 : 0);
}

void launch_gate_cnot(double2 *state, int n, int control, int target, hipStream_t s) {
    const uint64_t pairs = uint64_t{1} << (n - 1);
    gate_cnot_kernel<<<elementwise_grid(pairs), 256, 0, s>>>(state, pairs, control, target);
}

__global__ void ring_permute_kernel(const double2 *__restrict__ in, double2 *__restrict__ out,
                                    int n, uint64_t size, int inverse) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) {
        // gather form: (P x)[i] = x[P^-1 i], (P^-1 x)[i] = x[P i]
        const uint64_t src = inverse ? ring_fwd(i, n) : ring_inv(i, n);
        out[i] = in[src];
    }
}

void launch_ring_permute(const double2 *in, double2 *out, int n, bool inverse, hipStream_t s) {
    const uint64_t size = uint64_t{1} << n;
    ring_permute_kernel<<<elementwise_grid(size), 256, 0, s>>>(in, out, n, size, inverse ? 1 : 0);
}

__device__ __forceinline__ void block_partial(double v, double *__restrict__ partials) {
    __shared__ double red[4];
    v = wave_sum(v);
    const int t = threadIdx.x;
    if ((t & 63) == 0) {
        red[t >> 6] = v;
    }
    __syncthreads();
    if (t == 0) {
        partials[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    }
}

// psi, lam <- RY^dagger (psi, lam); partial of 2 Re<lam_after | dRY psi_before>
// (PennyLane adjoint step; dRY = [[-s/2, -c/2], [c/2, -s/2]]).
__global__ __launch_bounds__(256) void gate_ry_step_kernel(double2 *__restrict__ psi,
                                                           double2 *__restrict__ lam,
                                                           uint64_t pairs, int w,
                                                           const WireCoef *__restrict__ kp,
                                                           double *__restrict__ partials) {
    const double c = kp->c;
    const double s = kp->s;
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < pairs; p += stride) {
        const uint64_t i0 = pair_low(p, w);
        const uint64_t i1 = i0 | (uint64_t{1} << w);
        const double2 a0 = psi[i0];
        const double2 a1 = psi[i1];
        const double2 b0 = make_double2(c * a0.x + s * a1.x, c * a0.y + s * a1.y);
        const double2 b1 = make_double2(c * a1.x - s * a0.x, c * a1.y - s * a0.y);
        psi[i0] = b0;
        psi[i1] = b1;
        const double2 d0 = make_double2(-0.5 * (s * b0.x + c * b1.x), -0.5 * (s * b0.y + c * b1.y));
        const double2 d1 = make_double2(0.5 * (c * b0.x - s * b1.x), 0.5 * (c * b0.y - s * b1.y));
        const double2 l0 = lam[i0];
        const double2 l1 = lam[i1];
        acc += l0.x * d0.x + l0.y * d0.y + l1.x * d1.x + l1.y * d1.y;
        lam[i0] = make_double2(c * l0.x + s * l1.x, c * l0.y + s * l1.y);
        lam[i1] = make_double2(c * l1.x - s * l0.x, c * l1.y - s * l0.y);
    }
    block_partial(2.0 * acc, partials);
}

// dRZ = diag(-i/2 e^{-i phi/2}, +i/2 e^{+i phi/2}).
__global__ __launch_bounds__(256) void gate_rz_step_kernel(double2 *__restrict__ psi,
                                                           double2 *__restrict__ lam,
                                                           uint64_t pairs, int w,
                                                           const WireCoef *__restrict__ kp,
                                                           double *__restrict__ partials) {
    const double ch = kp->ch;
    const double sh = kp->sh;
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < pairs; p += stride) {
        const uint64_t i0 = pair_low(p, w);
        const uint64_t i1 = i0 | (uint64_t{1} << w);
        // before = after * e^{+-i phi/2} (inverse phase)
        const double2 a0 = psi[i0];
        const double2 a1 = psi[i1];
        const double2 b0 = make_double2(ch * a0.x - sh * a0.y, ch * a0.y + sh * a0.x);
        const double2 b1 = make_double2(ch * a1.x + sh * a1.y, ch * a1.y - sh * a1.x);
        psi[i0] = b0;
        psi[i1] = b1;
        // dRZ psi_before: bit0 -> (-i/2) e^{-i phi/2} b0 = (-i/2) a0 ; bit1 -> (+i/2) a1
        const double2 d0 = make_double2(0.5 * a0.y, -0.5 * a0.x);
        const double2 d1 = make_double2(-0.5 * a1.y, 0.5 * a1.x);
        const double2 l0 = lam[i0];
        const double2 l1 = lam[i1];
        acc += l0.x * d0.x + l0.y * d0.y + l1.x * d1.x + l1.y * d1.y;
        lam[i0] = make_double2(ch * l0.x - sh * l0.y, ch * l0.y + sh * l0.x);
        lam[i1] = make_double2(ch * l1.x + sh * l1.y, ch * l1.y - sh * l1.x);
    }
    block_partial(2.0 * acc, partials);
}

void launch_gate_ry_step(double2 *psi, double2 *lam, int n, int wire, const WireCoef *k,
                         double *partials, int grid, hipStream_t s) {
    const uint64_t pairs = uint64_t{1} << (n - 1);
    gate_ry_step_kernel<<<grid, 256, 0, s>>>(psi, lam, pairs, wire, k, partials);
}

void launch_gate_rz_step(double2 *psi, double2 *lam, int n, int wire, const WireCoef *k,
                         double *partials, int grid, hipStream_t s) {
    const uint64_t pairs = uint64_t{1} << (n - 1);
    gate_rz_step_kernel<<<grid, 256, 0, s>>>(psi, lam, pairs, wire, k, partials);
}

__global__ __launch_bounds__(256) void ham_gather_kernel(const double2 *__restrict__ psi,
                                                         double2 *__restrict__ lam, int n,
                                                         uint64_t size, double field,
                                                         double *__restrict__ partials) {
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) {
        const double2 v = psi[i];
        const double d = ham_diag(i, n);
        double fx = 0.0;
        double fy = 0.0;
        for (int w = 0; w < n; ++w) {
            const double2 p = psi[i ^ (uint64_t{1} << w)];
            fx += p.x;
            fy += p.y;
        }
        const double2 hv = make_double2(d * v.x - field * fx, d * v.y - field * fy);
        lam[i] = hv;
        acc += v.x * hv.x + v.y * hv.y;
    }
    block_partial(acc, partials);
}

void launch_ham_gather(const double2 *psi, double2 *lam, int n, double field, double *e_partials,
                       int grid, hipStream_t s) {
    const uint64_t size = uint64_t{1} << n;
    ham_gather_kernel<<<grid, 256, 0, s>>>(psi, lam, n, size, field, e_partials);
}

}  // namespace rie
