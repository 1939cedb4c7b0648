// Host-visible launch interface of the engine's HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rie_device.hpp"

namespace rie {

// A bit permutation of statevector indices: a lambda layout.  Segment s moves
// index bits [src[s], src[s] + len[s]) to address bits [dst[s], dst[s] + len[s]);
// bits no segment names are zero (indices are below 2^n).  count == 0: the
// natural order.
struct LamLayout {
    int count;
    int8_t src[8];
    int8_t len[8];
    int8_t dst[8];
};

// Geometry of one tile pass.  A tile holds 2^T amplitudes (T = 3R): its
// offset bits [0, a) are index bits [0, a) (contiguous in memory) and its
// offset bits [a, T) are index bits [h0, h0 + T - a).  The remaining index
// bits enumerate the tiles.  Only tile bits in [rot_lo, rot_hi) are
// processed; the others ride along for coalescing.  Either a == T (the low
// tile) or a <= R (a high tile), so that the register index k of the A and C
// mappings moves the global index by one uniform stride.
struct PassGeom {
    int n;
    int a;
    int h0;
    int rot_lo;
    int rot_hi;
    uint64_t num_tiles;
    uint64_t hi_mask;  // index bits [T, n): the ring-CNOT image of a low tile
    // Adjacent tile-bit pairs (j, j + 1), pair_lo <= j < pair_hi - 1, whose
    // transverse-field terms a fused adjoint pass adds before un-rotating.
    int pair_lo;
    int pair_hi;
    // Scattered tile (Hamiltonian pass only): when sc_n > 0, offset bits
    // [a, a + sc_n) are the index bits sc[0] < sc[1] < ... (all >= a), and h0
    // is unused.
    int sc_n;
    int sc[4];
    // Layouts of lambda in HBM for the lambda-only high passes and the
    // generating low pass (the natural order elsewhere): lay_in the layout
    // read, lay_out the layout written; see make_lam_layout.
    LamLayout lay_in;
    LamLayout lay_out;
    // Per-tile lambda addresses, tile_in[g] = lay_addr(tile_base(g), lay_in)
    // and tile_out[g] likewise (launch_lam_tile_table): the generated start's
    // phase-free high passes read them (tile_in, tile_out) and its low pass
    // stores through tile_out, so their tile loops do not evaluate a layout.
    const uint64_t *tile_in = nullptr;
    const uint64_t *tile_out = nullptr;
};

// Address of index i in layout L (a bit permutation, so lay_addr(x | y) =
// lay_addr(x) | lay_addr(y) for disjoint x, y, and strides map bit by bit).
// (Unrolled over all eight segments with constant indices, so a kernel keeps
// the layout in scalar registers instead of indexing a private copy.)
__host__ __device__ inline uint64_t lay_addr(uint64_t i, const LamLayout &L) {
    if (L.count == 0) {
        return i;
    }
    uint64_t r = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        if (s < L.count) {
            r |= ((i >> L.src[s]) & ((uint64_t{1} << L.len[s]) - 1)) << L.dst[s];
        }
    }
    return r;
}

// The layout that a high pass with tile index bits {0, 1, 2} and [h, h + 9)
// reads as one contiguous 2^12 block, written by a pass whose tile is index
// bits {0, 1, 2} and [w, w + 9) (w = 3: the low tile [0, 12)): address bits
// 0..2 and 3..11 hold the reader's tile bits in order, the writer's tile bits
// outside the reader's tile come next (ascending), then every other bit below
// n (ascending).  The writer's tile then spans 2^21 amplitudes (32 MB, 16 pages
// of 2 MB) however far apart its bits lie in the index, so its stores stay in
// few pages; with the writer's bits anywhere above, a top high group writing
// for the group below it touched 512 pages per tile and missed in the L1 TLB
// on a quarter of its translations.
inline LamLayout make_lam_layout(int n, int h, int w) {
    int order[64];
    int count = 0;
    bool placed[64] = {};
    auto put = [&](int b) {
        if (b >= 0 && b < n && !placed[b]) {
            placed[b] = true;
            order[count++] = b;
        }
    };
    for (int b = 0; b < 3; ++b) {
        put(b);
    }
    for (int b = h; b < h + 9; ++b) {
        put(b);
    }
    for (int b = w; b < w + 9; ++b) {
        put(b);
    }
    for (int b = 3; b < n; ++b) {
        put(b);
    }
    LamLayout L{};
    for (int d = 0; d < count; ++d) {
        const int b = order[d];
        if (L.count > 0 && L.src[L.count - 1] + L.len[L.count - 1] == b &&
            L.dst[L.count - 1] + L.len[L.count - 1] == d) {
            ++L.len[L.count - 1];
            continue;
        }
        if (L.count == 8) {
            return LamLayout{-1, {}, {}, {}};  // more pieces than a layout holds
        }
        L.src[L.count] = (int8_t)b;
        L.len[L.count] = 1;
        L.dst[L.count] = (int8_t)d;
        ++L.count;
    }
    if (L.count == 1) {
        L.count = 0;  // the identity: the natural order
    }
    return L;
}

// Whether layout L suits a pass whose load mapping holds the index bits
// [first, first + count) in its register index: L is a permutation of the bits
// below n and keeps those bits contiguous and in order.
inline bool lay_fits(const LamLayout &L, int n, int first, int count) {
    if (L.count == 0) {
        return true;
    }
    if (L.count < 0 || L.count > 8) {
        return false;
    }
    uint64_t image = 0;
    for (int b = 0; b < n; ++b) {
        image |= lay_addr(uint64_t{1} << b, L);
    }
    if (image != ((n >= 64) ? ~uint64_t{0} : ((uint64_t{1} << n) - 1))) {
        return false;
    }
    const uint64_t p0 = lay_addr(uint64_t{1} << first, L);
    for (int b = 1; b < count; ++b) {
        if (lay_addr(uint64_t{1} << (first + b), L) != (p0 << b)) {
            return false;
        }
    }
    return true;
}

// One independent circuit evaluated by a single workgroup (n <= 14).
struct BatchJob {
    int n;
    int layers;
    int chunk;             // wires per register chunk (1..4)
    int reserved;
    double field;
    const WireCoef *coef;  // layers * n entries
    double *energy;        // 1 value
    double *grad;          // 2 * n * layers values, or nullptr for energy only
    double *drift;         // || walked state - |0> ||^2, or nullptr
    const double *scale;   // optional chain-rule factor for grad (nullptr = 1)
};

constexpr int kBatchMaxLdsQubits = 12;  // psi and lambda in LDS
constexpr int kBatchMaxQubits = 14;     // global-scratch variant above that

// Widest rotation chunk per workgroup size: a backward group keeps 2 * 2^K
// amplitudes in registers, and the register budget per thread halves each
// time the workgroup doubles.
constexpr int batch_max_chunk(int threads) {
    return threads >= 1024 ? 2 : (threads >= 512 ? 3 : 4);
}

// Coefficient table from device parameters: coef[i] from params[2i], params[2i+1].
void launch_coef(const double *d_params, WireCoef *d_coef, int count, hipStream_t s);

// Persistent whole-circuit kernel: workgroup b runs the circuits
// d_order[d_wg_begin[b] .. d_wg_begin[b+1]) in turn (without d_wg_begin:
// circuit b of d_jobs, or `single` when d_jobs is null).  threads is 256,
// 512 or 1024; max_lds_n sizes the LDS (largest n <= 12 in the launch, 0 if
// none); circuits with n > 12 use scratch + b * scratch_stride (2 * 2^n
// amplitudes).
void launch_batch_structured(const BatchJob *d_jobs, const BatchJob &single, const int *d_order,
                             const int *d_wg_begin, int grid, int threads, int max_lds_n,
                             double2 *scratch, uint64_t scratch_stride, hipStream_t s);
uint32_t batch_lds_state_amps(int max_lds_n);

// Optimizer steps and helpers over packed arrays.
void launch_gd_update(double *p, const double *g, uint64_t count, double lr, hipStream_t s);
void launch_adam_update(double *p, const double *g, double *m, double *v, uint64_t count,
                        double lr, double b1, double b2, double eps, const int *step,
                        hipStream_t s);
void launch_record_energies(const double *energies, int count, double *history, const int *step,
                            int capacity, hipStream_t s);
void launch_step_advance(int *step, hipStream_t s);
void launch_fill(double *p, uint64_t count, double value, hipStream_t s);
void launch_set_int(int *p, int value, hipStream_t s);

// FP32 MLP kernels (learned warm start).  Epilogues of the GEMM:
constexpr int kMlpEpiNone = 0;      // C = op(A) op(B)
constexpr int kMlpEpiBias = 1;      // C = A B + bias
constexpr int kMlpEpiBiasGelu = 2;  // C = Z = A B + bias, act = GELU(Z)
constexpr int kMlpEpiDGelu = 3;     // C = (A B^T) * GELU'(zin)
// ta: A is stored K x M (use A^T); tb: B is stored N x K (use B^T).
// Supported: (ta, tb) = (0, 0) with None/Bias/BiasGelu, (0, 1) with None/DGelu, (1, 0) None.
void launch_mlp_gemm(bool ta, bool tb, int epi, const float *A, const float *B, float *C, int M,
                     int N, int K, const float *bias, float *act, const float *zin, hipStream_t s);
void launch_mlp_colsum(const float *X, int rows, int cols, float *out, hipStream_t s);
void launch_mlp_mse(const float *Y, const float *T, int count, float *dY, float *partials,
                    int blocks, float *loss_out, hipStream_t s);
void launch_mlp_adam(float *p, const float *g, float *m, float *v, int count, float lr, float b1,
                     float b2, float eps, int t, hipStream_t s);

// Layer 0 as a product state followed by the ring-CNOT permutation.
void launch_init_product_cnot(double2 *out, const WireCoef *coef_layer0, int n, hipStream_t s);

// Forward tile pass (rotations on the pass's tile wires, in place or out of
// place) or, with scatter, the low-tile pass followed by the ring-CNOT
// permutation into out.  gen_init: the input is not read but generated as
// the layer-0 product state after its ring CNOT (layer0_coef).
void launch_fwd_pass(int R, bool scatter, bool gen_init, const double2 *in, double2 *out,
                     const WireCoef *layer_coef, const WireCoef *layer0_coef, const PassGeom &g,
                     int grid, hipStream_t s);

// Layer-1 states from the parameters (chain_site in rie_device.hpp): the
// state at index i is Tr(run[i's run bits] Y(i)), where the run is the pass's
// register wires r0 .. r0 + R - 1 and Y(i) the product of the other site
// matrices around the ring from wire r0 + R.  In a pass's tile that product
// splits into tile-uniform segments and the thread's own bits:
//   forward (low tile, run at the tile's top):
//     Y = U_A . C0[bits 0..3] . C1[bits 4 .. r0 - 1]
//   adjoint high pass (tile bits [0, a) and [h0, h0 + T - a), run at the top):
//     Y = U_A . C0[bits 0 .. a - 1] . U_B . C1[bits h0 .. r0 - 1]
// with U_A the wires [r0 + R, n) and U_B the wires [a, h0), both fixed by the
// tile number.  A table set: tab[0, 2^R) the run products (register bit b is
// wire r0 + b), tab[16 + v] chunk C0 and tab[32 + v] chunk C1 at bit values
// v; utab[2g], utab[2g + 1] U_A and U_B of tile g (identity when empty).
// rot: the wires of layer 1 already rotated in the generated state.
constexpr int kChainC0 = 16;             // entry of chunk C0 (up to 4 bits)
constexpr int kChainC1 = 32;             // entry of chunk C1 (up to 6 bits)
constexpr int kChainTabEntries = 96;
constexpr int kChainMaxSets = 6;
constexpr int kChainMaxQubits = 31;
struct ChainSet {
    int r0;
    int R;
    uint64_t rot;
    uint64_t ry_only;  // rotated wires generated without their RZ (chain_site)
    int c0_first;  // chunk C0: wires [c0_first, c0_first + c0_len)
    int c0_len;
    int c1_first;  // chunk C1
    int c1_len;
    int ub_first;  // segment U_B (U_A is [r0 + R, n))
    int ub_len;
    int ta;        // the pass's tile geometry (PassGeom a and h0)
    int th0;
};
struct ChainSetup {
    int n;
    int count;
    uint64_t tiles;  // tiles per pass (every pass has 2^(n - 12))
    ChainSet set[kChainMaxSets];
};
// The X terms of H whose pre-CNOT flip P^-1 e_w leaves the low tile (bits
// [0, T)), as tile-uniform segments of table set 0 (U_A, wires [T, n)):
//   xtab[3g]     U_A at base ^ e_T            (wire T - 1: bits T - 1, T; the
//                                               register index's top bit flips)
//   xtab[3g + 1] U_A at base ^ e_{n-1}         (wire n - 1: bits n - 1, 0, 1;
//                                               the thread index flips by 3)
//   xtab[3g + 2] sum over w = T .. n - 2 of U_A at base ^ e_w ^ e_{w+1}
//                                              (both bits above the tile)
constexpr int kFlipSegments = 3;
void launch_chain_flips(const WireCoef *lc0, const WireCoef *lc1, const ChainSetup &cs,
                        CMat2 *xtab, hipStream_t s);
// Generating low pass (two-layer circuits, layer 1's backward start): per low
// tile the pre-CNOT state phi of layer 1 from chain table set 0 (tab, utab),
// lambda = P^-1 H psi (the ZZ diagonal, the in-tile pairs [g.pair_lo,
// g.pair_hi) and every flip leaving the tile, from xtab), then the adjoint
// sweep of wires 0..T-1 (gradients into partials) and lam_out = lambda
// un-rotated on them.  Reads nothing but the tables: 1 S written.
void launch_bwd_low_gen(int R, double2 *lam_out, const WireCoef *layer_coef, const PassGeom &g,
                        double *partials, int grid, hipStream_t s, double field,
                        const CMat2 *tab, const CMat2 *utab, const CMat2 *xtab);

// tabs: count * kChainTabEntries matrices, set i at tabs + i * kChainTabEntries;
// utab: count * tiles * 2 matrices, set i at utab + i * tiles * 2.
void launch_chain_tables(const WireCoef *lc0, const WireCoef *lc1, const ChainSetup &cs,
                         CMat2 *tabs, CMat2 *utab, hipStream_t s);
// Layer 1 (every wire rotated) and its ring CNOT as one write pass over the
// low tiles; tab, utab: a set with r0 = 12 - R.
void launch_gen_layer1(int R, double2 *out, const CMat2 *tab, const CMat2 *utab,
                       const PassGeom &g, int grid, hipStream_t s);

// Adjoint tile pass: un-rotates the pass's wires on (cur, lam) and writes
// per-workgroup gradient partials [grid][2T].  gather = apply the inverse
// ring-CNOT on load (low pass, out of place); write_cur / write_lam = store
// the un-rotated state / adjoint state.
// ham_pairs: before un-rotating, add the transverse-field terms of the
// tile-bit pairs [g.pair_lo, g.pair_hi) to lam (-field * X_j X_{j+1} cur).  In
// the pre-CNOT frame of the last layer these pairs are the X terms X_j of H;
// their energy is not summed here (the layer-0 contraction yields the whole
// energy once lambda = H psi is complete).
// gen_tab, gen_utab (layer 1, high pass on the lambda-only sweep): the state
// is not read from cur_in but generated from a chain table set whose run is
// mapping 0's register wires, and that set's per-tile segments.
// prefetch (high passes on the lambda-only sweep): each workgroup loads its
// next tile's lam into the registers the current tile's state frees.
void launch_bwd_pass(int R, bool gather, bool write_cur, bool write_lam, const double2 *cur_in,
                     const double2 *lam_in, double2 *cur_out, double2 *lam_out,
                     const WireCoef *layer_coef, const PassGeom &g, double *partials, int grid,
                     hipStream_t s, bool ham_pairs = false, double field = 0.0,
                     const CMat2 *gen_tab = nullptr, const CMat2 *gen_utab = nullptr,
                     bool prefetch = false, bool phase_free = false);

// Hamiltonian tile pass: adds the X terms of the pass's wires (and the ZZ
// diagonal if diag) of H psi, optionally into lam, and writes per-workgroup
// energy partials sum_i Re(conj(psi_i) * contribution_i).
void launch_ham_pass(int R, bool diag, bool read_lam, bool write_lam, const double2 *psi,
                     const double2 *lam_in, double2 *lam_out, const PassGeom &g, double field,
                     double *e_partials, int grid, hipStream_t s);
// The same with a scattered tile (g.sc_n > 0): the ZZ diagonal and the X terms
// of all 12 tile wires, written into lam; no energy partials (its callers take
// the energy from the layer-0 contraction).
void launch_ham_pass_scattered(int R, const double2 *psi, double2 *lam_out, const PassGeom &g,
                               double field, int grid, hipStream_t s);

// Layer-0 gradients by contracting the adjoint state with the layer-0 product
// state: one read pass over lambda (gathered through the inverse ring CNOT
// per low tile, R = 4), a reduction of the per-tile sums over the high wires,
// and the final per-wire assembly into grad[0 .. 2n).
// R: register shape of the contraction (4: 256 threads x 16, 3: 512 x 8).
void launch_layer0_contract(int R, const double2 *lam, const WireCoef *lc0, const PassGeom &g,
                            double2 *tile_sums, double *partials, int grid, hipStream_t s);
void launch_layer0_high_env(const double2 *tile_sums, uint64_t num_tiles, const WireCoef *lc0,
                            int n, double *partials, int chunks, hipStream_t s);
// energy_out (optional): Re <chi|mu>, the energy when lambda = H psi was carried
// back complete (fused Hamiltonian: the pair terms are all added by then).
// Waves of a 256-thread workgroup (register shape 4): the fused top pass
// writes one sum per wave and tile.
constexpr uint64_t Tile4Waves = 4;
// Layer-0 values per workgroup of the fused top pass (see launch_bwd_top_l0).
constexpr int kTopL0Cols = 52;
// The last high pass of a two-layer evaluation whose high passes run from the
// lowest group up (the top group's tiles, h0 = n - 9; register shape 4): the
// lambda-only sweep of its wires, then, instead of storing lambda, the layer-0
// contraction of its tiles (per-block partials: 2T layer-1 gradients in
// partials, kTopL0Cols = 52 layer-0 values in l0_partials; per tile 4 wave
// sums in tile_sums; phi_c and lam_tile: one value per tile, written first).
// out[g] = lay_addr(tile_base(g, g_, T), L) for every tile of pass geometry g_.
void launch_lam_tile_table(const PassGeom &g, const LamLayout &L, uint64_t *out, hipStream_t s);

void launch_bwd_top_l0(const double2 *lam_in, const WireCoef *layer_coef, const PassGeom &g,
                       double *partials, int grid, hipStream_t s, const CMat2 *gen_tab,
                       const CMat2 *gen_utab, const WireCoef *lc0, double2 *phi_c,
                       uint64_t *lam_tile, double2 *tile_sums, double *l0_partials);
// The fixed wires' environments from tile_sums and the layer-0 gradients (and
// energy) from the fused pass's partials.
void launch_layer0_top_finish(const double2 *tile_sums, uint64_t num_tiles, const double *blk_partials,
                              int blocks, const WireCoef *lc0, int n, double *fixed_partials,
                              int chunks, double *grad, const double *scale, hipStream_t s,
                              double *energy_out);
void launch_layer0_finish(const double *low_partials, int low_blocks, const double *high_partials,
                          int high_chunks, const WireCoef *lc0, int n, double *grad,
                          const double *scale, hipStream_t s, double *energy_out = nullptr);

// grad[(layer*n + wire)*2 + comp] = (sum over blocks) * scale for the pass's wires.
void launch_reduce_grad(const double *partials, int num_blocks, int T, const PassGeom &g,
                        int layer, double *grad, const double *scale, hipStream_t s);

// *out = sum of count values (fixed order).
void launch_reduce_sum(const double *values, int count, double *out, hipStream_t s);

// Squared distance of a state from |0...0>.
void launch_drift(const double2 *state, uint64_t size, double *partials, int grid, hipStream_t s);

// Gate-level kernels (PennyLane gate order).
void launch_zero_state(double2 *state, uint64_t size, hipStream_t s);
void launch_gate_ry(double2 *state, int n, int wire, const WireCoef *k, bool adjoint, hipStream_t s);
void launch_gate_rz(double2 *state, int n, int wire, const WireCoef *k, bool adjoint, hipStream_t s);
void launch_gate_cnot(double2 *state, int n, int control, int target, hipStream_t s);
void launch_ring_permute(const double2 *in, double2 *out, int n, bool inverse, hipStream_t s);
// Adjoint step of RY / RZ: psi, lam <- G^dagger (psi, lam) and per-block partials
// of 2 Re<lam_after | dG psi_before>.
void launch_gate_ry_step(double2 *psi, double2 *lam, int n, int wire, const WireCoef *k,
                         double *partials, int grid, hipStream_t s);
void launch_gate_rz_step(double2 *psi, double2 *lam, int n, int wire, const WireCoef *k,
                         double *partials, int grid, hipStream_t s);
// lam = H psi by gathering all n flips per amplitude, with energy partials.
void launch_ham_gather(const double2 *psi, double2 *lam, int n, double field, double *e_partials,
                       int grid, hipStream_t s);
// out[idx] = sum of partials[0..count) * scale (one value).
void launch_reduce_to(const double *partials, int count, double *out, const double *scale,
                      hipStream_t s);

int elementwise_grid(uint64_t work_items);

}  // namespace rie
