// Host orchestration and C ABI of the ring-Ising statevector engine.
//
// One engine owns the device buffers of one (n, L, h, strategy) problem
// (reference: RingIsingCudaBackend::Impl, cpp/ising_cuda_backend.cu:333-401)
// and issues the HIP kernels of rie_kernels.hip on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rie.h"
#include "rie_engine_internal.hpp"

#include <cstdlib>
#include "rie_kernels.hpp"

namespace rie {

thread_local std::string g_last_error;

Error::Error(rie_status st, const std::string &msg) : std::runtime_error(msg), status(st) {}

void fail(rie_status st, const std::string &msg) { throw Error(st, msg); }

void check(hipError_t e, const char *ctx) {
    if (e == hipSuccess) {
        return;
    }
    const rie_status st =
        (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? RIE_OUT_OF_MEMORY
                                                                     : RIE_DEVICE_ERROR;
    fail(st, std::string(ctx) + ": " + hipGetErrorString(e));
}

void check_launch(const char *ctx) { check(hipGetLastError(), ctx); }

void *device_alloc(size_t bytes, const char *what) {
    void *p = nullptr;
    if (bytes == 0) {
        return nullptr;
    }
    check(hipMalloc(&p, bytes), what);
    return p;
}

void device_free(void *p) {
    if (p != nullptr) {
        (void)hipFree(p);
    }
}

int parse_strategy(const std::string &name) {
    // reference: cpp/ising_cuda_backend.cu:249-264
    if (name == "inverse_walk_cuQuantum" || name == "inverse_walk_cuquantum") {
        return RIE_STRATEGY_INVERSE_WALK_GATE;
    }
    if (name == "structured_adjoint") {
        return RIE_STRATEGY_STRUCTURED_ADJOINT;
    }
    if (name == "dense_scan") {
        return RIE_STRATEGY_DENSE_SCAN;
    }
    fail(RIE_INVALID_ARGUMENT,
         "Unknown gradient strategy. Expected one of: inverse_walk_cuQuantum, "
         "structured_adjoint, dense_scan.");
}

// ---------------------------------------------------------------------------
// Plan: the tile passes of one rotation layer.
// ---------------------------------------------------------------------------

static int pass_grid(uint64_t tiles, int max_grid) {
    return (int)std::min<uint64_t>(tiles, (uint64_t)max_grid);
}

void Engine::build_plan() {
    passes.clear();
    if (n <= kBatchMaxLdsQubits) {
        small = true;
        return;
    }
    small = false;
    T = 12;  // tile bits of every pass (both register shapes)
    const uint64_t tiles = uint64_t{1} << (n - T);
    const uint64_t full = (n >= 64) ? ~uint64_t{0} : ((uint64_t{1} << n) - 1);
    const uint64_t tile_mask = (uint64_t{1} << T) - 1;
    PassGeom low{};
    low.n = n;
    low.a = T;
    low.h0 = T;
    low.rot_lo = 0;
    low.rot_hi = T;
    low.num_tiles = tiles;
    low.hi_mask = full & ~tile_mask;
    passes.push_back({low, pass_grid(tiles, max_grid), 0, T});
    // High wires [T, n) in groups of at most T - a (a = 3 contiguous low bits).
    const int a = 3;
    const int width = T - a;
    const int rest = n - T;
    const int groups = (rest + width - 1) / width;
    int g0 = T;
    for (int gi = 0; gi < groups; ++gi) {
        const int bg = rest / groups + (gi < rest % groups ? 1 : 0);
        PassGeom hg{};
        hg.n = n;
        hg.a = a;
        hg.h0 = std::min(g0, n - width);
        hg.rot_lo = a + (g0 - hg.h0);
        hg.rot_hi = hg.rot_lo + bg;
        hg.num_tiles = tiles;
        hg.hi_mask = 0;
        passes.push_back({hg, pass_grid(tiles, max_grid), g0, g0 + bg});
        g0 += bg;
    }
    // Fused Hamiltonian.  In the pre-CNOT frame of the last layer, X_w of H is
    // X_w X_{w+1} for w < n - 1 (X_{n-1} touches wires 0, 1 and n - 1).  The
    // pairs inside the low tile above wire c0 and inside each high group are
    // added by the adjoint passes; the crossing wires (11, the last wire below
    // each further high group, n - 1) and wires [0, c0) go to one psi-frame
    // pass over a scattered tile together with the ZZ diagonal.
    std::vector<int> sc{T - 1};
    for (size_t p = 2; p < passes.size(); ++p) {
        sc.push_back(passes[p].first_wire - 1);
    }
    sc.push_back(n - 1);
    std::sort(sc.begin(), sc.end());
    sc.erase(std::unique(sc.begin(), sc.end()), sc.end());
    // at most four crossing wires: the low pass's pairs then start at tile
    // bit T - 4 or above, which its kernel relies on
    fuse_ready = sc.size() <= 4 && n > T;
    if (fuse_ready) {
        PassGeom hf{};
        hf.n = n;
        hf.sc_n = (int)sc.size();
        hf.a = T - hf.sc_n;
        hf.h0 = 0;
        for (int i = 0; i < hf.sc_n; ++i) {
            hf.sc[i] = sc[(size_t)i];
        }
        hf.rot_lo = 0;
        hf.rot_hi = T;
        hf.num_tiles = tiles;
        hf.hi_mask = 0;
        ham_fused_g = hf;
        ham_fused_grid = pass_grid(tiles, max_grid);
        passes[0].g.pair_lo = hf.a;
        passes[0].g.pair_hi = T;
        for (size_t p = 1; p < passes.size(); ++p) {
            passes[p].g.pair_lo = passes[p].g.rot_lo;
            passes[p].g.pair_hi = passes[p].g.rot_hi;
        }
    }
}

Engine::Engine(size_t num_qubits, size_t num_layers, double field, const std::string &strat,
               size_t chunk_width, bool double_buffer_)
    : n((int)num_qubits), L((int)num_layers), h(field), strategy(parse_strategy(strat)),
      chunk_width(chunk_width), double_buffer(double_buffer_) {
    // reference: validate_and_get_state_size (cpp/ising_cuda_backend.cu:229-241),
    // Impl ctor width checks (:388-397), dense_scan limit (dense_modes.inc:353-355).
    if (num_qubits < 2) {
        fail(RIE_INVALID_ARGUMENT, "num_qubits must be at least 2.");
    }
    if (num_qubits > 34) {
        fail(RIE_INVALID_ARGUMENT, "num_qubits must be at most 34 on one device.");
    }
    if (num_layers < 1) {
        fail(RIE_INVALID_ARGUMENT, "layers must be at least 1.");
    }
    if (strategy == RIE_STRATEGY_STRUCTURED_ADJOINT) {
        if (chunk_width == 0) {
            fail(RIE_INVALID_ARGUMENT, "structured_rotation_chunk_width must be at least 1.");
        }
        if (chunk_width > 8) {
            fail(RIE_INVALID_ARGUMENT,
                 "structured_rotation_chunk_width exceeds supported maximum.");
        }
    }
    if (strategy == RIE_STRATEGY_DENSE_SCAN && num_qubits > 8) {
        fail(RIE_INVALID_ARGUMENT, "dense_scan requires num_qubits <= 8.");
    }
    N = uint64_t{1} << n;
    num_params = (size_t)2 * n * L;
    check(hipGetDevice(&device), "hipGetDevice");
    hipDeviceProp_t prop;
    check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    num_cus = prop.multiProcessorCount;
    // sixteen workgroups per CU (eight rounds of the two resident ones): the
    // tile passes balance better across CUs than with fewer, larger
    // workgroups (measured at 30 qubits: 2, 4, 8, 16, 32 per CU gave 28.2,
    // 27.6, 26.8, 26.4, 26.5 ms per evaluation), and the partial sums the
    // reductions read grow with the count
    max_grid = std::max(64, num_cus * 16);
    check(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking), "hipStreamCreate");
    check(hipEventCreate(&ev_a), "hipEventCreate");
    check(hipEventCreate(&ev_b), "hipEventCreate");
    check(hipEventCreateWithFlags(&ev_last, hipEventDisableTiming), "hipEventCreate");
    build_plan();

    d_params = (double *)device_alloc(sizeof(double) * num_params, "alloc params");
    d_grad = (double *)device_alloc(sizeof(double) * num_params, "alloc gradient");
    d_energy = (double *)device_alloc(sizeof(double) * 4, "alloc energy");
    d_coef = (WireCoef *)device_alloc(sizeof(WireCoef) * (size_t)n * L, "alloc coefficients");

    if (strategy == RIE_STRATEGY_STRUCTURED_ADJOINT && small) {
        // the whole circuit lives in one workgroup's LDS
    } else if (strategy == RIE_STRATEGY_STRUCTURED_ADJOINT) {
        // The stored schedule keeps every forward pass output; use it when it
        // fits comfortably in free device memory, else walk back by inverse gates.
        size_t free_b = 0;
        size_t total_b = 0;
        check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const double need = (double)stored_buffer_count() * (double)sizeof(double2) * (double)N;
        schedule = (L > 1 && need <= 0.85 * (double)free_b) ? RIE_SCHEDULE_STORED
                                                            : RIE_SCHEDULE_INVERSE_WALK;
        ensure_pool(schedule == RIE_SCHEDULE_STORED ? stored_buffer_count() : 4);
        size_t part = 0;
        for (const auto &p : passes) {
            part = std::max<size_t>(part, (size_t)p.grid * 2 * T);
        }
        size_t eparts = 0;
        for (const auto &p : passes) {
            eparts += (size_t)p.grid;
        }
        // the layer-0 contraction writes 48 values per workgroup of the low pass;
        // inside the top group's pass, 2T + kTopL0Cols per workgroup of that pass
        partial_count = std::max({part, eparts, (size_t)passes[0].grid * 48,
                                  (size_t)passes.back().grid * (2 * T + kTopL0Cols)});
        d_partials = (double *)device_alloc(sizeof(double) * partial_count, "alloc partials");
        d_epartials = (double *)device_alloc(sizeof(double) * eparts, "alloc energy partials");
        epartial_count = eparts;
        const uint64_t tiles = passes[0].g.num_tiles;
        // per-tile sums: one per tile for the low-tile contraction; for the
        // contraction inside the top group's pass one per wave of a tile plus
        // the tile's fixed-wire factor and lambda address (16 B each)
        tile_sum_count = L == 2 ? tiles * (Tile4Waves + 2) : tiles;
        d_tile_sums = (double2 *)device_alloc(sizeof(double2) * tile_sum_count, "alloc tile sums");
        if (L == 2) {
            // the generated start's per-tile lambda addresses (read and write)
            d_lam_tab = (uint64_t *)device_alloc(sizeof(uint64_t) * 2 * tiles, "alloc tile table");
        }
        d_chain = (CMat2 *)device_alloc(sizeof(CMat2) * kChainMaxSets * kChainTabEntries,
                                        "alloc chain tables");
        if (L >= 2 && n <= kChainMaxQubits && passes.size() <= (size_t)kChainMaxSets) {
            // per-tile uniform segments of each table set (128 B per tile and set)
            d_chain_u = (CMat2 *)device_alloc(sizeof(CMat2) * 2 * tiles * passes.size(),
                                              "alloc chain segments");
            if (L == 2 && fuse_ready) {
                // the out-of-tile flips of the generated start
                d_chain_x = (CMat2 *)device_alloc(sizeof(CMat2) * kFlipSegments * tiles,
                                                  "alloc chain flips");
            }
        }
        high_env_chunks = (int)std::min<uint64_t>(64, (tiles + 255) / 256);
        partial2_count = 4 * (size_t)high_env_chunks * (size_t)std::max(1, n - T);
        d_partials2 = (double *)device_alloc(sizeof(double) * partial2_count, "alloc partials");
    } else {
        // gate-level and dense strategies: psi, lambda
        ensure_pool(2);
        partial_count = (size_t)elementwise_grid(N);
        d_partials = (double *)device_alloc(sizeof(double) * partial_count, "alloc partials");
        d_epartials = (double *)device_alloc(sizeof(double) * partial_count, "alloc energy partials");
        epartial_count = partial_count;
    }
    if (strategy == RIE_STRATEGY_DENSE_SCAN) {
        dense_init();
    }
}

Engine::~Engine() {
    if (own_stream != nullptr) {
        (void)hipStreamSynchronize(own_stream);
    }
    for (auto &c : ckpts) {
        device_free(c);
    }
    for (auto p : pool) {
        device_free(p);
    }
    device_free(d_params);
    device_free(d_grad);
    device_free(d_energy);
    device_free(d_coef);
    device_free(d_partials);
    device_free(d_epartials);
    device_free(d_drift);
    device_free(d_tile_sums);
    device_free(d_lam_tab);
    device_free(d_chain);
    device_free(d_chain_u);
    device_free(d_chain_x);
    device_free(d_partials2);
    dense_free();
    for (auto &p : kpending) {
        kevent_pool.push_back(p.a);
        kevent_pool.push_back(p.b);
    }
    for (auto ev : kevent_pool) {
        (void)hipEventDestroy(ev);
    }
    if (ev_a != nullptr) {
        (void)hipEventDestroy(ev_a);
    }
    if (ev_b != nullptr) {
        (void)hipEventDestroy(ev_b);
    }
    if (ev_last != nullptr) {
        (void)hipEventDestroy(ev_last);
    }
    if (own_stream != nullptr) {
        (void)hipStreamDestroy(own_stream);
    }
}

// ---------------------------------------------------------------------------
// Stage timing (reference stage names: structured_adjoint_modes.inc:317-370)
// ---------------------------------------------------------------------------

void Engine::stage_begin(hipStream_t s) {
    if (profiling) {
        check(hipEventRecord(ev_a, s), "hipEventRecord");
    }
}

void Engine::stage_end(hipStream_t s, const char *name) {
    if (!profiling) {
        return;
    }
    check(hipEventRecord(ev_b, s), "hipEventRecord");
    check(hipEventSynchronize(ev_b), "hipEventSynchronize");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, ev_a, ev_b), "hipEventElapsedTime");
    timings.emplace_back(name, (double)ms);
}

static bool is_capturing(hipStream_t s) {
    hipStreamCaptureStatus status = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &status) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return status != hipStreamCaptureStatusNone;
}

// Successive calls on one engine may use different streams (its own for the
// host-pointer API, the caller's for the device API); they share buffers, so
// each call waits for the previous one.  Skipped inside stream capture.
void Engine::order_begin(hipStream_t s) {
    if (ev_last_valid && !is_capturing(s)) {
        check(hipStreamWaitEvent(s, ev_last, 0), "hipStreamWaitEvent");
    }
}

void Engine::order_end(hipStream_t s) {
    if (is_capturing(s)) {
        return;
    }
    check(hipEventRecord(ev_last, s), "hipEventRecord");
    ev_last_valid = true;
}

hipEvent_t Engine::kt_event() {
    if (!kevent_pool.empty()) {
        hipEvent_t e = kevent_pool.back();
        kevent_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    check(hipEventCreate(&e), "hipEventCreate");
    return e;
}

void Engine::kt_begin(hipStream_t s) {
    if (!ktime) {
        return;
    }
    kt_open = kt_event();
    check(hipEventRecord(kt_open, s), "hipEventRecord");
}

void Engine::kt_end(hipStream_t s, const char *name, double bytes) {
    if (!ktime) {
        return;
    }
    hipEvent_t b = kt_event();
    check(hipEventRecord(b, s), "hipEventRecord");
    kpending.push_back({name, bytes, kt_open, b});
    kt_open = nullptr;
}

void Engine::kt_flush() {
    for (auto &p : kpending) {
        check(hipEventSynchronize(p.b), "hipEventSynchronize");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
        KernelStat *st_ = nullptr;
        for (auto &k : kstats) {
            if (k.name == p.name) {
                st_ = &k;
            }
        }
        if (st_ == nullptr) {
            kstats.push_back({p.name, 0, 0.0, 0.0});
            st_ = &kstats.back();
        }
        st_->launches += 1;
        st_->total_ms += (double)ms;
        st_->total_bytes += p.bytes;
        kevent_pool.push_back(p.a);
        kevent_pool.push_back(p.b);
    }
    kpending.clear();
}

// ---------------------------------------------------------------------------
// Structured adjoint: tile-pass path (n > 12)
// ---------------------------------------------------------------------------

void Engine::structured_forward(hipStream_t s, bool save_ckpts) {
    const double S = (double)sizeof(double2) * (double)N;
    const int K = (int)passes.size() - 1;  // high passes per layer (>= 1 for n > 12)
    if (L == 1) {
        kt_begin(s);
        launch_init_product_cnot(pool[0], d_coef, n, s);
        check_launch("init_product_cnot");
        kt_end(s, "init_product_cnot", S);
        final_cur = 0;
        final_ptr = pool[0];
        return;
    }
    gen_fwd = gen_ready();
    if (gen_fwd) {
        build_chain_tables(s);
    }
    if (schedule == RIE_SCHEDULE_STORED) {
        // Every pass writes its own buffer; layer 1's first pass generates the
        // layer-0 state instead of reading it (or, with layer-1 generation,
        // layer 1 is one pass that writes only its output).
        const double2 *in = nullptr;
        for (int l = 1; l < L; ++l) {
            const WireCoef *lc = d_coef + (size_t)l * n;
            if (l == 1 && gen_fwd) {
                kt_begin(s);
                launch_gen_layer1(pass_r[RIE_PASS_FWD_GEN], fbuf(1, K), d_chain, d_chain_u,
                                  passes[0].g, passes[0].grid, s);
                check_launch("gen_layer1");
                kt_end(s, "fwd_layer1_gen", S);
                in = fbuf(1, K);
                continue;
            }
            for (int j = 0; j < K; ++j) {
                const bool gen = (l == 1 && j == 0);
                kt_begin(s);
                launch_fwd_pass(pass_r[gen ? RIE_PASS_FWD_GEN : RIE_PASS_FWD_HIGH], false, gen, in,
                                fbuf(l, j), lc, d_coef, passes[1 + j].g,
                                passes[1 + j].grid, s);
                check_launch("fwd_pass");
                kt_end(s, gen ? "fwd_pass_high_gen_init" : "fwd_pass_high", gen ? S : 2 * S);
                in = fbuf(l, j);
            }
            kt_begin(s);
            launch_fwd_pass(pass_r[RIE_PASS_FWD_LOW], true, false, in, fbuf(l, K), lc, d_coef,
                            passes[0].g,
                            passes[0].grid, s);
            check_launch("fwd_pass scatter");
            kt_end(s, "fwd_pass_low_cnot", 2 * S);
            in = fbuf(l, K);
        }
        final_cur = -1;
        final_ptr = fbuf(L - 1, K);
        return;
    }
    // Inverse-walk schedule: working buffers pool[0] and pool[1] (in place
    // where possible); checkpointed boundary states are written straight into
    // their checkpoint buffer by the low (scatter) pass and read from there
    // by the next layer's first pass.
    const double2 *cur = nullptr;
    int cur_buf = -1;  // -1: generated (layer 1, first pass) or a checkpoint buffer
    for (int l = 1; l < L; ++l) {
        const WireCoef *lc = d_coef + (size_t)l * n;
        const bool gen_layer = (l == 1 && gen_fwd);
        for (int j = 0; j < (gen_layer ? 0 : K); ++j) {
            const bool gen = (l == 1 && j == 0);
            double2 *out = (cur_buf >= 0) ? pool[cur_buf] : pool[0];
            kt_begin(s);
            launch_fwd_pass(pass_r[gen ? RIE_PASS_FWD_GEN : RIE_PASS_FWD_HIGH], false, gen, cur,
                            out, lc, d_coef, passes[1 + j].g,
                            passes[1 + j].grid, s);
            check_launch("fwd_pass");
            kt_end(s, gen ? "fwd_pass_high_gen_init" : "fwd_pass_high", gen ? S : 2 * S);
            if (cur_buf < 0) {
                cur_buf = 0;
            }
            cur = out;
        }
        const int boundary = l + 1;  // state after layer l
        const bool ckpt_due = save_ckpts && ckpt_every > 0 && boundary % (int)ckpt_every == 0 &&
                              boundary < L;
        double2 *out;
        int out_buf;
        if (ckpt_due) {
            out = ckpts[(size_t)boundary / ckpt_every - 1];
            out_buf = -1;
        } else {
            out_buf = (cur_buf == 0) ? 1 : 0;
            out = pool[out_buf];
        }
        kt_begin(s);
        if (gen_layer) {
            launch_gen_layer1(pass_r[RIE_PASS_FWD_GEN], out, d_chain, d_chain_u, passes[0].g,
                              passes[0].grid, s);
            check_launch("gen_layer1");
            kt_end(s, "fwd_layer1_gen", S);
        } else {
            launch_fwd_pass(pass_r[RIE_PASS_FWD_LOW], true, false, cur, out, lc, d_coef,
                            passes[0].g, passes[0].grid, s);
            check_launch("fwd_pass scatter");
            kt_end(s, "fwd_pass_low_cnot", 2 * S);
        }
        cur = out;
        cur_buf = out_buf;
    }
    final_cur = cur_buf;  // the last boundary (L) is never a checkpoint
    final_ptr = pool[final_cur];
}

const double2 *Engine::final_state() const { return final_ptr; }

size_t Engine::stored_buffer_count() const {
    const size_t per_layer = passes.empty() ? 1 : passes.size();
    return (size_t)std::max(0, L - 1) * per_layer + 2;
}

double2 *Engine::fbuf(int l, int j) const {
    return pool[(size_t)(l - 1) * passes.size() + (size_t)j];
}

double2 *Engine::lbuf(int i) const {
    return pool[(size_t)(L - 1) * passes.size() + (size_t)i];
}

void Engine::ensure_pool(size_t count) {
    while (pool.size() < count) {
        pool.push_back((double2 *)device_alloc(sizeof(double2) * N, "alloc statevector"));
    }
}

void Engine::set_schedule(int sched) {
    if (sched != RIE_SCHEDULE_INVERSE_WALK && sched != RIE_SCHEDULE_STORED) {
        fail(RIE_INVALID_ARGUMENT, "unknown schedule");
    }
    if (strategy != RIE_STRATEGY_STRUCTURED_ADJOINT || small) {
        return;  // the whole-circuit kernel and other strategies have one schedule
    }
    if (sched == RIE_SCHEDULE_STORED && L > 1) {
        ensure_pool(stored_buffer_count());
    }
    schedule = sched;
    resident = false;
}

void Engine::require_partials(size_t count) const {
    if (count > partial_count) {
        fail(RIE_DEVICE_ERROR, "internal: gradient partial buffer too small");
    }
}

int Engine::free_buffer(std::initializer_list<int> used) const {
    for (int i = 0; i < 4; ++i) {
        bool taken = false;
        for (int u : used) {
            taken = taken || (u == i);
        }
        if (!taken) {
            return i;
        }
    }
    fail(RIE_DEVICE_ERROR, "internal: no free state buffer");
}

void Engine::structured_hamiltonian(hipStream_t s, bool write_lam) {
    const double2 *psi = final_state();
    double2 *lam = nullptr;
    if (write_lam) {
        if (schedule == RIE_SCHEDULE_STORED && L > 1) {
            lam_idx = -1;
            lam = lbuf(0);
        } else {
            lam_idx = free_buffer({final_cur});
            lam = pool[lam_idx];
        }
    }
    size_t off = 0;
    const double S = (double)sizeof(double2) * (double)N;
    size_t total = 0;
    for (const auto &p : passes) {
        total += (size_t)p.grid;
    }
    if (total > epartial_count) {
        fail(RIE_DEVICE_ERROR, "internal: energy partial buffer too small");
    }
    for (size_t p = 0; p < passes.size(); ++p) {
        const bool read_lam = write_lam && p > 0;
        kt_begin(s);
        launch_ham_pass(pass_r[RIE_PASS_HAM], p == 0, read_lam, write_lam, psi, lam, lam,
                        passes[p].g, h,
                        d_epartials + off, passes[p].grid, s);
        check_launch("ham_pass");
        kt_end(s, p == 0 ? "ham_pass_low" : "ham_pass_high",
               S * (1.0 + (write_lam ? 1.0 : 0.0) + (read_lam ? 1.0 : 0.0)));
        off += (size_t)passes[p].grid;
    }
    energy_partials = off;
}

void Engine::launch_bwd(int l, bool gather, bool write_cur, const double2 *cur_in,
                        const double2 *lam_in, double2 *cur_out, double2 *lam_out, size_t p,
                        hipStream_t s, int gen_set, const PassGeom *geom, bool phase_free) {
    const WireCoef *lc = d_coef + (size_t)l * n;
    const PassPlan &pp = passes[p];
    const PassGeom &pg = geom != nullptr ? *geom : pp.g;
    const CMat2 *gen_tab = nullptr;
    const CMat2 *gen_utab = nullptr;
    if (gen_set >= 0) {
        if (l != 1 || write_cur || gather || gen_set >= (int)passes.size()) {
            fail(RIE_DEVICE_ERROR, "internal: generated state outside layer 1's lambda sweep");
        }
        gen_tab = d_chain + (size_t)gen_set * kChainTabEntries;
        gen_utab = d_chain_u + (size_t)gen_set * 2 * passes[0].g.num_tiles;
    }
    const bool pairs = l == fuse_layer;
    // double buffering: the lambda-only passes load each workgroup's next tile
    // (lam of a high pass, the state of the gather pass) behind the current
    // one; the walk's one-sweep passes do not
    const bool prefetch = double_buffer && !write_cur;
    launch_bwd_pass(pass_r[gather ? RIE_PASS_BWD_LOW : RIE_PASS_BWD_HIGH], gather, write_cur, true,
                    cur_in, lam_in, cur_out, lam_out, lc, pg, d_partials, pp.grid, s, pairs,
                    pairs ? h : 0.0, gen_tab, gen_utab, prefetch, phase_free);
}

// The generated start's lambda layouts.  The high passes run either from the
// lowest group up (bottom_up: then the last one holds the top group's tiles and
// can contract layer 0 itself) or from the top group down.  The low pass
// writes the layout of the first high pass, each high pass reads its own and
// writes the next one's, the last writes the natural order for the layer-0
// contraction (or nothing).  Each layout puts the reader's tile in one block
// and the writer's tile bits right above it (make_lam_layout).
LamLayout Engine::lam_layout_in(int j, bool bottom_up) const {
    const int K = (int)passes.size() - 1;
    const int prev = bottom_up ? j - 1 : j + 1;  // the pass that runs before j
    const int w = (prev < 0 || prev >= K) ? 3 : passes[1 + (size_t)prev].g.h0;
    return make_lam_layout(n, passes[1 + (size_t)j].g.h0, w);
}

LamLayout Engine::lam_layout_out(int j, bool bottom_up) const {
    const int K = (int)passes.size() - 1;
    const int next = bottom_up ? j + 1 : j - 1;
    return (next >= 0 && next < K) ? lam_layout_in(next, bottom_up) : LamLayout{};
}

bool Engine::lam_layouts_fit(bool bottom_up) const {
    const int K = (int)passes.size() - 1;
    if (K < 1) {
        return false;
    }
    const int r_low = pass_r[RIE_PASS_BWD_LOW];
    const int r_high = pass_r[RIE_PASS_BWD_HIGH];
    if (!lay_fits(lam_layout_in(bottom_up ? 0 : K - 1, bottom_up), n, T - r_low, r_low)) {
        return false;
    }
    for (int j = 0; j < K; ++j) {
        const PassGeom &g = passes[1 + (size_t)j].g;
        if (g.a != 3) {
            return false;
        }
        const int first = g.h0 + (T - r_high - g.a);
        if (!lay_fits(lam_layout_in(j, bottom_up), n, first, r_high) ||
            !lay_fits(lam_layout_out(j, bottom_up), n, first, r_high)) {
            return false;
        }
    }
    return true;
}

// The generated start runs its high passes from the lowest group up (and
// contracts layer 0 in the last) where that is possible and does not cost the
// layouts that the other order would have (at 27-29 and 31-34 qubits the top
// group's tile overlaps the one below it, and bottom-up the lower pass's
// register bits would straddle the layout it writes); otherwise from the top
// group down with the separate contraction.
bool Engine::gen_bottom_up() const {
    if (!l0_in_pass_ready()) {
        return false;
    }
    if (!lam_layouts_enabled || pool.size() < 2) {
        return true;
    }
    return lam_layouts_fit(true) || !lam_layouts_fit(false);
}

bool Engine::gen_layouts() const {
    return lam_layouts_enabled && pool.size() >= 2 && lam_layouts_fit(gen_bottom_up());
}

bool Engine::gen_ready() const {
    return gen_enabled && strategy == RIE_STRATEGY_STRUCTURED_ADJOINT && !small && L >= 2 &&
           n <= kChainMaxQubits && (int)passes.size() <= kChainMaxSets && d_chain_u != nullptr;
}

// Table sets of the current register shapes: 0 the forward (low tile, all of
// layer 1 rotated, run at mapping 0), 1 + j the adjoint high pass j, run at
// mapping 0's registers, with the wires still rotated when it runs: high
// groups 0 .. j when the high passes run from the top group down (the stored
// schedule and the walk), groups j .. K-1 when they run from the lowest group
// up (the generated start, bottom_up).
void Engine::build_chain_tables(hipStream_t s, int r_set0, bool bottom_up) {
    ChainSetup cs{};
    cs.n = n;
    cs.tiles = passes[0].g.num_tiles;
    const uint64_t all = (n >= 64) ? ~uint64_t{0} : ((uint64_t{1} << n) - 1);
    // set 0's register run: the forward generator's shape, or the caller's
    // (the generated start's low pass)
    const int rf = r_set0 > 0 ? r_set0 : pass_r[RIE_PASS_FWD_GEN];
    const int rh = pass_r[RIE_PASS_BWD_HIGH];
    // the forward's low tile: run at the tile's top, C0 = wires 0..3, C1 the
    // rest below the run, no U_B
    {
        const PassGeom &g = passes[0].g;
        ChainSet st{};
        st.r0 = T - rf;
        st.R = rf;
        st.rot = all;
        st.c0_first = 0;
        st.c0_len = 4;
        st.c1_first = 4;
        st.c1_len = T - rf - 4;
        st.ub_first = 0;
        st.ub_len = 0;
        st.ta = g.a;
        st.th0 = g.h0;
        cs.set[0] = st;
    }
    cs.count = 1;
    // adjoint high pass j: run at the top of the tile's high group, C0 = the
    // tile's low wires [0, a), U_B = [a, h0), C1 = [h0, r0)
    for (size_t p = 1; p < passes.size(); ++p) {
        const PassGeom &g = passes[p].g;
        ChainSet st{};
        st.r0 = g.h0 + (T - rh - g.a);
        st.R = rh;
        st.rot = bottom_up ? all & ~((uint64_t{1} << passes[p].first_wire) - 1)
                           : all & ~((uint64_t{1} << T) - 1) &
                                 ((uint64_t{1} << passes[p].end_wire) - 1);
        // from the lowest group up a pass takes its own wires' phases off
        // lambda once per tile, so its state is generated without them
        st.ry_only = bottom_up ? st.rot & ((uint64_t{1} << passes[p].end_wire) - 1) : 0;
        st.c0_first = 0;
        st.c0_len = g.a;
        st.c1_first = g.h0;
        st.c1_len = T - rh - g.a;
        st.ub_first = g.a;
        st.ub_len = g.h0 - g.a;
        st.ta = g.a;
        st.th0 = g.h0;
        cs.set[cs.count++] = st;
    }
    for (int i = 0; i < cs.count; ++i) {
        const ChainSet &st = cs.set[i];
        if (st.c0_len < 0 || st.c0_len > 4 || st.c1_len < 0 || st.c1_len > 6 || st.ub_len < 0 ||
            st.r0 + st.R > n || passes[i].g.num_tiles != cs.tiles) {
            fail(RIE_DEVICE_ERROR, "internal: chain tables do not fit this plan");
        }
    }
    kt_begin(s);
    launch_chain_tables(d_coef, d_coef + n, cs, d_chain, d_chain_u, s);
    check_launch("chain_tables");
    if (r_set0 > 0) {
        // U_A of set 0 at the out-of-tile flips
        launch_chain_flips(d_coef, d_coef + n, cs, d_chain_x, s);
        check_launch("chain_flips");
    }
    kt_end(s, "chain_tables", 0.0);
}

bool Engine::gen_start_ready() const {
    return gen_start_enabled && gen_ready() && L == 2 && fuse_ready && fuse_enabled &&
           d_chain_x != nullptr;
}

// Whether the generated start contracts layer 0 inside its last high pass
// (launch_bwd_top_l0): register shape 4 for the high passes, the top group's
// tiles whole in the layout that pass reads.
bool Engine::l0_in_pass_ready() const {
    const int K = (int)passes.size() - 1;
    if (!l0_in_pass_enabled || K < 1 || pass_r[RIE_PASS_BWD_HIGH] != 4) {
        return false;
    }
    const PassGeom &g = passes[(size_t)K].g;
    return g.a == 3 && g.h0 + 9 == n && g.h0 >= 4 && g.rot_lo >= 3;
}

// Gradient evaluation of a two-layer circuit without forward or Hamiltonian
// passes (see gen_start_enabled): 1 S (generating low pass) + 2 S per high
// pass but the last + 1 S for the last, which contracts layer 0 itself when
// the high passes run from the lowest group up (gen_bottom_up: the last one
// holds the top group's tiles); otherwise 2 S for it and the 1 S layer-0
// contraction.
void Engine::structured_eval_gen_start(hipStream_t s, const double *scale, double *energy_out,
                                       bool staged) {
    const double S = (double)sizeof(double2) * (double)N;
    const int r_low = pass_r[RIE_PASS_BWD_LOW];
    if (staged) {
        stage_begin(s);
    }
    const bool bottom_up = gen_bottom_up();
    build_chain_tables(s, r_low, bottom_up);
    gen_fwd = false;
    final_ptr = nullptr;  // no state is written
    if (staged) {
        stage_end(s, "forward_apply_ms");
        stage_begin(s);
        stage_end(s, "hamiltonian_ms");  // formed inside the low adjoint pass
        stage_begin(s);
        stage_end(s, "gradient_zero_ms");
        stage_begin(s);
    }
    PassGeom g0 = passes[0].g;
    g0.pair_lo = 0;  // every adjacent pair inside the low tile
    g0.pair_hi = T;
    double2 *lam = pool[0];
    const int K = (int)passes.size() - 1;
    // lambda layouts: high pass j reads its tiles as contiguous blocks.  A
    // tile's write addresses then differ from its read addresses, so the
    // high passes alternate between two buffers instead of working in place.
    const bool layouts = gen_layouts();
    double2 *lam_other = layouts ? pool[1] : lam;
    if (layouts) {
        g0.lay_out = lam_layout_in(bottom_up ? 0 : K - 1, bottom_up);
    }
    const bool l0_in_pass = bottom_up;
    // every X term is in lambda once the low pass is done: the high passes
    // add no pair terms (fuse_layer stays -1)
    fuse_layer = -1;
    l0_energy_out = energy_out;
    try {
        require_partials((size_t)passes[0].grid * 2 * T);
        if (d_lam_tab == nullptr) {
            fail(RIE_DEVICE_ERROR, "internal: no lambda tile table");
        }
        kt_begin(s);
        launch_lam_tile_table(g0, g0.lay_out, d_lam_tab, s);
        g0.tile_out = d_lam_tab;
        launch_bwd_low_gen(r_low, lam, d_coef + n, g0, d_partials, passes[0].grid, s, h, d_chain,
                           d_chain_u, d_chain_x);
        check_launch("bwd_pass low gen");
        kt_end(s, "bwd_pass_low_gen", S);
        kt_begin(s);
        launch_reduce_grad(d_partials, passes[0].grid, T, passes[0].g, 1, d_grad, scale, s);
        check_launch("reduce_grad");
        kt_end(s, "reduce_grad", 0.0);
        for (int step = 0; step < K; ++step) {
            const int j = bottom_up ? step : K - 1 - step;
            const PassPlan &pp = passes[1 + (size_t)j];
            PassGeom gj = pp.g;
            if (layouts) {
                gj.lay_in = lam_layout_in(j, bottom_up);
                gj.lay_out = lam_layout_out(j, bottom_up);
            }
            if (j == K - 1 && l0_in_pass) {
                const uint64_t tiles = pp.g.num_tiles;
                require_partials((size_t)pp.grid * (2 * T + kTopL0Cols));
                if ((size_t)high_env_chunks * 4 * (size_t)std::max(1, n - T) > partial2_count ||
                    tile_sum_count < tiles * (Tile4Waves + 2)) {
                    fail(RIE_DEVICE_ERROR, "internal: layer-0 partial buffer too small");
                }
                double *l0_part = d_partials + (size_t)pp.grid * 2 * T;
                double2 *phi_c = d_tile_sums + tiles * Tile4Waves;
                uint64_t *lam_tile = reinterpret_cast<uint64_t *>(phi_c + tiles);
                kt_begin(s);
                launch_bwd_top_l0(lam, d_coef + n, gj, d_partials, pp.grid, s,
                                  d_chain + (size_t)(1 + j) * kChainTabEntries,
                                  d_chain_u + (size_t)(1 + j) * 2 * passes[0].g.num_tiles, d_coef,
                                  phi_c, lam_tile, d_tile_sums, l0_part);
                check_launch("bwd_pass high gen layer0");
                kt_end(s, "bwd_pass_high_gen_layer0", S);
                kt_begin(s);
                launch_reduce_grad(d_partials, pp.grid, T, pp.g, 1, d_grad, scale, s);
                check_launch("reduce_grad");
                launch_layer0_top_finish(d_tile_sums, tiles, l0_part, pp.grid, d_coef, n,
                                         d_partials2, high_env_chunks, d_grad, scale, s,
                                         l0_energy_out);
                check_launch("layer0_top_finish");
                kt_end(s, "layer0_reduce", 0.0);
                break;
            }
            require_partials((size_t)pp.grid * 2 * T);
            kt_begin(s);
            if (bottom_up) {
                // the phase-free sweep reads its tiles' lambda addresses from tables
                launch_lam_tile_table(gj, gj.lay_in, d_lam_tab, s);
                launch_lam_tile_table(gj, gj.lay_out, d_lam_tab + pp.g.num_tiles, s);
                gj.tile_in = d_lam_tab;
                gj.tile_out = d_lam_tab + pp.g.num_tiles;
            }
            launch_bwd(1, false, false, nullptr, lam, nullptr, lam_other, (size_t)(1 + j), s, 1 + j,
                       &gj, bottom_up);
            check_launch("bwd_pass");
            kt_end(s, "bwd_pass_high_gen", 2 * S);
            std::swap(lam, lam_other);
            kt_begin(s);
            launch_reduce_grad(d_partials, pp.grid, T, pp.g, 1, d_grad, scale, s);
            check_launch("reduce_grad");
            kt_end(s, "reduce_grad", 0.0);
        }
        if (!l0_in_pass) {
            layer0_gradients(s, lam, scale);
        }
    } catch (...) {
        fuse_layer = -1;
        l0_energy_out = nullptr;
        throw;
    }
    fuse_layer = -1;
    l0_energy_out = nullptr;
    if (staged) {
        stage_end(s, "backward_sweep_ms");
        stage_begin(s);
        stage_end(s, "energy_dot_ms");  // written by the layer-0 contraction
    }
    energy_partials = 0;
    resident = false;
}

bool Engine::fused_eval() const {
    return fuse_enabled && fuse_ready && strategy == RIE_STRATEGY_STRUCTURED_ADJOINT && !small &&
           L >= 2;
}

void Engine::structured_hamiltonian_fused(hipStream_t s) {
    // lambda goes where the unfused Hamiltonian puts it
    double2 *lam = nullptr;
    if (schedule == RIE_SCHEDULE_STORED) {
        lam_idx = -1;
        lam = lbuf(0);
    } else {
        lam_idx = free_buffer({final_cur});
        lam = pool[lam_idx];
    }
    const double S = (double)sizeof(double2) * (double)N;
    kt_begin(s);
    launch_ham_pass_scattered(pass_r[RIE_PASS_HAM], final_state(), lam, ham_fused_g, h,
                              ham_fused_grid, s);
    check_launch("ham_pass scattered");
    kt_end(s, "ham_pass_fused", 2.0 * S);
}

// Energy and gradient with the fused Hamiltonian.  lambda = H psi is complete
// once the last layer's adjoint passes have added their pair terms, and the
// adjoint sweep is unitary, so at layer 0 <chi|mu> = <psi|H|psi>: the layer-0
// contraction writes the energy.
void Engine::structured_eval_fused(hipStream_t s, const double *scale, double *energy_out,
                                   bool staged) {
    if (gen_start_ready()) {
        structured_eval_gen_start(s, scale, energy_out, staged);
        return;
    }
    if (staged) {
        stage_begin(s);
    }
    structured_forward(s, ckpt_every > 0);
    if (staged) {
        stage_end(s, "forward_apply_ms");
        stage_begin(s);
    }
    structured_hamiltonian_fused(s);
    if (staged) {
        stage_end(s, "hamiltonian_ms");
        stage_begin(s);
        stage_end(s, "gradient_zero_ms");  // gradients are written, not accumulated
        stage_begin(s);
    }
    fuse_layer = L - 1;
    l0_energy_out = energy_out;
    try {
        structured_backward(s, scale, false);
    } catch (...) {
        fuse_layer = -1;
        l0_energy_out = nullptr;
        throw;
    }
    fuse_layer = -1;
    l0_energy_out = nullptr;
    if (staged) {
        stage_end(s, "backward_sweep_ms");
        stage_begin(s);
        stage_end(s, "energy_dot_ms");  // written by the layer-0 contraction
    }
    energy_partials = 0;
    resident = false;
}

void Engine::structured_backward(hipStream_t s, const double *scale, bool full_walk) {
    // Layers L-1 .. 1 (and 0 in a full walk) run the inverse walk in tile
    // passes: first the low pass, which applies the inverse ring CNOT on load
    // (out of place), then the high passes in place.  Layer 0 is a product
    // state, so unless a full walk is asked for its gradients come from one
    // contraction pass over lambda.
    const double S = (double)sizeof(double2) * (double)N;
    if (schedule == RIE_SCHEDULE_STORED && L > 1 && !full_walk) {
        structured_backward_stored(s, scale);
        return;
    }
    int cur = final_cur;  // -2 while the state lives in a checkpoint buffer
    const double2 *cur_in = final_state();
    int lam = lam_idx;
    const int last_walked = full_walk ? 0 : 1;
    for (int l = L - 1; l >= last_walked; --l) {
        // Is the state after un-rotating layer l read again?  Not when the
        // next layer is the contracted layer 0, nor when a checkpoint holds it.
        const bool ckpt_next = ckpt_every > 0 && l >= 2 && (l % (int)ckpt_every) == 0 &&
                               (size_t)(l / ckpt_every) <= ckpts.size();
        const bool need_cur_after = full_walk || (l >= 2 && !ckpt_next);
        const int ncur = free_buffer({cur, lam});
        const int nlam = free_buffer({cur, lam, ncur});
        const bool only_low = passes.size() == 1;
        // Layer 1 before the contracted layer 0, with layer-1 generation: its
        // high passes generate the state, so nothing of the walked state is
        // written (the low pass still reads the state after layer 1).
        const bool gen = l == 1 && gen_fwd && !full_walk;
        if (gen) {
            build_chain_tables(s);  // the register shapes may have changed since the forward
        }
        const bool low_cur = !gen && (only_low ? need_cur_after : true);
        require_partials((size_t)passes[0].grid * 2 * T);
        kt_begin(s);
        launch_bwd(l, true, low_cur, cur_in, pool[lam], pool[ncur], pool[nlam], 0, s);
        check_launch("bwd_pass gather");
        kt_end(s, "bwd_pass_low_cnot", S * (3.0 + (low_cur ? 1.0 : 0.0)));
        kt_begin(s);
        launch_reduce_grad(d_partials, passes[0].grid, T, passes[0].g, l, d_grad, scale, s);
        check_launch("reduce_grad");
        kt_end(s, "reduce_grad", 0.0);
        cur = ncur;
        lam = nlam;
        if (gen) {
            // in the stored schedule's order, so that table set 1 + j (high
            // groups 0 .. j rotated) holds the state each pass needs
            for (int j = (int)passes.size() - 2; j >= 0; --j) {
                const PassPlan &pp = passes[1 + (size_t)j];
                require_partials((size_t)pp.grid * 2 * T);
                kt_begin(s);
                launch_bwd(l, false, false, nullptr, pool[lam], nullptr, pool[lam],
                           (size_t)(1 + j), s, 1 + j);
                check_launch("bwd_pass");
                kt_end(s, "bwd_pass_high_gen", 2 * S);
                kt_begin(s);
                launch_reduce_grad(d_partials, pp.grid, T, pp.g, l, d_grad, scale, s);
                check_launch("reduce_grad");
                kt_end(s, "reduce_grad", 0.0);
            }
            continue;  // l == 1 is the last walked layer
        }
        for (size_t p = 1; p < passes.size(); ++p) {
            const bool write_cur = (p + 1 < passes.size()) || need_cur_after;
            require_partials((size_t)passes[p].grid * 2 * T);
            kt_begin(s);
            launch_bwd(l, false, write_cur, pool[cur], pool[lam], pool[cur], pool[lam], p, s);
            check_launch("bwd_pass");
            kt_end(s, "bwd_pass_high", S * (3.0 + (write_cur ? 1.0 : 0.0)));
            kt_begin(s);
            launch_reduce_grad(d_partials, passes[p].grid, T, passes[p].g, l, d_grad, scale, s);
            check_launch("reduce_grad");
            kt_end(s, "reduce_grad", 0.0);
        }
        if (ckpt_next) {
            cur_in = ckpts[(size_t)l / ckpt_every - 1];
            cur = -2;
        } else {
            cur_in = pool[cur];
        }
    }
    walked_cur = full_walk ? cur : -1;
    if (!full_walk) {
        layer0_gradients(s, pool[lam], scale);
    }
}

void Engine::layer0_gradients(hipStream_t s, const double2 *lam, const double *scale) {
    const double S = (double)sizeof(double2) * (double)N;
    const PassGeom &g0 = passes[0].g;
    const int grid0 = passes[0].grid;
    require_partials((size_t)grid0 * 48);
    if ((size_t)high_env_chunks * 4 * (size_t)std::max(1, n - T) > partial2_count) {
        fail(RIE_DEVICE_ERROR, "internal: layer-0 partial buffer too small");
    }
    kt_begin(s);
    launch_layer0_contract(pass_r[RIE_PASS_LAYER0], lam, d_coef, g0, d_tile_sums, d_partials,
                           grid0, s);
    check_launch("layer0_contract");
    kt_end(s, "layer0_contract", S);
    kt_begin(s);
    launch_layer0_high_env(d_tile_sums, g0.num_tiles, d_coef, n, d_partials2, high_env_chunks, s);
    check_launch("layer0_high_env");
    launch_layer0_finish(d_partials, grid0, d_partials2, high_env_chunks, d_coef, n, d_grad, scale,
                         s, l0_energy_out);
    check_launch("layer0_finish");
    kt_end(s, "layer0_reduce", 0.0);
}


void Engine::structured_backward_stored(hipStream_t s, const double *scale) {
    // Stored schedule: the forward kept every pass output, so the backward
    // reads the state at each pass from there and only rewrites lambda.  The
    // low pass (inverse ring CNOT on load) comes first; the high passes
    // follow in reverse forward order.
    // With layer-1 generation the forward wrote only layer 1's output: the low
    // pass of layer 1 gathers it as usual, its high passes generate the state
    // they need.
    const double S = (double)sizeof(double2) * (double)N;
    const int K = (int)passes.size() - 1;
    if (gen_fwd) {
        build_chain_tables(s);  // the register shapes may have changed since the forward
    }
    int a = 0;  // lambda lives in lbuf(a)
    for (int l = L - 1; l >= 1; --l) {
        const int b = 1 - a;
        const bool gen = (l == 1 && gen_fwd);
        require_partials((size_t)passes[0].grid * 2 * T);
        kt_begin(s);
        launch_bwd(l, true, false, fbuf(l, K), lbuf(a), nullptr, lbuf(b), 0, s);
        check_launch("bwd_pass gather");
        kt_end(s, "bwd_pass_low_cnot", 3 * S);
        kt_begin(s);
        launch_reduce_grad(d_partials, passes[0].grid, T, passes[0].g, l, d_grad, scale, s);
        check_launch("reduce_grad");
        kt_end(s, "reduce_grad", 0.0);
        for (int j = K - 1; j >= 0; --j) {
            const PassPlan &pp = passes[1 + j];
            require_partials((size_t)pp.grid * 2 * T);
            kt_begin(s);
            launch_bwd(l, false, false, gen ? nullptr : fbuf(l, j), lbuf(b), nullptr, lbuf(b),
                       (size_t)(1 + j), s, gen ? 1 + j : -1);
            check_launch("bwd_pass");
            kt_end(s, gen ? "bwd_pass_high_gen" : "bwd_pass_high", (gen ? 2 : 3) * S);
            kt_begin(s);
            launch_reduce_grad(d_partials, pp.grid, T, pp.g, l, d_grad, scale, s);
            check_launch("reduce_grad");
            kt_end(s, "reduce_grad", 0.0);
        }
        a = b;
    }
    layer0_gradients(s, lbuf(a), scale);
}

// ---------------------------------------------------------------------------
// Structured adjoint: whole-circuit path (n <= 12)
// ---------------------------------------------------------------------------

void Engine::small_run(hipStream_t s, bool want_grad, const double *scale, double *energy_out,
                       double *grad_out, double *drift_out) {
    BatchJob job{};
    job.n = n;
    job.layers = L;
    job.chunk = (int)std::min<size_t>(chunk_width, 4);
    job.field = h;
    job.coef = d_coef;
    job.energy = energy_out;
    job.grad = want_grad ? grad_out : nullptr;
    job.drift = drift_out;
    job.scale = scale;
    kt_begin(s);
    launch_batch_structured(nullptr, job, nullptr, nullptr, 1, 256, n, nullptr, 0, s);
    check_launch("small_circuit");
    kt_end(s, "small_circuit", 0.0);
}

// ---------------------------------------------------------------------------
// Gate-level adjoint (inverse_walk_cuQuantum)
// ---------------------------------------------------------------------------

void Engine::gate_forward(hipStream_t s) {
    launch_zero_state(pool[0], N, s);
    check_launch("zero_state");
    for (int l = 0; l < L; ++l) {
        for (int w = 0; w < n; ++w) {
            const WireCoef *k = d_coef + (size_t)l * n + w;
            launch_gate_ry(pool[0], n, w, k, false, s);
            launch_gate_rz(pool[0], n, w, k, false, s);
        }
        for (int w = 0; w < n; ++w) {
            launch_gate_cnot(pool[0], n, w, (w + 1) % n, s);
        }
    }
    check_launch("gate forward");
}

void Engine::gate_hamiltonian(hipStream_t s) {
    const int grid = elementwise_grid(N);
    if ((size_t)grid > epartial_count) {
        fail(RIE_DEVICE_ERROR, "internal: energy partial buffer too small");
    }
    launch_ham_gather(pool[0], pool[1], n, h, d_epartials, grid, s);
    check_launch("ham_gather");
    energy_partials = (size_t)grid;
}

void Engine::gate_backward(hipStream_t s, const double *scale) {
    const int grid = elementwise_grid(N / 2);
    require_partials((size_t)grid);
    for (int l = L - 1; l >= 0; --l) {
        for (int w = n - 1; w >= 0; --w) {
            launch_gate_cnot(pool[0], n, w, (w + 1) % n, s);
            launch_gate_cnot(pool[1], n, w, (w + 1) % n, s);
        }
        for (int w = n - 1; w >= 0; --w) {
            const size_t base = ((size_t)l * n + w) * 2;
            const WireCoef *k = d_coef + (size_t)l * n + w;
            launch_gate_rz_step(pool[0], pool[1], n, w, k, d_partials, grid, s);
            launch_reduce_to(d_partials, grid, d_grad + base + 1, scale, s);
            launch_gate_ry_step(pool[0], pool[1], n, w, k, d_partials, grid, s);
            launch_reduce_to(d_partials, grid, d_grad + base, scale, s);
        }
    }
    check_launch("gate backward");
}

// ---------------------------------------------------------------------------
// Device-resident drivers
// ---------------------------------------------------------------------------

void Engine::forward_device(const double *params, double *energy, hipStream_t s, bool keep_lam) {
    launch_coef(params, d_coef, n * L, s);
    check_launch("coef");
    if (strategy == RIE_STRATEGY_DENSE_SCAN) {
        dense_run(s, false, nullptr, energy, nullptr);
        return;
    }
    if (strategy == RIE_STRATEGY_INVERSE_WALK_GATE) {
        gate_forward(s);
        gate_hamiltonian(s);
        launch_reduce_sum(d_epartials, (int)energy_partials, energy, s);
        check_launch("energy reduce");
        resident = true;
        return;
    }
    if (small) {
        small_run(s, false, nullptr, energy, nullptr, nullptr);
        resident = false;
        return;
    }
    structured_forward(s, ckpt_every > 0);
    structured_hamiltonian(s, keep_lam);
    kt_begin(s);
    launch_reduce_sum(d_epartials, (int)energy_partials, energy, s);
    check_launch("energy reduce");
    kt_end(s, "reduce_energy", 0.0);
    resident = keep_lam;
}

void Engine::backward_device(const double *params, const double *scale, double *grad,
                             hipStream_t s) {
    if (strategy == RIE_STRATEGY_DENSE_SCAN) {
        launch_coef(params, d_coef, n * L, s);
        dense_run(s, true, scale, d_energy + 1, grad);
        return;
    }
    if (strategy == RIE_STRATEGY_INVERSE_WALK_GATE) {
        if (!resident) {
            forward_device(params, d_energy + 1, s, true);
        }
        gate_backward(s, scale);
        copy_grad_if_needed(grad, s);
        resident = false;
        return;
    }
    if (small) {
        launch_coef(params, d_coef, n * L, s);
        small_run(s, true, scale, d_energy + 1, grad, nullptr);
        return;
    }
    if (!resident) {
        forward_device(params, d_energy + 1, s, true);
    }
    structured_backward(s, scale, false);
    copy_grad_if_needed(grad, s);
    resident = false;
}

// Energy and gradient from device-resident params: the fused evaluation
// when it applies, else the forward and backward halves.
void Engine::eval_device(const double *params, double *energy, double *grad, hipStream_t s) {
    if (fused_eval()) {
        launch_coef(params, d_coef, n * L, s);
        check_launch("coef");
        structured_eval_fused(s, nullptr, energy, false);
        copy_grad_if_needed(grad, s);
        return;
    }
    forward_device(params, energy, s, true);
    backward_device(params, nullptr, grad, s);
}

void Engine::copy_grad_if_needed(double *grad, hipStream_t s) {
    if (grad != d_grad) {
        check(hipMemcpyAsync(grad, d_grad, sizeof(double) * num_params, hipMemcpyDeviceToDevice, s),
              "copy gradient");
    }
}

// ---------------------------------------------------------------------------
// Host-pointer driver with stage profiling
// ---------------------------------------------------------------------------

void Engine::energy_and_grad_host(const double *params, size_t count, bool want_grad,
                                  bool profile, double *energy_out, double *grad_out) {
    if (count != num_params) {
        fail(RIE_INVALID_ARGUMENT, "num_params does not match the circuit.");
    }
    if (params == nullptr || energy_out == nullptr || (want_grad && grad_out == nullptr)) {
        fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
    }
    hipStream_t s = own_stream;
    order_begin(s);
    timings.clear();
    profiling = profile;
    auto t0 = std::chrono::steady_clock::now();
    check(hipMemcpyAsync(d_params, params, sizeof(double) * num_params, hipMemcpyHostToDevice, s),
          "upload params");
    if (profile) {
        check(hipStreamSynchronize(s), "sync");
        const double host_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                .count();
        timings.emplace_back("build_ops_host_ms", host_ms);
    }
    if (strategy == RIE_STRATEGY_DENSE_SCAN) {
        stage_begin(s);
        launch_coef(d_params, d_coef, n * L, s);
        dense_run(s, want_grad, nullptr, d_energy, want_grad ? d_grad : nullptr);
        stage_end(s, "dense_scan_ms");
    } else if (strategy == RIE_STRATEGY_STRUCTURED_ADJOINT && small) {
        stage_begin(s);
        launch_coef(d_params, d_coef, n * L, s);
        stage_end(s, "zero_state_ms");
        stage_begin(s);
        small_run(s, want_grad, nullptr, d_energy, want_grad ? d_grad : nullptr, nullptr);
        stage_end(s, want_grad ? "fused_circuit_ms" : "forward_apply_ms");
    } else if (want_grad && fused_eval()) {
        stage_begin(s);
        launch_coef(d_params, d_coef, n * L, s);
        check_launch("coef");
        stage_end(s, "zero_state_ms");
        structured_eval_fused(s, nullptr, d_energy, true);
    } else {
        const bool gate = strategy == RIE_STRATEGY_INVERSE_WALK_GATE;
        stage_begin(s);
        launch_coef(d_params, d_coef, n * L, s);
        check_launch("coef");
        stage_end(s, "zero_state_ms");
        stage_begin(s);
        if (gate) {
            gate_forward(s);
        } else {
            structured_forward(s, want_grad && ckpt_every > 0);
        }
        stage_end(s, "forward_apply_ms");
        stage_begin(s);
        if (gate) {
            gate_hamiltonian(s);
        } else {
            structured_hamiltonian(s, want_grad);
        }
        stage_end(s, "hamiltonian_ms");
        stage_begin(s);
        launch_reduce_sum(d_epartials, (int)energy_partials, d_energy, s);
        check_launch("energy reduce");
        stage_end(s, "energy_dot_ms");
        if (want_grad) {
            stage_begin(s);
            stage_end(s, "gradient_zero_ms");
            stage_begin(s);
            if (gate) {
                gate_backward(s, nullptr);
            } else {
                structured_backward(s, nullptr, false);
            }
            stage_end(s, "backward_sweep_ms");
        }
    }
    stage_begin(s);
    check(hipMemcpyAsync(energy_out, d_energy, sizeof(double), hipMemcpyDeviceToHost, s),
          "copy energy");
    if (want_grad) {
        check(hipMemcpyAsync(grad_out, d_grad, sizeof(double) * num_params, hipMemcpyDeviceToHost,
                             s),
              "copy gradient");
    }
    order_end(s);
    check(hipStreamSynchronize(s), "synchronize");
    stage_end(s, "gradient_copy_ms");
    profiling = false;
    resident = false;
}

double Engine::inverse_walk_drift(const double *params, size_t count) {
    if (count != num_params) {
        fail(RIE_INVALID_ARGUMENT, "num_params does not match the circuit.");
    }
    if (strategy != RIE_STRATEGY_STRUCTURED_ADJOINT) {
        fail(RIE_INVALID_ARGUMENT, "inverse-walk drift requires structured_adjoint.");
    }
    hipStream_t s = own_stream;
    order_begin(s);
    check(hipMemcpyAsync(d_params, params, sizeof(double) * num_params, hipMemcpyHostToDevice, s),
          "upload params");
    launch_coef(d_params, d_coef, n * L, s);
    if (d_drift == nullptr) {
        d_drift = (double *)device_alloc(sizeof(double) * 8200, "alloc drift");
    }
    double drift2 = 0.0;
    if (small) {
        small_run(s, true, nullptr, d_energy, d_grad, d_drift);
        check(hipMemcpyAsync(&drift2, d_drift, sizeof(double), hipMemcpyDeviceToHost, s),
              "copy drift");
    } else {
        const int saved = schedule;
        schedule = RIE_SCHEDULE_INVERSE_WALK;
        structured_forward(s, ckpt_every > 0);
        structured_hamiltonian(s, true);
        structured_backward(s, nullptr, true);
        schedule = saved;
        const int grid = elementwise_grid(N);
        launch_drift(pool[walked_cur], N, d_drift, grid, s);
        launch_reduce_sum(d_drift, grid, d_drift + 8192, s);
        check_launch("drift");
        check(hipMemcpyAsync(&drift2, d_drift + 8192, sizeof(double), hipMemcpyDeviceToHost, s),
              "copy drift");
    }
    order_end(s);
    check(hipStreamSynchronize(s), "synchronize");
    resident = false;
    return std::sqrt(drift2);
}

void Engine::set_checkpoint_interval(size_t every) {
    for (auto &c : ckpts) {
        device_free(c);
    }
    ckpts.clear();
    ckpt_every = every;
    if (every == 0 || small || strategy != RIE_STRATEGY_STRUCTURED_ADJOINT) {
        return;
    }
    const size_t count = (size_t)(L - 1) / every;
    for (size_t i = 0; i < count; ++i) {
        ckpts.push_back((double2 *)device_alloc(sizeof(double2) * N, "alloc checkpoint"));
    }
}

// Tile register shapes and backward schedule for the caller's parameters.
// For each schedule that fits, energy_and_grad runs with every pass kind at
// shape 4 and then at shape 3 under per-launch kernel timing; each pass kind
// takes the shape whose kernels ran faster (the kinds are separate kernels),
// and that mixed choice is timed as well.  The fastest timed configuration
// wins.  Replaces fixed per-qubit-count kernel choices (reference:
// structured_register_threads_per_block,
// cpp/ising_cuda_rotation_kernels.cu:562-575).
namespace {
int pass_kind_of(const std::string &name) {
    if (name == "fwd_pass_high_gen_init" || name == "fwd_layer1_gen") {
        return RIE_PASS_FWD_GEN;
    }
    if (name == "fwd_pass_high") {
        return RIE_PASS_FWD_HIGH;
    }
    if (name == "fwd_pass_low_cnot") {
        return RIE_PASS_FWD_LOW;
    }
    if (name.rfind("ham_pass", 0) == 0) {
        return RIE_PASS_HAM;
    }
    if (name == "bwd_pass_low_cnot" || name == "bwd_pass_low_gen") {
        return RIE_PASS_BWD_LOW;
    }
    if (name == "bwd_pass_high" || name == "bwd_pass_high_gen") {
        return RIE_PASS_BWD_HIGH;
    }
    if (name == "layer0_contract") {
        return RIE_PASS_LAYER0;
    }
    return -1;
}
}  // namespace

std::string Engine::autotune(const double *params, size_t count, size_t repeats) {
    if (count != num_params) {
        fail(RIE_INVALID_ARGUMENT, "num_params does not match the circuit.");
    }
    if (repeats == 0) {
        fail(RIE_INVALID_ARGUMENT, "repeats must be at least 1.");
    }
    if (strategy != RIE_STRATEGY_STRUCTURED_ADJOINT || small) {
        return "{\"tuned\": false}";
    }
    hipStream_t s = own_stream;
    order_begin(s);
    check(hipMemcpyAsync(d_params, params, sizeof(double) * num_params, hipMemcpyHostToDevice, s),
          "upload params");
    std::vector<int> schedules{schedule};
    if (L > 1) {
        size_t free_b = 0;
        size_t total_b = 0;
        check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const double extra = (double)stored_buffer_count() - (double)pool.size();
        const bool stored_fits =
            extra <= 0.0 || extra * (double)sizeof(double2) * (double)N <= 0.85 * (double)free_b;
        schedules = {RIE_SCHEDULE_INVERSE_WALK};
        if (stored_fits) {
            schedules.push_back(RIE_SCHEDULE_STORED);
        }
    }
    const bool saved_ktime = ktime;
    kt_flush();
    // wall time of `repeats` evaluations; with per_kind, the kernel time per
    // pass kind as well
    auto time_config = [&](double *per_kind) -> double {
        eval_device(d_params, d_energy, d_grad, s);
        if (per_kind != nullptr) {
            kt_flush();
            kstats.clear();
            ktime = true;
        }
        check(hipEventRecord(ev_a, s), "event");
        for (size_t r = 0; r < repeats; ++r) {
            eval_device(d_params, d_energy, d_grad, s);
        }
        check(hipEventRecord(ev_b, s), "event");
        check(hipEventSynchronize(ev_b), "event sync");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, ev_a, ev_b), "elapsed");
        if (per_kind != nullptr) {
            kt_flush();
            ktime = false;
            for (int k = 0; k < RIE_PASS_KINDS; ++k) {
                per_kind[k] = 0.0;
            }
            for (const auto &st : kstats) {
                const int k = pass_kind_of(st.name);
                if (k >= 0) {
                    per_kind[k] += st.total_ms;
                }
            }
            kstats.clear();
        }
        return (double)ms / (double)repeats;
    };
    auto shapes_tag = [](const int *r) {
        std::string t;
        for (int k = 0; k < RIE_PASS_KINDS; ++k) {
            t += (char)('0' + r[k]);
        }
        return t;
    };
    std::string rep = "{\"tuned\": true, \"ms\": {";
    double best = -1.0;
    int best_sched = schedule;
    int best_r[RIE_PASS_KINDS];
    std::copy(pass_r, pass_r + RIE_PASS_KINDS, best_r);
    bool first = true;
    auto record = [&](int sched, const int *r, double ms) {
        char buf[96];
        std::snprintf(buf, sizeof(buf), "%s\"%s/%s\": %.3f", first ? "" : ", ",
                      sched == RIE_SCHEDULE_STORED ? "stored" : "inverse_walk",
                      shapes_tag(r).c_str(), ms);
        rep += buf;
        first = false;
        if (best < 0.0 || ms < best) {
            best = ms;
            best_sched = sched;
            std::copy(r, r + RIE_PASS_KINDS, best_r);
        }
    };
    try {
        for (int sched : schedules) {
            set_schedule(sched);
            double kind_ms[2][RIE_PASS_KINDS];
            for (int u = 0; u < 2; ++u) {
                std::fill(pass_r, pass_r + RIE_PASS_KINDS, u == 0 ? 4 : 3);
                record(sched, pass_r, time_config(kind_ms[u]));
            }
            int mixed[RIE_PASS_KINDS];
            bool differs = false;
            for (int k = 0; k < RIE_PASS_KINDS; ++k) {
                mixed[k] = kind_ms[1][k] < kind_ms[0][k] ? 3 : 4;
                differs = differs || mixed[k] != mixed[0];
            }
            if (differs) {
                std::copy(mixed, mixed + RIE_PASS_KINDS, pass_r);
                record(sched, pass_r, time_config(nullptr));
            }
        }
    } catch (...) {
        ktime = saved_ktime;
        throw;
    }
    ktime = saved_ktime;
    set_schedule(best_sched);
    std::copy(best_r, best_r + RIE_PASS_KINDS, pass_r);
    order_end(s);
    char tail[256];
    std::snprintf(tail, sizeof(tail),
                  "}, \"best_ms\": %.3f, \"schedule\": \"%s\", \"tile_shapes\": [%d, %d, %d], "
                  "\"pass_shapes\": [%d, %d, %d, %d, %d, %d, %d]}",
                  best, best_sched == RIE_SCHEDULE_STORED ? "stored" : "inverse_walk",
                  pass_r[RIE_PASS_FWD_HIGH], pass_r[RIE_PASS_HAM], pass_r[RIE_PASS_BWD_HIGH],
                  pass_r[0], pass_r[1], pass_r[2], pass_r[3], pass_r[4], pass_r[5], pass_r[6]);
    rep += tail;
    return rep;
}

void Engine::benchmark_forward_graph(const double *params, size_t count, size_t repeats,
                                     double *normal_ms, double *graph_ms) {
    // reference: benchmark_structured_forward_graph (cpp/ising_cuda_backend.cu:444-541)
    if (count != num_params) {
        fail(RIE_INVALID_ARGUMENT, "num_params does not match the circuit.");
    }
    if (strategy != RIE_STRATEGY_STRUCTURED_ADJOINT) {
        fail(RIE_INVALID_ARGUMENT, "CUDA graph forward benchmark requires structured_adjoint.");
    }
    if (repeats == 0) {
        fail(RIE_INVALID_ARGUMENT, "repeats must be at least 1.");
    }
    hipStream_t s = own_stream;
    order_begin(s);
    check(hipMemcpyAsync(d_params, params, sizeof(double) * num_params, hipMemcpyHostToDevice, s),
          "upload params");
    auto forward = [&]() {
        launch_coef(d_params, d_coef, n * L, s);
        if (small) {
            small_run_captured(s);
        } else {
            structured_forward(s, false);
        }
    };
    forward();
    check(hipStreamSynchronize(s), "warmup sync");
    auto time_it = [&](auto &&body) {
        check(hipEventRecord(ev_a, s), "event");
        for (size_t r = 0; r < repeats; ++r) {
            body();
        }
        check(hipEventRecord(ev_b, s), "event");
        check(hipEventSynchronize(ev_b), "event sync");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, ev_a, ev_b), "elapsed");
        return (double)ms / (double)repeats;
    };
    *normal_ms = time_it(forward);
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    check(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), "begin capture");
    forward();
    check(hipStreamEndCapture(s, &graph), "end capture");
    check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
    check(hipGraphLaunch(exec, s), "graph launch");
    check(hipStreamSynchronize(s), "graph warmup");
    *graph_ms = time_it([&]() { check(hipGraphLaunch(exec, s), "graph launch"); });
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    order_end(s);
    resident = false;
}

void Engine::small_run_captured(hipStream_t s) {
    small_run(s, false, nullptr, d_energy + 2, nullptr, nullptr);
}

void Engine::copy_state(double *host_out, size_t count) {
    if (count != 2 * N) {
        fail(RIE_INVALID_ARGUMENT, "state buffer must hold 2 * 2^n doubles.");
    }
    if (strategy != RIE_STRATEGY_STRUCTURED_ADJOINT || small) {
        fail(RIE_UNSUPPORTED, "copy_state requires structured_adjoint with num_qubits > 12.");
    }
    if (final_state() == nullptr) {
        fail(RIE_UNSUPPORTED, "no state is held: the last evaluation generated its states "
                              "(an energy-only evaluation keeps the final state).");
    }
    if (ev_last_valid) {
        check(hipEventSynchronize(ev_last), "hipEventSynchronize");
    }
    check(hipMemcpy(host_out, final_state(), sizeof(double2) * N, hipMemcpyDeviceToHost),
          "copy state");
}

}  // namespace rie

// ===========================================================================
// C ABI
// ===========================================================================

using rie::Engine;
using rie::guard;

namespace {

Engine *as_engine(rie_engine *e) {
    if (e == nullptr) {
        rie::fail(RIE_INVALID_ARGUMENT, "null engine");
    }
    return reinterpret_cast<Engine *>(e);
}

// Device-path calls run on the caller's stream; NULL is HIP's default
// (null) stream, which is also the handle torch reports for its default stream.
hipStream_t stream_of(Engine *e, void *stream) {
    (void)e;
    return static_cast<hipStream_t>(stream);
}

}  // namespace

extern "C" {

const char *rie_last_error(void) { return rie::g_last_error.c_str(); }

int rie_abi_version(void) { return RIE_ABI_VERSION; }

rie_status rie_parse_strategy(const char *name, int *strategy_out) {
    return guard([&]() {
        if (name == nullptr || strategy_out == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        *strategy_out = rie::parse_strategy(name);
    });
}

int rie_device_count(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return count;
}

rie_status rie_engine_create(size_t num_qubits, size_t num_layers, double field,
                             const char *gradient_strategy, size_t chunk_width, int double_buffer,
                             rie_engine **engine_out) {
    return guard([&]() {
        if (gradient_strategy == nullptr || engine_out == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        *engine_out = nullptr;
        auto engine = std::make_unique<Engine>(num_qubits, num_layers, field, gradient_strategy,
                                               chunk_width, double_buffer != 0);
        *engine_out = reinterpret_cast<rie_engine *>(engine.release());
    });
}

void rie_engine_destroy(rie_engine *engine) { delete reinterpret_cast<Engine *>(engine); }

rie_status rie_engine_energy_and_grad(rie_engine *engine, const double *params, size_t num_params,
                                      int compute_gradient, int profile, double *energy_out,
                                      double *gradient_out) {
    return guard([&]() {
        as_engine(engine)->energy_and_grad_host(params, num_params, compute_gradient != 0,
                                                profile != 0, energy_out, gradient_out);
    });
}

size_t rie_engine_timing_count(const rie_engine *engine) {
    if (engine == nullptr) {
        return 0;
    }
    return reinterpret_cast<const Engine *>(engine)->timings.size();
}

rie_status rie_engine_timing(const rie_engine *engine, size_t index, const char **name_out,
                             double *ms_out) {
    return guard([&]() {
        if (engine == nullptr || name_out == nullptr || ms_out == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        const auto *e = reinterpret_cast<const Engine *>(engine);
        if (index >= e->timings.size()) {
            rie::fail(RIE_INVALID_ARGUMENT, "timing index out of range.");
        }
        *name_out = e->timings[index].first.c_str();
        *ms_out = e->timings[index].second;
    });
}

rie_status rie_engine_benchmark_forward_graph(rie_engine *engine, const double *params,
                                              size_t num_params, size_t repeats,
                                              double *normal_forward_ms, double *graph_forward_ms) {
    return guard([&]() {
        if (params == nullptr || normal_forward_ms == nullptr || graph_forward_ms == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        as_engine(engine)->benchmark_forward_graph(params, num_params, repeats, normal_forward_ms,
                                                   graph_forward_ms);
    });
}

rie_status rie_energy_and_grad(size_t num_qubits, size_t num_layers, double field,
                               const char *gradient_strategy, const double *params,
                               size_t num_params, int compute_gradient, int profile,
                               size_t chunk_width, double *energy_out, double *gradient_out) {
    return guard([&]() {
        if (gradient_strategy == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        Engine engine(num_qubits, num_layers, field, gradient_strategy, chunk_width, false);
        engine.energy_and_grad_host(params, num_params, compute_gradient != 0, profile != 0,
                                    energy_out, gradient_out);
    });
}

size_t rie_engine_num_qubits(const rie_engine *engine) {
    return engine == nullptr ? 0 : (size_t)reinterpret_cast<const Engine *>(engine)->n;
}

size_t rie_engine_num_layers(const rie_engine *engine) {
    return engine == nullptr ? 0 : (size_t)reinterpret_cast<const Engine *>(engine)->L;
}

int rie_engine_strategy(const rie_engine *engine) {
    return engine == nullptr ? -1 : reinterpret_cast<const Engine *>(engine)->strategy;
}

rie_status rie_engine_forward_device(rie_engine *engine, const double *d_params, double *d_energy,
                                     void *stream) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        if (d_params == nullptr || d_energy == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        hipStream_t s = stream_of(e, stream);
        e->order_begin(s);
        e->forward_device(d_params, d_energy, s, true);
        e->order_end(s);
    });
}

rie_status rie_engine_backward_device(rie_engine *engine, const double *d_params,
                                      const double *d_scale, double *d_gradient, void *stream) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        if (d_params == nullptr || d_gradient == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        hipStream_t s = stream_of(e, stream);
        e->order_begin(s);
        e->backward_device(d_params, d_scale, d_gradient, s);
        e->order_end(s);
    });
}

rie_status rie_engine_energy_and_grad_device(rie_engine *engine, const double *d_params,
                                             double *d_energy, double *d_gradient, void *stream) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        if (d_params == nullptr || d_energy == nullptr || d_gradient == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        hipStream_t s = stream_of(e, stream);
        e->order_begin(s);
        e->eval_device(d_params, d_energy, d_gradient, s);
        e->order_end(s);
    });
}

rie_status rie_engine_autotune(rie_engine *engine, const double *params, size_t num_params,
                              size_t repeats, char *report_out, size_t report_len,
                              size_t *report_needed) {
    return guard([&]() {
        if (params == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        const std::string rep = as_engine(engine)->autotune(params, num_params, repeats);
        if (report_needed != nullptr) {
            *report_needed = rep.size() + 1;
        }
        if (report_out != nullptr && report_len > 0) {
            const size_t k = std::min(report_len - 1, rep.size());
            std::memcpy(report_out, rep.data(), k);
            report_out[k] = '\0';
        }
    });
}

rie_status rie_engine_set_tile_shapes(rie_engine *engine, int r_fwd, int r_ham, int r_bwd) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        for (int r : {r_fwd, r_ham, r_bwd}) {
            if (r != 3 && r != 4) {
                rie::fail(RIE_INVALID_ARGUMENT, "tile shapes must be 3 or 4.");
            }
        }
        e->pass_r[RIE_PASS_FWD_GEN] = r_fwd;
        e->pass_r[RIE_PASS_FWD_HIGH] = r_fwd;
        e->pass_r[RIE_PASS_FWD_LOW] = r_fwd;
        e->pass_r[RIE_PASS_HAM] = r_ham;
        e->pass_r[RIE_PASS_BWD_LOW] = r_bwd;
        e->pass_r[RIE_PASS_BWD_HIGH] = r_bwd;
    });
}

rie_status rie_engine_tile_shapes(const rie_engine *engine, int *shapes_out3) {
    return guard([&]() {
        if (engine == nullptr || shapes_out3 == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        const Engine *e = reinterpret_cast<const Engine *>(engine);
        shapes_out3[0] = e->pass_r[RIE_PASS_FWD_HIGH];
        shapes_out3[1] = e->pass_r[RIE_PASS_HAM];
        shapes_out3[2] = e->pass_r[RIE_PASS_BWD_HIGH];
    });
}

rie_status rie_engine_set_pass_shapes(rie_engine *engine, const int *shapes, size_t count) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        if (shapes == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        if (count != (size_t)RIE_PASS_KINDS) {
            rie::fail(RIE_INVALID_ARGUMENT, "count must be RIE_PASS_KINDS.");
        }
        for (size_t k = 0; k < count; ++k) {
            if (shapes[k] != 3 && shapes[k] != 4) {
                rie::fail(RIE_INVALID_ARGUMENT, "tile shapes must be 3 or 4.");
            }
        }
        std::copy(shapes, shapes + count, e->pass_r);
    });
}

rie_status rie_engine_pass_shapes(const rie_engine *engine, int *shapes_out, size_t count) {
    return guard([&]() {
        if (engine == nullptr || shapes_out == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        if (count != (size_t)RIE_PASS_KINDS) {
            rie::fail(RIE_INVALID_ARGUMENT, "count must be RIE_PASS_KINDS.");
        }
        const Engine *e = reinterpret_cast<const Engine *>(engine);
        std::copy(e->pass_r, e->pass_r + count, shapes_out);
    });
}

rie_status rie_engine_set_checkpoint_interval(rie_engine *engine, size_t every_k_layers) {
    return guard([&]() { as_engine(engine)->set_checkpoint_interval(every_k_layers); });
}

rie_status rie_engine_inverse_walk_drift(rie_engine *engine, const double *params,
                                         size_t num_params, double *drift_out) {
    return guard([&]() {
        if (params == nullptr || drift_out == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        *drift_out = as_engine(engine)->inverse_walk_drift(params, num_params);
    });
}

rie_status rie_engine_set_schedule(rie_engine *engine, int schedule) {
    return guard([&]() { as_engine(engine)->set_schedule(schedule); });
}

int rie_engine_schedule(const rie_engine *engine) {
    return engine == nullptr ? -1 : reinterpret_cast<const Engine *>(engine)->schedule;
}

rie_status rie_engine_set_fused_hamiltonian(rie_engine *engine, int enabled) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        if (enabled != 0 && enabled != 1) {
            rie::fail(RIE_INVALID_ARGUMENT, "enabled must be 0 or 1.");
        }
        e->fuse_enabled = enabled == 1;
        e->resident = false;
    });
}

int rie_engine_fused_hamiltonian(const rie_engine *engine) {
    return engine == nullptr ? -1 : (reinterpret_cast<const Engine *>(engine)->fused_eval() ? 1 : 0);
}

rie_status rie_engine_set_layer1_generation(rie_engine *engine, int enabled) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        if (enabled != 0 && enabled != 1) {
            rie::fail(RIE_INVALID_ARGUMENT, "enabled must be 0 or 1.");
        }
        e->gen_enabled = enabled == 1;
        e->resident = false;
    });
}

int rie_engine_layer1_generation(const rie_engine *engine) {
    return engine == nullptr ? -1 : (reinterpret_cast<const Engine *>(engine)->gen_ready() ? 1 : 0);
}

rie_status rie_engine_set_generated_start(rie_engine *engine, int enabled) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        if (enabled != 0 && enabled != 1) {
            rie::fail(RIE_INVALID_ARGUMENT, "enabled must be 0 or 1.");
        }
        e->gen_start_enabled = enabled == 1;
        e->resident = false;
    });
}

int rie_engine_generated_start(const rie_engine *engine) {
    return engine == nullptr ? -1
                             : (reinterpret_cast<const Engine *>(engine)->gen_start_ready() ? 1 : 0);
}

rie_status rie_engine_set_double_buffer(rie_engine *engine, int enabled) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        if (enabled != 0 && enabled != 1) {
            rie::fail(RIE_INVALID_ARGUMENT, "enabled must be 0 or 1.");
        }
        e->double_buffer = enabled == 1;
    });
}

rie_status rie_engine_set_lambda_layouts(rie_engine *engine, int enabled) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        if (enabled != 0 && enabled != 1) {
            rie::fail(RIE_INVALID_ARGUMENT, "enabled must be 0 or 1.");
        }
        e->lam_layouts_enabled = enabled == 1;
    });
}

rie_status rie_engine_set_layer0_in_pass(rie_engine *engine, int enabled) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        if (enabled != 0 && enabled != 1) {
            rie::fail(RIE_INVALID_ARGUMENT, "enabled must be 0 or 1.");
        }
        e->l0_in_pass_enabled = enabled == 1;
    });
}

int rie_engine_layer0_in_pass(const rie_engine *engine) {
    if (engine == nullptr) {
        return -1;
    }
    const Engine *e = reinterpret_cast<const Engine *>(engine);
    return e->gen_start_ready() && e->gen_bottom_up() ? 1 : 0;
}

int rie_engine_lambda_layouts(const rie_engine *engine) {
    if (engine == nullptr) {
        return -1;
    }
    const Engine *e = reinterpret_cast<const Engine *>(engine);
    return e->gen_start_ready() && e->gen_layouts() ? 1 : 0;
}

int rie_engine_double_buffer(const rie_engine *engine) {
    return engine == nullptr ? -1 : (reinterpret_cast<const Engine *>(engine)->double_buffer ? 1 : 0);
}

rie_status rie_engine_set_kernel_timing(rie_engine *engine, int enable) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        e->kt_flush();
        e->ktime = enable != 0;
    });
}

rie_status rie_engine_reset_kernel_stats(rie_engine *engine) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        e->kt_flush();
        e->kstats.clear();
    });
}

rie_status rie_engine_kernel_stat_count(rie_engine *engine, size_t *count_out) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        if (count_out == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        e->kt_flush();
        *count_out = e->kstats.size();
    });
}

rie_status rie_engine_kernel_stat(rie_engine *engine, size_t index, const char **name_out,
                                  long *launches_out, double *total_ms_out,
                                  double *total_bytes_out) {
    return guard([&]() {
        Engine *e = as_engine(engine);
        if (name_out == nullptr || launches_out == nullptr || total_ms_out == nullptr ||
            total_bytes_out == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        e->kt_flush();
        if (index >= e->kstats.size()) {
            rie::fail(RIE_INVALID_ARGUMENT, "kernel stat index out of range.");
        }
        const auto &k = e->kstats[index];
        *name_out = k.name.c_str();
        *launches_out = k.launches;
        *total_ms_out = k.total_ms;
        *total_bytes_out = k.total_bytes;
    });
}

rie_status rie_engine_copy_state(rie_engine *engine, double *host_out, size_t num_doubles) {
    return guard([&]() {
        if (host_out == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        as_engine(engine)->copy_state(host_out, num_doubles);
    });
}

}  // extern "C"
