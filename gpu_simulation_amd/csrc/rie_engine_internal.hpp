// Engine state shared by the host-side translation units of librie.so.
#pragma once

#include <hip/hip_runtime.h>

#include <initializer_list>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rie.h"
#include "rie_kernels.hpp"

namespace rie {

struct Error : std::runtime_error {
    Error(rie_status st, const std::string &msg);
    rie_status status;
};

[[noreturn]] void fail(rie_status st, const std::string &msg);
void check(hipError_t e, const char *ctx);
void check_launch(const char *ctx);
void *device_alloc(size_t bytes, const char *what);
void device_free(void *p);
int parse_strategy(const std::string &name);

extern thread_local std::string g_last_error;

// Runs f, mapping exceptions to an rie_status and the thread's last error.
template <typename F>
rie_status guard(F &&f) {
    try {
        f();
        return RIE_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return RIE_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return RIE_DEVICE_ERROR;
    }
}

struct PassPlan {
    PassGeom g;
    int grid;
    int first_wire;  // wires [first_wire, end_wire) are processed by the pass
    int end_wire;
};

class DensePlan;  // dense_scan buffers and launches (rie_dense.hpp)

struct Engine {
    Engine(size_t num_qubits, size_t num_layers, double field, const std::string &strategy,
           size_t chunk_width, bool double_buffer);
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    // problem
    int n;
    int L;
    double h;
    int strategy;
    size_t chunk_width;
    bool double_buffer;
    uint64_t N = 0;
    size_t num_params = 0;

    // device / plan
    int device = 0;
    int num_cus = 0;
    int max_grid = 0;
    hipStream_t own_stream = nullptr;
    bool small = false;
    // register shape of the tile passes per pass kind: 4 = 256 threads x 16
    // registers, 3 = 512 threads x 8 registers (both 4096-amplitude tiles)
    // (defaults measured on MI355X at 30 qubits: the 512-thread shape is
    // faster for the forward and H passes, the 256-thread one for the layer-1
    // generator and the adjoint passes)
    int pass_r[RIE_PASS_KINDS] = {4, 3, 3, 3, 4, 4, 4};
    int T = 12;
    std::vector<PassPlan> passes;  // [0]: low pass (a = T); [1..]: high passes
    // Fused Hamiltonian of gradient evaluations (L >= 2): one psi-frame pass
    // over a scattered tile (the ZZ diagonal and the X terms of its wires);
    // every other X term is an adjacent-wire pair in the pre-CNOT frame of the
    // last layer, added by that layer's adjoint passes (PassGeom pair ranges).
    PassGeom ham_fused_g{};
    int ham_fused_grid = 0;
    bool fuse_ready = false;   // the plan supports it
    bool fuse_enabled = true;  // caller's switch
    int fuse_layer = -1;       // layer whose adjoint passes add pair terms now
    // Layer-1 generation (L >= 2, n <= kChainMaxQubits): every state inside
    // layer 1 before its ring CNOT is a trace of a ring of 2x2 matrices over
    // the layer-0 product state, so layer 1's forward is one write pass and,
    // in both backward schedules, its adjoint high passes generate the state
    // instead of reading or walking it.  Chain table sets: 0 the forward,
    // 1 + j the adjoint high pass j.
    bool gen_enabled = true;    // caller's switch
    bool gen_fwd = false;       // the last forward generated layer 1
    // Generated start (L == 2 gradient evaluations with the fused
    // Hamiltonian): no forward and no Hamiltonian pass; layer 1's low adjoint
    // pass generates its state and forms lambda = P^-1 H psi for its tile
    // itself (ZZ diagonal, in-tile pairs, every other X term from flipped
    // chain segments), then the high passes (without pair terms) and the
    // layer-0 contraction.
    bool gen_start_enabled = true;  // caller's switch
    CMat2 *d_chain_x = nullptr;     // per tile: U_A at the out-of-tile flips
    // Lambda layouts of the generated start: each high pass reads its tile
    // as one contiguous block (the pass before it writes that layout).
    bool lam_layouts_enabled = true;  // caller's switch
    bool l0_in_pass_enabled = true;   // caller's switch: layer 0 inside the last high pass
    bool l0_in_pass_ready() const;
    bool lam_layouts_fit(bool bottom_up) const;
    LamLayout lam_layout_in(int j, bool bottom_up) const;   // read by high pass j
    LamLayout lam_layout_out(int j, bool bottom_up) const;  // written by high pass j
    bool gen_bottom_up() const;  // generated start: high passes from the lowest group up
    bool gen_layouts() const;    // generated start: lambda layouts in use
    CMat2 *d_chain = nullptr;   // kChainMaxSets * kChainTabEntries matrices
    CMat2 *d_chain_u = nullptr; // per set and tile: the uniform segments U_A, U_B

    // device buffers
    std::vector<double2 *> pool;  // statevector buffers (16 * 2^n bytes each)
    int schedule = RIE_SCHEDULE_INVERSE_WALK;
    double *d_params = nullptr;
    double *d_grad = nullptr;
    double *d_energy = nullptr;
    WireCoef *d_coef = nullptr;
    double *d_partials = nullptr;
    double *d_epartials = nullptr;
    double *d_drift = nullptr;
    double2 *d_tile_sums = nullptr;  // layer-0 contraction: per-tile sums (tile_sum_count)
    uint64_t tile_sum_count = 0;
    uint64_t *d_lam_tab = nullptr;  // generated start: per-tile lambda addresses (2 x tiles)
    double *d_partials2 = nullptr;   // layer-0 contraction: high-wire partials
    int high_env_chunks = 1;
    size_t partial_count = 0;    // capacity of d_partials (doubles)
    size_t epartial_count = 0;   // capacity of d_epartials
    size_t partial2_count = 0;   // capacity of d_partials2
    size_t energy_partials = 0;
    std::vector<double2 *> ckpts;
    size_t ckpt_every = 0;
    DensePlan *dense = nullptr;

    // resident-state bookkeeping
    int final_cur = 0;
    const double2 *final_ptr = nullptr;
    int lam_idx = -1;
    int walked_cur = -1;
    bool resident = false;

    // per-launch kernel timing (opt-in): HIP events around every launch on
    // the launch stream, with the launch's algorithmic HBM bytes
    struct KernelStat {
        std::string name;
        long launches = 0;
        double total_ms = 0.0;
        double total_bytes = 0.0;
    };
    struct PendingLaunch {
        const char *name;
        double bytes;
        hipEvent_t a;
        hipEvent_t b;
    };
    bool ktime = false;
    std::vector<PendingLaunch> kpending;
    std::vector<hipEvent_t> kevent_pool;
    std::vector<KernelStat> kstats;
    hipEvent_t kt_open = nullptr;
    void kt_begin(hipStream_t s);
    void kt_end(hipStream_t s, const char *name, double bytes);
    void kt_flush();
    hipEvent_t kt_event();

    // cross-stream ordering of successive calls on one engine
    hipEvent_t ev_last = nullptr;
    bool ev_last_valid = false;
    void order_begin(hipStream_t s);
    void order_end(hipStream_t s);

    // profiling
    bool profiling = false;
    hipEvent_t ev_a = nullptr;
    hipEvent_t ev_b = nullptr;
    std::vector<std::pair<std::string, double>> timings;

    void build_plan();
    const double2 *final_state() const;
    void require_partials(size_t count) const;
    int free_buffer(std::initializer_list<int> used) const;

    void stage_begin(hipStream_t s);
    void stage_end(hipStream_t s, const char *name);

    void structured_forward(hipStream_t s, bool save_ckpts);
    void structured_hamiltonian(hipStream_t s, bool write_lam);
    bool fused_eval() const;
    void structured_hamiltonian_fused(hipStream_t s);
    void structured_eval_fused(hipStream_t s, const double *scale, double *energy_out,
                               bool staged);
    void launch_bwd(int l, bool gather, bool write_cur, const double2 *cur_in,
                    const double2 *lam_in, double2 *cur_out, double2 *lam_out, size_t p,
                    hipStream_t s, int gen_set = -1, const PassGeom *geom = nullptr,
                    bool phase_free = false);
    bool gen_ready() const;
    bool gen_start_ready() const;
    void build_chain_tables(hipStream_t s, int r_set0 = -1, bool bottom_up = false);
    void structured_eval_gen_start(hipStream_t s, const double *scale, double *energy_out,
                                   bool staged);
    void structured_backward(hipStream_t s, const double *scale, bool full_walk);
    void structured_backward_stored(hipStream_t s, const double *scale);
    void layer0_gradients(hipStream_t s, const double2 *lam, const double *scale);
    double *l0_energy_out = nullptr;  // set by the fused evaluation: energy from layer 0
    size_t stored_buffer_count() const;
    double2 *fbuf(int l, int j) const;  // stored schedule: output of pass j of layer l
    double2 *lbuf(int i) const;         // stored schedule: lambda buffers 0 and 1
    void ensure_pool(size_t count);
    void set_schedule(int sched);
    void small_run(hipStream_t s, bool want_grad, const double *scale, double *energy_out,
                   double *grad_out, double *drift_out);
    void small_run_captured(hipStream_t s);
    void gate_forward(hipStream_t s);
    void gate_hamiltonian(hipStream_t s);
    void gate_backward(hipStream_t s, const double *scale);
    void dense_init();
    void dense_free();
    void dense_run(hipStream_t s, bool want_grad, const double *scale, double *energy_out,
                   double *grad_out);

    void forward_device(const double *params, double *energy, hipStream_t s, bool keep_lam);
    void backward_device(const double *params, const double *scale, double *grad, hipStream_t s);
    void copy_grad_if_needed(double *grad, hipStream_t s);
    void eval_device(const double *params, double *energy, double *grad, hipStream_t s);
    void energy_and_grad_host(const double *params, size_t count, bool want_grad, bool profile,
                              double *energy_out, double *grad_out);
    double inverse_walk_drift(const double *params, size_t count);
    void set_checkpoint_interval(size_t every);
    void benchmark_forward_graph(const double *params, size_t count, size_t repeats,
                                 double *normal_ms, double *graph_ms);
    void copy_state(double *host_out, size_t count);
    std::string autotune(const double *params, size_t count, size_t repeats);
};

}  // namespace rie
