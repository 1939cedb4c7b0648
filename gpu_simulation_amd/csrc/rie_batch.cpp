// Batched evaluation of many independent ring-Ising circuits of uneven qubit
// and layer counts, with an autotuner and HIP-graph-captured optimizer steps.
//
// Circuits with n <= dense_max_qubits take the dense_scan product tree
// (rie_dense.hip, one grouped FP64-MFMA GEMM launch per tree level and tile
// class for all of them); the others run in one persistent launch of the
// whole-circuit kernel (rie_batch.hip) whose workgroups take circuits from a
// longest-first (LPT) schedule built on the host from per-layer costs.  The
// two parts run concurrently on two streams (two branches of one graph when
// captured).  Parameters, energies and gradients of all circuits are packed
// in device arrays owned by the batch: circuit i's 2 n_i L_i parameters start
// at offset[i], in the reference's per-circuit layout (l n + w) 2 + {0, 1}
// (cpp/ising_cuda_structured_adjoint_modes.inc:73-77).
//
// The reference evaluates one circuit per call (RingIsingCudaBackend::
// energy_and_grad, cpp/ising_cuda_backend.cu:429-442) and picks its kernel
// shapes from fixed qubit-count thresholds (structured_register_threads_per_
// block, cpp/ising_cuda_rotation_kernels.cu:562-575; the chunk planner,
// cpp/ising_cuda_structured_adjoint_modes.inc:18-71).  Here the autotuner
// measures the candidates on the batch itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <memory>
#include <queue>
#include <sstream>
#include <string>
#include <vector>

#include "rie.h"
#include "rie_dense.hpp"
#include "rie_engine_internal.hpp"
#include "rie_kernels.hpp"

namespace rie {

namespace {

constexpr int kMaxN = kBatchMaxQubits;

int max_chunk_for(int threads) { return batch_max_chunk(threads); }

}  // namespace

struct Batch {
    Batch(size_t count, const int *nq, const int *nl, const double *fields, int dense_max_qubits);
    ~Batch();
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    // problem
    size_t count = 0;
    std::vector<int> n;
    std::vector<int> L;
    std::vector<double> h;
    std::vector<char> dense_flag;
    std::vector<size_t> poff;  // count + 1 parameter offsets
    size_t total = 0;
    int dense_max = 8;    // the caller's bound: n <= dense_max may take dense_scan
    int dense_split = 8;  // the split in force: n <= dense_split takes dense_scan
    bool tune_split = true;  // the autotuner chooses dense_split in [0, dense_max]

    // device
    int device = 0;
    int num_cus = 0;
    hipStream_t s_own = nullptr;
    hipStream_t s_aux = nullptr;
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join = nullptr;
    hipEvent_t ev_t0 = nullptr;
    hipEvent_t ev_t1 = nullptr;
    double *d_params = nullptr;
    double *d_grad = nullptr;
    double *d_energy = nullptr;
    WireCoef *d_coef = nullptr;
    double *d_m = nullptr;
    double *d_v = nullptr;
    int *d_step = nullptr;
    double *d_history = nullptr;
    size_t history_cap = 0;

    // structured part: persistent launch
    std::vector<int> sids;
    std::vector<BatchJob> jobs_g;  // with gradients, indexed like sids
    BatchJob *d_jobs_g = nullptr;
    BatchJob *d_jobs_e = nullptr;  // energy only
    int *d_order = nullptr;
    int *d_wg_begin = nullptr;
    double2 *d_scratch = nullptr;
    uint64_t scratch_stride = 0;
    int scratch_slots = 0;
    int grid = 0;
    int threads = 512;
    int max_lds_n = 0;
    int grid_limit = 0;  // 0: every CU
    int wg_per_cu = 1;
    int chunk_by_n[kMaxN + 1];
    double cost_per_layer[kMaxN + 1];  // measured (autotuner) or modelled
    bool measured = false;

    // dense part
    std::vector<int> dids;
    DensePlan dense;
    int dense_tile = 64;
    std::vector<std::unique_ptr<DensePlan>> dense_single;

    // graphs
    hipGraphExec_t graph_exec = nullptr;
    int graph_kind = -1;  // 0 energy, 1 energy+grad, 2 GD step, 3 Adam step
    double graph_lr = 0.0;
    int graph_capacity = 0;
    bool use_graph = false;

    std::string report = "{}";

    void build_structured();
    void build_dense();
    void repartition(int split);
    void release_structured();
    void invalidate_graph();
    void issue_eval(hipStream_t s, bool want_grad);
    void issue_step(hipStream_t s, int optimizer, double lr, int capacity);
    void launch_graph(int kind, double lr, int capacity, hipStream_t s);
    void run(hipStream_t s, bool want_grad);
    void run_per_circuit(hipStream_t s, bool want_grad);
    void optimize(int optimizer, double lr, size_t steps, bool graph, double *history_out,
                  hipStream_t s);
    void autotune(size_t repeats);
    double time_ms(size_t repeats, const std::function<void(hipStream_t)> &fn);
    void ensure_history(size_t entries);
    void ensure_adam();
    std::string config_json() const;
};

Batch::Batch(size_t count_, const int *nq, const int *nl, const double *fields, int dense_max_qubits)
    : count(count_), dense_max(dense_max_qubits), dense_split(dense_max_qubits) {
    if (count == 0) {
        fail(RIE_INVALID_ARGUMENT, "a batch needs at least one circuit.");
    }
    if (nq == nullptr || nl == nullptr || fields == nullptr) {
        fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
    }
    if (dense_max_qubits < 0 || dense_max_qubits > 8) {
        fail(RIE_INVALID_ARGUMENT, "dense_max_qubits must be in [0, 8].");
    }
    poff.assign(count + 1, 0);
    for (size_t i = 0; i < count; ++i) {
        if (nq[i] < 2 || nq[i] > kMaxN) {
            fail(RIE_INVALID_ARGUMENT, "batched circuits need 2 <= num_qubits <= 14.");
        }
        if (nl[i] < 1 || nl[i] > 65536) {
            fail(RIE_INVALID_ARGUMENT, "batched circuits need 1 <= layers <= 65536.");
        }
        if (!std::isfinite(fields[i])) {
            fail(RIE_INVALID_ARGUMENT, "field must be finite.");
        }
        n.push_back(nq[i]);
        L.push_back(nl[i]);
        h.push_back(fields[i]);
        dense_flag.push_back(nq[i] <= dense_max_qubits ? 1 : 0);
        poff[i + 1] = poff[i] + (size_t)2 * nq[i] * nl[i];
    }
    total = poff[count];
    if (total / 2 > 0x7fffffffULL) {
        fail(RIE_INVALID_ARGUMENT, "batch has too many parameters.");
    }
    for (int q = 0; q <= kMaxN; ++q) {
        chunk_by_n[q] = 3;
        // modelled per-layer cost: amplitude updates, with a floor for the
        // barrier-bound small states and a factor for the global-scratch ones
        const double amps = std::ldexp(1.0, q);
        cost_per_layer[q] = q * std::max(amps, 1024.0) * (q > kBatchMaxLdsQubits ? 3.0 : 1.0);
    }
    check(hipGetDevice(&device), "hipGetDevice");
    hipDeviceProp_t prop;
    check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    num_cus = prop.multiProcessorCount;
    check(hipStreamCreateWithFlags(&s_own, hipStreamNonBlocking), "hipStreamCreate");
    check(hipStreamCreateWithFlags(&s_aux, hipStreamNonBlocking), "hipStreamCreate");
    check(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreate(&ev_t0), "hipEventCreate");
    check(hipEventCreate(&ev_t1), "hipEventCreate");
    d_params = (double *)device_alloc(sizeof(double) * total, "alloc batch params");
    d_grad = (double *)device_alloc(sizeof(double) * total, "alloc batch gradient");
    d_energy = (double *)device_alloc(sizeof(double) * count, "alloc batch energies");
    d_coef = (WireCoef *)device_alloc(sizeof(WireCoef) * (total / 2), "alloc batch coefficients");
    d_step = (int *)device_alloc(sizeof(int), "alloc step counter");
    check(hipMemset(d_params, 0, sizeof(double) * total), "clear params");
    check(hipMemset(d_grad, 0, sizeof(double) * total), "clear gradient");
    check(hipMemset(d_energy, 0, sizeof(double) * count), "clear energies");
    check(hipMemset(d_step, 0, sizeof(int)), "clear step");
    for (size_t i = 0; i < count; ++i) {
        (dense_flag[i] ? dids : sids).push_back((int)i);
    }
    build_structured();
    build_dense();
}

// Circuits with n <= split on dense_scan, the others on the whole-circuit
// kernel; both parts rebuilt (the per-circuit comparison plans as well).
void Batch::repartition(int split) {
    dense_split = split;
    dids.clear();
    sids.clear();
    for (size_t i = 0; i < count; ++i) {
        dense_flag[i] = n[i] <= split ? 1 : 0;
        (dense_flag[i] ? dids : sids).push_back((int)i);
    }
    dense_single.clear();
    build_structured();
    build_dense();
}

Batch::~Batch() {
    if (s_own != nullptr) {
        (void)hipStreamSynchronize(s_own);
    }
    if (s_aux != nullptr) {
        (void)hipStreamSynchronize(s_aux);
    }
    invalidate_graph();
    release_structured();
    device_free(d_params);
    device_free(d_grad);
    device_free(d_energy);
    device_free(d_coef);
    device_free(d_m);
    device_free(d_v);
    device_free(d_step);
    device_free(d_history);
    for (hipEvent_t ev : {ev_fork, ev_join, ev_t0, ev_t1}) {
        if (ev != nullptr) {
            (void)hipEventDestroy(ev);
        }
    }
    if (s_own != nullptr) {
        (void)hipStreamDestroy(s_own);
    }
    if (s_aux != nullptr) {
        (void)hipStreamDestroy(s_aux);
    }
}

void Batch::invalidate_graph() {
    if (graph_exec != nullptr) {
        (void)hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
    }
    graph_kind = -1;
}

void Batch::release_structured() {
    device_free(d_jobs_g);
    device_free(d_jobs_e);
    device_free(d_order);
    device_free(d_wg_begin);
    device_free(d_scratch);
    d_jobs_g = d_jobs_e = nullptr;
    d_order = d_wg_begin = nullptr;
    d_scratch = nullptr;
    scratch_slots = 0;
    grid = 0;
}

// Longest-first schedule of the structured circuits over `grid` workgroups.
void Batch::build_structured() {
    invalidate_graph();
    release_structured();
    jobs_g.clear();
    if (sids.empty()) {
        return;
    }
    const int cap = grid_limit > 0 ? grid_limit : num_cus * wg_per_cu;
    grid = (int)std::min<size_t>(sids.size(), (size_t)std::max(1, cap));
    max_lds_n = 0;
    int max_global_n = 0;
    for (int i : sids) {
        BatchJob j{};
        j.n = n[(size_t)i];
        j.layers = L[(size_t)i];
        j.chunk = std::min(chunk_by_n[j.n], max_chunk_for(threads));
        j.field = h[(size_t)i];
        j.coef = d_coef + poff[(size_t)i] / 2;
        j.energy = d_energy + i;
        j.grad = d_grad + poff[(size_t)i];
        j.drift = nullptr;
        j.scale = nullptr;
        jobs_g.push_back(j);
        if (j.n <= kBatchMaxLdsQubits) {
            max_lds_n = std::max(max_lds_n, j.n);
        } else {
            max_global_n = std::max(max_global_n, j.n);
        }
    }
    // LPT: heaviest first onto the least-loaded workgroup
    std::vector<int> idx(sids.size());
    for (size_t k = 0; k < idx.size(); ++k) {
        idx[k] = (int)k;
    }
    auto cost = [&](int k) {
        const BatchJob &j = jobs_g[(size_t)k];
        return (double)j.layers * cost_per_layer[j.n];
    };
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost(a) > cost(b); });
    using Slot = std::pair<double, int>;
    std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
    for (int b = 0; b < grid; ++b) {
        heap.push({0.0, b});
    }
    std::vector<std::vector<int>> lists((size_t)grid);
    for (int k : idx) {
        Slot s = heap.top();
        heap.pop();
        lists[(size_t)s.second].push_back(k);
        s.first += cost(k);
        heap.push(s);
    }
    std::vector<int> order;
    std::vector<int> begin(1, 0);
    for (const auto &lst : lists) {
        order.insert(order.end(), lst.begin(), lst.end());
        begin.push_back((int)order.size());
    }
    std::vector<BatchJob> jobs_e = jobs_g;
    for (auto &j : jobs_e) {
        j.grad = nullptr;
    }
    d_jobs_g = (BatchJob *)device_alloc(sizeof(BatchJob) * jobs_g.size(), "alloc batch jobs");
    d_jobs_e = (BatchJob *)device_alloc(sizeof(BatchJob) * jobs_g.size(), "alloc batch jobs");
    d_order = (int *)device_alloc(sizeof(int) * order.size(), "alloc batch order");
    d_wg_begin = (int *)device_alloc(sizeof(int) * begin.size(), "alloc batch schedule");
    check(hipMemcpy(d_jobs_g, jobs_g.data(), sizeof(BatchJob) * jobs_g.size(),
                    hipMemcpyHostToDevice),
          "upload batch jobs");
    check(hipMemcpy(d_jobs_e, jobs_e.data(), sizeof(BatchJob) * jobs_e.size(),
                    hipMemcpyHostToDevice),
          "upload batch jobs");
    check(hipMemcpy(d_order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice),
          "upload batch order");
    check(hipMemcpy(d_wg_begin, begin.data(), sizeof(int) * begin.size(), hipMemcpyHostToDevice),
          "upload batch schedule");
    if (max_global_n > 0) {
        scratch_stride = (uint64_t)2 << max_global_n;
        // one slot per workgroup of the persistent grid, and at least as many
        // as the autotuner's one-circuit-per-workgroup probes use
        scratch_slots = std::max(grid, std::min((int)sids.size(), num_cus));
        d_scratch = (double2 *)device_alloc(sizeof(double2) * scratch_stride * scratch_slots,
                                            "alloc batch scratch");
    }
}

void Batch::build_dense() {
    invalidate_graph();
    dense_single.clear();
    std::vector<DenseSpec> specs;
    for (int i : dids) {
        specs.push_back({n[(size_t)i], L[(size_t)i], h[(size_t)i], d_coef + poff[(size_t)i] / 2,
                         d_energy + i, d_grad + poff[(size_t)i], nullptr});
    }
    dense.build(specs, dense_tile);
}

void Batch::issue_eval(hipStream_t s, bool want_grad) {
    launch_coef(d_params, d_coef, (int)(total / 2), s);
    check_launch("batch coef");
    const bool both = grid > 0 && !dids.empty();
    if (both) {
        check(hipEventRecord(ev_fork, s), "record fork");
        check(hipStreamWaitEvent(s_aux, ev_fork, 0), "wait fork");
    }
    if (grid > 0) {
        launch_batch_structured(want_grad ? d_jobs_g : d_jobs_e, BatchJob{}, d_order, d_wg_begin,
                                grid, threads, max_lds_n, d_scratch, scratch_stride, s);
        check_launch("batch structured");
    }
    if (!dids.empty()) {
        dense.run(both ? s_aux : s, want_grad);
        if (both) {
            check(hipEventRecord(ev_join, s_aux), "record join");
            check(hipStreamWaitEvent(s, ev_join, 0), "wait join");
        }
    }
}

void Batch::ensure_history(size_t entries) {
    if (entries <= history_cap) {
        return;
    }
    invalidate_graph();
    device_free(d_history);
    d_history = (double *)device_alloc(sizeof(double) * entries, "alloc energy history");
    history_cap = entries;
}

void Batch::ensure_adam() {
    if (d_m == nullptr) {
        d_m = (double *)device_alloc(sizeof(double) * total, "alloc adam m");
        d_v = (double *)device_alloc(sizeof(double) * total, "alloc adam v");
    }
}

// One optimizer step: energies and gradients at the current parameters,
// energies recorded at history[step], parameters updated, step advanced.
void Batch::issue_step(hipStream_t s, int optimizer, double lr, int capacity) {
    issue_eval(s, true);
    launch_record_energies(d_energy, (int)count, d_history, d_step, capacity, s);
    check_launch("record energies");
    if (optimizer == 0) {
        launch_gd_update(d_params, d_grad, total, lr, s);
    } else {
        launch_adam_update(d_params, d_grad, d_m, d_v, total, lr, 0.9, 0.999, 1e-8, d_step, s);
    }
    check_launch("optimizer update");
    launch_step_advance(d_step, s);
    check_launch("step advance");
}

void Batch::launch_graph(int kind, double lr, int capacity, hipStream_t s) {
    if (graph_exec == nullptr || graph_kind != kind || graph_lr != lr ||
        graph_capacity != capacity) {
        invalidate_graph();
        hipGraph_t graph = nullptr;
        check(hipStreamBeginCapture(s_own, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            if (kind <= 1) {
                issue_eval(s_own, kind == 1);
            } else {
                issue_step(s_own, kind - 2, lr, capacity);
            }
        } catch (...) {
            (void)hipStreamEndCapture(s_own, &graph);
            if (graph != nullptr) {
                (void)hipGraphDestroy(graph);
            }
            throw;
        }
        check(hipStreamEndCapture(s_own, &graph), "end capture");
        const hipError_t e = hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        check(e, "instantiate graph");
        graph_kind = kind;
        graph_lr = lr;
        graph_capacity = capacity;
    }
    check(hipGraphLaunch(graph_exec, s), "launch graph");
}

void Batch::run(hipStream_t s, bool want_grad) {
    if (use_graph) {
        launch_graph(want_grad ? 1 : 0, 0.0, 0, s);
    } else {
        issue_eval(s, want_grad);
    }
}

// The comparison mode: every circuit in its own launch(es), one after the
// other on one stream, with the same kernels.
void Batch::run_per_circuit(hipStream_t s, bool want_grad) {
    if (dense_single.empty() && !dids.empty()) {
        for (int i : dids) {
            auto plan = std::make_unique<DensePlan>();
            plan->build({{n[(size_t)i], L[(size_t)i], h[(size_t)i], d_coef + poff[(size_t)i] / 2,
                          d_energy + i, d_grad + poff[(size_t)i], nullptr}},
                        dense_tile);
            dense_single.push_back(std::move(plan));
        }
    }
    size_t di = 0;
    size_t si = 0;
    for (size_t i = 0; i < count; ++i) {
        launch_coef(d_params + poff[i], d_coef + poff[i] / 2, n[i] * L[i], s);
        if (dense_flag[i]) {
            dense_single[di++]->run(s, want_grad);
        } else {
            BatchJob j = jobs_g[si++];
            if (!want_grad) {
                j.grad = nullptr;
            }
            launch_batch_structured(nullptr, j, nullptr, nullptr, 1, threads,
                                    j.n <= kBatchMaxLdsQubits ? j.n : 0, d_scratch,
                                    scratch_stride, s);
        }
    }
    check_launch("per-circuit launches");
}

void Batch::optimize(int optimizer, double lr, size_t steps, bool graph, double *history_out,
                     hipStream_t s) {
    if (optimizer != 0 && optimizer != 1) {
        fail(RIE_INVALID_ARGUMENT, "optimizer must be 0 (gradient descent) or 1 (Adam).");
    }
    if (!std::isfinite(lr)) {
        fail(RIE_INVALID_ARGUMENT, "learning rate must be finite.");
    }
    if (steps == 0) {
        return;
    }
    if (steps > (size_t)1 << 20) {
        fail(RIE_INVALID_ARGUMENT, "at most 2^20 optimizer steps per call.");
    }
    ensure_history(steps * count);
    const int capacity = (int)steps;
    if (optimizer == 1) {
        ensure_adam();
        launch_fill(d_m, total, 0.0, s);
        launch_fill(d_v, total, 0.0, s);
    }
    launch_set_int(d_step, 0, s);
    check_launch("optimizer reset");
    if (graph) {
        // one captured step (recording into history rows [0, capacity)), replayed
        for (size_t k = 0; k < steps; ++k) {
            launch_graph(2 + optimizer, lr, capacity, s);
        }
    } else {
        for (size_t k = 0; k < steps; ++k) {
            issue_step(s, optimizer, lr, capacity);
        }
    }
    if (history_out != nullptr) {
        check(hipMemcpyAsync(history_out, d_history, sizeof(double) * steps * count,
                             hipMemcpyDeviceToHost, s),
              "download energy history");
        check(hipStreamSynchronize(s), "sync");
    }
}

double Batch::time_ms(size_t repeats, const std::function<void(hipStream_t)> &fn) {
    fn(s_own);
    check(hipStreamSynchronize(s_own), "sync");
    std::vector<float> t;
    for (size_t r = 0; r < repeats; ++r) {
        check(hipEventRecord(ev_t0, s_own), "record");
        fn(s_own);
        check(hipEventRecord(ev_t1, s_own), "record");
        check(hipEventSynchronize(ev_t1), "sync");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, ev_t0, ev_t1), "elapsed");
        t.push_back(ms);
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

void Batch::autotune(size_t repeats) {
    repeats = std::max<size_t>(1, std::min<size_t>(repeats, 50));
    invalidate_graph();
    // the split candidates: every qubit count <= dense_max present, and none
    std::vector<int> splits;
    if (tune_split) {
        for (int q = dense_max; q >= 1; --q) {
            if (std::find(n.begin(), n.end(), q) != n.end()) {
                splits.push_back(q);
            }
        }
        if (!splits.empty()) {
            splits.push_back(0);
            // every circuit on the whole-circuit kernel while its costs are probed
            repartition(0);
        }
    }
    if (splits.empty()) {
        splits.push_back(dense_split);
    }
    std::ostringstream rep;
    rep.precision(6);
    rep << "{\"per_layer_us\": {";
    const int thread_opts[3] = {256, 512, 1024};
    // c[n][ti][K]: microseconds per layer of the slowest probe circuit
    std::vector<std::vector<std::vector<double>>> c(
        kMaxN + 1, std::vector<std::vector<double>>(3, std::vector<double>(5, 0.0)));
    std::vector<double> layers_of_n(kMaxN + 1, 0.0);
    for (size_t k = 0; k < sids.size(); ++k) {
        layers_of_n[(size_t)jobs_g[k].n] += jobs_g[k].layers;
    }
    bool first_n = true;
    BatchJob *d_probe = nullptr;
    try {
        for (int q = 2; q <= kMaxN; ++q) {
            if (layers_of_n[(size_t)q] == 0.0) {
                continue;
            }
            std::vector<BatchJob> probe;
            int max_l = 0;
            for (size_t k = 0; k < sids.size() && (int)probe.size() < num_cus; ++k) {
                if (jobs_g[k].n == q) {
                    probe.push_back(jobs_g[k]);
                    max_l = std::max(max_l, jobs_g[k].layers);
                }
            }
            if ((int)probe.size() > scratch_slots && q > kBatchMaxLdsQubits) {
                probe.resize((size_t)scratch_slots);
            }
            device_free(d_probe);
            d_probe = (BatchJob *)device_alloc(sizeof(BatchJob) * probe.size(), "alloc probe");
            rep << (first_n ? "" : ", ") << "\"" << q << "\": {";
            first_n = false;
            for (int ti = 0; ti < 3; ++ti) {
                const int T = thread_opts[ti];
                rep << (ti ? ", " : "") << "\"" << T << "\": [";
                for (int K = 1; K <= max_chunk_for(T); ++K) {
                    for (auto &j : probe) {
                        j.chunk = K;
                    }
                    check(hipMemcpy(d_probe, probe.data(), sizeof(BatchJob) * probe.size(),
                                    hipMemcpyHostToDevice),
                          "upload probe");
                    const int lds_n = q <= kBatchMaxLdsQubits ? q : 0;
                    const double ms = time_ms(repeats, [&](hipStream_t s) {
                        launch_batch_structured(d_probe, BatchJob{}, nullptr, nullptr,
                                                (int)probe.size(), T, lds_n, d_scratch,
                                                scratch_stride, s);
                        check_launch("autotune probe");
                    });
                    c[(size_t)q][(size_t)ti][(size_t)K] = ms * 1e3 / max_l;
                    rep << (K > 1 ? ", " : "") << c[(size_t)q][(size_t)ti][(size_t)K];
                }
                rep << "]";
            }
            rep << "}";
        }
    } catch (...) {
        device_free(d_probe);
        throw;
    }
    device_free(d_probe);
    rep << "}";
    if (!sids.empty()) {
        // one thread count for the launch: least total modelled work
        double best = -1.0;
        int best_ti = 1;
        for (int ti = 0; ti < 3; ++ti) {
            double work = 0.0;
            for (int q = 2; q <= kMaxN; ++q) {
                if (layers_of_n[(size_t)q] == 0.0) {
                    continue;
                }
                double m = -1.0;
                for (int K = 1; K <= max_chunk_for(thread_opts[ti]); ++K) {
                    const double v = c[(size_t)q][(size_t)ti][(size_t)K];
                    m = (m < 0.0 || v < m) ? v : m;
                }
                work += layers_of_n[(size_t)q] * m;
            }
            if (best < 0.0 || work < best) {
                best = work;
                best_ti = ti;
            }
        }
        threads = thread_opts[best_ti];
        for (int q = 2; q <= kMaxN; ++q) {
            if (layers_of_n[(size_t)q] == 0.0) {
                continue;
            }
            int bk = 1;
            for (int K = 2; K <= max_chunk_for(threads); ++K) {
                if (c[(size_t)q][(size_t)best_ti][(size_t)K] <
                    c[(size_t)q][(size_t)best_ti][(size_t)bk]) {
                    bk = K;
                }
            }
            chunk_by_n[q] = bk;
            cost_per_layer[q] = c[(size_t)q][(size_t)best_ti][(size_t)bk];
        }
        measured = true;
    }
    // per split: the dense tile for D >= 64 and the structured launch's grid
    // (with a dense part running beside it, leave it CUs); the fastest whole
    // evaluation wins
    double best_ms = -1.0;
    int best_grid = 0;
    int best_split = splits.front();
    int best_tile = dense_tile;
    rep << ", \"split_ms\": {";
    for (size_t si = 0; si < splits.size(); ++si) {
        const int split = splits[si];
        if (split != dense_split || si == 0) {
            repartition(split);
        }
        bool big_dense = false;
        for (int i : dids) {
            big_dense = big_dense || n[(size_t)i] >= 6;
        }
        if (big_dense) {
            const int tiles[2] = {32, 64};
            double dense_ms[2] = {0.0, 0.0};
            for (int k = 0; k < 2; ++k) {
                dense_tile = tiles[k];
                build_dense();
                dense_ms[k] = time_ms(repeats, [&](hipStream_t s) {
                    launch_coef(d_params, d_coef, (int)(total / 2), s);
                    dense.run(s, true);
                });
            }
            dense_tile = dense_ms[0] < dense_ms[1] ? 32 : 64;
            build_dense();
        }
        std::vector<int> grid_opts;
        if (!sids.empty()) {
            if (!dids.empty()) {
                for (int f : {1, 2, 3, 4}) {
                    grid_opts.push_back(num_cus * f / 4);
                }
            } else {
                grid_opts.push_back(num_cus);
                if (max_lds_n <= 11) {
                    grid_opts.push_back(2 * num_cus);  // two workgroups per CU fit in LDS
                }
            }
        } else {
            grid_opts.push_back(0);
        }
        rep << (si ? ", " : "") << "\"" << split << "\": {";
        for (size_t k = 0; k < grid_opts.size(); ++k) {
            grid_limit = grid_opts[k];
            build_structured();
            const double ms = time_ms(repeats, [&](hipStream_t s) { issue_eval(s, true); });
            rep << (k ? ", " : "") << "\"" << grid_limit << "\": " << ms;
            if (best_ms < 0.0 || ms < best_ms) {
                best_ms = ms;
                best_grid = grid_limit;
                best_split = split;
                best_tile = dense_tile;
            }
        }
        rep << "}";
    }
    rep << "}";
    grid_limit = best_grid;
    dense_tile = best_tile;
    repartition(best_split);
    rep << ", \"dense_max_qubits\": " << best_split << ", \"batch_ms\": " << best_ms << "}";
    report = rep.str();
}

std::string Batch::config_json() const {
    std::ostringstream o;
    o << "{\"threads\": " << threads << ", \"grid\": " << grid << ", \"grid_limit\": " << grid_limit
      << ", \"dense_max_qubits\": " << dense_split << ", \"tune_split\": " << (tune_split ? "true" : "false")
      << ", \"dense_tile\": " << dense_tile << ", \"measured\": " << (measured ? "true" : "false")
      << ", \"structured_circuits\": " << sids.size() << ", \"dense_circuits\": " << dids.size()
      << ", \"chunk_by_n\": {";
    bool first = true;
    for (int q = 2; q <= kMaxN; ++q) {
        o << (first ? "" : ", ") << "\"" << q << "\": " << std::min(chunk_by_n[q], max_chunk_for(threads));
        first = false;
    }
    o << "}, \"autotune\": " << report << "}";
    return o.str();
}

}  // namespace rie

// ===========================================================================
// C ABI
// ===========================================================================

using rie::Batch;
using rie::guard;

namespace {

Batch *as_batch(rie_batch *b) {
    if (b == nullptr) {
        rie::fail(RIE_INVALID_ARGUMENT, "null batch");
    }
    return reinterpret_cast<Batch *>(b);
}

const Batch *as_batch(const rie_batch *b) {
    if (b == nullptr) {
        rie::fail(RIE_INVALID_ARGUMENT, "null batch");
    }
    return reinterpret_cast<const Batch *>(b);
}

}  // namespace

extern "C" {

rie_status rie_batch_create(size_t count, const int *num_qubits, const int *num_layers,
                            const double *fields, int dense_max_qubits, rie_batch **batch_out) {
    return guard([&]() {
        if (batch_out == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        *batch_out = nullptr;
        auto b = std::make_unique<Batch>(count, num_qubits, num_layers, fields, dense_max_qubits);
        *batch_out = reinterpret_cast<rie_batch *>(b.release());
    });
}

void rie_batch_destroy(rie_batch *batch) {
    if (batch != nullptr) {
        delete reinterpret_cast<Batch *>(batch);
    }
}

size_t rie_batch_size(const rie_batch *batch) {
    return batch == nullptr ? 0 : reinterpret_cast<const Batch *>(batch)->count;
}

size_t rie_batch_num_params(const rie_batch *batch) {
    return batch == nullptr ? 0 : reinterpret_cast<const Batch *>(batch)->total;
}

rie_status rie_batch_param_offsets(const rie_batch *batch, size_t *offsets_out, size_t len) {
    return guard([&]() {
        const Batch *b = as_batch(batch);
        if (offsets_out == nullptr || len != b->count + 1) {
            rie::fail(RIE_INVALID_ARGUMENT, "offsets_out must hold count + 1 values.");
        }
        std::copy(b->poff.begin(), b->poff.end(), offsets_out);
    });
}

rie_status rie_batch_strategies(const rie_batch *batch, int *strategy_out, size_t len) {
    return guard([&]() {
        const Batch *b = as_batch(batch);
        if (strategy_out == nullptr || len != b->count) {
            rie::fail(RIE_INVALID_ARGUMENT, "strategy_out must hold count values.");
        }
        for (size_t i = 0; i < b->count; ++i) {
            strategy_out[i] =
                b->dense_flag[i] ? RIE_STRATEGY_DENSE_SCAN : RIE_STRATEGY_STRUCTURED_ADJOINT;
        }
    });
}

rie_status rie_batch_device_buffers(rie_batch *batch, double **d_params, double **d_energies,
                                    double **d_gradient) {
    return guard([&]() {
        Batch *b = as_batch(batch);
        if (d_params != nullptr) {
            *d_params = b->d_params;
        }
        if (d_energies != nullptr) {
            *d_energies = b->d_energy;
        }
        if (d_gradient != nullptr) {
            *d_gradient = b->d_grad;
        }
    });
}

rie_status rie_batch_set_params(rie_batch *batch, const double *params, size_t num_params) {
    return guard([&]() {
        Batch *b = as_batch(batch);
        if (params == nullptr || num_params != b->total) {
            rie::fail(RIE_INVALID_ARGUMENT, "num_params does not match the batch.");
        }
        rie::check(hipMemcpyAsync(b->d_params, params, sizeof(double) * b->total,
                                  hipMemcpyHostToDevice, b->s_own),
                   "upload params");
        rie::check(hipStreamSynchronize(b->s_own), "sync");
    });
}

rie_status rie_batch_get_params(rie_batch *batch, double *params_out, size_t num_params) {
    return guard([&]() {
        Batch *b = as_batch(batch);
        if (params_out == nullptr || num_params != b->total) {
            rie::fail(RIE_INVALID_ARGUMENT, "num_params does not match the batch.");
        }
        rie::check(hipDeviceSynchronize(), "sync");
        rie::check(hipMemcpy(params_out, b->d_params, sizeof(double) * b->total,
                             hipMemcpyDeviceToHost),
                   "download params");
    });
}

rie_status rie_batch_get_results(rie_batch *batch, double *energies_out, double *gradient_out) {
    return guard([&]() {
        Batch *b = as_batch(batch);
        if (energies_out == nullptr) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        rie::check(hipDeviceSynchronize(), "sync");
        rie::check(hipMemcpy(energies_out, b->d_energy, sizeof(double) * b->count,
                             hipMemcpyDeviceToHost),
                   "download energies");
        if (gradient_out != nullptr) {
            rie::check(hipMemcpy(gradient_out, b->d_grad, sizeof(double) * b->total,
                                 hipMemcpyDeviceToHost),
                       "download gradient");
        }
    });
}

rie_status rie_batch_energy_and_grad(rie_batch *batch, const double *params, size_t num_params,
                                     int compute_gradient, double *energies_out,
                                     double *gradient_out) {
    return guard([&]() {
        Batch *b = as_batch(batch);
        if (params == nullptr || energies_out == nullptr ||
            (compute_gradient && gradient_out == nullptr)) {
            rie::fail(RIE_INVALID_ARGUMENT, "null pointer argument.");
        }
        if (num_params != b->total) {
            rie::fail(RIE_INVALID_ARGUMENT, "num_params does not match the batch.");
        }
        hipStream_t s = b->s_own;
        rie::check(hipMemcpyAsync(b->d_params, params, sizeof(double) * b->total,
                                  hipMemcpyHostToDevice, s),
                   "upload params");
        b->run(s, compute_gradient != 0);
        rie::check(hipMemcpyAsync(energies_out, b->d_energy, sizeof(double) * b->count,
                                  hipMemcpyDeviceToHost, s),
                   "download energies");
        if (compute_gradient) {
            rie::check(hipMemcpyAsync(gradient_out, b->d_grad, sizeof(double) * b->total,
                                      hipMemcpyDeviceToHost, s),
                       "download gradient");
        }
        rie::check(hipStreamSynchronize(s), "sync");
    });
}

rie_status rie_batch_run_device(rie_batch *batch, int compute_gradient, void *stream) {
    return guard([&]() { as_batch(batch)->run((hipStream_t)stream, compute_gradient != 0); });
}

rie_status rie_batch_run_per_circuit(rie_batch *batch, int compute_gradient, void *stream) {
    return guard(
        [&]() { as_batch(batch)->run_per_circuit((hipStream_t)stream, compute_gradient != 0); });
}

rie_status rie_batch_set_use_graph(rie_batch *batch, int use_graph) {
    return guard([&]() {
        Batch *b = as_batch(batch);
        b->use_graph = use_graph != 0;
        b->invalidate_graph();
    });
}

rie_status rie_batch_optimize(rie_batch *batch, int optimizer, double learning_rate, size_t steps,
                              int use_graph, double *energy_history_out, void *stream) {
    return guard([&]() {
        as_batch(batch)->optimize(optimizer, learning_rate, steps, use_graph != 0,
                                  energy_history_out, (hipStream_t)stream);
    });
}

rie_status rie_batch_autotune(rie_batch *batch, size_t repeats) {
    return guard([&]() {
        Batch *b = as_batch(batch);
        rie::check(hipDeviceSynchronize(), "sync");
        b->autotune(repeats);
    });
}

rie_status rie_batch_set_tune_split(rie_batch *batch, int enabled) {
    return guard([&]() { as_batch(batch)->tune_split = enabled != 0; });
}

rie_status rie_batch_set_config(rie_batch *batch, int threads, const int *chunk_by_n,
                                size_t chunk_len, int dense_tile, int structured_grid) {
    return guard([&]() {
        Batch *b = as_batch(batch);
        if (threads != 256 && threads != 512 && threads != 1024) {
            rie::fail(RIE_INVALID_ARGUMENT, "threads must be 256, 512 or 1024.");
        }
        if (dense_tile != 32 && dense_tile != 64) {
            rie::fail(RIE_INVALID_ARGUMENT, "dense_tile must be 32 or 64.");
        }
        if (structured_grid < 0) {
            rie::fail(RIE_INVALID_ARGUMENT, "structured_grid must be >= 0.");
        }
        if (chunk_by_n != nullptr) {
            if (chunk_len != (size_t)rie::kBatchMaxQubits + 1) {
                rie::fail(RIE_INVALID_ARGUMENT, "chunk_by_n must hold 15 values (index n).");
            }
            for (size_t q = 0; q < chunk_len; ++q) {
                if (chunk_by_n[q] < 1 || chunk_by_n[q] > 4) {
                    rie::fail(RIE_INVALID_ARGUMENT, "chunk widths must be in [1, 4].");
                }
            }
            std::copy(chunk_by_n, chunk_by_n + chunk_len, b->chunk_by_n);
        }
        rie::check(hipDeviceSynchronize(), "sync");
        b->threads = threads;
        b->grid_limit = structured_grid;
        const bool tile_changed = b->dense_tile != dense_tile;
        b->dense_tile = dense_tile;
        b->build_structured();
        if (tile_changed) {
            b->build_dense();
        }
    });
}

rie_status rie_batch_config(const rie_batch *batch, char *json_out, size_t len,
                            size_t *needed_out) {
    return guard([&]() {
        const std::string s = as_batch(batch)->config_json();
        if (needed_out != nullptr) {
            *needed_out = s.size() + 1;
        }
        if (json_out != nullptr && len > 0) {
            const size_t k = std::min(len - 1, s.size());
            std::copy(s.begin(), s.begin() + (long)k, json_out);
            json_out[k] = '\0';
        }
    });
}

}  // extern "C"
