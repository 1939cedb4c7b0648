/*
 * rie.h - C ABI of the MI355X ring-Ising statevector engine (librie.so).
 *
 * The engine computes E(theta) = <psi|H|psi> and dE/dtheta for the circuit
 * "RY(theta) RZ(phi) on every wire, then ring CNOTs i -> i+1 mod n" with
 * H = -sum_i Z_i Z_{i+1 mod n} - h sum_i X_i, in complex128 on one AMD
 * Instinct MI355X (gfx950).  Parameters are a flat float64 array of 2*n*L
 * values indexed (l*n + w)*2 + {0: theta_RY, 1: phi_RZ}
 * (reference: cpp/ising_cuda_structured_adjoint_modes.inc:73-77).
 *
 * Every entry point returns an rie_status; on failure rie_last_error()
 * returns a thread-local message.  RIE_INVALID_ARGUMENT corresponds to the
 * reference's std::invalid_argument (surfaced in Python as ValueError) and
 * RIE_DEVICE_ERROR to its std::runtime_error (RuntimeError)
 * (reference: cpp/ising_cuda_kernels.cu:18-44, cpp/ising_cuda_bindings.cpp).
 *
 * No torch or HIP types appear in these signatures: streams are passed as
 * void* (a hipStream_t; NULL is HIP's default stream) and device buffers as
 * plain pointers.  Host-pointer entry points run on a stream owned by the
 * engine and return after the results are copied back.
 */
#ifndef RIE_H
#define RIE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RIE_ABI_VERSION 1

typedef enum rie_status {
    RIE_OK = 0,
    RIE_INVALID_ARGUMENT = 1,
    RIE_DEVICE_ERROR = 2,
    RIE_OUT_OF_MEMORY = 3,
    RIE_UNSUPPORTED = 4
} rie_status;

/* Gradient strategies; names as accepted by the reference's parser
 * (cpp/ising_cuda_backend.cu:249-264, ring_ising/config.py:8-16). */
typedef enum rie_strategy {
    RIE_STRATEGY_INVERSE_WALK_GATE = 0, /* "inverse_walk_cuQuantum" / "inverse_walk_cuquantum" */
    RIE_STRATEGY_STRUCTURED_ADJOINT = 1, /* "structured_adjoint" */
    RIE_STRATEGY_DENSE_SCAN = 2          /* "dense_scan" (num_qubits <= 8) */
} rie_strategy;

typedef struct rie_engine rie_engine;

/* Thread-local message describing the last failure on this thread. */
const char *rie_last_error(void);

/* ABI version of the loaded library (RIE_ABI_VERSION it was built with). */
int rie_abi_version(void);

/* Parse a strategy name; replaces parse_gradient_strategy
 * (cpp/ising_cuda_backend.cu:249-264).  Unknown names -> RIE_INVALID_ARGUMENT. */
rie_status rie_parse_strategy(const char *name, int *strategy_out);

/* Number of visible HIP devices (0 when none); never fails. */
int rie_device_count(void);

/* Create an engine and allocate its device state buffers on the current
 * device.  Replaces RingIsingCudaBackend::RingIsingCudaBackend
 * (cpp/ising_cuda_backend.hpp:29-31, cpp/ising_cuda_backend.cu:374-419):
 * num_qubits in [2, 34], num_layers >= 1, structured_rotation_chunk_width in
 * [1, 8] for structured_adjoint, dense_scan requires num_qubits <= 8.
 * double_buffer: 1 makes the adjoint high passes double-buffered (see
 * rie_engine_set_double_buffer); results are identical either way. */
rie_status rie_engine_create(size_t num_qubits, size_t num_layers, double field,
                             const char *gradient_strategy,
                             size_t structured_rotation_chunk_width,
                             int double_buffer, rie_engine **engine_out);

/* Free an engine and its device buffers; NULL is ignored.  Replaces
 * RingIsingCudaBackend::~RingIsingCudaBackend (cpp/ising_cuda_backend.hpp:32). */
void rie_engine_destroy(rie_engine *engine);

/* Energy and, if compute_gradient, the 2*n*L gradient (host pointers,
 * synchronous).  gradient_out may be NULL when compute_gradient == 0.
 * profile != 0 records stage timings readable with rie_engine_timing*.
 * Replaces RingIsingCudaBackend::energy_and_grad
 * (cpp/ising_cuda_backend.hpp:39-41, cpp/ising_cuda_backend.cu:429-442). */
rie_status rie_engine_energy_and_grad(rie_engine *engine, const double *params,
                                      size_t num_params, int compute_gradient,
                                      int profile, double *energy_out,
                                      double *gradient_out);

/* Stage timings of the last profiled call, in the reference's stage names
 * (cpp/ising_cuda_structured_adjoint_modes.inc:317-370). */
size_t rie_engine_timing_count(const rie_engine *engine);
rie_status rie_engine_timing(const rie_engine *engine, size_t index,
                             const char **name_out, double *ms_out);

/* Time the structured forward pass issued normally vs. replayed from a HIP
 * graph.  Replaces RingIsingCudaBackend::benchmark_structured_forward_graph
 * (cpp/ising_cuda_backend.hpp:42-45, cpp/ising_cuda_backend.cu:444-541). */
rie_status rie_engine_benchmark_forward_graph(rie_engine *engine,
                                              const double *params,
                                              size_t num_params, size_t repeats,
                                              double *normal_forward_ms,
                                              double *graph_forward_ms);

/* One-shot helper constructing a temporary engine.  Replaces the free
 * function standalone_backend::energy_and_grad
 * (cpp/ising_cuda_backend.hpp:49-56, cpp/ising_cuda_backend.cu:543-554). */
rie_status rie_energy_and_grad(size_t num_qubits, size_t num_layers, double field,
                               const char *gradient_strategy, const double *params,
                               size_t num_params, int compute_gradient, int profile,
                               size_t structured_rotation_chunk_width,
                               double *energy_out, double *gradient_out);

/* Allocation-free energy_and_grad of one small circuit (structured_adjoint,
 * whole circuit in one workgroup).  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It calls no
 * hipMalloc, hipFree or allocating library routine: the coefficient table and
 * the state scratch (n = 13, 14) live in a file-scope device array of
 * RIE_NOALLOC_SCRATCH_BYTES per device, held until the process ends.  Every
 * call on a device shares that array, so calls that overlap on one device
 * race; serialize them (one stream, or wait between streams).
 * num_qubits in [2, RIE_NOALLOC_MAX_QUBITS], num_layers in
 * [1, RIE_NOALLOC_MAX_LAYERS]; larger problems get RIE_INVALID_ARGUMENT (use
 * rie_energy_and_grad or an engine).  d_params (2*n*L doubles), d_energy (1)
 * and d_gradient (2*n*L, or NULL for energy only) are device pointers; the
 * call is asynchronous on stream (a hipStream_t, NULL: the default stream).
 * Same results as rie_energy_and_grad with the structured_adjoint strategy
 * and the same structured_rotation_chunk_width (bitwise for n <= 12). */
#define RIE_NOALLOC_MAX_QUBITS 14
#define RIE_NOALLOC_MAX_LAYERS 1024
#define RIE_NOALLOC_SCRATCH_BYTES \
    (32u * RIE_NOALLOC_MAX_QUBITS * RIE_NOALLOC_MAX_LAYERS + 32u * (1u << RIE_NOALLOC_MAX_QUBITS))
rie_status rie_energy_and_grad_noalloc(size_t num_qubits, size_t num_layers, double field,
                                       const double *d_params, size_t num_params,
                                       size_t structured_rotation_chunk_width,
                                       double *d_energy, double *d_gradient, void *stream);

/* Engine properties. */
size_t rie_engine_num_qubits(const rie_engine *engine);
size_t rie_engine_num_layers(const rie_engine *engine);
int rie_engine_strategy(const rie_engine *engine);

/* ---- Device-resident entry points (no host synchronisation) ------------
 * These have no counterpart in the reference; they serve callers that keep
 * parameters on the GPU (the torch.autograd wrapper, captured optimizer
 * steps).  d_params, d_energy and d_gradient are device pointers; stream is
 * the hipStream_t to issue on (NULL: the default stream).  The calls are
 * asynchronous with respect to the host. */

/* Energy only; leaves the final state and lambda = H psi resident so that a
 * following rie_engine_backward_device can run the adjoint sweep without
 * recomputing the forward pass. */
rie_status rie_engine_forward_device(rie_engine *engine, const double *d_params,
                                     double *d_energy, void *stream);

/* Adjoint sweep on the state left by the last rie_engine_forward_device with
 * the same d_params; d_gradient receives 2*n*L values, scaled by *d_scale
 * when d_scale is non-NULL (the chain-rule factor of an autograd backward). */
rie_status rie_engine_backward_device(rie_engine *engine, const double *d_params,
                                      const double *d_scale, double *d_gradient,
                                      void *stream);

/* Forward + adjoint in one call. */
rie_status rie_engine_energy_and_grad_device(rie_engine *engine,
                                             const double *d_params,
                                             double *d_energy, double *d_gradient,
                                             void *stream);

/* How the backward sweep obtains the state at each pass (no counterpart in
 * the reference, whose sweep always walks back by inverse gates):
 *   RIE_SCHEDULE_INVERSE_WALK: recompute it by applying inverse gates
 *     (4 statevector buffers, plus checkpoints);
 *   RIE_SCHEDULE_STORED: keep every forward pass output and read it back
 *     ((L-1)(K+1)+2 buffers, K = high passes per layer; each backward pass
 *     then reads state and lambda and writes lambda only).
 * With layer-1 generation (rie_engine_set_layer1_generation) both schedules
 * finish layer 1 the same way: its low pass reads the state after layer 1,
 * its high passes generate the state they need.
 * A new engine uses STORED when it fits in 85% of free device memory. */
typedef enum rie_schedule {
    RIE_SCHEDULE_INVERSE_WALK = 0,
    RIE_SCHEDULE_STORED = 1
} rie_schedule;
rie_status rie_engine_set_schedule(rie_engine *engine, int schedule);
int rie_engine_schedule(const rie_engine *engine);

/* Fused Hamiltonian for gradient evaluations of the tile-pass path with at
 * least two layers (no counterpart in the reference; on by default): one
 * Hamiltonian pass over a scattered tile (ZZ diagonal and the X terms of 12
 * wires) replaces the per-group Hamiltonian passes, and the remaining X terms
 * are added, as adjacent-wire pairs of the pre-CNOT frame, by the last
 * layer's adjoint passes (cpp/ising_cuda_kernels.cu:313-358 computes all of
 * lambda = H psi in one kernel).  Energies and gradients agree with the
 * unfused evaluation to rounding.  enabled: 1 on, 0 off. */
rie_status rie_engine_set_fused_hamiltonian(rie_engine *engine, int enabled);
int rie_engine_fused_hamiltonian(const rie_engine *engine);

/* Layer-1 generation (no counterpart in the reference; on by default, tile
 * path with at least two layers and at most 31 qubits): layer 0 acts on
 * |0...0>, so every state inside layer 1 before its ring CNOT is a trace of
 * a ring of 2x2 matrices (one per wire) over the layer-0 product state.
 * Layer 1's forward becomes one write pass (generate, ring CNOT, store), and
 * in both backward schedules layer 1's adjoint high passes generate the state
 * they need instead of reading a stored or walked one (the reference generates only layer 0
 * as a product state and applies every later layer wire chunk by wire chunk:
 * cpp/ising_cuda_structured_adjoint_modes.inc:108-145).  Results agree
 * with the stored/walked evaluation to rounding.  enabled: 1 on, 0 off;
 * rie_engine_layer1_generation reports whether evaluations use it. */
rie_status rie_engine_set_layer1_generation(rie_engine *engine, int enabled);
int rie_engine_layer1_generation(const rie_engine *engine);

/* Generated start of two-layer gradient evaluations (no counterpart in the
 * reference, whose evaluations always run the forward sweep and H psi first:
 * cpp/ising_cuda_structured_adjoint_modes.inc:309-373).  With L == 2, the
 * fused Hamiltonian and layer-1 generation in use, energy_and_grad skips the
 * forward and Hamiltonian passes: layer 1's low adjoint pass generates the
 * pre-CNOT state of its tile and forms lambda = P^-1 H psi there (the ZZ
 * diagonal at P(i), the X terms as flips P^-1 e_w inside the tile or, for
 * flips leaving it, from the chain at the flipped tile), and the adjoint high
 * passes add no Hamiltonian terms, so the evaluation moves 6 state passes
 * instead of 11.  Results agree with the other paths to
 * rounding.  No final state is kept after such an evaluation
 * (rie_engine_copy_state then reports RIE_UNSUPPORTED; energy-only calls keep
 * it).  enabled: 1 on (default), 0 off; rie_engine_generated_start reports
 * whether gradient evaluations use it. */
rie_status rie_engine_set_generated_start(rie_engine *engine, int enabled);
int rie_engine_generated_start(const rie_engine *engine);

/* Lambda layouts of the generated start (no counterpart in the reference,
 * whose passes keep every vector in index order).  Between the passes of a
 * generated-start evaluation lambda is stored in the order that makes the
 * next high pass's tiles contiguous 64 KiB blocks (index bits 0..2 and that
 * pass's nine high bits at the bottom, the writing pass's tile bits right
 * above them), so every high pass reads whole blocks and writes 128-byte
 * runs within 32 MB per tile; the last high pass writes the natural order (or
 * nothing, see rie_engine_set_layer0_in_pass).  Results are bitwise identical
 * either way.  enabled: 1
 * on (default), 0 off; rie_engine_lambda_layouts reports whether gradient
 * evaluations use them (the generated start in use and the pass geometry
 * suited to them). */
rie_status rie_engine_set_lambda_layouts(rie_engine *engine, int enabled);
int rie_engine_lambda_layouts(const rie_engine *engine);

/* Layer 0 inside the last high pass of the generated start (no counterpart
 * in the reference, whose layer-0 gradients come from the backward sweep of
 * layer 0).  The generated start runs its high passes from the lowest group
 * up, so the last one holds the top group's tiles, which under the inverse
 * ring CNOT are product domains of the layer-0 coordinates (up to one parity
 * bit): that pass contracts lambda with the layer-0 product state instead of
 * storing it, and the separate 1 S layer-0 contraction pass is not run.
 * enabled: 1 on (default), 0 off; rie_engine_layer0_in_pass reports whether
 * gradient evaluations use it (the generated start in use, register shape 4
 * for the high passes). */
rie_status rie_engine_set_layer0_in_pass(rie_engine *engine, int enabled);
int rie_engine_layer0_in_pass(const rie_engine *engine);

/* Double-buffered adjoint high passes.  Replaces the double_buffer flag of
 * RingIsingCudaBackend (cpp/ising_cuda_backend.cu:374-419), which selects a
 * backward rotation-chunk kernel that stages the next chunk while computing
 * the current one (cpp/ising_cuda_structured_adjoint_modes.inc:174-185,
 * cpp/ising_cuda_statevector_grad_kernels.cu:1742-1925).  Here the
 * lambda-only high passes of the structured_adjoint backward sweep load each
 * workgroup's next tile of lambda into the registers that the current tile's
 * state frees after its gradients, so that load overlaps the rest of the
 * tile.  The arithmetic is unchanged: results are bitwise identical with and
 * without.  enabled: 1 on, 0 off; set at creation by double_buffer and
 * reported by rie_engine_double_buffer. */
rie_status rie_engine_set_double_buffer(rie_engine *engine, int enabled);
int rie_engine_double_buffer(const rie_engine *engine);

/* Register shape of the tile passes (no counterpart in the reference): each
 * of r_fwd, r_ham, r_bwd is 4 (256 threads x 16 registers per 4096-amplitude
 * tile, three register mappings) or 3 (512 threads x 8 registers, four
 * mappings: twice the waves per SIMD, one more LDS exchange).  Results agree
 * to rounding.  Defaults 3, 3, 4. */
rie_status rie_engine_set_tile_shapes(rie_engine *engine, int r_fwd, int r_ham, int r_bwd);
rie_status rie_engine_tile_shapes(const rie_engine *engine, int *shapes_out3);

/* The same per pass kind (no counterpart in the reference): shapes[i] is the
 * register shape (3 or 4) of kind i of rie_pass_kind, count = RIE_PASS_KINDS.
 * rie_engine_set_tile_shapes sets the three forward kinds to r_fwd, the
 * Hamiltonian kind to r_ham and the two backward kinds to r_bwd (the layer-0
 * contraction keeps its shape);
 * rie_engine_tile_shapes reports the high-pass forward, Hamiltonian and
 * high-pass backward kinds.  Defaults per kind: 4, 3, 3, 3, 4, 4, 4. */
typedef enum rie_pass_kind {
    RIE_PASS_FWD_GEN = 0,   /* layer 1 generated (one write pass), or without layer-1
                               generation layer 1's first pass (layer 0 generated) */
    RIE_PASS_FWD_HIGH = 1,  /* forward, high wires */
    RIE_PASS_FWD_LOW = 2,   /* forward, low wires + ring CNOT */
    RIE_PASS_HAM = 3,       /* Hamiltonian passes */
    RIE_PASS_BWD_LOW = 4,   /* adjoint, inverse ring CNOT + low wires */
    RIE_PASS_BWD_HIGH = 5,  /* adjoint, high wires */
    RIE_PASS_LAYER0 = 6,    /* layer-0 gradient contraction (a read pass) */
    RIE_PASS_KINDS = 7
} rie_pass_kind;
rie_status rie_engine_set_pass_shapes(rie_engine *engine, const int *shapes, size_t count);
rie_status rie_engine_pass_shapes(const rie_engine *engine, int *shapes_out, size_t count);

/* Autotuner for the tile-pass path: times energy_and_grad on `params` (host,
 * num_params values) `repeats` times for each tile shape combination and each
 * backward schedule that fits in memory, keeps the fastest, and writes a JSON
 * report (truncated to report_len; report_needed receives the full size).
 * Engines on the whole-circuit path (n <= 12) or other strategies report
 * {"tuned": false}.  Replaces the reference's fixed qubit-count heuristics
 * (cpp/ising_cuda_rotation_kernels.cu:562-575). */
rie_status rie_engine_autotune(rie_engine *engine, const double *params, size_t num_params,
                              size_t repeats, char *report_out, size_t report_len,
                              size_t *report_needed);

/* Store the statevector at every k-th layer boundary during the forward pass
 * and restore it there during the backward sweep instead of continuing the
 * inverse walk (k = 0 disables).  Costs one extra state buffer per stored
 * boundary. */
rie_status rie_engine_set_checkpoint_interval(rie_engine *engine, size_t every_k_layers);

/* Run energy_and_grad with a full inverse walk through layer 0 and report
 * || psi_walked - |0...0> ||_2, the accumulated drift of recomputing states
 * by inverse gates (host pointers, synchronous). */
rie_status rie_engine_inverse_walk_drift(rie_engine *engine, const double *params,
                                         size_t num_params, double *drift_out);

/* Per-launch kernel timing (no counterpart in the reference; for benchmarks).
 * When enabled, every kernel launch of the structured path is bracketed by
 * HIP events on its launch stream and accumulated per kernel name together
 * with the launch's algorithmic HBM bytes (state arrays read + written). */
rie_status rie_engine_set_kernel_timing(rie_engine *engine, int enable);
rie_status rie_engine_reset_kernel_stats(rie_engine *engine);
rie_status rie_engine_kernel_stat_count(rie_engine *engine, size_t *count_out);
rie_status rie_engine_kernel_stat(rie_engine *engine, size_t index, const char **name_out,
                                  long *launches_out, double *total_ms_out,
                                  double *total_bytes_out);

/* Copy the final statevector of the last forward pass to host memory
 * (2^n complex128 values, interleaved re/im). */
rie_status rie_engine_copy_state(rie_engine *engine, double *host_out, size_t num_doubles);

/* ---- Batches of independent circuits -----------------------------------
 * No counterpart in the reference, which evaluates one circuit per
 * RingIsingCudaBackend::energy_and_grad call (cpp/ising_cuda_backend.cu:
 * 429-442).  A batch holds `count` circuits with their own qubit count
 * (2..14), layer count and field; circuits with num_qubits <=
 * dense_max_qubits (0..8) use dense_scan, the others structured_adjoint.
 * All circuits are evaluated together: the structured ones in one persistent
 * launch, the dense ones level by level in grouped FP64-MFMA GEMM launches,
 * concurrently.  Parameters, energies and gradients live in device arrays
 * owned by the batch; circuit i's 2*n_i*L_i parameters (and gradients) start
 * at offsets[i] and follow the single-circuit layout. */
typedef struct rie_batch rie_batch;

rie_status rie_batch_create(size_t count, const int *num_qubits, const int *num_layers,
                            const double *fields, int dense_max_qubits, rie_batch **batch_out);
void rie_batch_destroy(rie_batch *batch);
size_t rie_batch_size(const rie_batch *batch);
size_t rie_batch_num_params(const rie_batch *batch);
/* offsets_out receives count + 1 values (the last is the total). */
rie_status rie_batch_param_offsets(const rie_batch *batch, size_t *offsets_out, size_t len);
/* strategy_out receives count values (RIE_STRATEGY_*). */
rie_status rie_batch_strategies(const rie_batch *batch, int *strategy_out, size_t len);
/* Device addresses of the batch's parameter, energy and gradient arrays
 * (any output may be NULL). */
rie_status rie_batch_device_buffers(rie_batch *batch, double **d_params, double **d_energies,
                                    double **d_gradient);
rie_status rie_batch_set_params(rie_batch *batch, const double *params, size_t num_params);
rie_status rie_batch_get_params(rie_batch *batch, double *params_out, size_t num_params);
/* Energies (count values) and, if gradient_out is non-NULL, the packed
 * gradient of the last evaluation, after synchronising the device. */
rie_status rie_batch_get_results(rie_batch *batch, double *energies_out, double *gradient_out);

/* Host pointers, synchronous: upload params, evaluate, download energies
 * (count values) and, if compute_gradient, the packed gradient. */
rie_status rie_batch_energy_and_grad(rie_batch *batch, const double *params, size_t num_params,
                                     int compute_gradient, double *energies_out,
                                     double *gradient_out);
/* Evaluate on the batch's device arrays, asynchronously on stream (NULL: the
 * default stream); replays a captured HIP graph when enabled. */
rie_status rie_batch_run_device(rie_batch *batch, int compute_gradient, void *stream);
rie_status rie_batch_set_use_graph(rie_batch *batch, int use_graph);
/* The same evaluation issued as separate launches per circuit, one after the
 * other (the one-launch-per-circuit comparison). */
rie_status rie_batch_run_per_circuit(rie_batch *batch, int compute_gradient, void *stream);

/* `steps` optimizer steps on every circuit at once, starting from the
 * batch's current parameters: optimizer 0 = gradient descent p -= lr g
 * (the reference's update, ring_ising/workflows/common.py:23-25), 1 = Adam
 * (beta 0.9 / 0.999, eps 1e-8, fresh moments, as torch.optim.Adam).  With
 * use_graph one step is captured in a HIP graph and replayed.  The energy of
 * every circuit before each step is recorded; energy_history_out (host,
 * steps * count values, may be NULL) receives it and the call then
 * synchronises. */
rie_status rie_batch_optimize(rie_batch *batch, int optimizer, double learning_rate, size_t steps,
                              int use_graph, double *energy_history_out, void *stream);

/* Autotuner: times the whole-circuit kernel per qubit count for every
 * workgroup size (256/512/1024) and rotation-chunk width (1..4), then, for
 * each split of the circuits between dense_scan (num_qubits <= s) and the
 * whole-circuit kernel, s from the dense_max_qubits given at creation down to
 * 0 (see rie_batch_set_tune_split), the dense GEMM tile (32/64) and the share
 * of CUs given to the structured launch; keeps the fastest whole evaluation
 * and rebuilds the schedule with the measured costs.  The reference fixes the
 * strategy per call and its kernel shapes by qubit-count thresholds
 * (cpp/ising_cuda_rotation_kernels.cu:562-575). */
rie_status rie_batch_autotune(rie_batch *batch, size_t repeats);
/* Whether rie_batch_autotune may move circuits between dense_scan and the
 * whole-circuit kernel (default 1; 0 keeps num_qubits <= dense_max_qubits on
 * dense_scan).  No counterpart in the reference, which has no batch. */
rie_status rie_batch_set_tune_split(rie_batch *batch, int enabled);
/* Manual configuration: threads 256/512/1024; chunk_by_n (15 values,
 * indexed by num_qubits, each 1..4) or NULL to keep; dense_tile 32/64;
 * structured_grid = workgroups of the persistent launch (0: one per CU). */
rie_status rie_batch_set_config(rie_batch *batch, int threads, const int *chunk_by_n,
                                size_t chunk_len, int dense_tile, int structured_grid);
/* Current configuration and the last autotuning measurements as JSON. */
rie_status rie_batch_config(const rie_batch *batch, char *json_out, size_t len,
                            size_t *needed_out);

/* ---- Warm-start MLP ------------------------------------------------------
 * No counterpart in the reference.  An FP32 perceptron
 * inputs -> GELU(hidden) -> GELU(hidden) -> outputs on hand-written HIP
 * kernels, used to predict starting parameters of a VQE run from
 * (h/J, n, L).  Parameters are packed W1 [in x hidden], b1, W2 [hidden x
 * hidden], b2, W3 [hidden x out], b3 (row-major; y = x W + b), initialised
 * U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from seed.  Inputs and outputs are
 * host arrays with one sample per row. */
typedef struct rie_mlp rie_mlp;

rie_status rie_mlp_create(int inputs, int hidden, int outputs, uint64_t seed, rie_mlp **mlp_out);
void rie_mlp_destroy(rie_mlp *mlp);
size_t rie_mlp_num_params(const rie_mlp *mlp);
rie_status rie_mlp_get_params(rie_mlp *mlp, float *params_out, size_t count);
rie_status rie_mlp_set_params(rie_mlp *mlp, const float *params, size_t count);
/* outputs: rows x outputs values. */
rie_status rie_mlp_forward(rie_mlp *mlp, const float *inputs, int rows, float *outputs);
/* Mean squared error over all rows x outputs values and its gradient with
 * respect to the packed parameters (grad_out may be NULL). */
rie_status rie_mlp_loss_and_grad(rie_mlp *mlp, const float *inputs, const float *targets,
                                 int rows, float *loss_out, float *grad_out);
/* `steps` full-batch Adam steps (beta 0.9 / 0.999, eps 1e-8, fresh moments)
 * on the mean squared error; loss_history_out (steps values, may be NULL)
 * receives the loss before each step. */
rie_status rie_mlp_train(rie_mlp *mlp, const float *inputs, const float *targets, int rows,
                         int steps, float learning_rate, float *loss_history_out);

#ifdef __cplusplus
}
#endif

#endif /* RIE_H */
