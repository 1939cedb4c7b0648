// pybind11 module `_cuda_backend`: the Python surface of the reference's
// extension (cpp/ising_cuda_bindings.cpp:71-143), implemented on top of the
// C ABI in include/rie.h only.  Built with the host compiler; no HIP headers.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rie.h"

namespace py = pybind11;

namespace {

using FlatArray = py::array_t<double, py::array::c_style | py::array::forcecast>;

// RIE_INVALID_ARGUMENT -> ValueError (std::invalid_argument in the
// reference), every other failure -> RuntimeError (std::runtime_error).
void check(rie_status st) {
    if (st == RIE_OK) {
        return;
    }
    const std::string msg = rie_last_error();
    if (st == RIE_INVALID_ARGUMENT) {
        throw py::value_error(msg);
    }
    throw std::runtime_error(msg);
}

struct ParamsView {
    const double *ptr;
    size_t size;
};

ParamsView view_params(const FlatArray &params) {
    const auto info = params.request();
    return {static_cast<const double *>(info.ptr), static_cast<size_t>(info.size)};
}

// reference: validate_params_shape (cpp/ising_cuda_bindings.cpp:27-37)
void validate_params_shape(size_t num_qubits, size_t num_layers, const FlatArray &params) {
    const auto info = params.request();
    if (info.ndim != 1) {
        throw py::value_error("params must be a flat 1D float64 array.");
    }
    if (static_cast<size_t>(info.size) != num_qubits * num_layers * 2) {
        throw py::value_error("params does not match num_qubits/layers.");
    }
}

// reference: make_energy_grad_dict (cpp/ising_cuda_bindings.cpp:39-55)
py::dict make_result(double energy, const std::vector<double> &grad,
                     const std::vector<std::pair<std::string, double>> &timings) {
    py::array_t<double> g(static_cast<py::ssize_t>(grad.size()));
    if (!grad.empty()) {
        std::memcpy(g.mutable_data(), grad.data(), grad.size() * sizeof(double));
    }
    py::dict out;
    out["energy"] = py::float_(energy);
    out["gradient"] = std::move(g);
    if (!timings.empty()) {
        py::dict t;
        for (const auto &kv : timings) {
            t[py::str(kv.first)] = py::float_(kv.second);
        }
        out["timings_ms"] = std::move(t);
    }
    return out;
}

std::vector<std::pair<std::string, double>> read_timings(const rie_engine *e) {
    std::vector<std::pair<std::string, double>> out;
    const size_t count = rie_engine_timing_count(e);
    for (size_t i = 0; i < count; ++i) {
        const char *name = nullptr;
        double ms = 0.0;
        if (rie_engine_timing(e, i, &name, &ms) == RIE_OK && name != nullptr) {
            out.emplace_back(name, ms);
        }
    }
    return out;
}

class RingIsingCudaBackend {
  public:
    RingIsingCudaBackend(size_t num_qubits, size_t num_layers, double field,
                         const std::string &gradient_strategy, size_t chunk_width,
                         bool double_buffer)
        : num_qubits_(num_qubits), num_layers_(num_layers) {
        check(rie_engine_create(num_qubits, num_layers, field, gradient_strategy.c_str(),
                                chunk_width, double_buffer ? 1 : 0, &engine_));
    }
    ~RingIsingCudaBackend() { rie_engine_destroy(engine_); }
    RingIsingCudaBackend(const RingIsingCudaBackend &) = delete;
    RingIsingCudaBackend &operator=(const RingIsingCudaBackend &) = delete;

    py::dict energy_and_grad(const FlatArray &params, bool compute_gradient, bool profile) {
        const auto view = view_params(params);
        double energy = 0.0;
        std::vector<double> grad(compute_gradient ? view.size : 0);
        rie_status st;
        {
            py::gil_scoped_release release;
            st = rie_engine_energy_and_grad(engine_, view.ptr, view.size, compute_gradient ? 1 : 0,
                                            profile ? 1 : 0, &energy,
                                            compute_gradient ? grad.data() : nullptr);
        }
        check(st);
        return make_result(energy, grad, profile ? read_timings(engine_) : decltype(read_timings(engine_)){});
    }

    py::dict benchmark_structured_forward_graph(const FlatArray &params, size_t repeats) {
        const auto view = view_params(params);
        double normal = 0.0;
        double graph = 0.0;
        rie_status st;
        {
            py::gil_scoped_release release;
            st = rie_engine_benchmark_forward_graph(engine_, view.ptr, view.size, repeats, &normal,
                                                    &graph);
        }
        check(st);
        py::dict out;
        out["normal_forward_ms"] = py::float_(normal);
        out["graph_forward_ms"] = py::float_(graph);
        out["speedup"] = py::float_(graph > 0.0 ? normal / graph : 0.0);
        return out;
    }

    // Device-resident entry points (pointers as integers, e.g. tensor.data_ptr()).
    void forward_device(std::uintptr_t d_params, std::uintptr_t d_energy, std::uintptr_t stream) {
        rie_status st;
        {
            py::gil_scoped_release release;
            st = rie_engine_forward_device(engine_, reinterpret_cast<const double *>(d_params),
                                           reinterpret_cast<double *>(d_energy),
                                           reinterpret_cast<void *>(stream));
        }
        check(st);
    }

    void backward_device(std::uintptr_t d_params, std::uintptr_t d_scale, std::uintptr_t d_grad,
                         std::uintptr_t stream) {
        rie_status st;
        {
            py::gil_scoped_release release;
            st = rie_engine_backward_device(engine_, reinterpret_cast<const double *>(d_params),
                                            reinterpret_cast<const double *>(d_scale),
                                            reinterpret_cast<double *>(d_grad),
                                            reinterpret_cast<void *>(stream));
        }
        check(st);
    }

    void energy_and_grad_device(std::uintptr_t d_params, std::uintptr_t d_energy,
                                std::uintptr_t d_grad, std::uintptr_t stream) {
        rie_status st;
        {
            py::gil_scoped_release release;
            st = rie_engine_energy_and_grad_device(
                engine_, reinterpret_cast<const double *>(d_params),
                reinterpret_cast<double *>(d_energy), reinterpret_cast<double *>(d_grad),
                reinterpret_cast<void *>(stream));
        }
        check(st);
    }

    void set_checkpoint_interval(size_t every) {
        check(rie_engine_set_checkpoint_interval(engine_, every));
    }

    double inverse_walk_drift(const FlatArray &params) {
        const auto view = view_params(params);
        double drift = 0.0;
        rie_status st;
        {
            py::gil_scoped_release release;
            st = rie_engine_inverse_walk_drift(engine_, view.ptr, view.size, &drift);
        }
        check(st);
        return drift;
    }

    py::array_t<std::complex<double>> final_state() {
        const size_t dim = size_t{1} << num_qubits_;
        py::array_t<std::complex<double>> out(static_cast<py::ssize_t>(dim));
        check(rie_engine_copy_state(engine_, reinterpret_cast<double *>(out.mutable_data()),
                                    2 * dim));
        return out;
    }

    void set_schedule(const std::string &name) {
        int sched;
        if (name == "inverse_walk") {
            sched = RIE_SCHEDULE_INVERSE_WALK;
        } else if (name == "stored") {
            sched = RIE_SCHEDULE_STORED;
        } else {
            throw py::value_error("schedule must be 'inverse_walk' or 'stored'");
        }
        check(rie_engine_set_schedule(engine_, sched));
    }

    std::string schedule() const {
        return rie_engine_schedule(engine_) == RIE_SCHEDULE_STORED ? "stored" : "inverse_walk";
    }

    void set_fused_hamiltonian(int enabled) {
        check(rie_engine_set_fused_hamiltonian(engine_, enabled));
    }

    bool fused_hamiltonian() const { return rie_engine_fused_hamiltonian(engine_) == 1; }

    void set_layer1_generation(int enabled) {
        check(rie_engine_set_layer1_generation(engine_, enabled));
    }

    bool layer1_generation() const { return rie_engine_layer1_generation(engine_) == 1; }

    void set_generated_start(int enabled) {
        check(rie_engine_set_generated_start(engine_, enabled));
    }

    bool generated_start() const { return rie_engine_generated_start(engine_) == 1; }

    void set_double_buffer(int enabled) { check(rie_engine_set_double_buffer(engine_, enabled)); }

    bool double_buffer() const { return rie_engine_double_buffer(engine_) == 1; }

    void set_lambda_layouts(int enabled) {
        check(rie_engine_set_lambda_layouts(engine_, enabled));
    }

    bool lambda_layouts() const { return rie_engine_lambda_layouts(engine_) == 1; }

    void set_layer0_in_pass(int enabled) {
        check(rie_engine_set_layer0_in_pass(engine_, enabled));
    }

    bool layer0_in_pass() const { return rie_engine_layer0_in_pass(engine_) == 1; }

    void set_kernel_timing(bool enable) { check(rie_engine_set_kernel_timing(engine_, enable)); }

    void reset_kernel_stats() { check(rie_engine_reset_kernel_stats(engine_)); }

    py::list kernel_stats() {
        size_t count = 0;
        check(rie_engine_kernel_stat_count(engine_, &count));
        py::list out;
        for (size_t i = 0; i < count; ++i) {
            const char *name = nullptr;
            long launches = 0;
            double ms = 0.0;
            double bytes = 0.0;
            check(rie_engine_kernel_stat(engine_, i, &name, &launches, &ms, &bytes));
            py::dict d;
            d["name"] = py::str(name);
            d["launches"] = py::int_(launches);
            d["total_ms"] = py::float_(ms);
            d["total_bytes"] = py::float_(bytes);
            out.append(std::move(d));
        }
        return out;
    }

    void set_tile_shapes(int r_fwd, int r_ham, int r_bwd) {
        check(rie_engine_set_tile_shapes(engine_, r_fwd, r_ham, r_bwd));
    }

    py::object autotune(const FlatArray &params, size_t repeats) {
        validate_params_shape(num_qubits_, num_layers_, params);
        const auto view = view_params(params);
        size_t need = 0;
        std::string buf(4096, '\0');
        rie_status st;
        {
            py::gil_scoped_release release;
            st = rie_engine_autotune(engine_, view.ptr, view.size, repeats, &buf[0], buf.size(),
                                     &need);
        }
        check(st);
        buf.resize(std::strlen(buf.c_str()));
        return py::module_::import("json").attr("loads")(buf);
    }

    std::vector<int> tile_shapes() const {
        std::vector<int> out(3);
        check(rie_engine_tile_shapes(engine_, out.data()));
        return out;
    }

    void set_pass_shapes(const std::vector<int> &shapes) {
        check(rie_engine_set_pass_shapes(engine_, shapes.data(), shapes.size()));
    }

    std::vector<int> pass_shapes() const {
        std::vector<int> out(RIE_PASS_KINDS);
        check(rie_engine_pass_shapes(engine_, out.data(), out.size()));
        return out;
    }

    size_t num_qubits() const { return num_qubits_; }
    size_t num_layers() const { return num_layers_; }

  private:
    rie_engine *engine_ = nullptr;
    size_t num_qubits_;
    size_t num_layers_;
};

// Many independent circuits evaluated together (C ABI rie_batch_*).
class RingIsingBatch {
  public:
    RingIsingBatch(const std::vector<int> &num_qubits, const std::vector<int> &num_layers,
                   const std::vector<double> &fields, int dense_max_qubits) {
        if (num_qubits.size() != num_layers.size() || num_qubits.size() != fields.size()) {
            throw py::value_error("num_qubits, layers and fields must have the same length.");
        }
        check(rie_batch_create(num_qubits.size(), num_qubits.data(), num_layers.data(),
                               fields.data(), dense_max_qubits, &batch_));
    }
    ~RingIsingBatch() { rie_batch_destroy(batch_); }
    RingIsingBatch(const RingIsingBatch &) = delete;
    RingIsingBatch &operator=(const RingIsingBatch &) = delete;

    size_t size() const { return rie_batch_size(batch_); }
    size_t num_params() const { return rie_batch_num_params(batch_); }

    py::array_t<uint64_t> param_offsets() const {
        std::vector<size_t> off(size() + 1);
        check(rie_batch_param_offsets(batch_, off.data(), off.size()));
        py::array_t<uint64_t> out(off.size());
        auto v = out.mutable_unchecked<1>();
        for (size_t i = 0; i < off.size(); ++i) {
            v(i) = off[i];
        }
        return out;
    }

    std::vector<std::string> strategies() const {
        std::vector<int> st(size());
        check(rie_batch_strategies(batch_, st.data(), st.size()));
        std::vector<std::string> out;
        for (int x : st) {
            out.push_back(x == RIE_STRATEGY_DENSE_SCAN ? "dense_scan" : "structured_adjoint");
        }
        return out;
    }

    py::tuple energy_and_grad(const FlatArray &params, bool compute_gradient) {
        const auto view = view_params(params);
        if (params.ndim() != 1 || view.size != num_params()) {
            throw py::value_error("params must be a flat array of num_params values.");
        }
        py::array_t<double> energies(size());
        py::array_t<double> grad(compute_gradient ? num_params() : 0);
        double *e = energies.mutable_data();
        double *g = compute_gradient ? grad.mutable_data() : nullptr;
        rie_status st;
        {
            py::gil_scoped_release release;
            st = rie_batch_energy_and_grad(batch_, view.ptr, view.size, compute_gradient ? 1 : 0,
                                           e, g);
        }
        check(st);
        return py::make_tuple(energies, grad);
    }

    void set_params(const FlatArray &params) {
        const auto view = view_params(params);
        check(rie_batch_set_params(batch_, view.ptr, view.size));
    }

    py::array_t<double> get_params() {
        py::array_t<double> out(num_params());
        check(rie_batch_get_params(batch_, out.mutable_data(), num_params()));
        return out;
    }

    py::tuple results() {
        py::array_t<double> energies(size());
        py::array_t<double> grad(num_params());
        check(rie_batch_get_results(batch_, energies.mutable_data(), grad.mutable_data()));
        return py::make_tuple(energies, grad);
    }

    py::dict device_buffers() {
        double *p = nullptr;
        double *e = nullptr;
        double *g = nullptr;
        check(rie_batch_device_buffers(batch_, &p, &e, &g));
        py::dict d;
        d["params"] = reinterpret_cast<uintptr_t>(p);
        d["energies"] = reinterpret_cast<uintptr_t>(e);
        d["gradient"] = reinterpret_cast<uintptr_t>(g);
        return d;
    }

    void run_device(bool compute_gradient, uintptr_t stream) {
        py::gil_scoped_release release;
        check(rie_batch_run_device(batch_, compute_gradient ? 1 : 0,
                                   reinterpret_cast<void *>(stream)));
    }

    void run_per_circuit(bool compute_gradient, uintptr_t stream) {
        py::gil_scoped_release release;
        check(rie_batch_run_per_circuit(batch_, compute_gradient ? 1 : 0,
                                        reinterpret_cast<void *>(stream)));
    }

    void set_use_graph(bool use) { check(rie_batch_set_use_graph(batch_, use ? 1 : 0)); }

    py::array_t<double> optimize(const std::string &optimizer, double lr, size_t steps,
                                 bool use_graph, uintptr_t stream) {
        int kind;
        if (optimizer == "gd" || optimizer == "gradient_descent") {
            kind = 0;
        } else if (optimizer == "adam") {
            kind = 1;
        } else {
            throw py::value_error("optimizer must be 'gd' or 'adam'.");
        }
        py::array_t<double> hist({steps, size()});
        double *h = hist.mutable_data();
        rie_status st;
        {
            py::gil_scoped_release release;
            st = rie_batch_optimize(batch_, kind, lr, steps, use_graph ? 1 : 0, h,
                                    reinterpret_cast<void *>(stream));
        }
        check(st);
        return hist;
    }

    py::object autotune(size_t repeats) {
        {
            py::gil_scoped_release release;
            check(rie_batch_autotune(batch_, repeats));
        }
        return config();
    }

    void set_tune_split(int enabled) { check(rie_batch_set_tune_split(batch_, enabled)); }

    py::object config() const {
        size_t need = 0;
        check(rie_batch_config(batch_, nullptr, 0, &need));
        std::string buf(need, '\0');
        check(rie_batch_config(batch_, &buf[0], buf.size(), nullptr));
        buf.resize(std::strlen(buf.c_str()));
        return py::module_::import("json").attr("loads")(buf);
    }

    void set_config(int threads, py::object chunk_by_n, int dense_tile, int structured_grid) {
        std::vector<int> chunks;
        const int *cp = nullptr;
        if (!chunk_by_n.is_none()) {
            chunks = chunk_by_n.cast<std::vector<int>>();
            cp = chunks.data();
        }
        check(rie_batch_set_config(batch_, threads, cp, chunks.size(), dense_tile,
                                   structured_grid));
    }

  private:
    rie_batch *batch_ = nullptr;
};

// FP32 warm-start perceptron (C ABI rie_mlp_*).
class WarmStartMLP {
  public:
    using F32Array = py::array_t<float, py::array::c_style | py::array::forcecast>;

    WarmStartMLP(int inputs, int hidden, int outputs, uint64_t seed)
        : inputs_(inputs), outputs_(outputs) {
        check(rie_mlp_create(inputs, hidden, outputs, seed, &mlp_));
    }
    ~WarmStartMLP() { rie_mlp_destroy(mlp_); }
    WarmStartMLP(const WarmStartMLP &) = delete;
    WarmStartMLP &operator=(const WarmStartMLP &) = delete;

    size_t num_params() const { return rie_mlp_num_params(mlp_); }

    py::array_t<float> get_params() {
        py::array_t<float> out(num_params());
        check(rie_mlp_get_params(mlp_, out.mutable_data(), num_params()));
        return out;
    }

    void set_params(const F32Array &p) {
        check(rie_mlp_set_params(mlp_, p.data(), (size_t)p.size()));
    }

    int rows_of(const F32Array &x, int width) const {
        if (x.ndim() != 2 || x.shape(1) != width) {
            throw py::value_error("expected a 2-D array with " + std::to_string(width) +
                                  " columns.");
        }
        return (int)x.shape(0);
    }

    py::array_t<float> forward(const F32Array &x) {
        const int rows = rows_of(x, inputs_);
        py::array_t<float> y({(py::ssize_t)rows, (py::ssize_t)outputs_});
        float *yp = y.mutable_data();
        rie_status st;
        {
            py::gil_scoped_release release;
            st = rie_mlp_forward(mlp_, x.data(), rows, yp);
        }
        check(st);
        return y;
    }

    py::tuple loss_and_grad(const F32Array &x, const F32Array &t) {
        const int rows = rows_of(x, inputs_);
        if (rows_of(t, outputs_) != rows) {
            throw py::value_error("inputs and targets need the same number of rows.");
        }
        float loss = 0.0f;
        py::array_t<float> g(num_params());
        check(rie_mlp_loss_and_grad(mlp_, x.data(), t.data(), rows, &loss, g.mutable_data()));
        return py::make_tuple(loss, g);
    }

    py::array_t<float> train(const F32Array &x, const F32Array &t, int steps, float lr) {
        const int rows = rows_of(x, inputs_);
        if (rows_of(t, outputs_) != rows) {
            throw py::value_error("inputs and targets need the same number of rows.");
        }
        py::array_t<float> hist(steps > 0 ? steps : 0);
        float *hp = steps > 0 ? hist.mutable_data() : nullptr;
        rie_status st;
        {
            py::gil_scoped_release release;
            st = rie_mlp_train(mlp_, x.data(), t.data(), rows, steps, lr, hp);
        }
        check(st);
        return hist;
    }

  private:
    rie_mlp *mlp_ = nullptr;
    int inputs_;
    int outputs_;
};

}  // namespace

PYBIND11_MODULE(_cuda_backend, m) {
    m.doc() = "MI355X (gfx950) ring-Ising statevector engine; C ABI: include/rie.h.";

    py::class_<RingIsingCudaBackend>(m, "RingIsingCudaBackend")
        .def(py::init<size_t, size_t, double, const std::string &, size_t, bool>(),
             py::arg("num_qubits"), py::arg("num_layers"), py::arg("field"),
             py::arg("gradient_strategy") = "structured_adjoint",
             py::arg("structured_rotation_chunk_width") = 8, py::arg("double_buffer") = false)
        .def("energy_and_grad", &RingIsingCudaBackend::energy_and_grad, py::arg("params"),
             py::arg("compute_gradient") = true, py::arg("profile") = false)
        .def("benchmark_structured_forward_graph",
             &RingIsingCudaBackend::benchmark_structured_forward_graph, py::arg("params"),
             py::arg("repeats") = 100)
        .def("forward_device", &RingIsingCudaBackend::forward_device, py::arg("params_ptr"),
             py::arg("energy_ptr"), py::arg("stream") = 0)
        .def("backward_device", &RingIsingCudaBackend::backward_device, py::arg("params_ptr"),
             py::arg("scale_ptr"), py::arg("grad_ptr"), py::arg("stream") = 0)
        .def("energy_and_grad_device", &RingIsingCudaBackend::energy_and_grad_device,
             py::arg("params_ptr"), py::arg("energy_ptr"), py::arg("grad_ptr"),
             py::arg("stream") = 0)
        .def("set_checkpoint_interval", &RingIsingCudaBackend::set_checkpoint_interval,
             py::arg("every_k_layers"))
        .def("inverse_walk_drift", &RingIsingCudaBackend::inverse_walk_drift, py::arg("params"))
        .def("final_state", &RingIsingCudaBackend::final_state)
        .def("set_schedule", &RingIsingCudaBackend::set_schedule, py::arg("schedule"))
        .def("set_fused_hamiltonian", &RingIsingCudaBackend::set_fused_hamiltonian,
             py::arg("enabled"))
        .def_property_readonly("fused_hamiltonian", &RingIsingCudaBackend::fused_hamiltonian)
        .def("set_layer1_generation", &RingIsingCudaBackend::set_layer1_generation,
             py::arg("enabled"))
        .def_property_readonly("layer1_generation", &RingIsingCudaBackend::layer1_generation)
        .def("set_generated_start", &RingIsingCudaBackend::set_generated_start,
             py::arg("enabled"))
        .def_property_readonly("generated_start", &RingIsingCudaBackend::generated_start)
        .def("set_double_buffer", &RingIsingCudaBackend::set_double_buffer, py::arg("enabled"))
        .def_property_readonly("double_buffer", &RingIsingCudaBackend::double_buffer)
        .def("set_lambda_layouts", &RingIsingCudaBackend::set_lambda_layouts, py::arg("enabled"))
        .def_property_readonly("lambda_layouts", &RingIsingCudaBackend::lambda_layouts)
        .def("set_layer0_in_pass", &RingIsingCudaBackend::set_layer0_in_pass, py::arg("enabled"))
        .def_property_readonly("layer0_in_pass", &RingIsingCudaBackend::layer0_in_pass)
        .def_property_readonly("schedule", &RingIsingCudaBackend::schedule)
        .def("set_kernel_timing", &RingIsingCudaBackend::set_kernel_timing, py::arg("enable"))
        .def("reset_kernel_stats", &RingIsingCudaBackend::reset_kernel_stats)
        .def("kernel_stats", &RingIsingCudaBackend::kernel_stats)
        .def("set_tile_shapes", &RingIsingCudaBackend::set_tile_shapes, py::arg("r_fwd"),
             py::arg("r_ham"), py::arg("r_bwd"))
        .def_property_readonly("tile_shapes", &RingIsingCudaBackend::tile_shapes)
        .def("set_pass_shapes", &RingIsingCudaBackend::set_pass_shapes, py::arg("shapes"))
        .def_property_readonly("pass_shapes", &RingIsingCudaBackend::pass_shapes)
        .def("autotune", &RingIsingCudaBackend::autotune, py::arg("params"),
             py::arg("repeats") = 2)
        .def_property_readonly("num_qubits", &RingIsingCudaBackend::num_qubits)
        .def_property_readonly("num_layers", &RingIsingCudaBackend::num_layers);

    m.def(
        "energy_and_grad",
        [](size_t num_qubits, size_t num_layers, double field, const std::string &strategy,
           const FlatArray &params, bool compute_gradient, bool profile, size_t chunk_width) {
            validate_params_shape(num_qubits, num_layers, params);
            const auto view = view_params(params);
            double energy = 0.0;
            std::vector<double> grad(compute_gradient ? view.size : 0);
            rie_status st;
            {
                py::gil_scoped_release release;
                st = rie_energy_and_grad(num_qubits, num_layers, field, strategy.c_str(), view.ptr,
                                         view.size, compute_gradient ? 1 : 0, profile ? 1 : 0,
                                         chunk_width, &energy,
                                         compute_gradient ? grad.data() : nullptr);
            }
            check(st);
            return make_result(energy, grad, {});
        },
        py::arg("num_qubits"), py::arg("num_layers"), py::arg("field"),
        py::arg("gradient_strategy") = "structured_adjoint", py::arg("params"),
        py::arg("compute_gradient") = true, py::arg("profile") = false,
        py::arg("structured_rotation_chunk_width") = 8);

    py::class_<RingIsingBatch>(m, "RingIsingBatch")
        .def(py::init<const std::vector<int> &, const std::vector<int> &,
                      const std::vector<double> &, int>(),
             py::arg("num_qubits"), py::arg("num_layers"), py::arg("fields"),
             py::arg("dense_max_qubits") = 8)
        .def_property_readonly("size", &RingIsingBatch::size)
        .def_property_readonly("num_params", &RingIsingBatch::num_params)
        .def("param_offsets", &RingIsingBatch::param_offsets)
        .def("strategies", &RingIsingBatch::strategies)
        .def("energy_and_grad", &RingIsingBatch::energy_and_grad, py::arg("params"),
             py::arg("compute_gradient") = true)
        .def("set_params", &RingIsingBatch::set_params, py::arg("params"))
        .def("get_params", &RingIsingBatch::get_params)
        .def("results", &RingIsingBatch::results)
        .def("device_buffers", &RingIsingBatch::device_buffers)
        .def("run_device", &RingIsingBatch::run_device, py::arg("compute_gradient") = true,
             py::arg("stream") = 0)
        .def("run_per_circuit", &RingIsingBatch::run_per_circuit,
             py::arg("compute_gradient") = true, py::arg("stream") = 0)
        .def("set_use_graph", &RingIsingBatch::set_use_graph, py::arg("use_graph"))
        .def("optimize", &RingIsingBatch::optimize, py::arg("optimizer"), py::arg("lr"),
             py::arg("steps"), py::arg("use_graph") = true, py::arg("stream") = 0)
        .def("autotune", &RingIsingBatch::autotune, py::arg("repeats") = 3)
        .def("set_tune_split", &RingIsingBatch::set_tune_split, py::arg("enabled"))
        .def("config", &RingIsingBatch::config)
        .def("set_config", &RingIsingBatch::set_config, py::arg("threads"),
             py::arg("chunk_by_n") = py::none(), py::arg("dense_tile") = 64,
             py::arg("structured_grid") = 0);

    py::class_<WarmStartMLP>(m, "WarmStartMLP")
        .def(py::init<int, int, int, uint64_t>(), py::arg("inputs"), py::arg("hidden"),
             py::arg("outputs"), py::arg("seed") = 0)
        .def_property_readonly("num_params", &WarmStartMLP::num_params)
        .def("get_params", &WarmStartMLP::get_params)
        .def("set_params", &WarmStartMLP::set_params, py::arg("params"))
        .def("forward", &WarmStartMLP::forward, py::arg("inputs"))
        .def("loss_and_grad", &WarmStartMLP::loss_and_grad, py::arg("inputs"), py::arg("targets"))
        .def("train", &WarmStartMLP::train, py::arg("inputs"), py::arg("targets"),
             py::arg("steps"), py::arg("lr"));

    m.def("device_count", &rie_device_count);
    m.def("abi_version", &rie_abi_version);
}
