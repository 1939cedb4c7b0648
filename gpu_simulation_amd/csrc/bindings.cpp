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

    size_t num_qubits() const { return num_qubits_; }
    size_t num_layers() const { return num_layers_; }

  private:
    rie_engine *engine_ = nullptr;
    size_t num_qubits_;
    size_t num_layers_;
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
        .def_property_readonly("schedule", &RingIsingCudaBackend::schedule)
        .def("set_kernel_timing", &RingIsingCudaBackend::set_kernel_timing, py::arg("enable"))
        .def("reset_kernel_stats", &RingIsingCudaBackend::reset_kernel_stats)
        .def("kernel_stats", &RingIsingCudaBackend::kernel_stats)
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

    m.def("device_count", &rie_device_count);
    m.def("abi_version", &rie_abi_version);
}
