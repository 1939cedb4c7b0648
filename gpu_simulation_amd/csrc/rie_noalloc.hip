// Allocation-free energy_and_grad for small circuits (structured_adjoint,
// 2 <= n <= 14), for callers that cannot allocate device memory: code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
//
// The whole circuit runs in one workgroup (batch_structured_kernel with a
// single job): psi and lambda in LDS for n <= 12, in a 2 * 2^n-amplitude
// global scratch for n = 13, 14.  Everything beyond the caller's inputs and
// outputs -- the per-(layer, wire) coefficient table and that state scratch --
// lives in the file-scope array `scratch` below, one copy per device, held
// until the process ends.  Its address on each device comes from
// hipGetSymbolAddress once per device (with that device current) and is the
// only form in which host code hands it to kernels.
//
// Every call on a device shares that device's array: calls that overlap on
// one device (different streams, or different host threads) race.  Nothing
// in it needs to start at zero: the coefficient kernel writes every entry the
// circuit reads, and the circuit kernel writes its initial state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "rie.h"
#include "rie_engine_internal.hpp"
#include "rie_kernels.hpp"

namespace rie {
namespace {

constexpr size_t kNoAllocMaxQubits = RIE_NOALLOC_MAX_QUBITS;
constexpr size_t kNoAllocMaxLayers = RIE_NOALLOC_MAX_LAYERS;
static_assert(kNoAllocMaxQubits <= (size_t)kBatchMaxQubits, "the whole-circuit kernel's limit");
constexpr size_t kCoefBytes = sizeof(WireCoef) * kNoAllocMaxQubits * kNoAllocMaxLayers;
constexpr size_t kStateBytes = sizeof(double2) * 2 * (size_t{1} << kNoAllocMaxQubits);
constexpr size_t kScratchBytes = kCoefBytes + kStateBytes;
static_assert(kCoefBytes % 256 == 0, "the state region starts 256-byte aligned");
static_assert(kScratchBytes == RIE_NOALLOC_SCRATCH_BYTES, "the size the header states");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the
// device's first call.
unsigned char *device_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        fail(RIE_DEVICE_ERROR, "no current HIP device");
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void *ptr = nullptr;
        const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (err != hipSuccess || ptr == nullptr) {
            fail(RIE_DEVICE_ERROR, std::string("hipGetSymbolAddress(scratch): ") +
                                       hipGetErrorString(err));
        }
        g_scratch_ptr[dev] = static_cast<unsigned char *>(ptr);
    }
    return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace rie

rie_status rie_energy_and_grad_noalloc(size_t num_qubits, size_t num_layers, double field,
                                       const double *params_device, size_t num_params,
                                       size_t structured_rotation_chunk_width,
                                       double *energy_device, double *gradient_device,
                                       void *stream_handle) {
    using namespace rie;
    return guard([&]() {
        if (num_qubits < 2 || num_qubits > kNoAllocMaxQubits) {
            fail(RIE_INVALID_ARGUMENT, "the allocation-free entry point takes 2 <= num_qubits <= " +
                                           std::to_string(kNoAllocMaxQubits) + ".");
        }
        if (num_layers < 1 || num_layers > kNoAllocMaxLayers) {
            fail(RIE_INVALID_ARGUMENT, "the allocation-free entry point takes 1 <= num_layers <= " +
                                           std::to_string(kNoAllocMaxLayers) + ".");
        }
        if (num_params != 2 * num_qubits * num_layers) {
            fail(RIE_INVALID_ARGUMENT, "num_params does not match the circuit.");
        }
        if (structured_rotation_chunk_width < 1 || structured_rotation_chunk_width > 8) {
            fail(RIE_INVALID_ARGUMENT, "structured_rotation_chunk_width must be in [1, 8].");
        }
        if (!std::isfinite(field)) {
            fail(RIE_INVALID_ARGUMENT, "field must be finite.");
        }
        if (params_device == nullptr || energy_device == nullptr) {
            fail(RIE_INVALID_ARGUMENT, "params and energy must be device pointers.");
        }
        const hipStream_t stream = static_cast<hipStream_t>(stream_handle);
        unsigned char *base = device_scratch();
        WireCoef *coef = reinterpret_cast<WireCoef *>(base);
        double2 *state = reinterpret_cast<double2 *>(base + kCoefBytes);
        const int n = (int)num_qubits;
        launch_coef(params_device, coef, n * (int)num_layers, stream);
        BatchJob job{};
        job.n = n;
        job.layers = (int)num_layers;
        job.chunk = (int)std::min<size_t>(structured_rotation_chunk_width, 4);
        job.field = field;
        job.coef = coef;
        job.energy = energy_device;
        job.grad = gradient_device;  // nullptr: energy only
        job.drift = nullptr;
        job.scale = nullptr;
        const bool in_lds = n <= kBatchMaxLdsQubits;
        launch_batch_structured(nullptr, job, nullptr, nullptr, 1, 256, in_lds ? n : 0,
                                in_lds ? nullptr : state, 0, stream);
        const hipError_t err = hipGetLastError();
        if (err != hipSuccess) {
            fail(RIE_DEVICE_ERROR, std::string("allocation-free launch: ") + hipGetErrorString(err));
        }
    });
}
