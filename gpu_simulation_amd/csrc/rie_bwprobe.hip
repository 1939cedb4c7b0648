// HBM calibration probe for gfx950: streaming kernels of known byte counts.
//
//   rie_bwprobe [GiB per buffer, default 8] [repeats, default 5]
//
// Prints one JSON object per kernel: algorithmic bytes per launch, average
// launch time (HIP events) and GB/s.  "copy" is the reference point for the
// achievable HBM rate and for scaling profiler FETCH_SIZE / WRITE_SIZE
// counters to bytes; "chunk128_copy" / "chunk64_copy" read and write
// 128-byte / 64-byte contiguous runs scattered with a large stride, the
// access shape of the engine's high-wire tile passes.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                  \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
            std::exit(1);                                                         \
        }                                                                         \
    } while (0)

__global__ __launch_bounds__(256) void copy_kernel(const double2 *__restrict__ in,
                                                   double2 *__restrict__ out, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        out[i] = in[i];
    }
}

// Each thread moves U consecutive-per-lane blocks of 16-byte elements with all
// U loads issued before the first store (U x 16 B in flight per lane).
template <int U, bool kNonTemporal>
__global__ __launch_bounds__(256) void copy_unrolled_kernel(const double2 *__restrict__ in,
                                                            double2 *__restrict__ out, size_t n) {
    const size_t block_elems = (size_t)blockDim.x * U;
    for (size_t base = (size_t)blockIdx.x * block_elems; base < n;
         base += (size_t)gridDim.x * block_elems) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            v[u] = in[base + (size_t)u * blockDim.x + threadIdx.x];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            double2 *dst = out + base + (size_t)u * blockDim.x + threadIdx.x;
            if constexpr (kNonTemporal) {
                __builtin_nontemporal_store(v[u].x, &dst->x);
                __builtin_nontemporal_store(v[u].y, &dst->y);
            } else {
                *dst = v[u];
            }
        }
    }
}

__global__ __launch_bounds__(256) void read_kernel(const double2 *__restrict__ in, size_t n,
                                                   double *__restrict__ sink) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double2 v = in[i];
        acc += v.x + v.y;
    }
    if (acc == 12345.678) {  // keeps the loads; never true for the zero-filled input
        sink[0] = acc;
    }
}

__global__ __launch_bounds__(256) void write_kernel(double2 *__restrict__ out, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        out[i] = make_double2((double)i, 0.0);
    }
}

// Element i of run r = (i / run_len) is placed at run index
// (r % runs_per_row) * rows + r / runs_per_row: contiguous runs of run_len
// amplitudes, consecutive runs 'rows' runs apart (a transposed layout).
__global__ __launch_bounds__(256) void chunk_copy_kernel(const double2 *__restrict__ in,
                                                         double2 *__restrict__ out, size_t n,
                                                         int run_shift, size_t rows) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t run_len = (size_t)1 << run_shift;
    const size_t runs = n >> run_shift;
    const size_t runs_per_row = runs / rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const size_t r = i >> run_shift;
        const size_t pr = (r % runs_per_row) * rows + r / runs_per_row;
        const size_t j = (pr << run_shift) | (i & (run_len - 1));
        out[j] = in[j];
    }
}

int main(int argc, char **argv) {
    const double gib = argc > 1 ? std::atof(argv[1]) : 8.0;
    const int repeats = argc > 2 ? std::atoi(argv[2]) : 5;
    const size_t n = (size_t)(gib * (double)(1ull << 30)) / sizeof(double2);
    double2 *a = nullptr;
    double2 *b = nullptr;
    double *sink = nullptr;
    CHECK(hipMalloc(&a, n * sizeof(double2)));
    CHECK(hipMalloc(&b, n * sizeof(double2)));
    CHECK(hipMalloc(&sink, sizeof(double)));
    CHECK(hipMemset(a, 0, n * sizeof(double2)));
    CHECK(hipMemset(b, 0, n * sizeof(double2)));
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int grid = prop.multiProcessorCount * 16;
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const size_t rows = 1024;  // runs of consecutive indices sit 1024 runs apart

    auto run = [&](const char *name, double bytes, auto &&launch) {
        launch();
        CHECK(hipDeviceSynchronize());
        CHECK(hipEventRecord(e0, 0));
        for (int r = 0; r < repeats; ++r) {
            launch();
        }
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double avg = ms / repeats;
        std::printf("{\"kernel\": \"%s\", \"bytes\": %.0f, \"avg_ms\": %.4f, \"GBps\": %.1f}\n", name,
                    bytes, avg, bytes / (avg * 1e-3) / 1e9);
        std::fflush(stdout);
    };
    const double S = (double)n * sizeof(double2);
    run("copy", 2 * S, [&]() { copy_kernel<<<grid, 256>>>(a, b, n); });
    // n is a multiple of 256 * 16 for every whole number of MiB
    run("copy_u8", 2 * S, [&]() { copy_unrolled_kernel<8, false><<<grid, 256>>>(a, b, n); });
    run("copy_u16", 2 * S, [&]() { copy_unrolled_kernel<16, false><<<grid, 256>>>(a, b, n); });
    run("copy_u16_nt", 2 * S, [&]() { copy_unrolled_kernel<16, true><<<grid, 256>>>(a, b, n); });
    run("copy_u16_grid2k", 2 * S,
        [&]() { copy_unrolled_kernel<16, false><<<prop.multiProcessorCount * 8, 256>>>(a, b, n); });
    run("read", S, [&]() { read_kernel<<<grid, 256>>>(a, n, sink); });
    run("write", S, [&]() { write_kernel<<<grid, 256>>>(b, n); });
    run("chunk128_copy", 2 * S, [&]() { chunk_copy_kernel<<<grid, 256>>>(a, b, n, 3, rows); });
    run("chunk64_copy", 2 * S, [&]() { chunk_copy_kernel<<<grid, 256>>>(a, b, n, 2, rows); });
    CHECK(hipGetLastError());
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    CHECK(hipFree(sink));
    return 0;
}
