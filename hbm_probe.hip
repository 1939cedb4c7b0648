// HBM calibration probe for gfx950 (a measurement tool beside bench.py, not
// part of the engine): streaming kernels of known byte counts.
//
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         hbm_probe.hip -o hbm_probe
//   ./hbm_probe [GiB per buffer, default 8] [repeats, default 5] [tiles]
//
// With "tiles" as the third argument only the tile-pass patterns run.
//
// Prints one JSON object per kernel variant: algorithmic bytes per launch,
// average launch time (HIP events) and GB/s.  Each variant keeps U
// independent 16-byte loads per lane in flight (all U issued before the first
// use), over a grid of G workgroups per CU; the sweep covers read-only
// (a reduction that keeps every load), write-only and copy, each at
// U in {1, 4, 8, 16} and G in {4, 8, 16, 32}, then the tile-pass patterns of
// the engine (tile_kernel below: runs of 128 B .. 512 B at large strides, and
// the contiguous 64 KB tile).  The largest rate of each kind is the
// achievable HBM rate for that access pattern; "copy" variants also
// calibrate profiler FETCH_SIZE / WRITE_SIZE counters to bytes.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <string>

#define CHECK(x)                                                                  \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
            std::exit(1);                                                         \
        }                                                                         \
    } while (0)

constexpr int kThreads = 256;

// Workgroup w handles the blocks w, w + grid, ... of kThreads * U elements;
// within a block lane t touches elements t, t + kThreads, ... (coalesced).
template <int U>
__global__ __launch_bounds__(kThreads) void read_kernel(const double2 *__restrict__ in, size_t n,
                                                        double *__restrict__ sink) {
    const size_t block = (size_t)kThreads * U;
    double acc = 0.0;
    for (size_t base = (size_t)blockIdx.x * block; base < n; base += (size_t)gridDim.x * block) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = in[base + (size_t)u * kThreads + threadIdx.x];
#pragma unroll
        for (int u = 0; u < U; ++u) acc += v[u].x + v[u].y;
    }
    if (acc == 12345.678) sink[0] = acc;  // never true for the zero-filled input
}

template <int U>
__global__ __launch_bounds__(kThreads) void write_kernel(double2 *__restrict__ out, size_t n) {
    const size_t block = (size_t)kThreads * U;
    for (size_t base = (size_t)blockIdx.x * block; base < n; base += (size_t)gridDim.x * block) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = base + (size_t)u * kThreads + threadIdx.x;
            out[i] = make_double2((double)i, 0.0);
        }
    }
}

template <int U>
__global__ __launch_bounds__(kThreads) void copy_kernel(const double2 *__restrict__ in,
                                                        double2 *__restrict__ out, size_t n) {
    const size_t block = (size_t)kThreads * U;
    for (size_t base = (size_t)blockIdx.x * block; base < n; base += (size_t)gridDim.x * block) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = in[base + (size_t)u * kThreads + threadIdx.x];
#pragma unroll
        for (int u = 0; u < U; ++u) out[base + (size_t)u * kThreads + threadIdx.x] = v[u];
    }
}

// Tile-pass access pattern: a workgroup of 256 threads moves 4096-amplitude
// tiles whose offset bits [0, a) are index bits [0, a) (runs of 16 * 2^a
// contiguous bytes) and offset bits [a, 12) are index bits [h0, h0 + 12 - a);
// thread t holds offsets t + 256 k, k = 0..15 (16 loads in flight per lane),
// as the engine's high passes (a = 3, 128-byte runs) and low passes (a = 12)
// load their tiles.
template <bool kCopy>
__global__ __launch_bounds__(kThreads) void tile_kernel(const double2 *__restrict__ in,
                                                        double2 *__restrict__ out, int nbits,
                                                        int a, int h0, double *__restrict__ sink) {
    const uint64_t tiles = uint64_t{1} << (nbits - 12);
    const int gap = h0 - a;
    double acc = 0.0;
    for (uint64_t g = blockIdx.x; g < tiles; g += gridDim.x) {
        const uint64_t base = ((g & ((uint64_t{1} << gap) - 1)) << a) | ((g >> gap) << (h0 + 12 - a));
        uint64_t idx[16];
        double2 v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t o = threadIdx.x + 256u * (uint32_t)k;
            idx[k] = base | (o & ((1u << a) - 1u)) | ((uint64_t)(o >> a) << h0);
            v[k] = in[idx[k]];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if constexpr (kCopy) {
                out[idx[k]] = v[k];
            } else {
                acc += v[k].x + v[k].y;
            }
        }
    }
    if (acc == 12345.678) sink[0] = acc;  // never true for the zero-filled input
}

// Read one tile pattern and write another: loads from tiles of (a_in, h_in),
// stores the same amplitudes to tiles of (a_out, h_out) -- the passes of a
// layout that keeps one side of every pass contiguous.
__global__ __launch_bounds__(kThreads) void tile_mixed_kernel(const double2 *__restrict__ in,
                                                              double2 *__restrict__ out, int nbits,
                                                              int a_in, int h_in, int a_out,
                                                              int h_out) {
    const uint64_t tiles = uint64_t{1} << (nbits - 12);
    const int gap_in = h_in - a_in;
    const int gap_out = h_out - a_out;
    for (uint64_t g = blockIdx.x; g < tiles; g += gridDim.x) {
        const uint64_t b_in = ((g & ((uint64_t{1} << gap_in) - 1)) << a_in) |
                              ((g >> gap_in) << (h_in + 12 - a_in));
        const uint64_t b_out = ((g & ((uint64_t{1} << gap_out) - 1)) << a_out) |
                               ((g >> gap_out) << (h_out + 12 - a_out));
        double2 v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t o = threadIdx.x + 256u * (uint32_t)k;
            v[k] = in[b_in | (o & ((1u << a_in) - 1u)) | ((uint64_t)(o >> a_in) << h_in)];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t o = threadIdx.x + 256u * (uint32_t)k;
            out[b_out | (o & ((1u << a_out) - 1u)) | ((uint64_t)(o >> a_out) << h_out)] = v[k];
        }
    }
}

int main(int argc, char **argv) {
    const double gib = argc > 1 ? std::atof(argv[1]) : 8.0;
    const int repeats = argc > 2 ? std::atoi(argv[2]) : 5;
    // whole multiples of the largest block (256 threads x 16 elements)
    const size_t quantum = (size_t)kThreads * 16;
    const size_t n = ((size_t)(gib * (double)(1ull << 30)) / sizeof(double2)) / quantum * quantum;
    if (n == 0 || repeats < 1) {
        std::fprintf(stderr, "usage: hbm_probe [GiB per buffer > 0] [repeats >= 1]\n");
        return 2;
    }
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
    const int cus = prop.multiProcessorCount;
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const double S = (double)n * sizeof(double2);

    auto run = [&](const char *kind, int u, int per_cu, double bytes, auto &&launch) {
        launch();
        CHECK(hipGetLastError());
        CHECK(hipDeviceSynchronize());
        CHECK(hipEventRecord(e0, 0));
        for (int r = 0; r < repeats; ++r) launch();
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double avg = ms / repeats;
        std::printf("{\"kind\": \"%s\", \"loads_in_flight_per_lane\": %d, \"workgroups_per_cu\": %d, "
                    "\"bytes\": %.0f, \"avg_ms\": %.4f, \"GBps\": %.1f}\n",
                    kind, u, per_cu, bytes, avg, bytes / (avg * 1e-3) / 1e9);
        std::fflush(stdout);
    };

    const bool tiles_only = argc > 3 && std::string(argv[3]) == "tiles";
    const int grids[] = {4, 8, 16, 32};
    for (int g : grids) {
        if (tiles_only) break;
        const int grid = cus * g;
        run("read", 1, g, S, [&]() { read_kernel<1><<<grid, kThreads>>>(a, n, sink); });
        run("read", 4, g, S, [&]() { read_kernel<4><<<grid, kThreads>>>(a, n, sink); });
        run("read", 8, g, S, [&]() { read_kernel<8><<<grid, kThreads>>>(a, n, sink); });
        run("read", 16, g, S, [&]() { read_kernel<16><<<grid, kThreads>>>(a, n, sink); });
        run("write", 1, g, S, [&]() { write_kernel<1><<<grid, kThreads>>>(b, n); });
        run("write", 4, g, S, [&]() { write_kernel<4><<<grid, kThreads>>>(b, n); });
        run("write", 8, g, S, [&]() { write_kernel<8><<<grid, kThreads>>>(b, n); });
        run("copy", 1, g, 2 * S, [&]() { copy_kernel<1><<<grid, kThreads>>>(a, b, n); });
        run("copy", 4, g, 2 * S, [&]() { copy_kernel<4><<<grid, kThreads>>>(a, b, n); });
        run("copy", 8, g, 2 * S, [&]() { copy_kernel<8><<<grid, kThreads>>>(a, b, n); });
        run("copy", 16, g, 2 * S, [&]() { copy_kernel<16><<<grid, kThreads>>>(a, b, n); });
    }
    // tile-pass patterns over a 2^nbits-amplitude state (nbits = floor(log2 n))
    int nbits = 0;
    while ((size_t{2} << nbits) <= n) ++nbits;
    for (int g : {2, 4, 8}) {
        const int grid = cus * g;
        for (int lo : {3, 4, 5, 12}) {
            for (int h0 : {12, nbits - 12 + lo}) {
                if (lo == 12 && h0 != 12) continue;  // the contiguous low tile
                if (h0 < 12 || h0 + 12 - lo > nbits) continue;
                const double Sb = (double)(size_t{1} << nbits) * sizeof(double2);
                char name[64];
                std::snprintf(name, sizeof(name), "tile_read_a%d_h%d", lo, h0);
                run(name, 16, g, Sb, [&]() {
                    tile_kernel<false><<<grid, kThreads>>>(a, b, nbits, lo, h0, sink);
                });
                std::snprintf(name, sizeof(name), "tile_copy_a%d_h%d", lo, h0);
                run(name, 16, g, 2 * Sb, [&]() {
                    tile_kernel<true><<<grid, kThreads>>>(a, b, nbits, lo, h0, sink);
                });
            }
        }
    }
    // one side contiguous, the other 128-byte runs (and both scattered, for reference)
    for (int g : {2, 4, 8}) {
        const int grid = cus * g;
        const double Sb = (double)(size_t{1} << nbits) * sizeof(double2);
        const int hs = nbits - 9;  // the top high group
        run("mixed_rd_contig_wr_a3", 16, g, 2 * Sb, [&]() {
            tile_mixed_kernel<<<grid, kThreads>>>(a, b, nbits, 12, 12, 3, hs);
        });
        run("mixed_rd_a3_wr_contig", 16, g, 2 * Sb, [&]() {
            tile_mixed_kernel<<<grid, kThreads>>>(a, b, nbits, 3, hs, 12, 12);
        });
        run("mixed_rd_a3_wr_a3", 16, g, 2 * Sb, [&]() {
            tile_mixed_kernel<<<grid, kThreads>>>(a, b, nbits, 3, hs, 3, hs);
        });
        // the scattered side's nine high bits right above the contiguous
        // block (h = 12): the order that keeps one side whole and puts the
        // other side's group at address bits 12..20
        run("mixed_rd_a3h12_wr_contig", 16, g, 2 * Sb, [&]() {
            tile_mixed_kernel<<<grid, kThreads>>>(a, b, nbits, 3, 12, 12, 12);
        });
        run("mixed_rd_contig_wr_a3h12", 16, g, 2 * Sb, [&]() {
            tile_mixed_kernel<<<grid, kThreads>>>(a, b, nbits, 12, 12, 3, 12);
        });
    }
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    CHECK(hipFree(sink));
    return 0;
}
