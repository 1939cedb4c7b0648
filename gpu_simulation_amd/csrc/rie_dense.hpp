// dense_scan product-tree plans: one circuit (an engine) or many (a batch).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "rie_device.hpp"

namespace rie {

// One dense_scan circuit as the kernels see it.  Node k of the product tree
// is tree + k * node_doubles (re plane of D*D doubles, then the im plane);
// level j starts at node 2P - 2 (P >> j), the root is node 2P - 2.
struct DenseJob {
    const WireCoef *coef;  // L * n coefficients
    double *tree;          // (2P - 1) nodes
    double2 *vec_a;        // P * 2D: down-sweep vectors (ping); root seed [e0 | eta]
    double2 *vec_b;        // (pong)
    double2 *head;         // psi, lambda, eta: 3D
    double *energy;
    double *grad;          // 2 n L values
    const double *scale;   // chain-rule factor for grad, or nullptr
    size_t node_doubles;   // 2 * D * D
    double field;
    int n;
    int L;
    int D;       // max(2^n, 16)
    int log_d;   // log2(D)
    int P;       // power of two >= L
    int levels;  // log2(P)
};

struct DenseSpec {
    int n;
    int L;
    double field;
    const WireCoef *coef;
    double *energy;
    double *grad;
    const double *scale;
};

// Device buffers, job table and per-launch work tables of a set of dense
// circuits.  Kernels of one launch find their job by binary search over the
// launch's (job, first block) table.
class DensePlan {
  public:
    DensePlan() = default;
    ~DensePlan();
    DensePlan(const DensePlan &) = delete;
    DensePlan &operator=(const DensePlan &) = delete;

    // max_tile: output tile edge for D >= 64 (32 or 64).
    void build(const std::vector<DenseSpec> &specs, int max_tile);
    // Issue every launch on s.  The overrides (single-circuit plans only)
    // redirect the energy / gradient / scale pointers of job 0.
    void run(hipStream_t s, bool want_grad, double *energy_override = nullptr,
             double *grad_override = nullptr, const double *scale_override = nullptr) const;
    size_t size() const { return jobs_.size(); }
    size_t device_bytes() const { return arena_bytes_; }
    static size_t bytes_for(int n, int L);

  private:
    struct Launch {
        int level = 0;
        int tile = 0;
        int blocks = 0;  // total blocks
        int count = 0;   // circuits in the launch
        size_t tab_off = 0;
    };
    std::vector<DenseJob> jobs_;
    DenseJob *d_jobs_ = nullptr;
    int2 *d_tab_ = nullptr;
    void *arena_ = nullptr;
    size_t arena_bytes_ = 0;
    Launch layers_;
    std::vector<Launch> gemm_;
    Launch head_;
    std::vector<Launch> down_;
    Launch tail_;
    void release();
};

}  // namespace rie
