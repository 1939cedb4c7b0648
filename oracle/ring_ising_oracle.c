/*
 * CPU oracle for the ring-Ising statevector energy / gradient path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP engine is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, and by nothing else.  The product path never
 * links, loads or calls it.
 *
 * Everything is FP64 / complex128.  Wire w is bit w of the basis index
 * (cpp/ising_cuda_kernel_common.cuh:20-23, bit_is_set).
 *
 * Two independent restatements are provided:
 *
 *  1. rio_gate_energy_and_grad: gate-by-gate adjoint differentiation in
 *     PennyLane's gate order (RY, RZ on every wire, then n CNOTs), i.e. the
 *     algorithm of PennyLane's diff_method="adjoint" that the reference's
 *     tests use as ground truth (tests/test_backends_parity.py:51-70) and that
 *     the reference's gate-level strategy follows
 *     (cpp/ising_cuda_backend.cu:150-174,
 *      cpp/ising_cuda_gate_level_adjoint_modes.inc:147-260).
 *     The Hamiltonian is applied term by term from its Pauli definition
 *     (ring_ising/models/ising.py:11-24), not from a closed form.
 *
 *  2. rio_structured_energy_and_grad: the reference's structured form
 *     (cpp/ising_cuda_structured_adjoint_modes.inc:309-373): fused RZ*RY per
 *     wire, the ring-CNOT layer as one closed-form index permutation
 *     (cpp/ising_cuda_kernels.cu:178-205), the Hamiltonian diagonal by
 *     domain-wall popcount (cpp/ising_cuda_kernels.cu:313-358) and the fused
 *     per-wire backward step (cpp/ising_cuda_statevector_grad_kernels.cu:223-319).
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   - end-to-end golden energies of the reference's 20-step gradient-descent
 *     runs (reference CSV, n = 8..24, L = 2), which fix init, ansatz, H,
 *     gradient and update together;
 *   - closed forms (theta = phi = 0 gives E = -n; RY(pi/2) everywhere gives
 *     E = -n*h; the n = 2 ring doubles its ZZ term);
 *   - the exact parameter-shift rule for RY/RZ and central finite differences;
 *   - a dense Kronecker-product simulation for n <= 6 (numpy);
 *   - brute-force composition of n sequential CNOTs against the closed-form
 *     ring permutation;
 *   - gate-level restatement == structured restatement.
 */

#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double re;
    double im;
} cd;

#define PAR_MIN_AMPS ((size_t)1 << 14)

static int g_num_threads = 0; /* 0: OpenMP default */

void rio_set_num_threads(int threads) { g_num_threads = threads; }

int rio_get_max_threads(void) {
#ifdef _OPENMP
    return g_num_threads > 0 ? g_num_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

static int nthreads_for(size_t work) {
#ifdef _OPENMP
    if (work < PAR_MIN_AMPS) {
        return 1;
    }
    return g_num_threads > 0 ? g_num_threads : omp_get_max_threads();
#else
    (void)work;
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Gate primitives in PennyLane's conventions.                               */
/* RY(t) = [[c, -s], [s, c]], c = cos(t/2), s = sin(t/2)                      */
/*   (cpp/ising_cuda_kernels.cu:78-100).                                     */
/* RZ(p) = diag(e^{-ip/2}, e^{+ip/2}) (cpp/ising_cuda_kernels.cu:128-141).   */
/* ------------------------------------------------------------------------ */

static void apply_ry(cd *s, int n, int wire, double theta, int adjoint) {
    const size_t pairs = (size_t)1 << (n - 1);
    const size_t hs = (size_t)1 << wire;
    const double c = cos(0.5 * theta);
    const double sn = adjoint ? -sin(0.5 * theta) : sin(0.5 * theta);
#pragma omp parallel for schedule(static) num_threads(nthreads_for(pairs))
    for (size_t p = 0; p < pairs; p++) {
        const size_t i0 = ((p >> wire) << (wire + 1)) | (p & (hs - 1));
        const size_t i1 = i0 | hs;
        const cd a0 = s[i0];
        const cd a1 = s[i1];
        s[i0].re = c * a0.re - sn * a1.re;
        s[i0].im = c * a0.im - sn * a1.im;
        s[i1].re = sn * a0.re + c * a1.re;
        s[i1].im = sn * a0.im + c * a1.im;
    }
}

static void apply_rz(cd *s, int n, int wire, double phi, int adjoint) {
    const size_t size = (size_t)1 << n;
    const double half = adjoint ? -0.5 * phi : 0.5 * phi;
    const double c = cos(half);
    const double sn = sin(half);
#pragma omp parallel for schedule(static) num_threads(nthreads_for(size))
    for (size_t i = 0; i < size; i++) {
        /* bit 0 -> e^{-i half}, bit 1 -> e^{+i half} */
        const double ss = ((i >> wire) & 1U) ? sn : -sn;
        const cd a = s[i];
        s[i].re = c * a.re - ss * a.im;
        s[i].im = c * a.im + ss * a.re;
    }
}

static void apply_cnot(cd *s, int n, int control, int target) {
    const size_t size = (size_t)1 << n;
    const size_t cm = (size_t)1 << control;
    const size_t tm = (size_t)1 << target;
#pragma omp parallel for schedule(static) num_threads(nthreads_for(size))
    for (size_t i = 0; i < size; i++) {
        if ((i & cm) && !(i & tm)) {
            const cd tmp = s[i];
            s[i] = s[i | tm];
            s[i | tm] = tmp;
        }
    }
}

/* out = H s with H = -sum_w Z_w Z_{w+1 mod n} - h sum_w X_w, applied one
 * Pauli term at a time exactly as ring_ising/models/ising.py:11-24 lists
 * them (so the n = 2 ring carries Z0Z1 twice). */
static void apply_hamiltonian_terms(const cd *s, cd *out, int n, double h) {
    const size_t size = (size_t)1 << n;
#pragma omp parallel for schedule(static) num_threads(nthreads_for(size))
    for (size_t i = 0; i < size; i++) {
        double re = 0.0;
        double im = 0.0;
        for (int w = 0; w < n; w++) {
            const int v = (w + 1) % n;
            const double zw = ((i >> w) & 1U) ? -1.0 : 1.0;
            const double zv = ((i >> v) & 1U) ? -1.0 : 1.0;
            re += -1.0 * zw * zv * s[i].re;
            im += -1.0 * zw * zv * s[i].im;
        }
        for (int w = 0; w < n; w++) {
            const size_t j = i ^ ((size_t)1 << w);
            re += -h * s[j].re;
            im += -h * s[j].im;
        }
        out[i].re = re;
        out[i].im = im;
    }
}

static double real_inner(const cd *a, const cd *b, size_t size) {
    double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc) num_threads(nthreads_for(size))
    for (size_t i = 0; i < size; i++) {
        acc += a[i].re * b[i].re + a[i].im * b[i].im;
    }
    return acc;
}

/* 2 Re <lam | dRY(theta) psi>, dRY = [[-s/2, -c/2], [c/2, -s/2]]
 * (cpp/ising_cuda_kernels.cu:102-126, cpp/ising_cuda_statevector_grad_kernels.cu:18-72). */
static double ry_grad(const cd *psi, const cd *lam, int n, int wire, double theta) {
    const size_t pairs = (size_t)1 << (n - 1);
    const size_t hs = (size_t)1 << wire;
    const double c = cos(0.5 * theta);
    const double s = sin(0.5 * theta);
    double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc) num_threads(nthreads_for(pairs))
    for (size_t p = 0; p < pairs; p++) {
        const size_t i0 = ((p >> wire) << (wire + 1)) | (p & (hs - 1));
        const size_t i1 = i0 | hs;
        const cd a0 = psi[i0];
        const cd a1 = psi[i1];
        const double d0r = -0.5 * s * a0.re - 0.5 * c * a1.re;
        const double d0i = -0.5 * s * a0.im - 0.5 * c * a1.im;
        const double d1r = 0.5 * c * a0.re - 0.5 * s * a1.re;
        const double d1i = 0.5 * c * a0.im - 0.5 * s * a1.im;
        acc += lam[i0].re * d0r + lam[i0].im * d0i + lam[i1].re * d1r + lam[i1].im * d1i;
    }
    return 2.0 * acc;
}

/* 2 Re <lam | dRZ(phi) psi>, dRZ = diag(-i/2 e^{-ip/2}, +i/2 e^{+ip/2})
 * (cpp/ising_cuda_kernels.cu:143-158). */
static double rz_grad(const cd *psi, const cd *lam, int n, int wire, double phi) {
    const size_t size = (size_t)1 << n;
    const double c = cos(0.5 * phi);
    const double s = sin(0.5 * phi);
    double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc) num_threads(nthreads_for(size))
    for (size_t i = 0; i < size; i++) {
        const int bit = (int)((i >> wire) & 1U);
        /* factor = e^{-+ip/2} times -+i/2 */
        const double fr = c;
        const double fi = bit ? s : -s;
        const double pr = 0.0;
        const double pi = bit ? 0.5 : -0.5;
        /* g = (pr + i pi) * (fr + i fi) */
        const double gr = pr * fr - pi * fi;
        const double gi = pr * fi + pi * fr;
        const double dr = gr * psi[i].re - gi * psi[i].im;
        const double di = gr * psi[i].im + gi * psi[i].re;
        acc += lam[i].re * dr + lam[i].im * di;
    }
    return 2.0 * acc;
}

static cd *alloc_state(size_t size) {
    cd *s = (cd *)malloc(sizeof(cd) * size);
    return s;
}

static void zero_state(cd *s, size_t size) {
    memset(s, 0, sizeof(cd) * size);
    s[0].re = 1.0;
}

/* Forward pass in PennyLane order: ring_ising/models/ising.py:27-44. */
static void gate_forward(cd *s, int n, int layers, const double *params) {
    for (int l = 0; l < layers; l++) {
        for (int w = 0; w < n; w++) {
            const size_t base = ((size_t)l * n + w) * 2;
            apply_ry(s, n, w, params[base], 0);
            apply_rz(s, n, w, params[base + 1], 0);
        }
        for (int w = 0; w < n; w++) {
            apply_cnot(s, n, w, (w + 1) % n);
        }
    }
}

int rio_gate_state(int n, int layers, const double *params, double *out_state) {
    if (n < 1 || n > 34 || layers < 0) {
        return 1;
    }
    const size_t size = (size_t)1 << n;
    cd *s = (cd *)out_state;
    zero_state(s, size);
    gate_forward(s, n, layers, params);
    return 0;
}

int rio_apply_hamiltonian(int n, double h, const double *state, double *out) {
    if (n < 1 || n > 34) {
        return 1;
    }
    apply_hamiltonian_terms((const cd *)state, (cd *)out, n, h);
    return 0;
}

/* Gate-level adjoint differentiation (PennyLane order). */
int rio_gate_energy_and_grad(int n, int layers, double h, const double *params,
                             int want_grad, double *energy, double *grad) {
    if (n < 2 || n > 34 || layers < 1) {
        return 1;
    }
    const size_t size = (size_t)1 << n;
    cd *psi = alloc_state(size);
    cd *lam = alloc_state(size);
    if (psi == NULL || lam == NULL) {
        free(psi);
        free(lam);
        return 2;
    }
    zero_state(psi, size);
    gate_forward(psi, n, layers, params);
    apply_hamiltonian_terms(psi, lam, n, h);
    *energy = real_inner(psi, lam, size);
    if (want_grad) {
        for (int l = layers - 1; l >= 0; l--) {
            for (int w = n - 1; w >= 0; w--) {
                apply_cnot(psi, n, w, (w + 1) % n);
                apply_cnot(lam, n, w, (w + 1) % n);
            }
            for (int w = n - 1; w >= 0; w--) {
                const size_t base = ((size_t)l * n + w) * 2;
                const double theta = params[base];
                const double phi = params[base + 1];
                apply_rz(psi, n, w, phi, 1);
                grad[base + 1] = rz_grad(psi, lam, n, w, phi);
                apply_rz(lam, n, w, phi, 1);
                apply_ry(psi, n, w, theta, 1);
                grad[base] = ry_grad(psi, lam, n, w, theta);
                apply_ry(lam, n, w, theta, 1);
            }
        }
    }
    free(psi);
    free(lam);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Structured restatement.                                                   */
/* ------------------------------------------------------------------------ */

/* Image of basis index i under the ring-CNOT layer
 * P = CNOT(n-1,0) ... CNOT(1,2) CNOT(0,1): bits k >= 1 become the prefix XOR
 * of bits 0..k, bit 0 becomes the parity of bits 1..n-1; the inverse map is
 * (i ^ (i << 1)) & mask, then ^3 when the top bit is set
 * (cpp/ising_cuda_kernels.cu:178-205). */
uint64_t rio_ring_cnot_index(uint64_t i, int n, int inverse) {
    const uint64_t mask = (n >= 64) ? ~(uint64_t)0 : (((uint64_t)1 << n) - 1);
    if (!inverse) {
        uint64_t p = i;
        p ^= p << 1;
        p ^= p << 2;
        p ^= p << 4;
        p ^= p << 8;
        p ^= p << 16;
        p ^= p << 32;
        p &= mask;
        uint64_t t = p & ~(uint64_t)1;
        t |= (uint64_t)(__builtin_popcountll(i >> 1) & 1);
        return t;
    }
    uint64_t t = (i ^ (i << 1)) & mask;
    if ((i >> (n - 1)) & 1U) {
        t ^= 3U;
    }
    return t;
}

/* out[P(i)] = in[i] (forward) or out[P^-1(i)] = in[i] (inverse). */
static void ring_cnot_layer(const cd *in, cd *out, int n, int inverse) {
    const size_t size = (size_t)1 << n;
#pragma omp parallel for schedule(static) num_threads(nthreads_for(size))
    for (size_t i = 0; i < size; i++) {
        out[rio_ring_cnot_index(i, n, inverse)] = in[i];
    }
}

/* Fused RZ(phi) RY(theta) on one wire or its inverse
 * (cpp/ising_cuda_kernels.cu:238-284). */
static void apply_ryrz(cd *s, int n, int wire, double theta, double phi, int inverse) {
    const size_t pairs = (size_t)1 << (n - 1);
    const size_t hs = (size_t)1 << wire;
    const double c = cos(0.5 * theta);
    const double sn = sin(0.5 * theta);
    const double ch = cos(0.5 * phi);
    const double sh = sin(0.5 * phi);
#pragma omp parallel for schedule(static) num_threads(nthreads_for(pairs))
    for (size_t p = 0; p < pairs; p++) {
        const size_t i0 = ((p >> wire) << (wire + 1)) | (p & (hs - 1));
        const size_t i1 = i0 | hs;
        const cd a0 = s[i0];
        const cd a1 = s[i1];
        if (!inverse) {
            const double b0r = c * a0.re - sn * a1.re;
            const double b0i = c * a0.im - sn * a1.im;
            const double b1r = sn * a0.re + c * a1.re;
            const double b1i = sn * a0.im + c * a1.im;
            s[i0].re = ch * b0r + sh * b0i;
            s[i0].im = ch * b0i - sh * b0r;
            s[i1].re = ch * b1r - sh * b1i;
            s[i1].im = ch * b1i + sh * b1r;
        } else {
            const double p0r = ch * a0.re - sh * a0.im;
            const double p0i = ch * a0.im + sh * a0.re;
            const double p1r = ch * a1.re + sh * a1.im;
            const double p1i = ch * a1.im - sh * a1.re;
            s[i0].re = c * p0r + sn * p1r;
            s[i0].im = c * p0i + sn * p1i;
            s[i1].re = -sn * p0r + c * p1r;
            s[i1].im = -sn * p0i + c * p1i;
        }
    }
}

/* lam = H psi with the diagonal from the domain-wall count
 * 2*popcount(i ^ rot(i)) - n (cpp/ising_cuda_kernels.cu:313-358). */
static double hamiltonian_closed_form(const cd *psi, cd *lam, int n, double h) {
    const size_t size = (size_t)1 << n;
    const uint64_t mask = (uint64_t)size - 1;
    double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc) num_threads(nthreads_for(size))
    for (size_t i = 0; i < size; i++) {
        const uint64_t rot = (((uint64_t)i << 1) & mask) | ((uint64_t)i >> (n - 1));
        const int walls = __builtin_popcountll(((uint64_t)i ^ rot) & mask);
        const double d = (double)(2 * walls) - (double)n;
        double re = d * psi[i].re;
        double im = d * psi[i].im;
        for (int w = 0; w < n; w++) {
            const size_t j = i ^ ((size_t)1 << w);
            re += -h * psi[j].re;
            im += -h * psi[j].im;
        }
        lam[i].re = re;
        lam[i].im = im;
        acc += psi[i].re * re + psi[i].im * im;
    }
    return acc;
}

/* One fused backward step on one wire: current <- U^dag current,
 * lam <- U^dag lam, with
 *   dphi   = 2 * sum 1/2 [Im(conj(l0) c0) - Im(conj(l1) c1)]   (l, c after U)
 *   dtheta = 2 * sum Re<lam_mid | dRY c_before>
 * (cpp/ising_cuda_statevector_grad_kernels.cu:223-319). */
static void structured_backward_wire(cd *cur, cd *lam, int n, int wire, double theta,
                                     double phi, double *dtheta, double *dphi) {
    const size_t pairs = (size_t)1 << (n - 1);
    const size_t hs = (size_t)1 << wire;
    const double rc = cos(0.5 * phi);
    const double rs = sin(0.5 * phi);
    const double yc = cos(0.5 * theta);
    const double ys = sin(0.5 * theta);
    double tsum = 0.0;
    double psum = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : tsum, psum) num_threads(nthreads_for(pairs))
    for (size_t p = 0; p < pairs; p++) {
        const size_t i0 = ((p >> wire) << (wire + 1)) | (p & (hs - 1));
        const size_t i1 = i0 | hs;
        const cd ca0 = cur[i0], ca1 = cur[i1];
        const cd la0 = lam[i0], la1 = lam[i1];
        const double cm0r = rc * ca0.re - rs * ca0.im;
        const double cm0i = rc * ca0.im + rs * ca0.re;
        const double cm1r = rc * ca1.re + rs * ca1.im;
        const double cm1i = rc * ca1.im - rs * ca1.re;
        const double lm0r = rc * la0.re - rs * la0.im;
        const double lm0i = rc * la0.im + rs * la0.re;
        const double lm1r = rc * la1.re + rs * la1.im;
        const double lm1i = rc * la1.im - rs * la1.re;
        psum += 0.5 * ((la0.re * ca0.im - la0.im * ca0.re) - (la1.re * ca1.im - la1.im * ca1.re));
        const double cb0r = yc * cm0r + ys * cm1r;
        const double cb0i = yc * cm0i + ys * cm1i;
        const double cb1r = -ys * cm0r + yc * cm1r;
        const double cb1i = -ys * cm0i + yc * cm1i;
        cur[i0].re = cb0r;
        cur[i0].im = cb0i;
        cur[i1].re = cb1r;
        cur[i1].im = cb1i;
        const double d0r = -0.5 * (ys * cb0r + yc * cb1r);
        const double d0i = -0.5 * (ys * cb0i + yc * cb1i);
        const double d1r = 0.5 * (yc * cb0r - ys * cb1r);
        const double d1i = 0.5 * (yc * cb0i - ys * cb1i);
        tsum += lm0r * d0r + lm0i * d0i + lm1r * d1r + lm1i * d1i;
        lam[i0].re = yc * lm0r + ys * lm1r;
        lam[i0].im = yc * lm0i + ys * lm1i;
        lam[i1].re = -ys * lm0r + yc * lm1r;
        lam[i1].im = -ys * lm0i + yc * lm1i;
    }
    *dtheta = 2.0 * tsum;
    *dphi = 2.0 * psum;
}

/* Structured energy_and_grad; on return *drift (if non-NULL) holds
 * || current - |0...0> ||_2 after the full inverse walk, the accumulated
 * round-off of recomputing states by inverse gates. */
int rio_structured_energy_and_grad(int n, int layers, double h, const double *params,
                                   int want_grad, double *energy, double *grad,
                                   double *drift) {
    if (n < 2 || n > 34 || layers < 1) {
        return 1;
    }
    const size_t size = (size_t)1 << n;
    cd *cur = alloc_state(size);
    cd *lam = alloc_state(size);
    cd *tmp = alloc_state(size);
    if (cur == NULL || lam == NULL || tmp == NULL) {
        free(cur);
        free(lam);
        free(tmp);
        return 2;
    }
    zero_state(cur, size);
    for (int l = 0; l < layers; l++) {
        for (int w = 0; w < n; w++) {
            const size_t base = ((size_t)l * n + w) * 2;
            apply_ryrz(cur, n, w, params[base], params[base + 1], 0);
        }
        ring_cnot_layer(cur, tmp, n, 0);
        cd *sw = cur;
        cur = tmp;
        tmp = sw;
    }
    *energy = hamiltonian_closed_form(cur, lam, n, h);
    if (want_grad) {
        for (int l = layers - 1; l >= 0; l--) {
            ring_cnot_layer(cur, tmp, n, 1);
            cd *sw = cur;
            cur = tmp;
            tmp = sw;
            ring_cnot_layer(lam, tmp, n, 1);
            sw = lam;
            lam = tmp;
            tmp = sw;
            for (int w = n - 1; w >= 0; w--) {
                const size_t base = ((size_t)l * n + w) * 2;
                structured_backward_wire(cur, lam, n, w, params[base], params[base + 1],
                                         &grad[base], &grad[base + 1]);
            }
        }
        if (drift != NULL) {
            double acc = 0.0;
            for (size_t i = 0; i < size; i++) {
                const double re = cur[i].re - (i == 0 ? 1.0 : 0.0);
                acc += re * re + cur[i].im * cur[i].im;
            }
            *drift = sqrt(acc);
        }
    }
    free(cur);
    free(lam);
    free(tmp);
    return 0;
}
