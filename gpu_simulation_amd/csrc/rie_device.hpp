// Device-side building blocks of the ring-Ising statevector engine (gfx950).
//
// Amplitudes are complex128 stored as double2 {re, im}; wire w is bit w of the
// basis index.  All arithmetic is IEEE FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rie {

// cos/sin of the half angles of one wire's RY(theta) and RZ(phi)
// (reference: cpp/ising_cuda_rotation_kernels.cu:661-675 computes the same
// quantities on the host; here they live in a device table so that captured
// graphs can re-read parameters).
struct WireCoef {
    double c;   // cos(theta / 2)
    double s;   // sin(theta / 2)
    double ch;  // cos(phi / 2)
    double sh;  // sin(phi / 2)
};

constexpr int kWave = 64;

__device__ __forceinline__ uint64_t low_mask(int n) {
    return n >= 64 ? ~uint64_t{0} : ((uint64_t{1} << n) - 1);
}

// Image of index i under the ring-CNOT layer P = CNOT(n-1,0)...CNOT(0,1):
// bits k >= 1 become the XOR of bits 0..k, bit 0 the parity of bits 1..n-1.
// (reference: cpp/ising_cuda_kernels.cu:178-205, forward branch)
__device__ __forceinline__ uint64_t ring_fwd(uint64_t i, int n) {
    uint64_t p = i;
    p ^= p << 1;
    p ^= p << 2;
    p ^= p << 4;
    p ^= p << 8;
    p ^= p << 16;
    p ^= p << 32;
    p &= low_mask(n);
    return (p & ~uint64_t{1}) | (uint64_t)(__popcll(i >> 1) & 1);
}

// P^-1(i) = (i ^ (i << 1)) & mask, with bits 0 and 1 flipped when the top bit
// of i is set (reference: cpp/ising_cuda_kernels.cu:178-205, inverse branch).
__device__ __forceinline__ uint64_t ring_inv(uint64_t i, int n) {
    uint64_t t = (i ^ (i << 1)) & low_mask(n);
    if ((i >> (n - 1)) & 1U) {
        t ^= 3U;
    }
    return t;
}

// 32-bit versions for tile-local offsets (n <= 31).
__device__ __forceinline__ uint32_t ring_fwd32(uint32_t i, int n) {
    uint32_t p = i;
    p ^= p << 1;
    p ^= p << 2;
    p ^= p << 4;
    p ^= p << 8;
    p ^= p << 16;
    p &= (1u << n) - 1u;
    return (p & ~1u) | (uint32_t)(__popc(i >> 1) & 1);
}

__device__ __forceinline__ uint32_t ring_inv32(uint32_t i, int n) {
    uint32_t t = (i ^ (i << 1)) & ((1u << n) - 1u);
    if ((i >> (n - 1)) & 1U) {
        t ^= 3U;
    }
    return t;
}

// Diagonal of -sum_w Z_w Z_{w+1 mod n}: 2 * (#domain walls) - n
// (reference: cpp/ising_cuda_kernels.cu:313-330).
__device__ __forceinline__ double ham_diag(uint64_t i, int n) {
    const uint64_t m = low_mask(n);
    const uint64_t rot = ((i << 1) & m) | (i >> (n - 1));
    const int walls = (int)__popcll((i ^ rot) & m);  // __popcll is unsigned
    return (double)(2 * walls - n);
}

// (a0, a1) <- RZ(phi) RY(theta) (a0, a1) on one wire pair.
__device__ __forceinline__ void fwd_pair(double2 &a0, double2 &a1, const WireCoef &k) {
    const double b0x = k.c * a0.x - k.s * a1.x;
    const double b0y = k.c * a0.y - k.s * a1.y;
    const double b1x = k.s * a0.x + k.c * a1.x;
    const double b1y = k.s * a0.y + k.c * a1.y;
    a0.x = k.ch * b0x + k.sh * b0y;
    a0.y = k.ch * b0y - k.sh * b0x;
    a1.x = k.ch * b1x - k.sh * b1y;
    a1.y = k.ch * b1y + k.sh * b1x;
}

// (a0, a1) <- (RZ(phi) RY(theta))^dagger (a0, a1).
__device__ __forceinline__ void inv_pair(double2 &a0, double2 &a1, const WireCoef &k) {
    const double p0x = k.ch * a0.x - k.sh * a0.y;
    const double p0y = k.ch * a0.y + k.sh * a0.x;
    const double p1x = k.ch * a1.x + k.sh * a1.y;
    const double p1y = k.ch * a1.y - k.sh * a1.x;
    a0.x = k.c * p0x + k.s * p1x;
    a0.y = k.c * p0y + k.s * p1y;
    a1.x = k.c * p1x - k.s * p0x;
    a1.y = k.c * p1y - k.s * p0y;
}

// One adjoint step of U = RZ(phi) RY(theta) on a pair.  On entry (c, l) are
// the state and the adjoint state just after U; on exit just before it.
//   dE/dphi   += Re<l_after | -iZ | c_after>   = Im(l0* c0) - Im(l1* c1)
//   dE/dtheta += Re<l_before| -iY | c_before>  = Re(l1* c0) - Re(l0* c1)
// (same quantities as cpp/ising_cuda_statevector_grad_kernels.cu:223-319,
//  which forms them from the half-rotated lambda.)
__device__ __forceinline__ void bwd_pair(double2 &c0, double2 &c1, double2 &l0, double2 &l1,
                                         const WireCoef &k, double &dt, double &dp) {
    dp += (l0.x * c0.y - l0.y * c0.x) - (l1.x * c1.y - l1.y * c1.x);
    inv_pair(c0, c1, k);
    inv_pair(l0, l1, k);
    dt += (l1.x * c0.x + l1.y * c0.y) - (l0.x * c1.x + l0.y * c1.y);
}

// RY(theta) / RY(theta)^dagger alone on a pair (the RZ factors of a whole
// register window are applied as one diagonal phase per amplitude).
__device__ __forceinline__ void ry_pair(double2 &a0, double2 &a1, const WireCoef &k) {
    const double b0x = k.c * a0.x - k.s * a1.x;
    const double b0y = k.c * a0.y - k.s * a1.y;
    a1.x = k.s * a0.x + k.c * a1.x;
    a1.y = k.s * a0.y + k.c * a1.y;
    a0.x = b0x;
    a0.y = b0y;
}

__device__ __forceinline__ void ry_inv_pair(double2 &a0, double2 &a1, const WireCoef &k) {
    const double b0x = k.c * a0.x + k.s * a1.x;
    const double b0y = k.c * a0.y + k.s * a1.y;
    a1.x = k.c * a1.x - k.s * a0.x;
    a1.y = k.c * a1.y - k.s * a0.y;
    a0.x = b0x;
    a0.y = b0y;
}

// RY^dagger in tangent form: [[c, s], [-s, c]] = c [[1, t], [-t, 1]] with
// t = s / c (form A, |c| >= |s|) or s [[t, 1], [-1, t]] with t = c / s
// (form B): one FMA per output component instead of a product and an FMA,
// the factor c or s taken out (a pass collects the factors of its wires into
// one scale and applies it once).  |t| <= 1 in both forms.
__device__ __forceinline__ void ry_inv_pair_tan_a(double2 &a0, double2 &a1, double t) {
    const double x0 = fma(t, a1.x, a0.x);
    const double y0 = fma(t, a1.y, a0.y);
    a1.x = fma(-t, a0.x, a1.x);
    a1.y = fma(-t, a0.y, a1.y);
    a0.x = x0;
    a0.y = y0;
}

__device__ __forceinline__ void ry_inv_pair_tan_b(double2 &a0, double2 &a1, double t) {
    const double x0 = fma(t, a0.x, a1.x);
    const double y0 = fma(t, a0.y, a1.y);
    a1.x = fma(t, a1.x, -a0.x);
    a1.y = fma(t, a1.y, -a0.y);
    a0.x = x0;
    a0.y = y0;
}

// The tangent form's choice for one wire: form B where |s| > |c|.
__device__ __forceinline__ bool ry_tan_form_b(const WireCoef &k) {
    return fabs(k.s) > fabs(k.c);
}

// Components of U|0> for one wire: u0 = c e^{-i phi/2}, u1 = s e^{+i phi/2}
// (reference: cpp/ising_cuda_rotation_kernels.cu:686-703).
__device__ __forceinline__ double2 u0_of(const WireCoef &k) {
    return make_double2(k.c * k.ch, -k.c * k.sh);
}
__device__ __forceinline__ double2 u1_of(const WireCoef &k) {
    return make_double2(k.s * k.ch, k.s * k.sh);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// acc + a * b
__device__ __forceinline__ double2 cfma(double2 acc, double2 a, double2 b) {
    return make_double2(acc.x + a.x * b.x - a.y * b.y, acc.y + a.x * b.y + a.y * b.x);
}

// 2x2 complex matrix [[a00, a01], [a10, a11]].
struct CMat2 {
    double2 a00, a01, a10, a11;
};

__device__ __forceinline__ CMat2 cmat_add(const CMat2 &p, const CMat2 &q) {
    return {make_double2(p.a00.x + q.a00.x, p.a00.y + q.a00.y),
            make_double2(p.a01.x + q.a01.x, p.a01.y + q.a01.y),
            make_double2(p.a10.x + q.a10.x, p.a10.y + q.a10.y),
            make_double2(p.a11.x + q.a11.x, p.a11.y + q.a11.y)};
}

__device__ __forceinline__ CMat2 cmat_mul(const CMat2 &p, const CMat2 &q) {
    CMat2 r;
    r.a00 = cfma(cmul(p.a00, q.a00), p.a01, q.a10);
    r.a01 = cfma(cmul(p.a00, q.a01), p.a01, q.a11);
    r.a10 = cfma(cmul(p.a10, q.a00), p.a11, q.a10);
    r.a11 = cfma(cmul(p.a10, q.a01), p.a11, q.a11);
    return r;
}

// Tr(p q)
__device__ __forceinline__ double2 cmat_trace_mul(const CMat2 &p, const CMat2 &q) {
    double2 r = cmul(p.a00, q.a00);
    r = cfma(r, p.a01, q.a10);
    r = cfma(r, p.a10, q.a01);
    return cfma(r, p.a11, q.a11);
}

// Layer-1 states as a ring of 2x2 matrices.  The layer-0 output after its
// ring CNOT is chi(P^-1 y) with chi = (x) u_w and (P^-1 y)_0 = y_0 ^ y_{n-1},
// (P^-1 y)_1 = y_1 ^ y_0 ^ y_{n-1}, (P^-1 y)_k = y_k ^ y_{k-1}.  Rotating a set
// S of wires by U_w = RZ(phi_w) RY(theta_w) gives
//   c(q) = sum_y prod_{w in S} U_w[q_w, y_w] prod_{w not in S} [q_w == y_w] chi(P^-1 y)
//        = Tr(M_0[q_0] M_1[q_1] ... M_{n-1}[q_{n-1}])
// with bond variables x_0 = y_0 ^ y_{n-1} between wires 0 and 1, y_{w} between
// w and w + 1, and z = y_{n-1} closing the ring:
//   M_0[q](z, x_0) = U_0[q, x_0 ^ z] u_0(x_0),
//   M_w[q](x, y)   = U_w[q, y] u_w(y ^ x)   (w >= 1).
// k0: the wire's layer-0 coefficients, k1: layer 1's; rot: w in S; rz: with
// the wire's RZ (false: U_w = RY(theta_w) alone, the state with that wire's
// phase e^{-+ i phi/2} divided out, as an adjoint pass uses it).
__device__ __forceinline__ CMat2 chain_site(int w, int q, const WireCoef &k0, const WireCoef &k1,
                                            bool rot, bool rz = true) {
    const double2 u[2] = {u0_of(k0), u1_of(k0)};
    double2 U[2];  // U_w[q, 0], U_w[q, 1]
    if (rot) {
        // RZ(phi) RY(theta) = [[c e^-, -s e^-], [s e^+, c e^+]], e^-+ = e^{-+ i phi/2}
        const double2 em = rz ? make_double2(k1.ch, -k1.sh) : make_double2(1.0, 0.0);
        const double2 ep = rz ? make_double2(k1.ch, k1.sh) : make_double2(1.0, 0.0);
        if (q == 0) {
            U[0] = make_double2(k1.c * em.x, k1.c * em.y);
            U[1] = make_double2(-k1.s * em.x, -k1.s * em.y);
        } else {
            U[0] = make_double2(k1.s * ep.x, k1.s * ep.y);
            U[1] = make_double2(k1.c * ep.x, k1.c * ep.y);
        }
    } else {
        U[0] = make_double2(q == 0 ? 1.0 : 0.0, 0.0);
        U[1] = make_double2(q == 1 ? 1.0 : 0.0, 0.0);
    }
    CMat2 m;
    if (w == 0) {
        m.a00 = cmul(U[0], u[0]);
        m.a01 = cmul(U[1], u[1]);
        m.a10 = cmul(U[1], u[0]);
        m.a11 = cmul(U[0], u[1]);
    } else {
        m.a00 = cmul(U[0], u[0]);
        m.a01 = cmul(U[1], u[1]);
        m.a10 = cmul(U[0], u[1]);
        m.a11 = cmul(U[1], u[0]);
    }
    return m;
}

// v from lane (this lane permuted by the DPP control kCtrl) of the same row
// of 16 lanes.  Every control used here permutes within a full row, so no
// source lane is out of range and no "old" value needs to be set up.
template <int kCtrl>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), kCtrl, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), kCtrl, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_value(double v, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Sum over this lane's row of 16 lanes, returned to every lane of the row,
// through DPP lane moves (quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror, row_mirror), which use no LDS.
__device__ __forceinline__ double row_sum(double v) {
    v += dpp_move<0xB1>(v);
    v += dpp_move<0x4E>(v);
    v += dpp_move<0x141>(v);
    v += dpp_move<0x140>(v);
    return v;
}

// Row sums of eight values at once (transpose and add): a lane holds u[m],
// the pair sum over lanes {s, s ^ 1} of v[2m + (s & 1)] -- the first step,
// row_sums8_first, run by the caller as each pair of values is ready -- and
// row_sums8_rest halves the values twice more (partner lane ^ 2 by
// quad_perm [2,3,0,1], then the lane 4 along by row_ror:4, which keeps lane
// bits 0 and 1) and adds the lane 8 along (row_ror:8).  Lane s of a row
// returns the sum over the row's 16 lanes of v[s & 7]: 52 VALU instructions
// against 96 for eight row_sum calls.
__device__ __forceinline__ double row_sums8_first(double v_even, double v_odd, bool b0) {
    const double keep = b0 ? v_odd : v_even;
    const double send = b0 ? v_even : v_odd;
    return keep + dpp_move<0xB1>(send);
}

__device__ __forceinline__ double row_sums8_rest(const double (&u)[4], int lane) {
    const bool b1 = (lane & 2) != 0;
    const bool b2 = (lane & 4) != 0;
    double w[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const double keep = b1 ? u[2 * p + 1] : u[2 * p];
        const double send = b1 ? u[2 * p] : u[2 * p + 1];
        w[p] = keep + dpp_move<0x4E>(send);
    }
    const double keep = b2 ? w[1] : w[0];
    const double send = b2 ? w[0] : w[1];
    const double x = keep + dpp_move<0x124>(send);
    return x + dpp_move<0x128>(x);
}

// Row sums of two values: lane s of a row returns the sum over the row's 16
// lanes of v0 (s even) or v1 (s odd).
__device__ __forceinline__ double row_sums2(double v0, double v1, int lane) {
    double x = row_sums8_first(v0, v1, (lane & 1) != 0);
    x += dpp_move<0x4E>(x);
    x += dpp_move<0x124>(x);
    x += dpp_move<0x128>(x);
    return x;
}

// Sum over the 64 lanes of a (fully active) wave, returned to every lane in a
// fixed order: the four row sums read from lanes 0, 16, 32, 48.
__device__ __forceinline__ double wave_sum(double v) {
    v = row_sum(v);
    return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

// Streaming (non-temporal) 16-byte loads and stores for statevector arrays
// that one pass reads or writes once and no later access of the same pass
// touches again.  Measured at 30 qubits: the generating low pass, which only
// writes, runs 12.9 -> 12.0 ms with them (its ordinary stores stall in the L2's
// write path to HBM); the lambda loads of the high passes and the layer-0
// contraction gain a little.
typedef double stream_v2d __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void store_stream(double2 *p, double2 v) {
    stream_v2d x;
    x.x = v.x;
    x.y = v.y;
    __builtin_nontemporal_store(x, reinterpret_cast<stream_v2d *>(p));
}

__device__ __forceinline__ double2 load_stream(const double2 *p) {
    const stream_v2d x = __builtin_nontemporal_load(reinterpret_cast<const stream_v2d *>(p));
    return make_double2(x.x, x.y);
}

}  // namespace rie
