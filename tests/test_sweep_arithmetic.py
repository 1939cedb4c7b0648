"""The adjoint sweep's rewritten arithmetic, restated in numpy and checked
against the direct forms it replaces.

* RY^dagger in tangent form (``ry_inv_pair_tan_a`` / ``ry_inv_pair_tan_b`` in
  ``gpu_simulation_amd/csrc/rie_device.hpp``, ``ry_inv_window_tan`` and
  ``ry_tan_tables`` in ``rie_kernels.hip``): per wire one FMA per output
  component, the factor c or s collected into one scale per pass.
* The gradient row sums by transpose and add (``row_sums8_first`` /
  ``row_sums8_rest`` / ``row_sums2``): the DPP lane moves emulated on 64
  lanes; lane s of each row must end with the row sum of value s & 7.
* The signed phase sums of a 4-register window from quad partials
  (``grad_window``).
* The generating low pass's H diagonal from two values per thread.
* Tile bit 3's gradient from the partner lane's lambda, the same lane move
  that its RY^dagger takes (``grad_ry_lane_bit3``).

The GPU kernels themselves are checked against the oracle in
``test_gpu_parity.py``; these tests pin the algebra on the CPU, and a planted
sign or index error in any of the restatements fails them.
"""

from __future__ import annotations

import numpy as np
import pytest


def _ry_inv(a, b, c, s):
    return c * a + s * b, -s * a + c * b


def _ry_inv_tan(a, b, c, s):
    """One wire in tangent form; returns the unscaled pair and the factor."""
    if abs(s) > abs(c):
        t = c / s
        return t * a + b, t * b - a, s
    t = s / c
    return a + t * b, b - t * a, c


def _sweep(x, thetas, tan):
    """RY^dagger of every wire on a statevector of len(thetas) qubits."""
    x = x.copy()
    sigma = 1.0
    n = len(thetas)
    idx = np.arange(x.size)
    for w in range(n):
        c, s = np.cos(thetas[w] / 2), np.sin(thetas[w] / 2)
        lo = idx[((idx >> w) & 1) == 0]
        hi = lo | (1 << w)
        if tan:
            x[lo], x[hi], f = _ry_inv_tan(x[lo], x[hi], c, s)
            sigma *= f
        else:
            x[lo], x[hi] = _ry_inv(x[lo], x[hi], c, s)
    return x * sigma


def test_tangent_form_equals_the_rotation():
    rng = np.random.default_rng(3)
    n = 12
    x = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    # random angles plus the ones where a form changes or c, s vanish
    thetas = rng.uniform(0.0, 4.0 * np.pi, n)
    thetas[:6] = [0.0, np.pi, 2 * np.pi, np.pi / 2, 3 * np.pi / 2, -np.pi / 2]
    want = _sweep(x, thetas, tan=False)
    got = _sweep(x, thetas, tan=True)
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))


def test_tangent_form_keeps_the_tangent_bounded():
    # the form is chosen so |t| <= 1, and the scale stays >= 2^(-w/2)
    for theta in np.linspace(0.0, 4.0 * np.pi, 401):
        c, s = np.cos(theta / 2), np.sin(theta / 2)
        t = c / s if abs(s) > abs(c) else s / c
        assert abs(t) <= 1.0 + 1e-15
        assert max(abs(c), abs(s)) >= np.sqrt(0.5) - 1e-15


# --- DPP lane moves on one wave (64 lanes, rows of 16) ---------------------

def _quad_perm(v, perm):
    lane = np.arange(64)
    return v[(lane & ~3) | np.asarray(perm)[lane & 3]]


def _row_ror(v, n, direction):
    lane = np.arange(64)
    return v[(lane & ~15) | ((lane + direction * n) & 15)]


def _row_sums8(v, direction):
    """v: (8, 64) per-lane values; returns the per-lane result."""
    lane = np.arange(64)
    b0 = (lane & 1) != 0
    b1 = (lane & 2) != 0
    b2 = (lane & 4) != 0
    u = []
    for m in range(4):
        keep = np.where(b0, v[2 * m + 1], v[2 * m])
        send = np.where(b0, v[2 * m], v[2 * m + 1])
        u.append(keep + _quad_perm(send, [1, 0, 3, 2]))
    w = []
    for p in range(2):
        keep = np.where(b1, u[2 * p + 1], u[2 * p])
        send = np.where(b1, u[2 * p], u[2 * p + 1])
        w.append(keep + _quad_perm(send, [2, 3, 0, 1]))
    keep = np.where(b2, w[1], w[0])
    send = np.where(b2, w[0], w[1])
    x = keep + _row_ror(send, 4, direction)
    return x + _row_ror(x, 8, direction)


def _row_sums2(v0, v1, direction):
    lane = np.arange(64)
    b0 = (lane & 1) != 0
    x = np.where(b0, v1, v0) + _quad_perm(np.where(b0, v0, v1), [1, 0, 3, 2])
    x = x + _quad_perm(x, [2, 3, 0, 1])
    x = x + _row_ror(x, 4, direction)
    return x + _row_ror(x, 8, direction)


@pytest.mark.parametrize("direction", [1, -1])
def test_transposed_row_sums(direction):
    # whichever way row_ror rotates, lane s of a row holds the row sum of
    # value s & 7 (s & 1 for two values)
    rng = np.random.default_rng(11)
    v = rng.normal(size=(8, 64))
    got = _row_sums8(v, direction)
    lane = np.arange(64)
    want = np.array([v[l & 7, (l & ~15):(l & ~15) + 16].sum() for l in lane])
    assert np.allclose(got, want, rtol=0, atol=1e-12)
    g2 = _row_sums2(v[0], v[1], direction)
    want2 = np.array([v[l & 1, (l & ~15):(l & ~15) + 16].sum() for l in lane])
    assert np.allclose(g2, want2, rtol=0, atol=1e-12)


def test_phase_sums_from_quad_partials():
    # dps[b] = sum_q (-1)^{q_b} im_q over 16 registers, as grad_window forms it
    rng = np.random.default_rng(12)
    im = rng.normal(size=16)
    dps = np.zeros(4)
    for x in range(4):
        q = im[4 * x:4 * x + 4]
        a0, a1 = q[0] + q[1], q[2] + q[3]
        d = (q[0] - q[1]) + (q[2] - q[3])
        s, e = a0 + a1, a0 - a1
        dps[0] += d
        dps[1] += e
        dps[2] += -s if x & 1 else s
        dps[3] += -s if x & 2 else s
    k = np.arange(16)
    want = [np.sum(np.where((k >> b) & 1, -im, im)) for b in range(4)]
    assert np.allclose(dps, want, rtol=0, atol=1e-12)


def _ham_diag_pre_ring(i, n):
    """H's diagonal at P(i) from the pre-CNOT index i, per index."""
    hi = bin(i >> 2).count("1")
    b0 = i & 1
    return 2 * (hi + (b0 ^ (hi & 1)) + b0) - n


def _ring_fwd(i, n):
    """P(i): bit k >= 1 the parity of bits 0..k, bit 0 the parity of bits 1..n-1."""
    p = i
    for sh in (1, 2, 4, 8, 16, 32):
        p ^= p << sh
    p &= (1 << n) - 1
    return (p & ~1) | (bin(i >> 1).count("1") & 1)


def test_pre_ring_diagonal_equals_the_walls_of_the_image():
    # -sum_w Z_w Z_{w+1 mod n} at j = P(i) is 2 (domain walls of j) - n
    rng = np.random.default_rng(15)
    for _ in range(3000):
        n = int(rng.integers(3, 35))
        i = int(rng.integers(0, 1 << n))
        j = _ring_fwd(i, n)
        rot = ((j << 1) | (j >> (n - 1))) & ((1 << n) - 1)
        assert _ham_diag_pre_ring(i, n) == 2 * bin(j ^ rot).count("1") - n


def test_diagonal_from_two_values_per_thread():
    # the generating low pass forms H's diagonal of register k (index bits
    # 8..11) from two per-thread values: d = dpar[popcount(k) & 1] + 2 popcount(k)
    rng = np.random.default_rng(13)
    for _ in range(2000):
        n = int(rng.integers(13, 35))
        i0 = int(rng.integers(0, 1 << n)) & ~(0xF << 8)
        hi0 = bin(i0 >> 2).count("1")
        b0 = i0 & 1
        e0 = b0 ^ (hi0 & 1)
        dpar = (2 * (hi0 + b0 + e0) - n, 2 * (hi0 + b0 + (e0 ^ 1)) - n)
        for k in range(16):
            pk = bin(k).count("1")
            assert dpar[pk & 1] + 2 * pk == _ham_diag_pre_ring(i0 | (k << 8), n)


def test_bit3_gradient_from_the_partner_lambda():
    # tile bit 3 is lane bit 3: a pair's lanes hold (c0, l0) and (c1, l1).
    # The sweep moves lam only: Re(p* c) on the lane with bit 3 clear and
    # -Re(p* c) on the other (p the partner's lam) sum to the pair's
    # Re(l1* c0) - Re(l0* c1), and p is what RY^dagger of bit 3 takes.
    rng = np.random.default_rng(14)
    c = rng.normal(size=(64, 16)) + 1j * rng.normal(size=(64, 16))
    l = rng.normal(size=(64, 16)) + 1j * rng.normal(size=(64, 16))
    lane = np.arange(64)
    partner = (lane & ~15) | ((lane + 8) & 15)
    hi = ((lane >> 3) & 1) != 0
    p = l[partner]
    re = (np.conj(p) * c).real
    got = np.sum(np.where(hi[:, None], -re, re))
    lo = lane[~hi]
    want = np.sum((np.conj(l[lo | 8]) * c[lo]).real - (np.conj(l[lo]) * c[lo | 8]).real)
    assert abs(got - want) <= 1e-11 * max(1.0, abs(want))
    # RY^dagger across the lanes equals the pair rotation
    th = 1.234
    cc, ss = np.cos(th / 2), np.sin(th / 2)
    s_lane = np.where(hi, -ss, ss)[:, None]
    out = cc * l + s_lane * p
    a, b = l[lo], l[lo | 8]
    assert np.allclose(out[lo], cc * a + ss * b) and np.allclose(out[lo | 8], -ss * a + cc * b)
