"""Layer-0 contraction inside the last high pass, on the CPU.

When a two-layer evaluation runs its high passes from the lowest group up, the
last pass holds the top group's tiles: index bits {0, 1, 2} and [h, n) of the
frame after layer 0's ring CNOT (h = n - 9), the other bits fixed per tile.
Under P^-1 such a tile is, in layer-0 coordinates i = P^-1 j, a product domain:
i_0, i_1, i_2 and i_h .. i_{n-1} run over all 2^12 values, i_4 .. i_{h-1} are
fixed by the tile, and i_3 = j_3 ^ i_0 ^ i_1 ^ i_2 couples three of the free
bits.  So the layer-0 environments

    E_v[b] = sum_{i : i_v = b} conj(lambda[P(i)]) prod_{u != v} u_u(i_u)

follow from per-tile work shaped like the separate contraction pass: lambda
permuted through LDS to q = (i_0, i_1, i_2, i_h, .., i_{n-1}), a product
contraction over the registers, per-thread sums split by j_3 ^ parity(q_0..q_2),
a per-tile sum for the fixed wires.  These tests emulate that arithmetic step
by step (``fused_environments``, in the order the kernel uses) and compare it
with the direct sums above for random lambda and layer-0 coefficients, at sizes
with one, two and three high groups, the tile's top group overlapping the one
below it or not.
"""

from __future__ import annotations

import numpy as np
import pytest

T = 12


def ring_fwd(i, n):
    """P(i): bit k >= 1 the parity of bits 0..k, bit 0 the parity of bits 1..n-1."""
    p = i.copy()
    for s in (1, 2, 4, 8, 16, 32):
        p ^= p << s
    p &= (1 << n) - 1
    par = np.zeros_like(i)
    x = i >> 1
    while np.any(x):
        par ^= x & 1
        x >>= 1
    return (p & ~1) | par


def direct_environments(lam, u, n):
    """E[v, b] from the definition (u[v, b] = u_v(b))."""
    i = np.arange(1 << n, dtype=np.int64)
    mu = lam[ring_fwd(i, n)]
    bits = (i[:, None] >> np.arange(n)) & 1
    fac = u[np.arange(n)[None, :], bits]  # (N, n)
    E = np.zeros((n, 2), dtype=complex)
    for v in range(n):
        rest = np.prod(np.delete(fac, v, axis=1), axis=1)
        for b in range(2):
            sel = bits[:, v] == b
            E[v, b] = np.sum(np.conj(mu[sel]) * rest[sel])
    return E


def q_of_o(o, c):
    """Layer-0 coordinates of top-tile offset o (bits 0..2: j_0..j_2, bits
    3..11: j_h..j_{n-1}); c = j_{h-1}."""
    ob = [(o >> b) & 1 for b in range(T)]
    q = [ob[0] ^ ob[11], ob[0] ^ ob[1] ^ ob[11], ob[1] ^ ob[2], ob[3] ^ c]
    q += [ob[m] ^ ob[m - 1] for m in range(4, T)]
    return sum(q[b] << b for b in range(T))


def fused_environments(lam, u, n):
    """The kernel's arithmetic: per top tile g and thread t (mapping 0:
    thread bits 0..7, register bits 8..11), lambda permuted to q, the register
    product contraction pk and environments env, the thread factor a_t and
    parity par_t, the tile factor Phi_C(g) over the fixed wires 4..h-1, per
    thread sums zc[x] (x = j_3 ^ par_t), the register wires' environments with
    the thread factor applied per tile, per tile S(g); then the fixed wires'
    environments over the tiles."""
    h = n - 9
    wire = [0, 1, 2] + list(range(h, n))  # wire of q bit b
    o = np.arange(1 << T)
    t = np.arange(256)
    k = np.arange(16)
    tiles = 1 << (n - T)
    # per-thread (tile-independent) factors
    a_t = np.ones(256, dtype=complex)
    for b in range(8):
        a_t *= u[wire[b], (t >> b) & 1]
    par_t = (t ^ (t >> 1) ^ (t >> 2)) & 1
    kb = (k[:, None] >> np.arange(4)) & 1  # (16, 4)
    ufk = u[np.array(wire[8:])[None, :], kb]  # (16, 4): u_{n-4+e}(k_e)
    zc = np.zeros((256, 2), dtype=complex)
    benv = np.zeros((256, 4, 2), dtype=complex)
    S = np.zeros(tiles, dtype=complex)
    for g in range(tiles):
        base = g << 3  # j bits [3, h) = tile bits (no bits above n - 1)
        c = (base >> (h - 1)) & 1
        j3 = (base >> 3) & 1
        idx = base | (o & 7) | ((o >> 3) << h)
        m = np.zeros(1 << T, dtype=complex)
        m[np.array([q_of_o(int(x), c) for x in o])] = lam[idx]
        mq = m.reshape(16, 256).T  # [t, k]: q = t | k << 8
        cm = np.conj(mq)
        pk = cm @ np.prod(ufk, axis=1)
        env = np.zeros((256, 4, 2), dtype=complex)
        for e in range(4):
            skip = np.prod(np.delete(ufk, e, axis=1), axis=1)
            for b in range(2):
                env[:, e, b] = cm[:, kb[:, e] == b] @ skip[kb[:, e] == b]
        phi_c = 1.0 + 0j
        for v in range(4, h):
            phi_c *= u[v, ((base >> v) ^ (base >> (v - 1))) & 1]
        x = j3 ^ par_t
        zc[np.arange(256), x] += phi_c * pk
        # register wires: a_t u_3(x) phi_c env, summed over the threads per tile
        benv += (a_t * u[3, x] * phi_c)[:, None, None] * env
        S[g] = np.sum(a_t * u[3, x] * pk)
    E = np.zeros((n, 2), dtype=complex)
    # thread-held wires (q bits 0..7)
    zsum = u[3, 0] * zc[:, 0] + u[3, 1] * zc[:, 1]
    for b_w in range(8):
        skip = np.ones(256, dtype=complex)
        for b in range(8):
            if b != b_w:
                skip *= u[wire[b], (t >> b) & 1]
        for b in range(2):
            sel = ((t >> b_w) & 1) == b
            E[wire[b_w], b] = np.sum(skip[sel] * zsum[sel])
    # wire 3
    for b in range(2):
        E[3, b] = np.sum(a_t * zc[:, b])
    # register wires
    for e in range(4):
        for b in range(2):
            E[wire[8 + e], b] = np.sum(benv[:, e, b])
    # fixed wires 4..h-1 from the tile sums
    g = np.arange(tiles)
    base = g << 3
    for v in range(4, h):
        f = S.copy()
        for w_ in range(4, h):
            if w_ != v:
                f = f * u[w_, ((base >> w_) ^ (base >> (w_ - 1))) & 1]
        iv = ((base >> v) ^ (base >> (v - 1))) & 1
        for b in range(2):
            E[v, b] = np.sum(f[iv == b])
    return E


def random_u(n, rng):
    th = rng.uniform(0, 2 * np.pi, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    # u_w = RZ(phi) RY(theta) |0> = (cos(theta/2) e^{-i phi/2}, sin(theta/2) e^{i phi/2})
    return np.stack([np.cos(th / 2) * np.exp(-0.5j * ph), np.sin(th / 2) * np.exp(0.5j * ph)], axis=1)


@pytest.mark.parametrize("n", [13, 14, 16, 17])
def test_fused_environments_equal_the_direct_sums(n):
    rng = np.random.default_rng(700 + n)
    lam = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    u = random_u(n, rng)
    E_dir = direct_environments(lam, u, n)
    E_fus = fused_environments(lam, u, n)
    assert np.max(np.abs(E_fus - E_dir)) <= 1e-9 * np.max(np.abs(E_dir))


def test_fused_environments_at_special_angles():
    # zeros in the layer-0 factors (theta = 0 or pi): no division anywhere
    n = 14
    rng = np.random.default_rng(77)
    lam = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    u = random_u(n, rng)
    u[3] = [1.0, 0.0]
    u[0] = [0.0, 1.0]
    u[n - 1] = [1.0, 0.0]
    u[7] = [0.0, 1.0j]
    E_dir = direct_environments(lam, u, n)
    E_fus = fused_environments(lam, u, n)
    assert np.max(np.abs(E_fus - E_dir)) <= 1e-9 * np.max(np.abs(E_dir))


def test_q_map_is_a_bijection_and_inverts_the_ring_cnot():
    # q_of_o lists, for each tile offset, the free layer-0 bits of
    # P^-1(base | offset); checked against ring_fwd at 16 qubits
    n, h = 16, 7
    rng = np.random.default_rng(5)
    for g in [0, 1, 5, 15] + [int(x) for x in rng.integers(0, 1 << (n - T), 4)]:
        base = g << 3
        c = (base >> (h - 1)) & 1
        o = np.arange(1 << T)
        q = np.array([q_of_o(int(x), c) for x in o])
        assert sorted(q.tolist()) == list(range(1 << T))
        j = base | (o & 7) | ((o >> 3) << h)
        # i = P^-1 j has free bits q at wires 0,1,2,h..n-1 and fixed bits elsewhere
        i = np.zeros_like(j)
        for b, w in enumerate([0, 1, 2] + list(range(h, n))):
            i |= ((q >> b) & 1) << w
        fixed = 0
        for v in range(4, h):
            fixed |= (((base >> v) ^ (base >> (v - 1))) & 1) << v
        j3 = (base >> 3) & 1
        i3 = j3 ^ (q & 1) ^ ((q >> 1) & 1) ^ ((q >> 2) & 1)
        i |= fixed | (i3 << 3)
        assert np.array_equal(ring_fwd(i, n), j)
