"""Layer-1 generation, checked on the CPU against the oracle.

The engine generates every state inside layer 1 (before its ring CNOT) as
Tr(M_0[q_0] M_1[q_1] ... M_{n-1}[q_{n-1}]), 2x2 complex matrices per wire built
from the layer-0 and layer-1 coefficients (``chain_site`` in
``gpu_simulation_amd/csrc/rie_device.hpp``), and evaluates the product per
tile as a register-run table times Y = U_A C0 U_B C1: two tile-uniform
segments (``chain_uniform_kernel``, U_B the identity in the forward's tile)
and two tables over the thread's own index bits (``chain_tables_kernel``,
``chain_y_low`` / ``chain_y_high``, ``chain_fill``).  These tests
restate both in numpy and compare with the oracle's layer-1 states: the state
after layer 0 (oracle, one layer) rotated on a wire subset, and P^-1 applied
to the oracle's two-layer state.  The GPU kernels themselves are checked
against the oracle and the stored path in ``test_gpu_parity.py``.
"""

from __future__ import annotations

import numpy as np
import pytest

T = 12


def _coefs(layer):
    th, ph = layer[:, 0], layer[:, 1]
    return np.cos(th / 2), np.sin(th / 2), np.cos(ph / 2), np.sin(ph / 2)


def _site(w, q, l0, l1, rot):
    c0, s0, ch0, sh0 = (x[w] for x in l0)
    u = (c0 * (ch0 - 1j * sh0), s0 * (ch0 + 1j * sh0))
    if rot:
        c, s, ch, sh = (x[w] for x in l1)
        em, ep = ch - 1j * sh, ch + 1j * sh
        U = np.array([[c * em, -s * em], [s * ep, c * ep]])
    else:
        U = np.eye(2)
    m = np.empty((2, 2), complex)
    for x in range(2):
        for y in range(2):
            m[x, y] = U[q, y ^ x] * u[y] if w == 0 else U[q, y] * u[y ^ x]
    return m


def _rotated_reference(oracle, n, params, rot):
    """Oracle state after layer 0 and its ring CNOT, then RZ RY of layer 1 on
    the wires of ``rot``."""
    v = oracle.state(n, 1, params[:1]).copy()
    c, s, ch, sh = _coefs(params[1])
    idx = np.arange(1 << n)
    for w in range(n):
        if (rot >> w) & 1:
            lo = idx[((idx >> w) & 1) == 0]
            hi = lo | (1 << w)
            a0, a1 = v[lo].copy(), v[hi].copy()
            v[lo] = (c[w] * a0 - s[w] * a1) * (ch[w] - 1j * sh[w])
            v[hi] = (s[w] * a0 + c[w] * a1) * (ch[w] + 1j * sh[w])
    return v


@pytest.mark.parametrize("n", [3, 4, 5, 7])
def test_chain_trace_equals_layer1_states(oracle, n):
    rng = np.random.default_rng(40 + n)
    params = rng.uniform(0.0, 2.0 * np.pi, size=(2, n, 2))
    l0, l1 = _coefs(params[0]), _coefs(params[1])
    full = (1 << n) - 1
    psi2 = oracle.state(n, 2, params)
    phi = np.array([psi2[oracle.ring_cnot_index(q, n)] for q in range(1 << n)])
    for rot in (0, 1, full & ~3, 1 << (n - 1), full):
        want = _rotated_reference(oracle, n, params, rot)
        got = np.empty(1 << n, complex)
        for i in range(1 << n):
            m = np.eye(2)
            for w in range(n):
                m = m @ _site(w, (i >> w) & 1, l0, l1, (rot >> w) & 1)
            got[i] = np.trace(m)
        assert np.max(np.abs(got - want)) <= 1e-14, rot
        if rot == full:
            # every wire of layer 1 rotated: P^-1 of the two-layer state
            assert np.max(np.abs(got - phi)) <= 1e-14


def _omap(R, M, t, k):
    lo = T - R if M == 0 else (M - 1) * R
    return ((t >> lo) << (lo + R)) | (k << lo) | (t & ((1 << lo) - 1))


def _prod(sites, first, length, v):
    """Product of the site matrices of wires first .. first + length - 1 at the
    bits of v (bit b is wire first + b)."""
    m = np.eye(2)
    for b in range(length):
        m = m @ sites[first + b][(v >> b) & 1]
    return m


def _set(n, R, ta, th0, rot, l0, l1):
    """A table set as the engine lays it out for a pass whose tile is the index
    bits [0, ta) and [th0, th0 + T - ta), run at the tile's top wires: the run
    products, the thread chunks C0 and C1 (as (first, length, table)), and the
    uniform segments U_A = [r0 + R, n) and U_B (as (first, length))."""
    sites = [[_site(w, q, l0, l1, (rot >> w) & 1) for q in (0, 1)] for w in range(n)]
    if ta == T:  # the forward's low tile
        r0 = T - R
        c0, c1, ub = (0, 4), (4, T - R - 4), (0, 0)
    else:  # an adjoint high pass
        r0 = th0 + (T - R - ta)
        c0, c1, ub = (0, ta), (th0, T - R - ta), (ta, th0 - ta)
    run = [_prod(sites, r0, R, v) for v in range(1 << R)]
    chunks = [(f, ln, [_prod(sites, f, ln, v) for v in range(1 << ln)]) for f, ln in (c0, c1)]
    return sites, r0, run, chunks, (r0 + R, n - r0 - R), ub


def _uniform(sites, seg, base):
    first, length = seg
    return _prod(sites, first, length, base >> first)


def _chain_y(sites, chunks, ua, ub, i, base):
    """Y(i) = U_A . C0 . U_B . C1 (U_B the identity for the forward's tile)."""
    (f0, n0, t0), (f1, n1, t1) = chunks
    y = _uniform(sites, ua, base) @ t0[(i >> f0) & ((1 << n0) - 1)]
    y = y @ _uniform(sites, ub, base)
    return y @ t1[(i >> f1) & ((1 << n1) - 1)]


@pytest.mark.parametrize("n", [13, 16, 22])
@pytest.mark.parametrize("R", [3, 4])
def test_tile_tables_reproduce_states(oracle, n, R):
    # the forward's low tile (every wire rotated) and the adjoint high passes
    # (the high group's wires rotated), on sampled tiles and threads
    rng = np.random.default_rng(60 + n + R)
    params = rng.uniform(0.0, 2.0 * np.pi, size=(2, n, 2))
    l0, l1 = _coefs(params[0]), _coefs(params[1])
    full = (1 << n) - 1
    a = 3
    cases = [(T, T, full)]  # (tile low bits, first high index bit, rotated wires)
    # the high groups as the engine plans them: the n - 12 high wires split
    # into ceil((n - 12) / 9) groups of near-equal size, each tile's high
    # part 9 wires from h0 = min(g0, n - 9)
    rest, width = n - T, T - a
    groups = (rest + width - 1) // width
    g0 = T
    for gi in range(groups):
        bg = rest // groups + (1 if gi < rest % groups else 0)
        h0 = min(g0, n - width)
        cases.append((a, h0, full & ~((1 << g0) - 1) & ((1 << (g0 + bg)) - 1)))
        g0 += bg
    for ta, th0, rot in cases:
        want = _rotated_reference(oracle, n, params, rot)
        sites, r0, run, chunks, ua, ub = _set(n, R, ta, th0, rot, l0, l1)
        gap = th0 - ta
        err = 0.0
        for g in rng.integers(0, 1 << (n - T), 4):
            g = int(g)
            base = ((g & ((1 << gap) - 1)) << ta) | ((g >> gap) << (th0 + T - ta))
            for t in rng.integers(0, 1 << (T - R), 24):
                t = int(t)

                def index(o):
                    return base | (o & ((1 << ta) - 1)) | ((o >> ta) << th0)

                y = _chain_y(sites, chunks, ua, ub, index(_omap(R, 0, t, 0)), base)
                for k in range(1 << R):
                    err = max(err, abs(np.trace(run[k] @ y) - want[index(_omap(R, 0, t, k))]))
        assert err <= 1e-14, (ta, th0, r0)
