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


def _plan_cases(n):
    """(tile low bits, first high index bit, rotated wires) of every table set
    as the engine plans them: the forward's low tile (all wires rotated), then
    one adjoint high pass per high group.  The n - 12 high wires split into
    ceil((n - 12) / 9) groups of near-equal size; group j's tile takes 9 wires
    from h0 = min(g0, n - 9), and its state still carries the layer-1
    rotations of groups 0 .. j (wires [12, end of group j)), the low wires and
    the groups above j being un-rotated already."""
    full = (1 << n) - 1
    a = 3
    cases = [(T, T, full)]
    rest, width = n - T, T - a
    groups = (rest + width - 1) // width
    g0 = T
    for gi in range(groups):
        bg = rest // groups + (1 if gi < rest % groups else 0)
        h0 = min(g0, n - width)
        cases.append((a, h0, full & ~((1 << T) - 1) & ((1 << (g0 + bg)) - 1)))
        g0 += bg
    return cases


def _direct_trace(sites, n, i):
    """Tr(M_0[i_0] ... M_{n-1}[i_{n-1}]): the layer-1 amplitude at index i,
    wire by wire (tied to the oracle by test_chain_trace_equals_layer1_states)."""
    m = np.eye(2)
    for w in range(n):
        m = m @ sites[w][(i >> w) & 1]
    return np.trace(m)


def test_plan_cases_follow_the_engine_groups():
    # n = 30: two full 9-wire groups, h0 == g0 (wires 12..20 and 21..29)
    c = _plan_cases(30)
    assert [(ta, h0) for ta, h0, _ in c] == [(12, 12), (3, 12), (3, 21)]
    assert c[1][2] == ((1 << 21) - 1) & ~((1 << 12) - 1)
    assert c[2][2] == ((1 << 30) - 1) & ~((1 << 12) - 1)
    # n = 31: three groups of 7, 6, 6 wires; the last two tiles overlap lower wires
    c = _plan_cases(31)
    assert [(ta, h0) for ta, h0, _ in c] == [(12, 12), (3, 12), (3, 19), (3, 22)]
    assert c[2][2] == ((1 << 25) - 1) & ~((1 << 12) - 1)


def _check_sampled(rng, n, R, params, want_at, tiles=4, threads=24):
    l0, l1 = _coefs(params[0]), _coefs(params[1])
    for ta, th0, rot in _plan_cases(n):
        want = want_at(rot)
        sites, r0, run, chunks, ua, ub = _set(n, R, ta, th0, rot, l0, l1)
        gap = th0 - ta
        err = 0.0
        for g in rng.integers(0, 1 << (n - T), tiles):
            g = int(g)
            base = ((g & ((1 << gap) - 1)) << ta) | ((g >> gap) << (th0 + T - ta))
            for t in rng.integers(0, 1 << (T - R), threads):
                t = int(t)

                def index(o):
                    return base | (o & ((1 << ta) - 1)) | ((o >> ta) << th0)

                y = _chain_y(sites, chunks, ua, ub, index(_omap(R, 0, t, 0)), base)
                for k in range(1 << R):
                    err = max(err, abs(np.trace(run[k] @ y) - want(index(_omap(R, 0, t, k)))))
        assert err <= 1e-14, (ta, th0, r0)


@pytest.mark.parametrize("n", [13, 16, 22])
@pytest.mark.parametrize("R", [3, 4])
def test_tile_tables_reproduce_states(oracle, n, R):
    # sampled tiles and threads of every table set against the oracle's states
    rng = np.random.default_rng(60 + n + R)
    params = rng.uniform(0.0, 2.0 * np.pi, size=(2, n, 2))

    def want_at(rot):
        v = _rotated_reference(oracle, n, params, rot)
        return lambda i: v[i]

    _check_sampled(rng, n, R, params, want_at)


@pytest.mark.parametrize("n", [30, 31])
@pytest.mark.parametrize("R", [3, 4])
def test_tile_tables_at_the_largest_geometries(n, R):
    # n = 30 (two full groups, h0 == g0) and n = 31 (three groups, overlapping
    # tiles): sampled tiles and threads against the direct per-wire trace
    rng = np.random.default_rng(90 + n + R)
    params = rng.uniform(0.0, 2.0 * np.pi, size=(2, n, 2))
    l0, l1 = _coefs(params[0]), _coefs(params[1])

    def want_at(rot):
        sites = [[_site(w, q, l0, l1, (rot >> w) & 1) for q in (0, 1)] for w in range(n)]
        return lambda i: _direct_trace(sites, n, i)

    _check_sampled(rng, n, R, params, want_at, tiles=3, threads=6)
