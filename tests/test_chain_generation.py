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


def _ring_fwd(i, n):
    """P(i): bits k >= 1 the prefix XOR of bits 0..k, bit 0 the parity of bits 1..n-1."""
    j = 0
    acc = 0
    for k in range(n):
        acc ^= (i >> k) & 1
        if k >= 1:
            j |= acc << k
    j |= (bin(i >> 1).count("1") & 1)
    return j


def _walls_after_ring(i):
    """Domain walls of P(i) around the ring, in closed form from i."""
    hi = bin(i >> 2).count("1")
    b0 = i & 1
    return hi + (b0 ^ (hi & 1)) + b0


def _flip_mask(w, n):
    """P^-1 e_w: bits w, w + 1 (w < n - 1), or n - 1, 0, 1 (w = n - 1)."""
    return (1 << w) | (1 << (w + 1)) if w < n - 1 else (1 << (n - 1)) | 3


def _crossing_wires(n):
    """The X terms the generating low pass takes besides its in-tile pairs:
    wire T - 1, the last wire of every high group but the last, and n - 1."""
    rest, width = n - T, T - 3
    groups = (rest + width - 1) // width
    g0, firsts = T, []
    for gi in range(groups):
        firsts.append(g0)
        g0 += rest // groups + (1 if gi < rest % groups else 0)
    return sorted({T - 1, *(f - 1 for f in firsts[1:]), n - 1})


def _high_group_pairs(n):
    """X_w the high passes take: pairs w, w + 1 inside one high group (the
    group's wires from its first to the one before its last)."""
    starts = [T] + [c + 1 for c in _crossing_wires(n) if T <= c < n - 1]
    ends = [c for c in _crossing_wires(n) if c >= T]
    return [w for s, e in zip(starts, ends) for w in range(s, e)]


@pytest.mark.parametrize("n", [13, 16, 22, 25, 30, 31])
def test_hamiltonian_terms_are_split_exactly_once(n):
    # in the pre-CNOT frame X_w of H is the flip P^-1 e_w; the generating low
    # pass takes w = 0 .. T - 2 (pairs inside its tile) and the crossing wires,
    # the high passes the pairs inside their groups: every w exactly once
    taken = list(range(T - 1)) + _crossing_wires(n) + _high_group_pairs(n)
    assert sorted(taken) == list(range(n))
    # the groups' pairs lie inside the groups the engine plans
    for w in _high_group_pairs(n):
        assert any((rot >> w) & 1 and (rot >> (w + 1)) & 1 for _, _, rot in _plan_cases(n)[1:])
    for w in range(n):
        assert _ring_fwd(_flip_mask(w, n), n) == 1 << w  # P^-1 e_w is X_w's flip


@pytest.mark.parametrize("n", [5, 9, 13])
def test_diagonal_after_ring_closed_form(n):
    for i in range(1 << n):
        j = _ring_fwd(i, n)
        rot = (j >> 1) | ((j & 1) << (n - 1))
        assert bin(j ^ rot).count("1") == _walls_after_ring(i), i


@pytest.mark.parametrize("n", [13, 16, 22])
@pytest.mark.parametrize("R", [3, 4])
def test_generating_low_pass_start_equals_hamiltonian(oracle, n, R):
    # lambda at the start of layer 1's backward sweep, in the pre-CNOT frame:
    # P^-1 H psi.  The generating low pass forms it for its tile as
    # -ZZ(P i) phi(i) - h sum_{w < T-1} phi(i ^ m_w) - h sum_{w crossing} phi(i ^ m_w),
    # the flipped amplitudes of crossing wires from the chain with U_A taken at
    # base ^ (m_w above the tile) and the register / thread index flipped by m_w
    # inside it; the high passes add -h phi(i ^ m_w) for the pairs inside
    # their groups.  Sampled tiles and threads against the oracle.
    rng = np.random.default_rng(130 + n + R)
    field = 0.85
    params = rng.uniform(0.0, 2.0 * np.pi, size=(2, n, 2))
    l0, l1 = _coefs(params[0]), _coefs(params[1])
    full = (1 << n) - 1
    psi = oracle.state(n, 2, params)
    hpsi = oracle.apply_hamiltonian(n, field, psi)
    sites, r0, run, chunks, ua, ub = _set(n, R, T, T, full, l0, l1)
    (f0, n0, t0), (f1, n1, t1) = chunks
    hi_pairs = _high_group_pairs(n)

    def phi_at(i):
        return _direct_trace(sites, n, i)

    err = 0.0
    for g in rng.integers(0, 1 << (n - T), 3):
        base = int(g) << T
        for t in rng.integers(0, 1 << (T - R), 8):
            t = int(t)
            for k in range(1 << R):
                o = (k << (T - R)) | t
                i = base | o
                y = _uniform(sites, ua, base) @ t0[t & 15] @ t1[(t >> 4) & ((1 << n1) - 1)]
                ph = np.trace(run[k] @ y)
                lam = -(n - 2 * _walls_after_ring(i)) * ph
                for w in range(T - 1):
                    lam -= field * phi_at(i ^ _flip_mask(w, n))
                for w in _crossing_wires(n):
                    m = _flip_mask(w, n)
                    hi_m, lo_m = m & ~((1 << T) - 1), m & ((1 << T) - 1)
                    k2 = k ^ ((lo_m >> (T - R)) & ((1 << R) - 1))
                    t2 = t ^ (lo_m & ((1 << (T - R)) - 1))
                    y2 = (_uniform(sites, ua, base ^ hi_m) @ t0[t2 & 15] @
                          t1[(t2 >> 4) & ((1 << n1) - 1)])
                    lam -= field * np.trace(run[k2] @ y2)
                for w in hi_pairs:
                    lam -= field * phi_at(i ^ _flip_mask(w, n))
                want = hpsi[oracle.ring_cnot_index(i, n)]
                err = max(err, abs(ph - psi[oracle.ring_cnot_index(i, n)]), abs(lam - want))
    assert err <= 1e-12
