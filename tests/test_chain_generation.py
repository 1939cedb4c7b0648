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


def _site(w, q, l0, l1, rot, rz=True):
    """rz = False: the wire's RY alone (the RZ phase left out of the state)."""
    c0, s0, ch0, sh0 = (x[w] for x in l0)
    u = (c0 * (ch0 - 1j * sh0), s0 * (ch0 + 1j * sh0))
    if rot:
        c, s, ch, sh = (x[w] for x in l1)
        em, ep = (ch - 1j * sh, ch + 1j * sh) if rz else (1.0, 1.0)
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


@pytest.mark.parametrize("n", [4, 5, 7])
def test_chain_trace_without_the_rz_of_a_wire_subset(oracle, n):
    # a high pass's table set leaves the RZ of the pass's own wires out: the
    # trace is then the layer-1 state with those wires' phases divided out
    rng = np.random.default_rng(60 + n)
    params = rng.uniform(0.0, 2.0 * np.pi, size=(2, n, 2))
    l0, l1 = _coefs(params[0]), _coefs(params[1])
    full = (1 << n) - 1
    idx = np.arange(1 << n)
    for rot, ry in ((full, 0b110), (full & ~1, full & ~1), (full & ~3, 0b1100), (full, full)):
        ph = params[1, :, 1]
        div = np.ones(1 << n, complex)
        for w in range(n):
            if (ry >> w) & 1:
                div *= np.where((idx >> w) & 1, np.exp(0.5j * ph[w]), np.exp(-0.5j * ph[w]))
        want = _rotated_reference(oracle, n, params, rot) / div
        got = np.empty(1 << n, complex)
        for i in range(1 << n):
            m = np.eye(2)
            for w in range(n):
                m = m @ _site(w, (i >> w) & 1, l0, l1, (rot >> w) & 1, not (ry >> w) & 1)
            got[i] = np.trace(m)
        assert np.max(np.abs(got - want)) <= 1e-14, (rot, ry)


def _sweep_grads(c, l, phases, order):
    """The gradient inputs a lambda-only sweep takes per rotated tile bit b:
    sum_x (-1)^{x_b} Im(conj(l[x]) c[x]) and sum_{x_b = 0} Re(conj(l[x | b])
    c[x]) - Re(conj(l[x]) c[x | b]), after the RZ^dagger of the bit groups in
    ``order`` (applied to both arrays group by group; ``phases[b]`` the bit's
    e^{-i phi/2}, e^{+i phi/2}).  Returns the per-bit values and the final l."""
    c, l = c.copy(), l.copy()
    x = np.arange(c.size)
    out = {}
    for group in order:
        for b in group:
            z = np.where((x >> b) & 1, phases[b][1], phases[b][0])
            c *= np.conj(z)
            l *= np.conj(z)
        for b in group:
            sgn = np.where((x >> b) & 1, -1.0, 1.0)
            im = np.sum(sgn * (np.conj(l) * c).imag)
            lo = x[((x >> b) & 1) == 0]
            hi = lo | (1 << b)
            re = np.sum((np.conj(l[hi]) * c[lo]).real - (np.conj(l[lo]) * c[hi]).real)
            out[b] = (im, re)
    return out, l


def test_phase_free_sweep_equals_the_phased_one():
    # the state generated without the pass's phases and lambda times the
    # conjugate of the tile's combined phase (once) give the gradients and the
    # lambda of the window-by-window RZ^dagger (window 0: tile bits 8..11,
    # window 2: bits 4..7, then bit 3; bits 0..2 ride along unrotated)
    rng = np.random.default_rng(91)
    x = np.arange(1 << T)
    c_true = rng.normal(size=1 << T) + 1j * rng.normal(size=1 << T)
    l = rng.normal(size=1 << T) + 1j * rng.normal(size=1 << T)
    phi = rng.uniform(0.0, 2.0 * np.pi, T)
    phases = {b: (np.exp(-0.5j * phi[b]), np.exp(0.5j * phi[b])) for b in range(3, T)}
    order = [range(8, 12), range(4, 8), [3]]
    want, l_want = _sweep_grads(c_true, l, phases, order)
    tile_phase = np.ones(1 << T, complex)
    for b in range(3, T):
        tile_phase *= np.where((x >> b) & 1, phases[b][1], phases[b][0])
    c_gen = c_true / tile_phase
    # the phase-free windows take the same gradient formulas with no phase work
    got, l_got = _sweep_grads(c_gen, l * np.conj(tile_phase),
                              {b: (1.0, 1.0) for b in range(3, T)}, order)
    for b in range(3, T):
        assert np.allclose(got[b], want[b], rtol=0, atol=1e-10), b
    assert np.max(np.abs(l_got - l_want)) <= 1e-13


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


def _plan_cases(n, bottom_up=False):
    """(tile low bits, first high index bit, rotated wires) of every table set
    as the engine plans them: the forward's low tile (all wires rotated), then
    one adjoint high pass per high group.  The n - 12 high wires split into
    ceil((n - 12) / 9) groups of near-equal size; group j's tile takes 9 wires
    from h0 = min(g0, n - 9).  With the high passes running from the top group
    down (the stored schedule, the walk) group j's state still carries the
    layer-1 rotations of groups 0 .. j (wires [12, end of group j)), the low
    wires and the groups above j being un-rotated already; with the passes
    running from the lowest group up (the generated start, bottom_up) it
    carries groups j .. K-1 (wires [g0 of group j, n))."""
    full = (1 << n) - 1
    a = 3
    cases = [(T, T, full)]
    rest, width = n - T, T - a
    groups = (rest + width - 1) // width
    g0 = T
    for gi in range(groups):
        bg = rest // groups + (1 if gi < rest % groups else 0)
        h0 = min(g0, n - width)
        if bottom_up:
            cases.append((a, h0, full & ~((1 << g0) - 1)))
        else:
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
    # the generated start's order: the top group last, with only its own wires
    c = _plan_cases(30, bottom_up=True)
    assert c[1][2] == ((1 << 30) - 1) & ~((1 << 12) - 1)
    assert c[2][2] == ((1 << 30) - 1) & ~((1 << 21) - 1)
    c = _plan_cases(31, bottom_up=True)
    assert [(ta, h0) for ta, h0, _ in c] == [(12, 12), (3, 12), (3, 19), (3, 22)]
    assert c[3][2] == ((1 << 31) - 1) & ~((1 << 25) - 1)


def _check_sampled(rng, n, R, params, want_at, tiles=4, threads=24, bottom_up=False):
    l0, l1 = _coefs(params[0]), _coefs(params[1])
    for ta, th0, rot in _plan_cases(n, bottom_up):
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
@pytest.mark.parametrize("bottom_up", [False, True])
def test_tile_tables_reproduce_states(oracle, n, R, bottom_up):
    # sampled tiles and threads of every table set against the oracle's states
    rng = np.random.default_rng(60 + n + R)
    params = rng.uniform(0.0, 2.0 * np.pi, size=(2, n, 2))

    def want_at(rot):
        v = _rotated_reference(oracle, n, params, rot)
        return lambda i: v[i]

    _check_sampled(rng, n, R, params, want_at, bottom_up=bottom_up)


@pytest.mark.parametrize("n", [30, 31])
@pytest.mark.parametrize("R", [3, 4])
@pytest.mark.parametrize("bottom_up", [False, True])
def test_tile_tables_at_the_largest_geometries(n, R, bottom_up):
    # n = 30 (two full groups, h0 == g0) and n = 31 (three groups, overlapping
    # tiles): sampled tiles and threads against the direct per-wire trace
    rng = np.random.default_rng(90 + n + R)
    params = rng.uniform(0.0, 2.0 * np.pi, size=(2, n, 2))
    l0, l1 = _coefs(params[0]), _coefs(params[1])

    def want_at(rot):
        sites = [[_site(w, q, l0, l1, (rot >> w) & 1) for q in (0, 1)] for w in range(n)]
        return lambda i: _direct_trace(sites, n, i)

    _check_sampled(rng, n, R, params, want_at, tiles=3, threads=6, bottom_up=bottom_up)


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


@pytest.mark.parametrize("n", [13, 16, 22, 25, 30, 31])
@pytest.mark.parametrize("R", [3, 4])
def test_out_of_tile_flips_split(n, R):
    # in the pre-CNOT frame X_w of H is the flip m_w = P^-1 e_w.  The
    # generating low pass takes w = 0 .. T - 2 inside its tile, and the rest
    # from flipped segments of U_A (the wires above the tile): wire T - 1
    # flips the register index's top bit and U_A's first wire, wire n - 1
    # flips thread bits 0, 1 and U_A's last wire, and w = T .. n - 2 flip two
    # wires of U_A and nothing inside the tile
    tile = (1 << T) - 1
    for w in range(n):
        m = _flip_mask(w, n)
        assert _ring_fwd(m, n) == 1 << w
        lo, hi = m & tile, m & ~tile
        if w <= T - 2:
            assert hi == 0
        elif w == T - 1:
            assert hi == 1 << T and lo == 1 << (T - 1)
            assert (lo >> (T - R)) == 1 << (R - 1) and lo & ((1 << (T - R)) - 1) == 0
        elif w == n - 1:
            assert hi == 1 << (n - 1) and lo == 3 and (lo >> (T - R)) == 0
        else:
            assert lo == 0 and hi == (1 << w) | (1 << (w + 1))


def _summed_pair_segment(sites, first, n, base):
    """Sum over w = first .. n - 2 of the segment prod_{v >= first} site[v]
    at base ^ e_w ^ e_{w+1}, by one walk with three running products: A the
    plain prefix, C the prefix whose last site is flipped, B the prefixes with
    one completed flipped pair."""
    bit = lambda v: (base >> v) & 1
    a = sites[first][bit(first)]
    c = sites[first][bit(first) ^ 1]
    b = np.zeros((2, 2), dtype=complex)
    for v in range(first + 1, n):
        s, f = sites[v][bit(v)], sites[v][bit(v) ^ 1]
        b, c, a = b @ s + c @ f, a @ f, a @ s
    return b


@pytest.mark.parametrize("n", [13, 14, 17])
def test_summed_pair_segment_matches_the_direct_sum(n):
    rng = np.random.default_rng(40 + n)
    params = rng.uniform(0.0, 2.0 * np.pi, size=(2, n, 2))
    l0, l1 = _coefs(params[0]), _coefs(params[1])
    full = (1 << n) - 1
    sites = [[_site(w, q, l0, l1, True) for q in (0, 1)] for w in range(n)]
    for base in [0, full & ~((1 << T) - 1)] + [int(x) << T for x in rng.integers(0, 1 << (n - T), 4)]:
        want = np.zeros((2, 2), dtype=complex)
        for w in range(T, n - 1):
            want += _prod(sites, T, n - T, (base ^ _flip_mask(w, n)) >> T)
        got = _summed_pair_segment(sites, T, n, base)
        assert np.max(np.abs(got - want)) <= 1e-13 * max(1.0, np.max(np.abs(want)))
    # one wire above the tile has no pair: the sum is empty
    assert not np.any(_summed_pair_segment(sites[:T + 1], T, T + 1, 0))


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
    # P^-1 H psi.  The generating low pass forms all of it for its tile as
    # -ZZ(P i) phi(i) - h sum_{w < T-1} phi(i ^ m_w) - h Tr(run[k ^ 2^(R-1)] Y_0)
    # - h Tr(run[k] Y_sum), with Y_0 from U_A at base ^ e_T (wire T - 1) and
    # Y_sum from U_A at base ^ e_{n-1} on thread row t ^ 3 (wire n - 1) plus
    # the summed pair segment on row t (w = T .. n - 2).  Sampled tiles and
    # threads against the oracle.
    rng = np.random.default_rng(130 + n + R)
    field = 0.85
    params = rng.uniform(0.0, 2.0 * np.pi, size=(2, n, 2))
    l0, l1 = _coefs(params[0]), _coefs(params[1])
    full = (1 << n) - 1
    psi = oracle.state(n, 2, params)
    hpsi = oracle.apply_hamiltonian(n, field, psi)
    sites, r0, run, chunks, ua, ub = _set(n, R, T, T, full, l0, l1)
    (f0, n0, t0), (f1, n1, t1) = chunks
    assert ua[0] == T and ua[0] + ua[1] == n

    def phi_at(i):
        return _direct_trace(sites, n, i)

    def y_of(u, t):
        return u @ t0[t & 15] @ t1[(t >> 4) & ((1 << n1) - 1)]

    err = 0.0
    for g in rng.integers(0, 1 << (n - T), 3):
        base = int(g) << T
        u0 = _prod(sites, T, n - T, (base ^ (1 << T)) >> T)
        u1 = _prod(sites, T, n - T, (base ^ (1 << (n - 1))) >> T)
        usum = _summed_pair_segment(sites, T, n, base)
        for t in rng.integers(0, 1 << (T - R), 8):
            t = int(t)
            y = y_of(_uniform(sites, ua, base), t)
            y0 = y_of(u0, t)
            ysum = y_of(u1, t ^ 3) + y_of(usum, t)
            for k in range(1 << R):
                i = base | (k << (T - R)) | t
                ph = np.trace(run[k] @ y)
                lam = -(n - 2 * _walls_after_ring(i)) * ph
                for w in range(T - 1):
                    lam -= field * phi_at(i ^ _flip_mask(w, n))
                lam -= field * np.trace(run[k ^ (1 << (R - 1))] @ y0)
                lam -= field * np.trace(run[k] @ ysum)
                want = hpsi[oracle.ring_cnot_index(i, n)]
                err = max(err, abs(ph - psi[oracle.ring_cnot_index(i, n)]), abs(lam - want))
    assert err <= 1e-12
