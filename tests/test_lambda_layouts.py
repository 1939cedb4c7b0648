"""Lambda layouts of the generated start, on the CPU.

Between the passes of a generated-start evaluation lambda sits in the order
that makes the next high pass's tile one contiguous 2^12 block.  These tests
restate the layout (``make_lam_layout`` / ``lay_addr`` / ``lay_fits`` of
``gpu_simulation_amd/csrc/rie_kernels.hpp``) and the engine's pass plan and
check what the kernels rely on: each layout is a permutation of the index
bits below n; under its own layout a high pass's tile is an aligned block
whose offsets appear in order; the register bits of every pass's load mapping
stay contiguous in the layouts it reads and writes, wherever the engine turns
the layouts on (all sizes but 18..20 qubits); each writer's tile stays within
2^21 amplitudes; and the chain of layouts hands every amplitude from the pass
that writes it to the pass that reads it.  The kernels themselves are compared
with and without layouts on the GPU.
"""

from __future__ import annotations

import numpy as np
import pytest

T = 12


def make_lam_layout(n, h, w):
    """make_lam_layout: address bits 0..11 hold the reader's tile bits
    ({0, 1, 2} and [h, h + 9)), the writer's tile bits outside it ({0, 1, 2}
    and [w, w + 9), w = 3 for the low tile) come next, then every other bit
    below n, each group ascending.  Returned as segments (src, len, dst);
    [] is the natural order."""
    order = []
    for b in list(range(3)) + list(range(h, h + 9)) + list(range(w, w + 9)) + list(range(3, n)):
        if 0 <= b < n and b not in order:
            order.append(b)
    segs = []
    for d, b in enumerate(order):
        if segs and segs[-1][0] + segs[-1][1] == b and segs[-1][2] + segs[-1][1] == d:
            segs[-1][1] += 1
        else:
            segs.append([b, 1, d])
    if len(segs) > 8:
        return None
    return [] if len(segs) == 1 else [tuple(x) for x in segs]


def lay_addr(i, L):
    r = 0
    if not L:
        return i
    for src, ln, dst in L:
        r |= ((i >> src) & ((1 << ln) - 1)) << dst
    return r


def lay_fits(L, n, first, count):
    if L is None:
        return False
    if not L:
        return True
    image = 0
    for b in range(n):
        image |= lay_addr(1 << b, L)
    if image != (1 << n) - 1:
        return False
    p0 = lay_addr(1 << first, L)
    return all(lay_addr(1 << (first + b), L) == p0 << b for b in range(1, count))


def high_groups(n):
    """h0 of every high pass as the engine plans them (a = 3, groups of at
    most 9 wires above the 12-bit low tile, h0 = min(g0, n - 9))."""
    rest, width = n - T, T - 3
    groups = (rest + width - 1) // width
    g0, out = T, []
    for gi in range(groups):
        bg = rest // groups + (1 if gi < rest % groups else 0)
        out.append(min(g0, n - width))
        g0 += bg
    return out


def layout_in(n, j, bottom_up=True):
    """Engine::lam_layout_in: the high passes run j = 0 .. K-1 (bottom_up,
    lowest group first) or K-1 .. 0; pass j's input was written by the pass
    that runs before it (the low pass for the first)."""
    h0s = high_groups(n)
    prev = j - 1 if bottom_up else j + 1
    w = 3 if prev < 0 or prev >= len(h0s) else h0s[prev]
    return make_lam_layout(n, h0s[j], w)


def layout_out(n, j, bottom_up=True):
    nxt = j + 1 if bottom_up else j - 1
    return layout_in(n, nxt, bottom_up) if 0 <= nxt < len(high_groups(n)) else []


def layouts_fit(n, r_low=4, r_high=4, bottom_up=True):
    """Engine::lam_layouts_fit."""
    h0s = high_groups(n)
    K = len(h0s)
    if not lay_fits(layout_in(n, 0 if bottom_up else K - 1, bottom_up), n, T - r_low, r_low):
        return False
    for j in range(K):
        first = h0s[j] + (T - r_high - 3)
        if not (lay_fits(layout_in(n, j, bottom_up), n, first, r_high)
                and lay_fits(layout_out(n, j, bottom_up), n, first, r_high)):
            return False
    return True


def engine_bottom_up(n):
    """Engine::gen_bottom_up at the default register shapes (4): from the
    lowest group up unless that loses layouts the other order has."""
    return layouts_fit(n, bottom_up=True) or not layouts_fit(n, bottom_up=False)


def tile_base(g, a, h0):
    gap = h0 - a
    return ((g & ((1 << gap) - 1)) << a) | ((g >> gap) << (h0 + T - a))


def tile_index(base, o, a, h0):
    return base | (o & ((1 << a) - 1)) | ((o >> a) << h0)


def boundaries(n, bottom_up):
    """(reader h0, writer tile first bit, layout) of every layout the chain uses."""
    h0s = high_groups(n)
    K = len(h0s)
    out = []
    for j in range(K):
        prev = j - 1 if bottom_up else j + 1
        w = 3 if prev < 0 or prev >= K else h0s[prev]
        out.append((h0s[j], w, layout_in(n, j, bottom_up)))
    return out


ORDERS = [True, False]


@pytest.mark.parametrize("bottom_up", ORDERS)
@pytest.mark.parametrize("n", list(range(13, 35)))
def test_layout_is_a_bit_permutation(n, bottom_up):
    # every index bit below n lands on a distinct address bit below n, and the
    # map is linear over disjoint bits
    rng = np.random.default_rng(n)
    for h, w, L in boundaries(n, bottom_up):
        images = [lay_addr(1 << b, L) for b in range(n)]
        assert all(v and v & (v - 1) == 0 for v in images)
        assert sorted(images) == [1 << b for b in range(n)]
        for x in rng.integers(0, 1 << n, 50):
            x = int(x)
            want = 0
            for b in range(n):
                if (x >> b) & 1:
                    want |= images[b]
            assert lay_addr(x, L) == want


@pytest.mark.parametrize("bottom_up", ORDERS)
@pytest.mark.parametrize("n", [13, 16, 21, 25, 30, 31])
def test_own_layout_makes_each_high_tile_one_block(n, bottom_up):
    rng = np.random.default_rng(n)
    tiles = 1 << (n - T)
    for h0, w, L in boundaries(n, bottom_up):
        for g in [0, tiles - 1] + [int(x) for x in rng.integers(0, tiles, 6)]:
            base = tile_base(g, 3, h0)
            start = lay_addr(base, L)
            assert start % (1 << T) == 0
            offs = [lay_addr(tile_index(base, o, 3, h0), L) - start for o in range(1 << T)]
            assert offs == list(range(1 << T))


@pytest.mark.parametrize("bottom_up", ORDERS)
@pytest.mark.parametrize("n", list(range(13, 35)))
def test_writer_tiles_stay_within_2_21_amplitudes(n, bottom_up):
    # the writer of each boundary stores its tile within a 2^21-amplitude span
    # (32 MB: 16 pages of 2 MB), the point of placing its bits right above the
    # reader's block
    rng = np.random.default_rng(n + 100)
    tiles = 1 << (n - T)
    for h0, w, L in boundaries(n, bottom_up):
        for g in [0, tiles - 1] + [int(x) for x in rng.integers(0, tiles, 4)]:
            if w == 3:
                idx = [(g << T) | o for o in range(1 << T)]
            else:
                base = tile_base(g, 3, w)
                idx = [tile_index(base, o, 3, w) for o in range(1 << T)]
            addr = [lay_addr(i, L) for i in idx]
            assert max(addr) - min(addr) < (1 << 21)


def test_layouts_are_on_at_the_planned_sizes():
    assert high_groups(30) == [12, 21]
    assert high_groups(31) == [12, 19, 22]
    # the writer's bits sit right above the reader's block, so a high pass's
    # register bits stay contiguous in the layouts it reads and writes, except
    # where a tile's top register bits run into the next reader's block: the
    # single group [n - 9, n) against the low tile's bits 8..11 (18..20
    # qubits) and, from the lowest group up, a lower group's tile overlapping
    # the top group's (27..29 and 31..34 qubits)
    top_down = [13, 14, 15, 16, 17] + list(range(21, 35))
    assert [n for n in range(13, 35) if layouts_fit(n, bottom_up=False)] == top_down
    assert [n for n in range(13, 35) if layouts_fit(n, bottom_up=True)] == [
        13, 14, 15, 16, 17, 21, 22, 23, 24, 25, 26, 30]
    assert [n for n in range(13, 35) if layouts_fit(n, 3, 3, bottom_up=False)] == [
        13, 14, 15, 16, 17, 18] + list(range(21, 35))
    # the engine's order: bottom-up (layer 0 in the last pass) unless that
    # loses layouts; layouts wherever its order has them
    assert [n for n in range(13, 35) if engine_bottom_up(n)] == [
        13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 30]
    assert [n for n in range(13, 35) if layouts_fit(n, bottom_up=engine_bottom_up(n))] == top_down
    # 30 qubits: the group [12, 21) writes for the top group [21, 30): its own
    # bits go to address bits 12..20, the low tile's bits 3..11 above them
    assert layout_out(30, 0) == [(0, 3, 0), (21, 9, 3), (12, 9, 12), (3, 9, 21)]
    # from the top down, the top group writes for [12, 21)
    assert layout_out(30, 1, bottom_up=False) == [(0, 3, 0), (12, 18, 3), (3, 9, 21)]


@pytest.mark.parametrize("bottom_up", ORDERS)
@pytest.mark.parametrize("n", list(range(13, 35)))
def test_register_bits_stay_contiguous_where_layouts_fit(n, bottom_up):
    # lay_fits is what the launchers check; here its verdict is recomputed by
    # brute force over the register index for every pass the plan makes
    if not layouts_fit(n, bottom_up=bottom_up):
        return
    h0s = high_groups(n)
    K = len(h0s)

    def contiguous(L, bits):
        pos = [lay_addr(1 << b, L).bit_length() - 1 for b in bits]
        return pos == list(range(pos[0], pos[0] + len(bits)))

    assert contiguous(layout_in(n, 0 if bottom_up else K - 1, bottom_up), range(T - 4, T))
    for j in range(K):
        first = h0s[j] + (T - 4 - 3)
        assert contiguous(layout_in(n, j, bottom_up), range(first, first + 4))
        assert contiguous(layout_out(n, j, bottom_up), range(first, first + 4))


def _lay_vec(i, L):
    if not L:
        return i
    r = np.zeros_like(i)
    for src, ln, dst in L:
        r |= ((i >> src) & ((1 << ln) - 1)) << dst
    return r


@pytest.mark.parametrize("n,bottom_up", [(16, True), (21, True), (21, False), (25, True),
                                         (30, True), (30, False), (31, False)])
def test_kernel_address_arithmetic_lands_on_the_layout(n, bottom_up):
    # The kernels form an address as lay(base) + lay(thread offset) + k *
    # lay(register step) (lay_tile_addr, store_lam); reader and writer of a
    # boundary agree when that equals lay(index) for every (tile, thread,
    # register).  Low pass: offset t | k << 8 (mapping 0 of the low tile);
    # high pass with first high bit h0: offset (t & 7) | (t >> 3) << h0,
    # register step 2^(h0 + 5).  All tiles for n <= 21, a sample above.
    assert layouts_fit(n, bottom_up=bottom_up)
    h0s = high_groups(n)
    K = len(h0s)
    tiles = 1 << (n - T)
    if n <= 21:
        gs = np.arange(tiles, dtype=np.int64)
    else:
        rng = np.random.default_rng(n)
        gs = np.unique(np.concatenate([[0, tiles - 1], rng.integers(0, tiles, 64)])).astype(np.int64)
    t = np.arange(256, dtype=np.int64)
    k = np.arange(16, dtype=np.int64)
    G, Tt, Kk = np.meshgrid(gs, t, k, indexing="ij")

    def check(base, off, step, L):
        index = base + off + Kk * step
        addr = _lay_vec(base, L) + _lay_vec(off, L) + Kk * _lay_vec(np.int64(step), L)
        assert np.array_equal(addr, _lay_vec(index, L))
        if len(gs) == tiles:  # the pass covers every index once
            assert np.array_equal(np.sort(index.ravel()), np.arange(1 << n))
            assert np.array_equal(np.sort(addr.ravel()), np.arange(1 << n))

    # low pass store into the layout of the first high pass to run
    check(G << T, Tt, 1 << 8, layout_in(n, 0 if bottom_up else K - 1, bottom_up))
    for j in range(K):
        h0 = h0s[j]
        base = ((G & ((1 << (h0 - 3)) - 1)) << 3) | ((G >> (h0 - 3)) << (h0 + 9))
        off = (Tt & 7) | ((Tt >> 3) << h0)
        check(base, off, 1 << (h0 + 5), layout_in(n, j, bottom_up))    # its read
        check(base, off, 1 << (h0 + 5), layout_out(n, j, bottom_up))   # its write
